"""The reference entry points mc_gpu_ / get_results_ (mc_gpu_noint.cu:660,
:876) driven exactly as the Fortran DNS does (channel.f:239-257): launch,
harvest, relaunch; outputs in the idx_F layout with local k."""
import os

import numpy as np
import pytest

import radchan_amd as R
from radchan_amd import fields as F
from radchan_amd import params as P
from radchan_amd.tables import synthetic_nb

pytestmark = pytest.mark.gpu


def test_mc_gpu_get_results_roundtrip(tmp_path):
    tab = synthetic_nb()
    tab.write(str(tmp_path), 0.5)
    p = P.base_params(16, 12, 16, photmax=32, nVar=3, fact_mesh=1.5, seed=4)
    p.p_col = 2           # one z slab of two, as one node of a two-node run
    p.calcwalls = 1
    p.chanHeight = 0.5
    R.set_params(p)
    R.set_table_dir(str(tmp_path))
    T = F.channel_field(p, 300.0, 1500.0, noise=0.03, seed=2)
    kloc = p.kmax // p.p_col
    n = (p.imax + 2) * (p.jmax + 2) * (kloc + 2)
    res = np.full(n, -7.0, np.float32)
    var = np.full(n, -7.0, np.float32)
    for rnd in range(2):          # two DNS radiative steps
        R.mc_gpu(T, zstart=kloc + 1, rank=0)
        R.get_results(res, var, rank=0)
    # same solve through the session API on the tables as read back
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, 0.5)
    s = R.Solver(p, t2)
    s.set_fields([T], zstart=kloc + 1)
    s.solve(R.VARIANT_BATCHED)
    mean, sd = s.results()
    wm, ws, _ = s.surface()
    r3 = res.reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    v3 = var.reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    assert np.array_equal(r3[1:kloc + 1, 1:p.jmax + 1, 1:p.imax + 1], mean)
    assert np.array_equal(v3[1:kloc + 1, 1:p.jmax + 1, 1:p.imax + 1], sd)
    assert np.array_equal(r3[1:kloc + 1, 1:p.jmax + 1, 0], wm[0])
    assert np.array_equal(r3[1:kloc + 1, 1:p.jmax + 1, p.imax + 1], wm[1])
    # untouched entries: ghost j/k planes
    assert np.all(r3[0] == -7.0) and np.all(r3[:, 0, :] == -7.0)
    assert np.all(np.isfinite(sd)) and (sd > 0).all()
    s.close()


def test_get_results_without_launch_is_an_error():
    p = P.base_params(8, 8, 8, photmax=8, grey=True)
    R.set_params(p)
    res = np.zeros(10 * 10 * 10, np.float32)
    var = np.zeros_like(res)
    before = res.copy()
    R.get_results(res, var)
    assert np.array_equal(res, before)
    assert "without a preceding mc_gpu_" in R.lib.last_error()


def test_mc_gpu_with_exact_estimator_and_spline_restriction(tmp_path, nb_tables):
    """The drop-in session with both options on equals a session with the
    same options (three grid levels, so the restriction matters)."""
    from tests import cases
    p, tab, T = cases.mg_case(nb_tables)
    p.calcwalls = 0
    tab.write(str(tmp_path), p.chanHeight)
    R.set_params(p)
    R.set_table_dir(str(tmp_path))
    R.set_dropin_modes(1, 1)
    n = (p.imax + 2) * (p.jmax + 2) * (p.kmax + 2)
    res = np.zeros(n, np.float32)
    var = np.zeros(n, np.float32)
    try:
        R.mc_gpu(T, zstart=1, rank=0)
        R.get_results(res, var, rank=0)
    finally:
        R.set_dropin_modes(0, 0)
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, p.chanHeight)
    s = R.Solver(p, t2)
    s.set_estimator(1)
    s.set_restriction(1)
    s.set_fields([T])
    s.solve(R.VARIANT_BATCHED)
    mean, _ = s.results()
    s.close()
    r3 = res.reshape(p.kmax + 2, p.jmax + 2, p.imax + 2)
    assert np.array_equal(r3[1:-1, 1:-1, 1:-1], mean)


DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "dropin_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="C driver not built (radchan_amd.build.build_dropin_driver())")
def test_c_driver_as_the_dns_calls(tmp_path):
    """A C program (tests/dropin/dropin_driver.c) calls get_results_ before
    any launch (outputs untouched, error reported), then mc_gpu_ /
    get_results_ for two radiative steps, by reference, on one slab of
    two: its harvest equals the session API's solve of the same slab."""
    import subprocess
    tab = synthetic_nb()
    tab.write(str(tmp_path), 0.5)
    p = P.base_params(16, 12, 16, photmax=32, nVar=3, fact_mesh=1.5, seed=9)
    p.p_col = 2
    p.calcwalls = 1
    p.chanHeight = 0.5
    T = F.channel_field(p, 300.0, 1500.0, noise=0.03, seed=6)
    T.astype(np.float32).tofile(tmp_path / "field.bin")
    kloc = p.kmax // p.p_col
    print("c driver: 16x12x16 slab 2 of 2", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path), str(tmp_path / "field.bin"), str(tmp_path / "out.bin"),
                        str(p.imax), str(p.jmax), str(p.kmax), str(p.p_col), str(kloc + 1), str(p.photmax),
                        str(p.nVar), str(p.calcwalls), str(p.seed), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = (p.imax + 2) * (p.jmax + 2) * (kloc + 2)
    out = np.fromfile(tmp_path / "out.bin", np.float32)
    res, var = out[:n].reshape(kloc + 2, p.jmax + 2, p.imax + 2), out[n:].reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, 0.5)
    s = R.Solver(p, t2)
    s.set_fields([T], zstart=kloc + 1)
    s.solve(R.VARIANT_BATCHED)
    mean, sd = s.results()
    wm, _, _ = s.surface()
    assert np.array_equal(res[1:kloc + 1, 1:p.jmax + 1, 1:p.imax + 1], mean)
    assert np.array_equal(var[1:kloc + 1, 1:p.jmax + 1, 1:p.imax + 1], sd)
    assert np.array_equal(res[1:kloc + 1, 1:p.jmax + 1, 0], wm[0])
    assert np.array_equal(res[1:kloc + 1, 1:p.jmax + 1, p.imax + 1], wm[1])
    s.close()
