"""The full-FP64 variant (param.h floatpoint 1, rc_params.fp64 = 1) against
the CPU oracle's full-FP64 mode, through the C ABI.

Both sides run the same IEEE double operations (the project-written
expm1/exp and sin/cos of 2 pi u included), so each (cell, run) tally, an
FP64 sum of FP64 ray tallies in ray order, must equal the oracle's bit for
bit, and the step counts must agree.  Cases: grey, narrow band sorted and
unsorted, three-level multigrid with step caps, adiabatic walls, wall fluxes
and specular z walls, fixed-temperature and bd 4 / bd 5 walls, Sobol
directions, the exact estimator and the spline restriction.
"""
import numpy as np
import pytest

import radchan_amd as R
from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def _fp64(p):
    q = p.copy()
    q.fp64 = 1
    return q


def _parity(p, tab, T, surface=False, estimator=0, restriction=0):
    print(f"fp64 parity: {p.imax}x{p.jmax}x{p.kmax} grey={p.grey} srt={p.srt} grid_num={p.grid_num} "
          f"sampler={p.sampler} surface={surface} estimator={estimator} restriction={restriction}", flush=True)
    s = Solver(p, tab)
    if estimator:
        s.set_estimator(estimator)
    if restriction:
        s.set_restriction(restriction)
    s.set_fields([T])
    st = s.solve(VARIANT_PLAIN)
    o = O.Oracle(p, tab, T, estimator=estimator, restriction=restriction)
    if surface:
        got = s.surface()[2].reshape(-1, p.nVar)
        _, want, _, steps = o.solve(O.all_wall_cells(p), surface=True, precision=1)
    else:
        got = s.tallies()
        _, want, _, steps = o.solve(O.all_cells(p), precision=1)
        if not p.calcwalls:
            assert st["steps"] == steps.sum()
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()} of {np.abs(want).max()}"
    assert np.abs(want).max() > 0
    s.close()
    o.close()


def test_fp64_grey():
    p, tab, T = cases.grey_case()
    _parity(_fp64(p), tab, T)


@pytest.mark.parametrize("srt", [1, 0])
def test_fp64_narrow_band(nb_tables, srt):
    p, tab, T = cases.nb_case(nb_tables, srt=srt)
    _parity(_fp64(p), tab, T)


def test_fp64_multigrid_walls_specular(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    q = _fp64(p)
    _parity(q, tab, T)
    _parity(q, tab, T, surface=True)


def test_fp64_boundary_conditions(nb_tables):
    p, tab, T = cases.bc_case(nb_tables)
    _parity(_fp64(p), tab, T)


def test_fp64_sobol_directions(nb_tables):
    p, tab, T = cases.nb_case(nb_tables)
    q = _fp64(p)
    q.sampler = 1
    _parity(q, tab, T)


def test_fp64_exact_estimator_and_spline_restriction(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    q = _fp64(p)
    _parity(q, tab, T, estimator=1)
    _parity(q, tab, T, restriction=1)
    _parity(q, tab, T, surface=True, restriction=1)


def test_fp64_differs_from_fp32_by_rounding_only(nb_tables):
    """The FP64 and FP32 solves of one field share every stream: their
    fields agree to the FP32 path's rounding, far below the Monte Carlo
    noise, and are not identical."""
    p, tab, T = cases.nb_case(nb_tables, photmax=48)
    s32 = Solver(p, tab)
    s32.set_fields([T])
    s32.solve(VARIANT_PLAIN)
    a = s32.tallies()
    s64 = Solver(_fp64(p), tab)
    s64.set_fields([T])
    s64.solve(VARIANT_PLAIN)
    b = s64.tallies()
    scale = np.abs(b).max()
    assert not np.array_equal(a, b)
    assert np.abs(a - b).max() <= 1e-3 * scale
    s32.close()
    s64.close()


def test_fp64_runs_in_the_plain_variant_only(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, n=8, photmax=16)
    s = Solver(_fp64(p), tab)
    s.set_fields([T])
    with pytest.raises(RuntimeError, match="plain variant"):
        s.solve(VARIANT_BATCHED)
    s.close()


def test_fp64_dropin_entry_points(tmp_path, nb_tables):
    """mc_gpu_ / get_results_ with floatpoint 1 run the FP64 kernels: the
    harvested means equal the FP64 session's."""
    p, tab, T = cases.nb_case(nb_tables, n=12, photmax=32, nVar=3)
    q = _fp64(p)
    q.chanHeight = 0.5
    tab.write(str(tmp_path), 0.5)
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, 0.5)
    R.set_params(q)
    R.set_table_dir(str(tmp_path))
    n = (q.imax + 2) * (q.jmax + 2) * (q.kmax + 2)
    res = np.zeros(n, np.float32)
    var = np.zeros(n, np.float32)
    R.mc_gpu(T, zstart=1)
    R.get_results(res, var)
    s = Solver(q, t2)
    s.set_fields([T])
    s.solve(VARIANT_PLAIN)
    mean, sd = s.results()
    r3 = res.reshape(q.kmax + 2, q.jmax + 2, q.imax + 2)
    assert np.array_equal(r3[1:q.kmax + 1, 1:q.jmax + 1, 1:q.imax + 1], mean)
    s.close()
