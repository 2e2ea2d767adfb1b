"""The estimator's exact mode (rc_set_estimator(ctx, 1)): band/g pdf and
normalisation from the tables interpolated at Tmax, instead of the
reference's CDFs of the nearest table row and its tabulated Planck mean
(mc_gpu_noint.cu:1110-1131, read_tables.cpp:35-38; quirks A10, A11).

CPU (oracle): at a temperature between table rows, with a kP column that
disagrees with the k tables by 25 % (as the generator's line-by-line kPr
can), the exact mode reproduces the narrow-band slab closed form while the
reference mode is off by the 25 %; the exact mode's tallies do not depend
on kP at all.  GPU: the library's exact mode equals the oracle's (fluid and
walls, both variants)."""
import numpy as np
import pytest
from scipy.special import expn

from oracle import oracle as O
from radchan_amd import fields as F
from radchan_amd import params as P
from radchan_amd.tables import WQ16
from tests import cases


def _slab(t, f=0.5):
    Tg = float(t.Tnb[12] + f * (t.Tnb[13] - t.Tnb[12]))
    p = P.base_params(16, 4, 4, photmax=3000, nVar=12, fact_mesh=1.5, seed=31)
    p.adiabw = p.adiabe = 0
    p.Tw[0] = p.Tw[1] = 0.0
    T = F.uniform_field(p, Tg, Tg)
    cells = np.array([[i, 2, 3] for i in (1, 4, 8, 13, 16)], np.int32)
    x = O.grid(p, 0)["x"][cells[:, 0]].astype(np.float64)
    # the gas as the estimator sees it: k and Ib interpolated linearly in T
    a = (np.float64(Tg) - t.Tnb[12]) / (np.float64(t.Tnb[13]) - t.Tnb[12])
    k = t.kq[:, 12, :].astype(np.float64) + a * (t.kq[:, 13, :].astype(np.float64) - t.kq[:, 12, :])
    ib0 = np.array([O.planck(float(t.Tnb[12]), float(w)) for w in t.wvc], np.float64)
    ib1 = np.array([O.planck(float(t.Tnb[13]), float(w)) for w in t.wvc], np.float64)
    ib = ib0 + a * (ib1 - ib0)
    dnu = (t.wvr - t.wvl).astype(np.float64) * 100.0
    wq = WQ16.astype(np.float64)
    ana = np.array([np.sum(2 * np.pi * k * wq[None, :] * (ib * dnu)[:, None]
                           * (expn(2, k * xc) + expn(2, k * (2.0 - xc)))) for xc in x])
    return p, T, cells, ana


def test_exact_mode_is_unbiased_off_row_and_ignores_kP(nb_tables):
    t2 = nb_tables.copy()
    t2.kP[:] = (nb_tables.kP * 1.25).astype(np.float32)
    p, T, cells, ana = _slab(t2)
    _, _, ex, _ = O.Oracle(p, t2, T, estimator=1).solve(cells)
    m, se = ex.mean(axis=1), ex.std(axis=1, ddof=1) / np.sqrt(p.nVar)
    assert np.all(np.abs((m - ana) / se) < 4.5), (m - ana) / se
    _, _, rf, _ = O.Oracle(p, t2, T).solve(cells)
    ratio = rf.mean(axis=1).sum() / ana.sum()
    assert abs(ratio - 1.25) < 0.05, ratio
    _, _, ex0, _ = O.Oracle(p, nb_tables, T, estimator=1).solve(cells)
    assert np.array_equal(ex0, ex)


def test_exact_mode_equals_reference_mode_where_they_agree(nb_tables):
    """At a tabulated temperature of self-consistent tables the two modes
    sample the same pdf up to rounding: the tallies agree statistically."""
    p, T, cells, ana = _slab(nb_tables, f=0.0)
    _, _, ex, _ = O.Oracle(p, nb_tables, T, estimator=1).solve(cells)
    _, _, rf, _ = O.Oracle(p, nb_tables, T).solve(cells)
    se = np.hypot(ex.std(axis=1, ddof=1), rf.std(axis=1, ddof=1)) / np.sqrt(p.nVar)
    assert np.all(np.abs(ex.mean(axis=1) - rf.mean(axis=1)) < 4.5 * se)


def _gpu_mode_parity(p, tab, T):
    from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
    o = O.Oracle(p, tab, T, estimator=1)
    de32, de32d, _, _ = o.solve(O.all_cells(p))
    oi = o.info()
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"exact estimator: variant {variant} {p.imax}x{p.jmax}x{p.kmax}", flush=True)
        s = Solver(p, tab)
        s.set_estimator(1)
        s.set_fields([T])
        s.solve(variant)
        got = s.tallies()
        info = s.info()
        s.close()
        assert info["kappamax"] == np.float32(oi["kappamax"]) and info["Ibmax"] == np.float32(oi["Ibmax"])
        if variant == VARIANT_PLAIN:
            assert np.array_equal(got, de32.astype(np.float64))
        else:
            assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()


@pytest.mark.gpu
def test_gpu_exact_mode_narrow_band(nb_tables):
    _gpu_mode_parity(*cases.nb_case(nb_tables))


@pytest.mark.gpu
def test_gpu_exact_mode_multigrid_walls(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    _gpu_mode_parity(p, tab, T)
    from radchan_amd import VARIANT_BATCHED, Solver
    s = Solver(p, tab)
    s.set_estimator(1)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    got = s.surface()[2].reshape(-1, p.nVar)
    s.close()
    _, de32d, _, _ = O.Oracle(p, tab, T, estimator=1).solve(O.all_wall_cells(p), surface=True)
    assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
