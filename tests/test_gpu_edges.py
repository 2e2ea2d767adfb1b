"""Edge cases of the solver against the CPU oracle on the GPU: the smallest
grid and ray count (one run: variance 0), odd grid sizes with wall fluxes,
and the largest photmax the band sort takes (16384 keys: the radix sort
and the plain variant's LDS sort at their limits)."""
import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
from radchan_amd import fields as F
from radchan_amd import params as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _both(p, tab, T, surface=False):
    o = O.Oracle(p, tab, T)
    cells = O.all_wall_cells(p) if surface else O.all_cells(p)
    de32, de32d, _, _ = o.solve(cells, surface=surface)
    assert np.abs(de32d).max() > 0
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"edge: {p.imax}x{p.jmax}x{p.kmax} photmax {p.photmax} nVar {p.nVar} variant {variant} "
              f"surface {surface}", flush=True)
        s = Solver(p, tab)
        s.set_fields([T])
        s.solve(variant)
        got = s.surface()[2].reshape(-1, p.nVar) if surface else s.tallies()
        if variant == VARIANT_PLAIN:
            assert np.array_equal(got, de32.astype(np.float64))
        else:
            assert np.abs(got - de32d).max() <= 1e-12 * max(np.abs(de32d).max(), 1e-300)
        if not surface:
            mean, sd = s.results()
            assert np.all(np.isfinite(mean)) and np.all(np.isfinite(sd))
            if p.nVar == 1:
                assert np.all(sd == 0.0)
        s.close()
    o.close()


def test_smallest_grid_one_ray_one_run(nb_tables):
    """2x1x1 cells, one ray per cell, one run (seed 31: a ray leaves its
    cell, so the tally is not trivially 0).  Wall and centre temperatures
    close together keep the extrapolated ghost planes positive."""
    p = P.base_params(2, 1, 1, photmax=1, nVar=1, fact_mesh=1.5, seed=31)
    T = F.channel_field(p, 700.0, 900.0, noise=0.05, seed=1)
    _both(p, nb_tables, T)


def test_odd_sizes_with_walls(nb_tables):
    p = P.base_params(6, 3, 5, photmax=17, nVar=3, fact_mesh=1.2, seed=32)
    p.calcwalls = 1
    T = F.channel_field(p, 350.0, 1300.0, noise=0.04, seed=2)
    _both(p, nb_tables, T)
    _both(p, nb_tables, T, surface=True)


def test_largest_sorted_photmax(nb_tables):
    p = P.base_params(2, 2, 2, photmax=16384, nVar=1, fact_mesh=1.5, seed=33)
    T = F.channel_field(p, 800.0, 1000.0, noise=0.03, seed=3)
    _both(p, nb_tables, T)


def test_field_check_and_tmax_on_the_device(nb_tables):
    """rc_set_fields checks the field and finds Tmax on the GPU: a zero in a
    ghost plane, a negative or a NaN anywhere is refused with the message;
    Tmax (interior cells and x-wall midpoints) equals the oracle's, also
    when the hottest value is an x-wall midpoint."""
    p = P.base_params(5, 3, 4, photmax=8, nVar=2, fact_mesh=1.2, seed=7)
    T = F.channel_field(p, 700.0, 900.0, noise=0.03, seed=7)
    for q, val in ((0, 0.0), (T.size - 1, -1.0), (T.size // 2, np.nan), (17, np.inf)):
        bad = T.copy()
        bad[q] = val
        s = Solver(p, nb_tables)
        with pytest.raises(RuntimeError, match="non-positive or non-finite"):
            s.set_fields([bad])
        s.close()
    A = T.reshape(p.kmax + 2, p.jmax + 2, p.imax + 2).copy()
    A[2, 2, 0] = 1500.0              # a hot west ghost: the midpoint (T_0 + T_1)/2 is the maximum
    hot = A.reshape(-1)
    for field in (T, hot):
        s = Solver(p, nb_tables)
        s.set_fields([field])
        o = O.Oracle(p, nb_tables, field)
        assert s.info()["Tmax"] == o.info()["Tmax"]
        o.close()
        s.close()
    assert O.Oracle(p, nb_tables, hot).info()["Tmax"] == (np.float32(1500.0) + A[2, 2, 1]) * np.float32(0.5)


def test_many_runs_per_cell(nb_tables):
    """nVar 37 runs of 5 rays per cell (units are taken a cell's runs at a
    time): both variants against the oracle, and the scattered mean and
    the reference's std-error formula from the tallies."""
    p = P.base_params(6, 4, 5, photmax=5, nVar=37, fact_mesh=1.2, seed=41)
    T = F.channel_field(p, 700.0, 1000.0, noise=0.03, seed=41)
    _both(p, nb_tables, T)
    s = Solver(p, nb_tables)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    de = s.tallies()
    mean, sd = s.results()
    m = de.mean(axis=1)
    want_sd = np.sqrt(((de - m[:, None]) ** 2).sum(axis=1) / ((37 - 0.5) * 36))
    assert np.allclose(mean.reshape(-1), m, rtol=1e-6, atol=1e-6 * np.abs(m).max())
    assert np.allclose(sd.reshape(-1), want_sd, rtol=1e-5, atol=1e-6 * np.abs(m).max())
    s.close()
