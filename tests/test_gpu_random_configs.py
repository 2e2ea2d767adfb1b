"""Seeded random configurations (tests/random_configs.py: grid sizes and
levels, step caps, photmax across the register and radix band sorts, nVar,
sorted / unsorted narrow bands, banded grey, every wall condition of
param.h, wall fluxes, Philox / Sobol directions, the exact estimator), each
solved on the GPU by the plain and batched variants and compared with the
CPU oracle over every cell (and every wall cell when wall fluxes are on):
plain tallies equal the oracle's FP32 sums bit for bit, batched FP64 sums
to 1e-12 of the largest.  Every fifth configuration also runs the
full-FP64 plain kernels against the oracle's FP64 mode, bit for bit.
Wall fluxes exist for the adiabatic x walls only (param.h's boundini /
boundend, mc_gpu_noint.cu:1343): the other wall's entries stay 0."""
import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
from oracle import oracle as O
from tests.random_configs import random_case

pytestmark = pytest.mark.gpu


def _solve(p, tab, T, est, variant):
    s = Solver(p, tab)
    if est:
        s.set_estimator(est)
    s.set_fields([T])
    s.solve(variant)
    got = s.tallies()
    wall = s.surface()[2].reshape(-1, p.nVar) if p.calcwalls else None
    s.close()
    return got, wall


@pytest.mark.parametrize("i", range(48))
def test_random_configuration(nb_tables, i):
    p, tab, T, est = random_case(i, nb_tables)
    print(f"random configuration {i}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} photmax {p.photmax} "
          f"nVar {p.nVar} grey {p.grey} srt {p.srt} bd {list(p.bd)} adiab {p.adiabw}{p.adiabe} "
          f"calcwalls {p.calcwalls} sampler {p.sampler} estimator {est}", flush=True)
    o = O.Oracle(p, tab, T, estimator=est)
    de32, de32d, _, _ = o.solve(O.all_cells(p))
    w32 = w32d = None
    if p.calcwalls:
        w32, w32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
    assert np.abs(de32d).max() > 0
    if p.calcwalls:
        # rows of the walls solved (bound-major: west, then east)
        half = w32.shape[0] // 2
        solved = np.repeat(np.array([p.adiabw == 1, p.adiabe == 1]), half)
        w32, w32d = w32[solved], w32d[solved]
    got, wall = _solve(p, tab, T, est, VARIANT_PLAIN)
    assert np.array_equal(got, de32.astype(np.float64))
    if p.calcwalls:
        assert np.array_equal(wall[solved], w32.astype(np.float64)) and np.all(wall[~solved] == 0.0)
    got, wall = _solve(p, tab, T, est, VARIANT_BATCHED)
    assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
    if p.calcwalls:
        assert np.abs(wall[solved] - w32d).max() <= 1e-12 * np.abs(w32d).max() and np.all(wall[~solved] == 0.0)
    o.close()
    if i % 5 == 0:
        q = p.copy()
        q.fp64 = 1
        o64 = O.Oracle(q, tab, T, estimator=est)
        _, want, _, _ = o64.solve(O.all_cells(q), precision=1)
        o64.close()
        got, _ = _solve(q, tab, T, est, VARIANT_PLAIN)
        assert np.array_equal(got, want)
