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


@pytest.mark.parametrize("i", range(16))
def test_random_session_features(nb_tables, i):
    """Random configurations with the session's options drawn too: one z
    slab of p_col = 1..3 (zstart of a random rank), 1..3 snapshots, the
    spline restriction, the deterministic reduction, BF16 k tables (the
    oracle fed the dequantized tables) and hipGraph replays; batched
    against the oracle snapshot by snapshot."""
    from radchan_amd import fields as F
    from radchan_amd.tables import bf16_dequantized
    p, tab, T0, est = random_case(200 + i, nb_tables)
    rng = np.random.default_rng(5000 + i)
    p.p_col = int(rng.choice([d for d in (1, 2, 3) if p.kmax % d == 0]))
    kloc = p.kmax // p.p_col
    z0 = int(rng.integers(0, p.p_col)) * kloc + 1
    nsnap = int(rng.integers(1, 4))
    tw = float(T0[(p.imax + 2) * (p.jmax + 2) + 1] + T0[(p.imax + 2) * (p.jmax + 2)]) / 2
    Ts = [T0] + [F.channel_field(p, tw, tw * 1.5, noise=0.03, seed=700 + 10 * i + s) for s in range(1, nsnap)]
    restriction = int(p.grid_num > 1 and min(p.imax, p.jmax, p.kmax) >= 2 and rng.random() < 0.5)
    reduction = int(rng.random() < 0.5)
    bf16 = bool(tab is not None and rng.random() < 0.5)
    graph = bool(rng.random() < 0.5)
    print(f"session features {i}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} photmax {p.photmax} "
          f"nVar {p.nVar} p_col {p.p_col} zstart {z0} snapshots {nsnap} restriction {restriction} "
          f"reduction {reduction} bf16 {bf16} graph {graph} estimator {est}", flush=True)
    s = Solver(p, tab)
    if est:
        s.set_estimator(est)
    if restriction:
        s.set_restriction(1)
    s.set_fields(Ts, zstart=z0)
    s.set_table_precision(bf16)
    s.set_reduction(reduction)
    s.set_graph(graph)
    otab = bf16_dequantized(tab) if bf16 else tab
    for rep in range(2 if graph else 1):
        s.solve(VARIANT_BATCHED)
        for k, T in enumerate(Ts):
            o = O.Oracle(p, otab, T, zstart=z0, snapshot=k, restriction=restriction, estimator=est)
            _, de32d, _, _ = o.solve(O.all_cells(p, z0))
            o.close()
            got = s.tallies(k)
            assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max(), (rep, k)
    s.close()
