"""Seeded random configurations (tests/random_configs.py: grid sizes and
levels, step caps, photmax across the register and radix band sorts, nVar,
sorted / unsorted narrow bands, banded grey, every wall condition of
param.h, wall fluxes, Philox / Sobol directions, the exact estimator), each
solved on the GPU by the plain and batched variants and compared with the
CPU oracle over every cell (and every wall cell when wall fluxes are on):
plain tallies (and, sorted narrow band, the per-band launches) equal the
oracle's FP32 sums bit for bit, batched FP64 sums to 1e-12 of the largest.  Every fifth configuration also runs the
full-FP64 plain kernels against the oracle's FP64 mode, bit for bit.
Wall fluxes exist for the adiabatic x walls only (param.h's boundini /
boundend, mc_gpu_noint.cu:1343): the other wall's entries stay 0."""
import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, VARIANT_PLAIN_BANDS, Group, Solver
from radchan_amd.tables import bf16_dequantized
from oracle import oracle as O
from tests.random_configs import deep_case, field_for, random_case, regrid

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
    if p.srt == 1 and not p.grey:
        got, _ = _solve(p, tab, T, est, VARIANT_PLAIN_BANDS)
        assert np.array_equal(got, de32.astype(np.float64))
    o.close()
    if i % 5 == 0:
        q = p.copy()
        q.fp64 = 1
        o64 = O.Oracle(q, tab, T, estimator=est)
        _, want, _, _ = o64.solve(O.all_cells(q), precision=1)
        wwant = o64.solve(O.all_wall_cells(q), surface=True, precision=1)[1] if q.calcwalls else None
        o64.close()
        got, wall = _solve(q, tab, T, est, VARIANT_PLAIN)
        assert np.array_equal(got, want)
        if q.calcwalls:
            assert np.array_equal(wall[solved], wwant[solved]) and np.all(wall[~solved] == 0.0)


@pytest.mark.parametrize("i", range(16))
def test_random_session_features(nb_tables, i):
    """Random configurations with the session's options drawn too: one z
    slab of p_col = 1..3 (zstart of a random rank), 1..3 snapshots, the
    spline restriction, the deterministic reduction, BF16 k tables (the
    oracle fed the dequantized tables) and hipGraph replays; batched
    against the oracle snapshot by snapshot."""
    from radchan_amd import fields as F
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


@pytest.mark.parametrize("i", range(12))
def test_random_groups(nb_tables, i):
    """Two sessions of one random configuration on two random grids (same
    number of levels; own slab, step caps, stretching, wall temperatures,
    estimator and snapshot count) solved by one group launch; each
    session's tallies and wall fluxes against the oracle."""
    p0, tab, _, _ = random_case(300 + i, nb_tables)
    rng = np.random.default_rng(7000 + i)
    bf16 = bool(tab is not None and rng.random() < 0.5)
    sessions, want = [], []
    for m, p in enumerate((p0, regrid(p0, rng))):
        p.p_col = int(rng.choice([d for d in (1, 2, 3) if p.kmax % d == 0]))
        z0 = int(rng.integers(0, p.p_col)) * (p.kmax // p.p_col) + 1
        est = int(tab is not None and rng.random() < 0.4)
        Ts = [field_for(p, rng, 900 + 10 * i + 3 * m + k) for k in range(int(rng.integers(1, 3)))]
        s = Solver(p, tab)
        if est:
            s.set_estimator(est)
        s.set_fields(Ts, zstart=z0)
        s.set_table_precision(bf16)
        sessions.append(s)
        otab = bf16_dequantized(tab) if bf16 else tab
        for k, T in enumerate(Ts):
            o = O.Oracle(p, otab, T, zstart=z0, snapshot=k, estimator=est)
            _, de32d, _, _ = o.solve(O.all_cells(p, z0))
            w32d = o.solve(O.all_wall_cells(p, z0), surface=True)[1] if p.calcwalls else None
            o.close()
            want.append((m, k, p, de32d, w32d))
        print(f"group {i} session {m}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} p_col {p.p_col} "
              f"zstart {z0} snapshots {len(Ts)} estimator {est} bf16 {bf16} calcwalls {p.calcwalls}", flush=True)
    g = Group(sessions)
    g.solve()
    for m, k, p, de32d, w32d in want:
        got = sessions[m].tallies(k)
        assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max(), (m, k)
        if p.calcwalls:
            half = w32d.shape[0] // 2
            solved = np.repeat(np.array([p.adiabw == 1, p.adiabe == 1]), half)
            wall = sessions[m].surface(k)[2].reshape(-1, p.nVar)
            assert np.abs(wall[solved] - w32d[solved]).max() <= 1e-12 * np.abs(w32d[solved]).max(), (m, k)
    g.close()
    for s in sessions:
        s.close()


@pytest.mark.parametrize("i", range(12))
def test_random_dropin(tmp_path, nb_tables, i):
    """mc_gpu_ / get_results_ as the Fortran DNS calls them, on random
    narrow-band configurations and slabs: the scattered mean and the
    reference's std-error (idx_F layout, local k, wall fluxes in the i = 0
    and imax+1 planes of the adiabatic walls) against the oracle's."""
    import radchan_amd as R
    p, tab, T, est = random_case(400 + i, nb_tables)
    if tab is None:
        pytest.skip("grey configuration: the drop-in reads narrow-band tables")
    rng = np.random.default_rng(8000 + i)
    p.p_col = int(rng.choice([d for d in (1, 2, 3) if p.kmax % d == 0]))
    kloc = p.kmax // p.p_col
    z0 = int(rng.integers(0, p.p_col)) * kloc + 1
    p.chanHeight = 0.5
    tab.write(str(tmp_path), 0.5)
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, 0.5)
    print(f"drop-in {i}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} p_col {p.p_col} zstart {z0} "
          f"nVar {p.nVar} calcwalls {p.calcwalls} estimator {est}", flush=True)
    o = O.Oracle(p, t2, T, zstart=z0, estimator=est)
    _, f32d, _, _ = o.solve(O.all_cells(p, z0))
    w32d = o.solve(O.all_wall_cells(p, z0), surface=True)[1] if p.calcwalls else None
    o.close()
    R.set_params(p)
    R.set_table_dir(str(tmp_path))
    R.set_dropin_modes(est, 0)
    n = (p.imax + 2) * (p.jmax + 2) * (kloc + 2)
    res = np.full(n, -7.0, np.float32)
    var = np.full(n, -7.0, np.float32)
    R.mc_gpu(T, zstart=z0, rank=0)
    R.get_results(res, var, rank=0)
    R.set_dropin_modes(0, 0)
    r3 = res.reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    v3 = var.reshape(kloc + 2, p.jmax + 2, p.imax + 2)

    def mean_sd(de):
        nv = de.shape[1]
        m = de.mean(axis=1)
        sd = np.sqrt(((de - m[:, None]) ** 2).sum(axis=1) / ((nv - 0.5) * (nv - 1))) if nv > 1 else None
        return m, sd

    m, sd = mean_sd(f32d)
    m = m.reshape(kloc, p.jmax, p.imax)
    scale = np.abs(m).max()
    assert np.abs(r3[1:-1, 1:-1, 1:-1] - m).max() <= 1e-6 * scale
    if sd is not None:
        assert np.abs(v3[1:-1, 1:-1, 1:-1] - sd.reshape(kloc, p.jmax, p.imax)).max() <= 1e-6 * scale
    if p.calcwalls:
        wm, _ = mean_sd(w32d)
        wm = wm.reshape(2, kloc, p.jmax)
        for b, (ii, on) in enumerate(((0, p.adiabw == 1), (p.imax + 1, p.adiabe == 1))):
            if on:
                assert np.abs(r3[1:-1, 1:-1, ii] - wm[b]).max() <= 1e-6 * np.abs(wm[b]).max()
    assert np.all(r3[0] == -7.0) and np.all(r3[-1] == -7.0)


@pytest.mark.parametrize("i", range(8))
def test_random_radix_widths(nb_tables, i):
    """Random configurations at random band-sort widths above 256 keys (the
    radix sort through HBM), up to param.h's 6000, sorted bands and wall
    fluxes on wherever an x wall is adiabatic: both variants against the
    oracle on every cell and wall cell."""
    p, tab, T, est = random_case(500 + i, nb_tables)
    rng = np.random.default_rng(9000 + i)
    p.photmax = int(rng.integers(257, 6001))
    if p.imax * p.jmax * p.kmax * p.photmax * p.nVar > 4_000_000:
        p.nVar = 1
    print(f"radix width {i}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} photmax {p.photmax} nVar {p.nVar} "
          f"srt {p.srt} grey {p.grey} sampler {p.sampler}", flush=True)
    if tab is not None and p.srt == 0:
        p.srt = 1                      # the radix sort is the sorted narrow band's
    p.calcwalls = int(p.adiabw == 1 or p.adiabe == 1)
    o = O.Oracle(p, tab, T, estimator=est)
    de32, de32d, _, _ = o.solve(O.all_cells(p))
    w32 = w32d = None
    if p.calcwalls:
        w32, w32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
        half = w32.shape[0] // 2
        solved = np.repeat(np.array([p.adiabw == 1, p.adiabe == 1]), half)
    o.close()
    got, wall = _solve(p, tab, T, est, VARIANT_PLAIN)
    assert np.array_equal(got, de32.astype(np.float64))
    if p.calcwalls:
        assert np.array_equal(wall[solved], w32[solved].astype(np.float64))
    got, wall = _solve(p, tab, T, est, VARIANT_BATCHED)
    assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
    if p.calcwalls:
        assert np.abs(wall[solved] - w32d[solved]).max() <= 1e-12 * np.abs(w32d[solved]).max()


@pytest.mark.parametrize("i", range(16))
def test_random_ray_traces(nb_tables, i):
    """Single rays of random configurations traced on the GPU and by the
    oracle: band, g, end cell, level, step count, end status and the hash
    of the visited cells equal; direction, final transmissivity and FP32
    tally bit for bit; fluid rays, and wall rays when wall fluxes are on."""
    from tests.test_gpu_parity import _check_trace
    p, tab, T, _ = random_case(600 + i, nb_tables)
    _check_trace(p, tab, T, n=256, seed=i)
    if p.calcwalls:
        _check_trace(p, tab, T, surface=True, n=128, seed=100 + i)


@pytest.mark.parametrize("i", range(12))
def test_random_deep_levels(nb_tables, i):
    """Three to six grid levels coarsened by ratios up to 7 per axis (the
    index map's multiplicative inverses), tight step caps, optically thin
    banded grey among them so that rays reach the coarse levels: single-ray
    traces (level, cell sequence, steps, tally bits) and both variants'
    tallies against the oracle."""
    from tests.test_gpu_parity import _check_trace
    p, tab, T = deep_case(i, nb_tables)
    print(f"deep levels {i}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num} "
          f"levels {[(p.maxi[g], p.maxj[g], p.maxk[g]) for g in range(p.grid_num)]} grey {p.grey}", flush=True)
    _check_trace(p, tab, T, n=256, seed=i)
    o = O.Oracle(p, tab, T)
    de32, de32d, _, _ = o.solve(O.all_cells(p))
    o.close()
    got, _ = _solve(p, tab, T, 0, VARIANT_PLAIN)
    assert np.array_equal(got, de32.astype(np.float64))
    got, _ = _solve(p, tab, T, 0, VARIANT_BATCHED)
    assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
