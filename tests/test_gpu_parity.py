"""GPU parity: the HIP library against the CPU oracle on identical Philox
streams, called through the C ABI.

Bar: integer traversal (band, g, cell sequence hash, end cell, level,
steps, status) bit-exact; per-ray FP32 tallies bit-exact (same operation
sequence); per-ray FP32 vs the oracle's FP64 energy path within
1e-5 * sum over terms of (|absorbed part| + |emitted part|), the scale of
the tally before the reciprocity difference cancels; plain-variant per-(cell, run) tallies equal to the
oracle's FP32 sums bit for bit; batched-variant FP64 tallies equal to the
FP64 sum of the oracle's per-ray FP32 tallies within 1e-12 relative.
"""
import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

INT_FIELDS = ("nb", "g", "ic", "jc", "kc", "level", "steps", "status", "path_hash")


def _rays(p, n, seed, surface=False):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2, n) if surface else rng.integers(1, p.imax + 1, n)
    j = rng.integers(1, p.jmax + 1, n)
    k = rng.integers(1, p.kmax + 1, n)
    v = rng.integers(0, p.nVar, n)
    h = rng.integers(0, p.photmax, n)
    return np.stack([a, j, k, v, h], axis=1).astype(np.int32)


def _check_trace(p, tab, T, surface=False, n=512, seed=0):
    print(f"trace parity: {p.imax}x{p.jmax}x{p.kmax} grey={p.grey} surface={surface} n={n}", flush=True)
    s = Solver(p, tab)
    s.set_fields([T])
    o = O.Oracle(p, tab, T)
    rays = _rays(p, n, seed, surface)
    g = s.trace(rays, surface)
    c = o.trace(rays, surface)
    for f in INT_FIELDS:
        bad = np.nonzero(g[f] != c[f])[0]
        assert bad.size == 0, f"{f} differs for rays {rays[bad[:5]]}: gpu {g[f][bad[:5]]} cpu {c[f][bad[:5]]}"
    for f in ("sx", "sy", "sz", "tra", "de"):
        assert np.array_equal(g[f].view(np.uint32), c[f].view(np.uint32)), f"{f} not bit-exact"
    tol = 1e-5 * np.maximum(c["sumabs"], 1e-300)
    assert np.all(np.abs(g["de"].astype(np.float64) - c["de64"]) <= tol + 1e-30)
    s.close()
    return g


def test_trace_grey(nb_tables):
    p, tab, T = cases.grey_case()
    _check_trace(p, tab, T)


def test_trace_narrow_band(nb_tables):
    p, tab, T = cases.nb_case(nb_tables)
    g = _check_trace(p, tab, T, n=1024)
    assert len(np.unique(g["nb"])) > 20


def test_trace_narrow_band_unsorted(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, srt=0)
    _check_trace(p, tab, T)


def test_trace_multigrid_walls(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    g = _check_trace(p, tab, T, n=1024)
    assert set(np.unique(g["level"])) >= {0, 1, 2}
    _check_trace(p, tab, T, surface=True, n=512, seed=1)


def test_trace_boundary_conditions(nb_tables):
    p, tab, T = cases.bc_case(nb_tables)
    g = _check_trace(p, tab, T, n=1024)
    assert (g["status"] == 1).any() and (g["status"] == 2).any()


def _field_parity(p, tab, T, variant, surface=False):
    print(f"field parity: variant={variant} {p.imax}x{p.jmax}x{p.kmax} grid_num={p.grid_num}", flush=True)
    s = Solver(p, tab)
    s.set_fields([T])
    st = s.solve(variant)
    o = O.Oracle(p, tab, T)
    if not surface:
        got = s.tallies()
        de32, de32d, de64, steps = o.solve(O.all_cells(p))
    else:
        got = s.surface()[2].reshape(-1, p.nVar)
        de32, de32d, de64, steps = o.solve(O.all_wall_cells(p), surface=True)
    if variant == VARIANT_PLAIN:
        assert np.array_equal(got, de32.astype(np.float64)), "plain tallies differ from the oracle FP32 sums"
    else:
        scale = np.abs(de32d).max()
        assert np.abs(got - de32d).max() <= 1e-12 * scale
    if not surface and not p.calcwalls:
        assert st["steps"] == steps.sum()
        assert st["rays"] == p.imax * p.jmax * p.kmax * p.photmax * p.nVar
    s.close()


@pytest.mark.parametrize("variant", [VARIANT_PLAIN, VARIANT_BATCHED])
def test_field_grey(variant):
    _field_parity(*cases.grey_case(), variant)


@pytest.mark.parametrize("variant", [VARIANT_PLAIN, VARIANT_BATCHED])
def test_field_narrow_band(nb_tables, variant):
    _field_parity(*cases.nb_case(nb_tables), variant)


@pytest.mark.parametrize("photmax", [100, 200, 300])
@pytest.mark.parametrize("variant", [VARIANT_PLAIN, VARIANT_BATCHED])
def test_field_narrow_band_sort_widths(nb_tables, variant, photmax):
    """photmax above 64: the batched kernel's band keys sorted in registers
    2 per lane (<= 128), 4 per lane (<= 256) and by radix passes through
    HBM (above)."""
    _field_parity(*cases.nb_case(nb_tables, n=8, photmax=photmax), variant)


@pytest.mark.parametrize("photmax", [257, 1000])
def test_radix_sorted_keys_multigrid_and_walls(nb_tables, photmax):
    """The radix-sorted band keys (photmax above 256, not a multiple of 64)
    in the multigrid kernel, fluid and wall units."""
    p, tab, T = cases.mg_case(nb_tables, photmax=photmax)
    _field_parity(p, tab, T, VARIANT_BATCHED)
    _field_parity(p, tab, T, VARIANT_BATCHED, surface=True)


@pytest.mark.parametrize("variant", [VARIANT_PLAIN, VARIANT_BATCHED])
def test_field_multigrid_and_walls(nb_tables, variant):
    p, tab, T = cases.mg_case(nb_tables)
    _field_parity(p, tab, T, variant)
    _field_parity(p, tab, T, variant, surface=True)


@pytest.mark.parametrize("variant", [VARIANT_PLAIN, VARIANT_BATCHED])
def test_field_boundary_conditions(nb_tables, variant):
    _field_parity(*cases.bc_case(nb_tables), variant)


def test_snapshot_batch_one_launch(nb_tables):
    """Three temperature snapshots solved in one batched launch: each matches
    the oracle run on that snapshot's own Philox streams."""
    p, tab, _ = cases.nb_case(nb_tables, n=12, photmax=48, nVar=2)
    from radchan_amd import fields as F
    Ts = [F.channel_field(p, 300.0, 1500.0 - 200.0 * s, noise=0.03, seed=40 + s) for s in range(3)]
    print("snapshot batch: 3 x 12^3", flush=True)
    s = Solver(p, tab)
    s.set_fields(Ts)
    s.solve(VARIANT_BATCHED)
    for q, T in enumerate(Ts):
        o = O.Oracle(p, tab, T, snapshot=q)
        de32, de32d, de64, steps = o.solve(O.all_cells(p))
        got = s.tallies(q)
        assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max(), q
        info, oi = s.info(q), o.info()
        for k in ("Tmax", "kappamax", "Ibmax", "Ibmax_w", "Ibmax_e"):
            assert np.float32(info[k]) == np.float32(oi[k]), k
        assert (info["tm"], info["tstar"]) == (oi["tm"], oi["tstar"])
    s.solve(VARIANT_PLAIN)
    for q, T in enumerate(Ts):
        o = O.Oracle(p, tab, T, snapshot=q)
        de32, de32d, de64, steps = o.solve(O.all_cells(p))
        assert np.array_equal(s.tallies(q), de32.astype(np.float64)), q
    s.close()


def test_group_of_grids_one_launch(nb_tables):
    """Snapshots of two sessions with different grids (12^3 with two
    snapshots, 16x8x20 with one) in one batched launch: every snapshot's
    tallies equal the oracle's on its own grid and Philox streams, and the
    group's step count is the sum of the oracle's."""
    from radchan_amd import fields as F
    from radchan_amd import params as P
    from radchan_amd.solver import Group
    p1, tab, _ = cases.nb_case(nb_tables, n=12, photmax=48, nVar=2)
    T1 = [F.channel_field(p1, 300.0, 1500.0 - 300.0 * q, noise=0.03, seed=60 + q) for q in range(2)]
    p2 = P.base_params(16, 8, 20, photmax=48, nVar=2, fact_mesh=1.2, seed=p1.seed, srt=1)
    T2 = [F.channel_field(p2, 300.0, 1400.0, noise=0.03, seed=70)]
    print("group: 2 x 12^3 + 1 x 16x8x20", flush=True)
    s1, s2 = Solver(p1, tab), Solver(p2, tab)
    s1.set_fields(T1)
    s2.set_fields(T2)
    g = Group([s1, s2])
    g.set_schedule(4, 8, 2)
    st = g.solve()
    total = 0
    for s, p, Ts in ((s1, p1, T1), (s2, p2, T2)):
        for q, T in enumerate(Ts):
            o = O.Oracle(p, tab, T, snapshot=q)
            de32, de32d, de64, steps = o.solve(O.all_cells(p))
            got = s.tallies(q)
            assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max(), (p.imax, q)
            total += int(steps.sum())
            m, _ = s.results(q)
            assert np.allclose(m.reshape(-1), got.mean(axis=1), rtol=1e-6, atol=0)
    assert st["steps"] == total
    assert st["rays"] == (p1.imax * p1.jmax * p1.kmax * 2 + p2.imax * p2.jmax * p2.kmax) * 48 * 2
    g.close()
    # sessions that differ in more than the grid are refused
    p3 = P.base_params(8, 8, 8, photmax=32, nVar=2, fact_mesh=1.5, seed=p1.seed, srt=1)
    s3 = Solver(p3, tab)
    s3.set_fields([F.channel_field(p3, 300.0, 1500.0, noise=0.02, seed=1)])
    with pytest.raises(RuntimeError):
        Group([s1, s3])
    for s in (s1, s2, s3):
        s.close()


def test_results_are_mean_and_reference_sdev(nb_tables):
    """mean over runs, and sqrt of sum (x-m)^2 / ((n-1/2)(n-1)) as the
    reference reports it (mc_gpu_noint.cu:234-239, :984)."""
    p, tab, T = cases.nb_case(nb_tables, n=8, photmax=32, nVar=5)
    s = Solver(p, tab)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    de = s.tallies()
    mean, sd = s.results()
    m = de.mean(axis=1)
    var = ((de - m[:, None]) ** 2).sum(axis=1) / ((p.nVar - 0.5) * (p.nVar - 1))
    assert np.allclose(mean.reshape(-1), m.astype(np.float32), rtol=1e-6, atol=0)
    assert np.allclose(sd.reshape(-1), np.sqrt(var).astype(np.float32), rtol=1e-6, atol=0)
    s.close()


def test_graph_replay_matches_eager(nb_tables):
    """The batched solve recorded as a hipGraph: several replays, with eager
    solves and host synchronisation between them, each equal to the eager
    result (FP64 atomic sums: equal to 1e-12 relative)."""
    p, tab, T = cases.mg_case(nb_tables)
    s = Solver(p, tab)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    ref = s.tallies()
    wref = s.surface()[2]
    s.set_graph(True)
    for rep in range(4):
        st = s.solve(VARIANT_BATCHED)
        got = s.tallies()
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), rep
        assert np.abs(s.surface()[2] - wref).max() <= 1e-12 * np.abs(wref).max(), rep
        s.set_graph(False) if rep == 1 else None
        s.solve(VARIANT_PLAIN)                     # eager work between replays
        s.set_graph(True) if rep == 1 else None
    s.close()


def test_autotune_keeps_results(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, n=12, photmax=40, nVar=2)
    s = Solver(p, tab)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    ref = s.tallies()
    best = s.autotune()
    assert best["wg_per_cu"] in (8, 12, 16, 20, 24, 32, 40) and best["threshold"] in (4, 8, 16, 32)
    assert best["steps_per_iter"] in (1, 2, 3, 4, 6, 8) and best["refill"] in (0, 1)
    assert best["refill"] == 1 or best["steps_per_iter"] <= 4
    s.solve(VARIANT_BATCHED)
    assert np.abs(s.tallies() - ref).max() <= 1e-12 * np.abs(ref).max()
    s.close()


def test_group_with_wall_fluxes(nb_tables):
    """A group whose sessions solve wall fluxes (adiabatic black x walls):
    the walls of all three snapshots go through one second launch, and every
    snapshot's fluid and wall tallies equal the oracle's."""
    from radchan_amd import fields as F
    from radchan_amd import params as P
    from radchan_amd.solver import Group
    p1, tab, _ = cases.nb_case(nb_tables, n=12, photmax=48, nVar=2)
    p1.calcwalls = 1
    T1 = [F.channel_field(p1, 300.0, 1500.0 - 300.0 * q, noise=0.03, seed=80 + q) for q in range(2)]
    p2 = P.base_params(16, 8, 20, photmax=48, nVar=2, fact_mesh=1.2, seed=p1.seed, srt=1)
    p2.calcwalls = 1
    T2 = [F.channel_field(p2, 300.0, 1400.0, noise=0.03, seed=90)]
    print("group with walls: 2 x 12^3 + 1 x 16x8x20", flush=True)
    s1, s2 = Solver(p1, tab), Solver(p2, tab)
    s1.set_fields(T1)
    s2.set_fields(T2)
    g = Group([s1, s2])
    g.set_schedule(4, 8, 2)
    g.solve()
    for s, p, Ts in ((s1, p1, T1), (s2, p2, T2)):
        for q, T in enumerate(Ts):
            o = O.Oracle(p, tab, T, snapshot=q)
            _, de32d, _, _ = o.solve(O.all_cells(p))
            got = s.tallies(q)
            assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max(), (p.imax, q)
            _, ds32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
            wall = s.surface(q)[2].reshape(-1, p.nVar)
            assert np.abs(wall - ds32d).max() <= 1e-12 * np.abs(ds32d).max(), (p.imax, q, "walls")
    g.close()
    s1.close()
    s2.close()


def test_plain_one_launch_per_band_equals_plain(nb_tables):
    """The plain variant launched once per sorted band (and run) adds every
    cell's rays in the same order as the one-launch plain variant: equal
    tallies bit for bit (single grid, and three levels with walls)."""
    from radchan_amd import VARIANT_PLAIN_BANDS
    for p, tab, T in (cases.nb_case(nb_tables, n=12, photmax=48), cases.mg_case(nb_tables)):
        print(f"per-band plain: {p.imax}x{p.jmax}x{p.kmax} grid_num={p.grid_num}", flush=True)
        s = Solver(p, tab)
        s.set_fields([T])
        st0 = s.solve(VARIANT_PLAIN)
        a, aw = s.tallies(), s.surface()[2]
        st1 = s.solve(VARIANT_PLAIN_BANDS)
        assert np.array_equal(s.tallies(), a) and np.array_equal(s.surface()[2], aw)
        assert st1["rays"] == st0["rays"] and st1["steps"] == st0["steps"]
        assert st1["launches"] > p.nB * p.nVar
        s.close()


@pytest.mark.parametrize("refill", [0, 1])
def test_refill_modes(nb_tables, refill):
    """Refills once per iteration and between the march steps, on a single
    grid and on three levels with walls: the tallies equal the oracle's
    (FP64 sums of its per-ray FP32 tallies) either way."""
    for p, tab, T in (cases.nb_case(nb_tables, n=12, photmax=64), cases.mg_case(nb_tables)):
        print(f"refill {refill}: {p.imax}x{p.jmax}x{p.kmax} grid_num {p.grid_num}", flush=True)
        s = Solver(p, tab)
        s.set_schedule(16, 4, 6, refill)
        s.set_fields([T])
        s.solve(VARIANT_BATCHED)
        got = s.tallies()
        o = O.Oracle(p, tab, T)
        _, de32d, _, _ = o.solve(O.all_cells(p))
        assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
        if p.calcwalls:
            _, w32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
            gw = s.surface()[2].reshape(-1, p.nVar)
            assert np.abs(gw - w32d).max() <= 1e-12 * np.abs(w32d).max()
        s.close()


@pytest.mark.parametrize("maxs0", [12, 40])
def test_long_iterations_on_several_levels(nb_tables, maxs0):
    """Steps per iteration above 8: the default on several levels is the
    finest level's step cap (12; 40 clamped to 16), and 16 steps with
    refills between them; tallies equal the oracle's either way."""
    p, tab, T = cases.mg_case(nb_tables)
    p.maxs[0] = maxs0
    o = O.Oracle(p, tab, T)
    _, de32d, _, _ = o.solve(O.all_cells(p))
    _, w32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
    o.close()
    for sched in ((0, 0, 0, -1), (20, 4, 16, 1)):
        print(f"maxs[0] {maxs0}, schedule {sched}", flush=True)
        s = Solver(p, tab)
        s.set_schedule(*sched)
        s.set_fields([T])
        s.solve(VARIANT_BATCHED)
        assert np.abs(s.tallies() - de32d).max() <= 1e-12 * np.abs(de32d).max()
        gw = s.surface()[2].reshape(-1, p.nVar)
        assert np.abs(gw - w32d).max() <= 1e-12 * np.abs(w32d).max()
        s.close()
