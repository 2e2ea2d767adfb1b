"""The workloads at their stated sizes, and the one-GPU share of a
multi-node run, against the CPU oracle on the same Philox streams.

* The slab of a two-node run (p_col = 2, zstart = kloc + 1: the z share
  mc_gpu_noint.cu:95-98 and :1356-1358 give the second node): fluid and
  wall tallies of both variants, and the drop-in scatter of get_results_
  (mc_gpu_noint.cu:976-1003), against the oracle of that slab.
* Grey 32^3 (photmax 100): every cell; the true-grey slab at 1000 K
  between 500 K black walls against 2 kappa sigma (Tg^4 - Tw^4)
  [E2(kappa x) + E2(kappa (L - x))], column by column.
* Narrow band 64^3 (photmax 200), the 192x128x192 three-level field with
  wall fluxes and specular z walls, and the snapshot ensemble (8 x
  64x96x64 + 4 x 256x128x256 in one launch, BF16 k tables): the GPU
  solves the whole field, the oracle sampled cells.
* The reference's production parameters (param.h: photmax 6000, nine
  grid levels, 5 steps per level but the last, bd 4 z walls) on a reduced
  grid with the same coarsening pattern, and as they stand: one GPU's z
  slab (rank 24 of p_col = 48) of the 406x576x2304 field, batched solver
  against sampled oracle cells.

Bar (as the parity tests): plain tallies equal to the oracle's FP32 sums
bit for bit, batched FP64 tallies to 1e-12 of the largest.
"""
import numpy as np
import pytest
from scipy.special import expn

import radchan_amd as R
from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Group, Solver, config
from radchan_amd import fields as F
from radchan_amd import params as P
from radchan_amd.tables import bf16_dequantized
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIGMA = 5.670373e-08


def _check(variant, got, de32, de32d):
    if variant == VARIANT_PLAIN:
        assert np.array_equal(got, de32.astype(np.float64)), "plain tallies differ from the oracle FP32 sums"
    else:
        assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()


def _sample(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False))


def _mean_sd(de):
    """mean over the runs and the reference's sqrt-variance
    (mc_gpu_noint.cu:234-239, :984)"""
    n = de.shape[1]
    m = de.mean(axis=1)
    var = ((de - m[:, None]) ** 2).sum(axis=1) / ((n - 0.5) * (n - 1))
    return m, np.sqrt(var)


def test_slab_share_against_oracle(tmp_path, nb_tables):
    tab = nb_tables
    tab.write(str(tmp_path), 0.5)
    p = P.base_params(16, 12, 16, photmax=32, nVar=3, fact_mesh=1.5, seed=4)
    p.p_col = 2
    p.calcwalls = 1
    p.chanHeight = 0.5
    kloc = p.kmax // p.p_col
    z0 = kloc + 1
    T = F.channel_field(p, 300.0, 1500.0, noise=0.03, seed=2)
    t2 = R.Tables.read(str(tmp_path), tab.nB, tab.nT, tab.nQ, 0.5)
    o = O.Oracle(p, t2, T, zstart=z0)
    cells, walls = O.all_cells(p, z0), O.all_wall_cells(p, z0)
    assert cells[:, 2].min() == z0 and cells[:, 2].max() == p.kmax
    f32, f32d, _, _ = o.solve(cells)
    w32, w32d, _, _ = o.solve(walls, surface=True)
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"slab share: variant {variant}, zstart {z0}", flush=True)
        s = Solver(p, t2)
        s.set_fields([T], zstart=z0)
        s.solve(variant)
        _check(variant, s.tallies(), f32, f32d)
        _check(variant, s.surface()[2].reshape(-1, p.nVar), w32, w32d)
        s.close()
    # the drop-in entry points, as node 2 of 2 calls them
    print("slab share: mc_gpu_ / get_results_", flush=True)
    R.set_params(p)
    R.set_table_dir(str(tmp_path))
    n = (p.imax + 2) * (p.jmax + 2) * (kloc + 2)
    res = np.full(n, -7.0, np.float32)
    var = np.full(n, -7.0, np.float32)
    R.mc_gpu(T, zstart=z0, rank=0)
    R.get_results(res, var, rank=0)
    r3 = res.reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    v3 = var.reshape(kloc + 2, p.jmax + 2, p.imax + 2)
    m, sd = _mean_sd(f32d)
    m, sd = m.reshape(kloc, p.jmax, p.imax), sd.reshape(kloc, p.jmax, p.imax)
    scale = np.abs(m).max()
    assert np.abs(r3[1:-1, 1:-1, 1:-1] - m).max() <= 1e-6 * scale
    assert np.abs(v3[1:-1, 1:-1, 1:-1] - sd).max() <= 1e-6 * scale
    wm, wsd = _mean_sd(w32d)
    wm, wsd = wm.reshape(2, kloc, p.jmax), wsd.reshape(2, kloc, p.jmax)
    wscale = np.abs(wm).max()
    for b, i in ((0, 0), (1, p.imax + 1)):
        assert np.abs(r3[1:-1, 1:-1, i] - wm[b]).max() <= 1e-6 * wscale
        assert np.abs(v3[1:-1, 1:-1, i] - wsd[b]).max() <= 1e-6 * wscale
    assert np.all(r3[0] == -7.0) and np.all(r3[-1] == -7.0) and np.all(r3[:, 0, :] == -7.0)


def test_grey32_every_cell():
    p = config("grey32")
    p.nVar, p.seed = 2, 0
    T = F.channel_field(p, 500.0, 1000.0, noise=0.02, seed=0)
    o = O.Oracle(p, None, T)
    de32, de32d, _, steps = o.solve(O.all_cells(p))
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"grey 32^3: variant {variant}", flush=True)
        s = Solver(p, None)
        s.set_fields([T])
        st = s.solve(variant)
        _check(variant, s.tallies(), de32, de32d)
        assert st["steps"] == steps.sum() and st["rays"] == 32 ** 3 * p.photmax * p.nVar
        s.close()


def test_true_grey_slab_32_closed_form():
    p = config("slab32")
    p.nVar, p.seed = 8, 7
    T = F.uniform_field(p, 1000.0, 500.0)
    # the x walls are fixed at 500 K (adiab 0): their ghost planes are not
    # read, but the library takes only positive temperatures
    A = T.reshape(p.kmax + 2, p.jmax + 2, p.imax + 2)
    A[:, :, 0] = A[:, :, -1] = 500.0
    print("true-grey slab 32^3, photmax 100, nVar 8", flush=True)
    s = Solver(p, None)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    de = s.tallies().reshape(p.kmax, p.jmax, p.imax, p.nVar)
    s.close()
    col = de.mean(axis=(0, 1))                      # [i][run]: 1024 cells per column and run
    m = col.mean(axis=1)
    se = col.std(axis=1, ddof=1) / np.sqrt(p.nVar)
    x = O.grid(p, 0)["x"][1:p.imax + 1].astype(np.float64)
    ana = 2 * SIGMA * (1000.0 ** 4 - 500.0 ** 4) * (expn(2, x) + expn(2, 2.0 - x))
    z = (m - ana) / se
    zall = (m - ana).sum() / np.sqrt((se ** 2).sum())
    print("slab: max |z| %.2f, pooled z %.2f, mean ratio %.5f" % (np.abs(z).max(), zall, np.mean(m / ana)),
          flush=True)
    assert np.all(np.abs(z) < 4.5), z
    assert abs(zall) < 4.5
    assert abs(np.mean(m / ana) - 1.0) < 1e-2


def test_nb64_sampled_cells(nb_tables):
    p = config("nb64")
    p.nVar, p.seed = 2, 1
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=0)
    cells = O.all_cells(p)
    idx = _sample(len(cells), 512, 1)
    de32, de32d, _, _ = O.Oracle(p, nb_tables, T).solve(cells[idx])
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"nb 64^3 photmax 200: variant {variant}", flush=True)
        s = Solver(p, nb_tables)
        s.set_fields([T])
        s.solve(variant)
        _check(variant, s.tallies()[idx], de32, de32d)
        s.close()


def test_mg192_sampled_cells(nb_tables):
    p = config("mg192")
    p.nVar, p.seed = 2, 2
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=3)
    cells, walls = O.all_cells(p), O.all_wall_cells(p)
    ic, iw = _sample(len(cells), 512, 2), _sample(len(walls), 256, 3)
    o = O.Oracle(p, nb_tables, T)
    f32, f32d, _, _ = o.solve(cells[ic])
    w32, w32d, _, _ = o.solve(walls[iw], surface=True)
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"192x128x192, 3 levels, walls: variant {variant}", flush=True)
        s = Solver(p, nb_tables)
        s.set_fields([T])
        s.solve(variant)
        _check(variant, s.tallies()[ic], f32, f32d)
        _check(variant, s.surface()[2].reshape(-1, p.nVar)[iw], w32, w32d)
        s.close()


def test_snapshot_ensemble_bf16_sampled_cells(nb_tables):
    """8 x 64x96x64 + 4 x 256x128x256 snapshots (300-1800 K, 5% noise),
    photmax 100, BF16 k tables, one batched launch of a two-session group;
    the oracle runs on the dequantized tables."""
    q = bf16_dequantized(nb_tables)
    sessions, checks = [], []
    for name, n, seed0 in (("big256", 4, 100), ("ens64x96", 8, 200)):
        p = config(name)
        p.nVar, p.seed = 2, 9
        Ts = [F.channel_field(p, 300.0, 1800.0, noise=0.05, seed=seed0 + i) for i in range(n)]
        s = Solver(p, nb_tables)
        s.set_table_precision(True)
        s.set_fields(Ts)
        sessions.append(s)
        cells = O.all_cells(p)
        for i in (0, n - 1):
            idx = _sample(len(cells), 96, seed0 + i)
            _, de32d, _, _ = O.Oracle(p, q, Ts[i], snapshot=i).solve(cells[idx])
            checks.append((len(sessions) - 1, i, idx, de32d))
    print("snapshot ensemble: 8 x 64x96x64 + 4 x 256x128x256, BF16 tables, one launch", flush=True)
    g = Group(sessions)
    st = g.solve()
    assert st["rays"] == (4 * 256 * 128 * 256 + 8 * 64 * 96 * 64) * 100 * 2
    for si, i, idx, de32d in checks:
        got = sessions[si].tallies(i)[idx]
        assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()
    g.close()
    for s in sessions:
        s.close()


def test_reference_production_parameters(nb_tables):
    """param.h's solver settings -- photmax 6000, nine grid levels, maxs 5
    on levels 0-7 and 100000 on the last, bd 4 (black at the adjacent
    cell's T) z walls, adiabatic black x walls, srt -- on 14x48x64 cells
    coarsened as param.h coarsens 406x576x2304 (x kept over five levels,
    y by 3 first, z halved every level)."""
    p = P.reference_params()
    levels = [(14, 48, 64), (14, 16, 32), (14, 8, 16), (14, 4, 8), (14, 2, 4), (2, 1, 2), (2, 1, 1), (2, 1, 1),
              (1, 1, 1)]
    q = P.base_params(14, 48, 64, photmax=p.photmax, nVar=2, fact_mesh=p.fact_mesh, seed=5, levels=levels,
                      maxs=[p.maxs[g] for g in range(9)])
    assert q.photmax == 6000 and q.grid_num == 9 and list(q.maxs) == [5] * 8 + [100000]
    q.bd[:] = list(p.bd)
    q.Lx, q.Ly, q.Lz = p.Lx, p.Ly, p.Lz
    T = F.channel_field(q, 300.0, 1500.0, noise=0.03, seed=5)
    cells = O.all_cells(q)
    idx = _sample(len(cells), 48, 5)
    de32, de32d, _, _ = O.Oracle(q, nb_tables, T).solve(cells[idx])
    for variant in (VARIANT_PLAIN, VARIANT_BATCHED):
        print(f"production parameters on 14x48x64: variant {variant}", flush=True)
        s = Solver(q, nb_tables)
        s.set_fields([T])
        st = s.solve(variant)
        _check(variant, s.tallies()[idx], de32, de32d)
        assert st["rays"] == len(cells) * 6000 * 2
        s.close()


def test_production_share_at_full_size():
    """param.h unchanged: the 406x576x2304 field (544 M cells with ghosts,
    2.2 GB) resident, the 48-plane z slab of rank 24 of p_col = 48 solved
    (11.2 M cells x 6000 rays, nine levels, Sobol directions, 999 bands x
    135 T x 16 g), batched solver against 24 sampled oracle cells.  The
    sizes reach the library's 32-bit index ranges (table positions of all
    levels 0.63 G of the 2^31 it accepts) and 64-bit ray and step counts.
    The plain variant is checked at these settings on the reduced grid
    above (here its sorted keys alone would take 135 GB)."""
    from radchan_amd.tables import synthetic_nb
    p = P.reference_params()
    assert (p.imax, p.jmax, p.kmax, p.p_col, p.photmax, p.grid_num) == (406, 576, 2304, 48, 6000, 9)
    kloc = p.kmax // p.p_col
    z0 = 24 * kloc + 1
    print("production share: field and tables", flush=True)
    T = F.wavy_channel_field(p, 300.0, 1500.0, amp=0.03)
    tab = synthetic_nb(p.nB, p.nT, p.nQ)
    print("production share: oracle set-up (all nine levels of the whole field)", flush=True)
    o = O.Oracle(p, tab, T, zstart=z0)
    cells = O.all_cells(p, z0)
    idx = _sample(len(cells), 24, 7)
    de32, de32d, _, _ = o.solve(cells[idx], nthreads=8)
    o.close()
    print(f"production share: batched solve of {len(cells)} cells", flush=True)
    s = Solver(p, tab)
    s.set_fields([T], zstart=z0)
    st = s.solve(VARIANT_BATCHED)
    _check(VARIANT_BATCHED, s.tallies()[idx], de32, de32d)
    assert st["rays"] == len(cells) * p.photmax * p.nVar
    assert st["steps"] > 2 ** 32
    s.close()
