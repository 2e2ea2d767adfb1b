"""BF16 k tables (mixed precision): bit parity against the oracle fed the
dequantized tables, and the precision study -- per-ray error of the FP32
march with FP32 or BF16 k tables against the FP64 oracle on the same
streams, by band group (optically thin / medium / thick thirds of the
sorted bands)."""
import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver
from radchan_amd.tables import bf16_dequantized
from oracle import oracle as O
from tests import cases
from tests.test_gpu_parity import INT_FIELDS, _rays

pytestmark = pytest.mark.gpu


def test_bf16_tables_parity(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    q = bf16_dequantized(tab)
    print("bf16 parity: mg case", flush=True)
    s = Solver(p, tab)
    s.set_table_precision(True)
    s.set_fields([T])
    o = O.Oracle(p, q, T)
    rays = _rays(p, 1024, 3)
    g, c = s.trace(rays), o.trace(rays)
    for f in INT_FIELDS:
        assert np.array_equal(g[f], c[f]), f
    for f in ("tra", "de"):
        assert np.array_equal(g[f].view(np.uint32), c[f].view(np.uint32)), f
    s.solve(VARIANT_PLAIN)
    de32, de32d, de64, steps = o.solve(O.all_cells(p))
    assert np.array_equal(s.tallies(), de32.astype(np.float64))
    s.solve(VARIANT_BATCHED)
    assert np.abs(s.tallies() - de32d).max() <= 1e-12 * np.abs(de32d).max()
    s.close()


def test_precision_study_by_band_group(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, n=16, photmax=64, nVar=2)
    rays = _rays(p, 6000, 11)
    o64 = O.Oracle(p, tab, T)
    ref = o64.trace(rays)                       # FP64 energy on FP32 geometry
    out = {}
    for bf in (False, True):
        s = Solver(p, tab)
        s.set_table_precision(bf)
        s.set_fields([T])
        out[bf] = s.trace(rays)
        s.close()
    scale = ref["sumabs"]
    groups = np.minimum(3 * ref["nb"] // p.nB, 2)
    rows = []
    for gi, name in enumerate(("thin", "medium", "thick")):
        m = groups == gi
        e32 = np.abs(out[False]["de"][m] - ref["de64"][m]) / scale[m]
        e16 = np.abs(out[True]["de"][m] - ref["de64"][m]) / scale[m]
        rows.append((name, int(m.sum()), float(np.sqrt(np.mean(e32 ** 2))), float(e32.max()),
                     float(np.sqrt(np.mean(e16 ** 2))), float(e16.max())))
    for r in rows:
        print("group %-6s rays %5d  fp32 tables: rms %.2e max %.2e  bf16 tables: rms %.2e max %.2e" % r, flush=True)
    assert all(r[3] <= 1e-5 for r in rows)          # FP32 march: 1e-5 of the tally scale
    assert all(r[5] <= 2e-2 for r in rows)          # BF16 k: bounded by the k rounding (2^-9)


def test_precision_study_field_128(nb_tables):
    """The field-level precision study on a 128^3 narrow-band field
    (photmax 200, nVar 4, FP64 tallies): div q_r of each band group
    (thin / medium / thick thirds of the sorted bands) from the
    mixed-precision solver -- FP32 march with FP32 or BF16 k tables --
    against the full-FP64 oracle on the same Philox streams, on 2048
    sampled cells.  The GPU solves the whole field; its per-cell totals
    equal the oracle's FP32 path (FP32 tables) and the oracle fed the
    dequantized tables (BF16), so the group errors are those of the GPU
    results.  Printed: L2 and max error per group, relative to the group's
    FP64 div q_r, beside the statistical standard error."""
    from radchan_amd import fields as F
    from radchan_amd import params as P
    p = P.base_params(128, 128, 128, photmax=200, nVar=4, fact_mesh=1.5, seed=21)
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=21)
    print("precision study: 128^3, photmax 200, nVar 4", flush=True)
    s = Solver(p, nb_tables)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    g32 = s.tallies()
    s.set_table_precision(True)
    s.solve(VARIANT_BATCHED)
    g16 = s.tallies()
    s.close()
    cells = O.all_cells(p)
    idx = np.random.default_rng(3).choice(len(cells), 2048, replace=False)
    gor = np.minimum(3 * np.arange(p.nB) // p.nB, 2)
    o = O.Oracle(p, nb_tables, T)
    f64 = o.solve_groups(cells[idx], gor, precision=1)
    m32 = o.solve_groups(cells[idx], gor, precision=0)
    m16 = O.Oracle(p, bf16_dequantized(nb_tables), T).solve_groups(cells[idx], gor, precision=0)
    # the GPU's per-cell totals are those of the oracle's FP32 paths
    for g, m in ((g32, m32), (g16, m16)):
        tot = m.sum(axis=2)
        assert np.abs(g[idx] - tot).max() <= 1e-11 * np.abs(tot).max()
    rows = []
    for gi, name in enumerate(("thin", "medium", "thick", "all")):
        sel = slice(None) if name == "all" else gi
        ref = f64[:, :, sel].sum(axis=2) if name == "all" else f64[:, :, sel]
        a32 = m32[:, :, sel].sum(axis=2) if name == "all" else m32[:, :, sel]
        a16 = m16[:, :, sel].sum(axis=2) if name == "all" else m16[:, :, sel]
        q, q32, q16 = ref.mean(axis=1), a32.mean(axis=1), a16.mean(axis=1)
        nrm = np.sqrt((q ** 2).sum())
        se = np.sqrt(((ref.std(axis=1, ddof=1) / np.sqrt(p.nVar)) ** 2).sum()) / nrm
        rows.append((name, np.sqrt(((q32 - q) ** 2).sum()) / nrm, np.abs(q32 - q).max() / np.abs(q).max(),
                     np.sqrt(((q16 - q) ** 2).sum()) / nrm, np.abs(q16 - q).max() / np.abs(q).max(), se))
    for r in rows:
        print("group %-6s  FP32 tables: L2 %.2e max %.2e   BF16 tables: L2 %.2e max %.2e   "
              "statistical std-error %.2e" % r, flush=True)
    for r in rows:
        assert r[1] <= 1e-4 and r[2] <= 1e-3         # FP32 march against FP64, far below the noise
        assert r[3] <= 2e-2                          # BF16 k: bounded by the k rounding
