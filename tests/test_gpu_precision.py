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
