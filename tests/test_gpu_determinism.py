"""The deterministic reduction of the batched variant (rc_set_reduction 1):
every ray's FP32 tally stored at its own entry, then summed per (cell,
run) in a fixed order with compensated FP64 arithmetic.  The tallies are
the same bits for different schedules, and they equal the exactly rounded
sum (math.fsum) of the oracle's per-ray FP32 tallies to within the
compensated sum's error bound."""
import math

import numpy as np
import pytest

from radchan_amd import VARIANT_BATCHED, Solver
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu


def test_deterministic_sums_are_exactly_rounded(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, n=8, photmax=32, nVar=2)
    print("deterministic reduction: nb 8^3, photmax 32, nVar 2", flush=True)
    s = Solver(p, tab)
    s.set_fields([T])
    s.set_reduction(1)
    s.set_schedule(16, 4, 1)
    s.solve(VARIANT_BATCHED)
    a = s.tallies()
    s.set_schedule(40, 16, 3)
    s.solve(VARIANT_BATCHED)
    b = s.tallies()
    s.set_reduction(0)
    s.solve(VARIANT_BATCHED)
    c = s.tallies()
    s.close()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    cells = O.all_cells(p)
    rays = np.array([[i, j, k, v, h] for (i, j, k) in cells for v in range(p.nVar) for h in range(p.photmax)],
                    np.int32)
    de = O.Oracle(p, tab, T).trace(rays)["de"].astype(np.float64).reshape(len(cells), p.nVar, p.photmax)
    exact = np.array([[math.fsum(de[n, v]) for v in range(p.nVar)] for n in range(len(cells))])
    bound = 2.0 ** -50 * np.abs(de).sum(axis=2)
    assert np.all(np.abs(a - exact) <= bound)
    # the default (atomic) summation agrees to FP64 reordering
    assert np.abs(c - exact).max() <= 1e-12 * np.abs(exact).max()


def test_deterministic_across_schedules_with_walls(nb_tables):
    p, tab, T = cases.mg_case(nb_tables)
    print("deterministic reduction: three levels with walls", flush=True)
    s = Solver(p, tab)
    s.set_fields([T])
    s.set_reduction(1)
    out = []
    for sched in ((8, 4, 1), (32, 16, 4)):
        s.set_schedule(*sched)
        s.solve(VARIANT_BATCHED)
        out.append((s.tallies(), s.surface()[2]))
    s.close()
    assert np.array_equal(out[0][0].view(np.uint64), out[1][0].view(np.uint64))
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))
    o = O.Oracle(p, tab, T)
    _, f32d, _, _ = o.solve(O.all_cells(p))
    _, w32d, _, _ = o.solve(O.all_wall_cells(p), surface=True)
    assert np.abs(out[0][0] - f32d).max() <= 1e-12 * np.abs(f32d).max()
    assert np.abs(out[0][1].reshape(-1, p.nVar) - w32d).max() <= 1e-12 * np.abs(w32d).max()
