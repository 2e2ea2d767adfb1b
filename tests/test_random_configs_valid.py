"""The seeded random configurations of the GPU parity test are valid inputs:
the library's parameter check accepts every one and every field (ghost
planes included) is positive, so the GPU test exercises solves, not the
input checks."""
import ctypes as C

import numpy as np

from radchan_amd import lib as L
from tests.random_configs import random_case


def test_random_configurations_are_valid(nb_tables):
    lib = L.load()
    seen = set()
    for i in range(48):
        p, tab, T, est = random_case(i, nb_tables)
        assert lib.rc_set_params(C.byref(p)) == 0, (i, L.last_error())
        assert np.isfinite(T).all() and T.min() > 0.0, i
        seen.add((p.grid_num, p.grey, p.srt, p.calcwalls, p.sampler, p.photmax > 256))
    # the set spans levels, grey and narrow band, sorted and not, wall
    # fluxes, both samplers and both band sorts
    assert {g for g, *_ in seen} == {1, 2, 3}
    for k in range(1, 6):
        assert {s[k] for s in seen} == {0, 1} or {s[k] for s in seen} == {False, True}
