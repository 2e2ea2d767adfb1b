"""rc_spline_restrict_noalloc: the cubic-spline restriction on device
arrays with its working space in the library's fixed device array (no
hipMalloc), against rc_spline_restrict (host arrays, allocates) on the
same fields: equal bit for bit, on two calls in a row, the second on a
smaller field (the array keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

from radchan_amd import fields as F
from radchan_amd import lib as L
from radchan_amd import params as P
from radchan_amd.solver import spline_restrict


def _case(n, seed):
    """a stretched field of (2n, n, 2n) interior cells with two coarse
    levels and wall planes of their own"""
    im, jm, km = 2 * n, n, 2 * n
    p = P.base_params(im, jm, km, photmax=8, nVar=2, fact_mesh=1.5, seed=seed,
                      levels=[(im, jm, km), (im // 2, jm // 2, km // 2), (im // 4, jm // 4, km // 4)],
                      maxs=[4, 8, 64])
    T = F.channel_field(p, 300.0, 1500.0, noise=0.05, seed=seed)
    A = T.reshape(p.kmax + 2, p.jmax + 2, p.imax + 2)
    rng = np.random.default_rng(seed)
    A[:, :, 0] = 400.0 + 50.0 * rng.random((p.kmax + 2, p.jmax + 2))
    A[:, :, -1] = 600.0 + 80.0 * rng.random((p.kmax + 2, p.jmax + 2))
    return p, T


class _Dev:
    """float32 device buffers through the HIP runtime the library links"""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n):
        out = np.zeros(n, np.float32)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def test_noalloc_refuses_fields_beyond_its_array():
    """a 256x128x256 field needs more than the fixed array: refused before
    any device work (runs without a GPU)"""
    p = P.base_params(256, 128, 256, photmax=8, nVar=2, fact_mesh=1.5,
                      levels=[(256, 128, 256), (128, 64, 128)], maxs=[4, 64])
    rc = L.load().rc_spline_restrict_noalloc(C.byref(p), 1, C.c_void_p(16), C.c_void_p(16), None, None)
    assert rc != 0 and "scratch" in L.last_error()
    q, _ = _case(8, 1)
    assert L.load().rc_spline_restrict_noalloc(C.byref(q), 0, C.c_void_p(16), C.c_void_p(16), None, None) != 0
    assert "level" in L.last_error()


@pytest.mark.gpu
def test_noalloc_equals_the_allocating_entry_point():
    lib = L.load()
    d = _Dev()
    try:
        for n, seed, level in ((32, 3, 1), (8, 5, 2), (8, 5, 1)):
            p, T = _case(n, seed)
            im, jm, km = p.maxi[level], p.maxj[level], p.maxk[level]
            print(f"rc_spline_restrict_noalloc: {p.imax}x{p.jmax}x{p.kmax} level {level}", flush=True)
            want_c, want_w = spline_restrict(p, level, T)
            dT = d.put(T)
            dc = d.put(np.full(im * jm * km, -1.0, np.float32))
            dw = d.put(np.full(2 * jm * km, -1.0, np.float32))
            L.check(lib.rc_spline_restrict_noalloc(C.byref(p), level, dT, dc, dw, None), "rc_spline_restrict_noalloc")
            got_c, got_w = d.get(dc, im * jm * km), d.get(dw, 2 * jm * km)
            assert np.array_equal(got_c.view(np.uint32), want_c.reshape(-1).view(np.uint32))
            assert np.array_equal(got_w.view(np.uint32), want_w.reshape(-1).view(np.uint32))
            # without the wall output
            dc2 = d.put(np.zeros(im * jm * km, np.float32))
            L.check(lib.rc_spline_restrict_noalloc(C.byref(p), level, dT, dc2, None, None), "rc_spline_restrict_noalloc")
            assert np.array_equal(d.get(dc2, im * jm * km).view(np.uint32), want_c.reshape(-1).view(np.uint32))
    finally:
        d.free()
