"""Spectral tables: container, synthetic narrow-band generator, file I/O.

The container holds the arrays readT produces (src/montecarlo/read_tables.cpp
:21-95) in the original band order: Tnb[nT], kP[nT] and kq[nB][nT][nQ]
(already scaled by chanHeight, as after reading), wvl/wvr/wvc/kavg[nB],
prob[nT][nB] (band CDF per temperature row), probg[nT][nB][nQ] (g CDF per
band and row) and the quadrature weights wq[nQ].

The synthetic generator builds self-consistent H2O-like tables: the band and
g CDFs are the normalised cumulative sums of k*wq*Ib*dnu on every row, and
kP is the matching Planck mean, so that the importance-sampled estimator is
unbiased up to the reference's nearest-row band pdf.  File I/O goes through
the library's reader/writer of the reference text formats
(nbcorrk/write_tables.cpp:18-109).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L

SIGMA = 5.670373e-08
PI = 3.14159265358979323846
C1 = 3.741771790075259e-16
C2 = 0.014387741858429

# NarrowBand.h:21-27 (nQ = 16); the generator holds the same digits as
# doubles (nbcorrk/param.h:99-105)
_GQ16 = [0.04830766, 0.14447196, 0.23928736, 0.33186860, 0.42135127, 0.50689990, 0.58771575, 0.66304426,
         0.73218211, 0.79448379, 0.84936761, 0.89632115, 0.93490607, 0.96476225, 0.98561151, 0.99726386]
_WQ16 = [0.09654009, 0.09563872, 0.09384440, 0.09117387, 0.08765209, 0.08331193, 0.07819389, 0.07234580,
         0.06582222, 0.05868409, 0.05099806, 0.04283589, 0.03427387, 0.02539207, 0.01627439, 0.00701862]
GQ16 = np.array(_GQ16, dtype=np.float32)
WQ16 = np.array(_WQ16, dtype=np.float32)
GQ16_F64 = np.array(_GQ16, dtype=np.float64)
WQ16_F64 = np.array(_WQ16, dtype=np.float64)


def planck(T, nu):
    """Spectral blackbody intensity per unit wavenumber (nu in cm^-1), FP64."""
    w = 100.0 * np.asarray(nu, dtype=np.float64)
    return 1.0 / PI * C1 * w * w * w / np.expm1(C2 * w / np.asarray(T, dtype=np.float64))


class Tables:
    def __init__(self, nB: int, nT: int, nQ: int):
        self.nB, self.nT, self.nQ = nB, nT, nQ
        f = np.float32
        self.Tnb = np.zeros(nT, f)
        self.kP = np.zeros(nT, f)
        self.wvl = np.zeros(nB, f)
        self.wvr = np.zeros(nB, f)
        self.wvc = np.zeros(nB, f)
        self.kavg = np.zeros(nB, f)
        self.kq = np.zeros((nB, nT, nQ), f)
        self.prob = np.zeros((nT, nB), f)
        self.probg = np.zeros((nT, nB, nQ), f)
        self.wq = np.zeros(nQ, f)
        self._c = None

    def cview(self) -> L.TablesC:
        """rc_tables / orc_tables view (the numpy arrays stay owned here)."""
        t = L.TablesC()
        t.nB, t.nT, t.nQ, t.pad = self.nB, self.nT, self.nQ, 0
        for name in ("Tnb", "kP", "wvl", "wvr", "wvc", "kavg", "kq", "prob", "probg", "wq"):
            arr = getattr(self, name)
            assert arr.dtype == np.float32 and arr.flags.c_contiguous
            setattr(t, name, arr.ctypes.data_as(C.POINTER(C.c_float)))
        self._c = t
        return t

    def copy(self) -> "Tables":
        t = Tables(self.nB, self.nT, self.nQ)
        for name in ("Tnb", "kP", "wvl", "wvr", "wvc", "kavg", "kq", "prob", "probg", "wq"):
            setattr(t, name, getattr(self, name).copy())
        return t

    # ---- file I/O through the library (reference text formats) ----------
    def write(self, directory: str, chanHeight: float) -> None:
        L.check(L.load().rc_write_tables(directory.encode(), C.byref(self.cview()), chanHeight), "rc_write_tables")

    @classmethod
    def read(cls, directory: str, nB: int, nT: int, nQ: int, chanHeight: float, wq=None) -> "Tables":
        t = cls(nB, nT, nQ)
        L.check(L.load().rc_read_tables(directory.encode(), chanHeight, C.byref(t.cview())), "rc_read_tables")
        t.wq[:] = WQ16 if (wq is None and nQ == 16) else (np.ones(nQ, np.float32) / nQ if wq is None else wq)
        return t


def synthetic_nb(nB: int = 360, nT: int = 20, nQ: int = 16, T_lo: float = 300.0, T_hi: float = 2000.0,
                 wv_lo: float = 150.0, wv_hi: float = 9300.0, k_lo: float = 1e-3, k_hi: float = 1e2,
                 seed: int = 7) -> Tables:
    """H2O-like narrow-band correlated-k tables.

    Bands split [wv_lo, wv_hi] evenly.  Per band a reference absorption
    coefficient is log-uniform in [k_lo, k_hi]; the g-points spread it by a
    seeded factor of up to 10^+-1 monotonically in the quadrature point gq,
    and every k depends log-linearly on temperature with a seeded slope.
    Values are final (as read, i.e. after the chanHeight scaling)."""
    if nQ != 16:
        raise ValueError("synthetic tables use the 16-point quadrature of NarrowBand.h")
    rng = np.random.default_rng(seed)
    t = Tables(nB, nT, nQ)
    t.Tnb[:] = np.linspace(T_lo, T_hi, nT).astype(np.float32)
    edges = np.linspace(wv_lo, wv_hi, nB + 1)
    t.wvl[:] = edges[:-1]
    t.wvr[:] = edges[1:]
    t.wvc[:] = 0.5 * (edges[:-1] + edges[1:])
    t.wq[:] = WQ16
    lk_ref = rng.uniform(np.log10(k_lo), np.log10(k_hi), nB)           # log10 k at the band centre g
    spread = rng.uniform(0.2, 1.0, nB)                                 # decades across g
    slope = rng.uniform(-1.0, 1.0, nB)                                 # decades per 1000 K
    g = GQ16.astype(np.float64)
    T = t.Tnb.astype(np.float64)
    lk = (lk_ref[:, None, None] + spread[:, None, None] * 2.0 * (g[None, None, :] - 0.5)
          + slope[:, None, None] * (T[None, :, None] - 1000.0) / 1000.0)
    kq = np.power(10.0, lk)                                            # [nB][nT][nQ]
    t.kq[:] = kq.astype(np.float32)
    kq32 = t.kq.astype(np.float64)
    dnu = (t.wvr - t.wvl).astype(np.float64)
    ib = planck(T[:, None], t.wvc.astype(np.float64)[None, :])         # [nT][nB]
    wq = WQ16.astype(np.float64)
    w = kq32.transpose(1, 0, 2) * wq[None, None, :] * (ib * dnu[None, :] * 100.0)[:, :, None]  # [nT][nB][nQ]
    band = w.sum(axis=2)                                               # [nT][nB]
    cdf = np.cumsum(band, axis=1)
    t.prob[:] = (cdf / cdf[:, -1:]).astype(np.float32)
    gc = np.cumsum(w, axis=2)
    t.probg[:] = (gc / gc[:, :, -1:]).astype(np.float32)
    t.kP[:] = (band.sum(axis=1) * PI / (SIGMA * T ** 4)).astype(np.float32)
    # kavg: quadrature mean on the LAST temperature row (read_tables.cpp:48-52)
    t.kavg[:] = (kq32[:, -1, :] * WQ16.astype(np.float64)[None, :]).sum(axis=1).astype(np.float32)
    return t


def grey_sizes(true_grey: bool):
    """(nB, nT, nQ) of the grey tables (NarrowBand.h:35-37)."""
    return (1, 65, 1) if true_grey else (1000, 65, 1)


def bf16_dequantized(t: Tables) -> Tables:
    """The tables as the device holds them in BF16 mode: every k rounded to
    BF16 (nearest even) after division by the band's power-of-two scale
    2^floor(log2 max k), then scaled back (both scalings exact).  Feeding
    these tables to an FP32 solver reproduces the BF16 solver exactly."""
    q = t.copy()
    k = t.kq.astype(np.float32)                                    # [nB][nT][nQ]
    mx = k.reshape(t.nB, -1).max(axis=1)
    mx = np.where(mx > 0, mx, np.float32(1.0))
    _, ex = np.frexp(mx)
    sc = np.ldexp(np.float32(1.0), ex - 1).astype(np.float32)
    inv = np.ldexp(np.float32(1.0), 1 - ex).astype(np.float32)
    x = (k * inv[:, None, None]).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    q.kq[:] = (u.astype(np.uint32).view(np.float32) * sc[:, None, None]).astype(np.float32)
    return q
