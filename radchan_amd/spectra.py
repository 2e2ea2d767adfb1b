"""Synthetic line-by-line absorption spectra for the correlated-k
generator (rc_ckdist): per temperature and narrow band, k(nu) on a uniform
wavenumber grid of spacing dwv (nbcorrk/param.h:20) made of a continuum and
Lorentz lines whose positions, widths and strengths come from a seeded
generator; each line's strength has its own power law in temperature, so
the k-distribution changes shape from one temperature to the next.  This is
the shape of an H2O spectrum, not its physics.

Band layout follows the reference generator (nbcorrk/ckdist_hitran_part.cpp
:120-160): every point belongs to exactly one narrow band, wv_left and
wv_right are the band's first and last wavenumber, and the default band
width is Deltawv = 10 cm^-1 (param.h:21-23).
"""
from __future__ import annotations

import numpy as np


def synthetic_lbl(wv0: float = 150.0, wv1: float = 9300.0, deltawv: float = 10.0, temps=None,
                  dwv: float = 0.01, lines_per_band: int = 12, seed: int = 0):
    """Spectra ordered temperature-major (spectrum t*nB + j).  Returns a dict
    with off[nspec+1] (int64 point offsets), wvn, kvn (FP64 points, k in
    m^-1), wvl, wvr (band edges), T (per spectrum), nB and temps."""
    if temps is None:
        temps = np.arange(52) * 50.0 + 300.0
    temps = np.asarray(temps, np.float64)
    rng = np.random.default_rng(seed)
    npt_all = int(round((wv1 - wv0) / dwv))
    per = int(round(deltawv / dwv))
    nB = max(1, npt_all // per)
    nu_all = wv0 + dwv * np.arange(nB * per)
    nu = nu_all.reshape(nB, per)                                     # [nB][npt]
    x = (nu - nu[:, :1]) / (nu[:, -1:] - nu[:, :1])
    pos = rng.uniform(0.0, 1.0, (nB, lines_per_band))
    width = 10.0 ** rng.uniform(-3.0, -1.0, (nB, lines_per_band))     # half width / band width
    s0 = 10.0 ** rng.uniform(-2.0, 3.0, (nB, lines_per_band))        # peak strength at 1000 K, m^-1
    expo = rng.uniform(-2.0, 1.0, (nB, lines_per_band))              # strength ~ (T/1000)^expo
    cont = 10.0 ** rng.uniform(-3.0, -1.0, nB)
    shape = width[:, :, None] ** 2 / ((x[:, None, :] - pos[:, :, None]) ** 2 + width[:, :, None] ** 2)
    ks = []
    for T in temps:
        s = s0 * (T / 1000.0) ** expo                                  # [nB][lines]
        ks.append(cont[:, None] * (T / 1000.0) ** -0.5 + np.einsum("jq,jqx->jx", s, shape))
    kvn = np.concatenate([k.reshape(-1) for k in ks])
    wvn = np.tile(nu.reshape(-1), len(temps))
    nspec = nB * len(temps)
    off = np.arange(nspec + 1, dtype=np.int64) * per
    wvl = np.tile(nu[:, 0], len(temps))
    wvr = np.tile(nu[:, -1], len(temps))
    return dict(off=off, wvn=wvn, kvn=kvn, wvl=wvl, wvr=wvr, T=np.repeat(temps, nB), nB=nB, temps=temps)


def ckdist(off, wvn, kvn, wvl, wvr, gq, n_pwrk: int = 20000, pwr: float = 0.1):
    """Correlated-k distributions on the GPU (rc_ckdist): returns a dict
    with kq [nspec][nq], kavg, kmin, kmax and kernel_ms."""
    import ctypes as C

    from . import lib as L

    d = lambda a: np.ascontiguousarray(a, np.float64)
    off = np.ascontiguousarray(off, np.int64)
    wvn, kvn, wvl, wvr, gq = d(wvn), d(kvn), d(wvl), d(wvr), d(gq)
    ns, nq = len(off) - 1, len(gq)
    kq = np.zeros((ns, nq))
    kavg, kmin, kmax = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    ms = C.c_double()
    L.check(L.load().rc_ckdist(ns, off.ctypes.data_as(C.POINTER(C.c_int64)), L.dptr(wvn), L.dptr(kvn),
                               L.dptr(wvl), L.dptr(wvr), n_pwrk, pwr, nq, L.dptr(gq), L.dptr(kq), L.dptr(kavg),
                               L.dptr(kmin), L.dptr(kmax), C.byref(ms)), "rc_ckdist")
    return dict(kq=kq, kavg=kavg, kmin=kmin, kmax=kmax, kernel_ms=ms.value)


def tables_from_distributions(T, wvl, wvr, kq, kavg, wq, chanHeight: float = 1.0):
    """The table step (rc_ck_tables, calcProbT): kq [nT][nB][nQ] and kavg
    [nT][nB] in FP64 with the band edges wvl/wvr [nB] -> Tables as the
    reader leaves them (FP32, k values times chanHeight)."""
    import ctypes as C

    from . import lib as L
    from .tables import Tables

    kq = np.ascontiguousarray(kq, np.float64)
    nT, nB, nQ = kq.shape
    d = lambda a: np.ascontiguousarray(a, np.float64)
    T, wvl, wvr, kavg, wq = d(T), d(wvl), d(wvr), d(kavg), d(wq)
    assert T.shape == (nT,) and wvl.shape == (nB,) and wvr.shape == (nB,) and kavg.shape == (nT, nB)
    t = Tables(nB, nT, nQ)
    L.check(L.load().rc_ck_tables(nB, nT, nQ, L.dptr(T), L.dptr(wvl), L.dptr(wvr), L.dptr(kq), L.dptr(kavg),
                                  L.dptr(wq), chanHeight, C.byref(t.cview())), "rc_ck_tables")
    return t


def generate_tables(spec: dict, gq, wq, chanHeight: float = 1.0, n_pwrk: int = 20000, pwr: float = 0.1):
    """The generator on the GPU: line-by-line spectra (synthetic_lbl's
    layout) -> correlated-k distributions (rc_ckdist) -> tables
    (rc_ck_tables).  Returns (Tables, the rc_ckdist outputs)."""
    g = ckdist(spec["off"], spec["wvn"], spec["kvn"], spec["wvl"], spec["wvr"], gq, n_pwrk=n_pwrk, pwr=pwr)
    nB, nT, nQ = spec["nB"], len(spec["temps"]), len(gq)
    t = tables_from_distributions(spec["temps"], spec["wvl"][:nB], spec["wvr"][:nB], g["kq"].reshape(nT, nB, nQ),
                                  g["kavg"].reshape(nT, nB), wq, chanHeight)
    return t, g



# ---------------------------------------------------------------------------
# the generator's front end on line-by-line spectrum files

REF_TOLL, REF_WVTOLL, REF_DWV, REF_DELTAWV = 1e-6, 25000.0, 0.01, (10.0, 10.0, 10.0)   # nbcorrk/param.h:18-23


def read_lbl(path: str):
    """A line-by-line spectrum file as the generator reads it
    (nbcorrk/read_files.cpp:29: "%lf %lf" pairs, wavenumber in cm^-1 and
    absorption coefficient in cm^-1): returns (wv, kv) float64 arrays."""
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1])


def lbl_bands(wv, kv, toll: float = REF_TOLL, wvtoll: float = REF_WVTOLL, dwv: float = REF_DWV,
              deltawv=REF_DELTAWV) -> dict:
    """Wide and narrow bands of the reference spectrum (rc_lbl_bands):
    returns first, count (point ranges), wvl, wvr, wvc per narrow band and
    nwide."""
    import ctypes as C

    from . import lib as L
    wv = np.ascontiguousarray(wv, np.float64)
    kv = np.ascontiguousarray(kv, np.float64)
    dw = np.ascontiguousarray(deltawv, np.float64)
    lib = L.load()
    pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    pi = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64))
    nw = C.c_int32()
    nb = lib.rc_lbl_bands(len(wv), pd(wv), pd(kv), toll, wvtoll, dwv, pd(dw), 0, None, None, None, None, None,
                          C.byref(nw))
    if nb < 0:
        raise RuntimeError("rc_lbl_bands: " + L.last_error())
    out = dict(first=np.zeros(nb, np.int64), count=np.zeros(nb, np.int64), wvl=np.zeros(nb), wvr=np.zeros(nb),
               wvc=np.zeros(nb))
    nb2 = lib.rc_lbl_bands(len(wv), pd(wv), pd(kv), toll, wvtoll, dwv, pd(dw), nb, pi(out["first"]),
                           pi(out["count"]), pd(out["wvl"]), pd(out["wvr"]), pd(out["wvc"]), C.byref(nw))
    assert nb2 == nb
    out["nwide"] = nw.value
    return out


def lbl_fill(wv, kv, T: float, wvn, dwv: float = REF_DWV):
    """One temperature's spectrum into the narrow bands' points wvn
    (rc_lbl_fill): returns (kvn in m^-1, kPr in m^-1, points filled)."""
    import ctypes as C

    from . import lib as L
    wv = np.ascontiguousarray(wv, np.float64)
    kv = np.ascontiguousarray(kv, np.float64)
    wvn = np.ascontiguousarray(wvn, np.float64)
    kvn = np.zeros(len(wvn))
    kpr = C.c_double()
    pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    q = L.load().rc_lbl_fill(len(wv), pd(wv), pd(kv), T, dwv, len(wvn), pd(wvn), pd(kvn), C.byref(kpr))
    if q < 0:
        raise RuntimeError("rc_lbl_fill: " + L.last_error())
    return kvn, kpr.value, int(q)


def spectra_from_lbl(ref, temps, spectra, dwv: float = REF_DWV, **band_args) -> dict:
    """The generator's front end end to end: the narrow bands of the
    reference spectrum ref = (wv, kv) (ckdist_hitran_part.cpp:31-114), then
    every temperature's spectrum (temps[i], spectra[i] = (wv, kv)) filled
    into them (allocT + readFillT, alloc_bands.cpp:17-51,
    read_files.cpp:76-114).  Returns the layout rc_ckdist and
    generate_tables take (spectrum i nB + j), plus kPr per temperature."""
    wv0, kv0 = ref
    b = lbl_bands(wv0, kv0, dwv=dwv, **band_args)
    nB = len(b["first"])
    idx = np.concatenate([np.arange(f, f + c) for f, c in zip(b["first"], b["count"])])
    wvn1 = np.asarray(wv0, np.float64)[idx]
    temps = np.asarray(temps, np.float64)
    kvn, kpr = [], []
    for T, (wv, kv) in zip(temps, spectra):
        k, pr, q = lbl_fill(wv, kv, T, wvn1, dwv)
        if q != len(wvn1):
            raise RuntimeError(f"spectrum at {T} K lacks wavenumber {wvn1[q]} of the bands (filled {q} of {len(wvn1)})")
        kvn.append(k)
        kpr.append(pr)
    per_off = np.concatenate([[0], np.cumsum(b["count"])]).astype(np.int64)
    nT = len(temps)
    off = np.concatenate([per_off[:-1] + t * per_off[-1] for t in range(nT)] + [[nT * per_off[-1]]]).astype(np.int64)
    return dict(off=off, wvn=np.tile(wvn1, nT), kvn=np.concatenate(kvn), wvl=np.tile(b["wvl"], nT),
                wvr=np.tile(b["wvr"], nT), T=np.repeat(temps, nB), nB=nB, temps=temps, kPr=np.array(kpr),
                bands=b)
