"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.  It shares the parameter and table structures of the
library binding (the C layouts are identical).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RADCHAN_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    from radchan_amd import lib as L
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    f = C.c_float
    sig = {
        "orc_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "orc_u01": (f, [C.c_uint32]),
        "orc_sincos2pi": (None, [f, P(f), P(f)]),
        "orc_expm1_exp": (None, [f, P(f), P(f)]),
        "orc_sincos2pi_d": (None, [f, P(C.c_double), P(C.c_double)]),
        "orc_expm1_exp_d": (None, [C.c_double, P(C.c_double), P(C.c_double)]),
        "orc_planck": (f, [f, f]),
        "orc_upper_bound": (C.c_int, [P(f), C.c_int, f]),
        "orc_sort_idx": (None, [P(f), C.c_int, P(C.c_int32)]),
        "orc_iso_dir": (None, [f, f, P(f)]),
        "orc_sobol2": (None, [C.c_uint32, P(C.c_uint32)]),
        "orc_sobol_uniforms": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P(f)]),
        "orc_lamb_dir": (None, [f, f, C.c_int, P(f)]),
        "orc_grid": (C.c_int, [P(L.Params), C.c_int] + [P(f)] * 6),
        "orc_tmax": (f, [P(L.Params), P(f)]),
        "orc_read_tables": (C.c_int, [C.c_char_p, C.c_double, P(L.TablesC)]),
        "orc_grey_tables": (C.c_int, [P(L.Params), f, P(L.TablesC)]),
        "orc_create": (vp, [P(L.Params), P(L.TablesC), P(f), C.c_int, C.c_char_p, C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_snapshot": (None, [vp, C.c_int]),
        "orc_plant": (None, [C.c_int]),
        "orc_tex3": (f, [vp, f, f, f]),
        "orc_level_T": (C.c_int, [vp, C.c_int, P(f), P(f), P(f)]),
        "orc_get_info": (None, [vp, C.c_void_p]),
        "orc_band_sequence": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     P(C.c_int32), P(C.c_int32)]),
        "orc_trace_fluid": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(L.Ray)]),
        "orc_trace_surface": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(L.Ray)]),
        "orc_solve_fluid": (C.c_int, [vp, C.c_int, P(C.c_int32), C.c_int, P(f), P(C.c_double), P(C.c_double),
                                      P(C.c_int64), C.c_int]),
        "orc_solve_surface": (C.c_int, [vp, C.c_int, P(C.c_int32), C.c_int, P(f), P(C.c_double), P(C.c_double),
                                        P(C.c_int64), C.c_int]),
        "orc_solve_fluid_groups": (C.c_int, [vp, C.c_int, P(C.c_int32), C.c_int, P(C.c_int32), C.c_int,
                                             P(C.c_double), C.c_int]),
        "orc_interp3d": (C.c_int, [C.c_int] * 3 + [P(f)] * 4 + [C.c_int] * 3 + [P(f)] * 5),
        "orc_set_restriction": (C.c_int, [vp, C.c_int]),
        "orc_set_estimator": (C.c_int, [vp, C.c_int]),
        "orc_spline": (C.c_int, [P(f), P(f), C.c_int, P(f), P(f), C.c_int, P(f)]),
        "orc_ck_tables": (C.c_int, [C.c_int, C.c_int, P(C.c_int64)] + [P(C.c_double)] * 5 + [C.c_int, C.c_double, C.c_int]
                          + [P(C.c_double)] * 7 + [C.c_int]),
        "orc_ckdist": (C.c_int, [C.c_int, P(C.c_int64), P(C.c_double), P(C.c_double), P(C.c_double),
                                 P(C.c_double), C.c_int, C.c_double, C.c_int, P(C.c_double), P(C.c_double),
                                 P(C.c_double), P(C.c_double), P(C.c_double), C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Info(C.Structure):
    _fields_ = [("Tmax", C.c_float), ("kappamax", C.c_float), ("Ibmax", C.c_float), ("Ibmax_w", C.c_float),
                ("Ibmax_e", C.c_float), ("tm", C.c_int32), ("tstar", C.c_int32), ("n_emit", C.c_int32),
                ("kappamax64", C.c_double), ("Ibmax64", C.c_double)]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ---- building blocks ---------------------------------------------------------

def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def u01(x: int) -> float:
    return load().orc_u01(x)


def sincos2pi(u: float):
    s, c = C.c_float(), C.c_float()
    load().orc_sincos2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


def expm1_exp(x: float):
    a, b = C.c_float(), C.c_float()
    load().orc_expm1_exp(x, C.byref(a), C.byref(b))
    return a.value, b.value


def sincos2pi_d(u: float):
    s, c = C.c_double(), C.c_double()
    load().orc_sincos2pi_d(u, C.byref(s), C.byref(c))
    return s.value, c.value


def expm1_exp_d(x: float):
    a, b = C.c_double(), C.c_double()
    load().orc_expm1_exp_d(x, C.byref(a), C.byref(b))
    return a.value, b.value


def planck(T: float, nu: float) -> float:
    return load().orc_planck(T, nu)


def upper_bound(cdf: np.ndarray, R: float) -> int:
    cdf = np.ascontiguousarray(cdf, np.float32)
    return load().orc_upper_bound(_f(cdf), len(cdf), R)


def sort_idx(kavg: np.ndarray) -> np.ndarray:
    kavg = np.ascontiguousarray(kavg, np.float32)
    out = np.zeros(len(kavg), np.int32)
    load().orc_sort_idx(_f(kavg), len(kavg), out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def sobol2(h: int):
    """Point h of the two-dimensional Sobol sequence as 32-bit fractions."""
    x = (C.c_uint32 * 2)()
    load().orc_sobol2(h, x)
    return int(x[0]), int(x[1])


def sobol_uniforms(seed: int, h: int, v: int, cell: int, tag: int = 0):
    """(phi, theta) uniforms of fluid ray h of run v of cell id `cell` under
    the Sobol sampler."""
    u = (C.c_float * 2)()
    load().orc_sobol_uniforms(seed, h, v, cell, tag, u)
    return float(u[0]), float(u[1])


def iso_dir(u0: float, u1: float):
    s = (C.c_float * 3)()
    load().orc_iso_dir(u0, u1, s)
    return tuple(s)


def lamb_dir(u0: float, u1: float, bound: int):
    s = (C.c_float * 3)()
    load().orc_lamb_dir(u0, u1, bound, s)
    return tuple(s)


def grid(p, level: int):
    im, jm, km = p.maxi[level], p.maxj[level], p.maxk[level]
    a = [np.zeros(n, np.float32) for n in (im + 2, im + 1, jm + 2, jm + 1, km + 2, km + 1)]
    rc = load().orc_grid(C.byref(p), level, *[_f(x) for x in a])
    assert rc == 0
    return dict(zip(("x", "xu", "y", "yu", "z", "zu"), a))


def interp3d(p, level: int, T: np.ndarray, walls: bool = True):
    """interp3D of the field T to grid level `level`: (Tc [km][jm][im],
    Tw [2][km][jm] or None)."""
    g0, gc = grid(p, 0), grid(p, level)
    im, jm, km = p.maxi[level], p.maxj[level], p.maxk[level]
    T = np.ascontiguousarray(T, np.float32).reshape(-1)
    tc = np.zeros(im * jm * km, np.float32)
    tw = np.zeros(2 * jm * km, np.float32) if walls else None
    rc = load().orc_interp3d(p.imax, p.jmax, p.kmax, _f(g0["x"]), _f(g0["y"]), _f(g0["z"]), _f(T), im, jm, km,
                             _f(gc["x"]), _f(gc["y"]), _f(gc["z"]), _f(tc), _f(tw) if walls else None)
    assert rc == 0
    return tc.reshape(km, jm, im), (tw.reshape(2, km, jm) if walls else None)


def spline(x, y, xs):
    """The spline routines on one line: (y2, values at xs), FP32."""
    x, y, xs = (np.ascontiguousarray(a, np.float32) for a in (x, y, xs))
    y2 = np.zeros(len(x), np.float32)
    ys = np.zeros(len(xs), np.float32)
    assert load().orc_spline(_f(x), _f(y), len(x), _f(y2), _f(xs), len(xs), _f(ys)) == 0
    return y2, ys


def tmax(p, T: np.ndarray) -> float:
    T = np.ascontiguousarray(T, np.float32)
    return load().orc_tmax(C.byref(p), _f(T))


def read_tables(directory: str, nB: int, nT: int, nQ: int, chanHeight: float):
    from radchan_amd.tables import Tables
    t = Tables(nB, nT, nQ)
    rc = load().orc_read_tables(directory.encode(), chanHeight, C.byref(t.cview()))
    if rc != 0:
        raise RuntimeError(f"orc_read_tables failed ({rc})")
    return t


def grey_tables(p, Tmax: float):
    from radchan_amd.tables import Tables
    t = Tables(p.nB, p.nT, p.nQ)
    rc = load().orc_grey_tables(C.byref(p), Tmax, C.byref(t.cview()))
    if rc != 0:
        raise RuntimeError(f"orc_grey_tables failed ({rc})")
    t.wq[:] = 1.0
    return t


class planted:
    """Context manager: the oracle commits planted mistake `pid` inside the
    block (test-only; see PLANT_* in oracle.cpp)."""

    GREY_ROW, GREY_WVL, WALL_ROW, WALL_G_UNIFORM, TEX_NO_SHIFT, TMAP_NO_GHOST, XWALL_GHOST = range(1, 8)

    def __init__(self, pid: int):
        self.pid = pid

    def __enter__(self):
        load().orc_plant(self.pid)
        return self

    def __exit__(self, *exc):
        load().orc_plant(0)
        return False


# ---- solver ------------------------------------------------------------------

class Oracle:
    """CPU restatement of the solver for one temperature field."""

    def __init__(self, p, tables, T: np.ndarray, zstart: int = 1, snapshot: int = 0, restriction: int = 0,
                 estimator: int = 0):
        from radchan_amd import lib as L
        self.L = L
        self.lib = load()
        self.p = p.copy()
        if self.p.grey:
            self.p.srt = 0
        self.T = np.ascontiguousarray(T, np.float32).reshape(-1)
        if tables is None:
            tables = grey_tables(self.p, tmax(self.p, self.T))
        self.tables = tables
        err = C.create_string_buffer(256)
        self.ctx = self.lib.orc_create(C.byref(self.p), C.byref(tables.cview()), _f(self.T), zstart, err, 256)
        if not self.ctx:
            raise RuntimeError("orc_create: " + err.value.decode())
        if snapshot:
            self.lib.orc_set_snapshot(self.ctx, snapshot)
        if restriction:
            assert self.lib.orc_set_restriction(self.ctx, restriction) == 0
        if estimator:
            assert self.lib.orc_set_estimator(self.ctx, estimator) == 0

    def close(self):
        if self.ctx:
            self.lib.orc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = Info()
        self.lib.orc_get_info(self.ctx, C.byref(i))
        return {k: getattr(i, k) for k, _ in Info._fields_}

    def tex3(self, tx: float, ty: float, tz: float) -> float:
        return self.lib.orc_tex3(self.ctx, tx, ty, tz)

    def level_T(self, g: int):
        """Temperatures the march reads on level g: cells [km][jm][im],
        adiabatic x walls [2][km][jm], bd=4 z walls [2][jm][im]."""
        im, jm, km = self.p.maxi[g], self.p.maxj[g], self.p.maxk[g]
        tc = np.zeros(im * jm * km, np.float32)
        tw = np.zeros(2 * jm * km, np.float32)
        tz = np.zeros(2 * im * jm, np.float32)
        assert self.lib.orc_level_T(self.ctx, g, _f(tc), _f(tw), _f(tz)) == 0
        return tc.reshape(km, jm, im), tw.reshape(2, km, jm), tz.reshape(2, jm, im)

    def band_sequence(self, i: int, j: int, k: int, v: int, surface: bool = False, bound: int = 0):
        P = self.p.photmax
        nb = np.zeros(P, np.int32)
        g = np.zeros(P, np.int32)
        self.lib.orc_band_sequence(self.ctx, int(surface), bound, i, j, k, v,
                                   nb.ctypes.data_as(C.POINTER(C.c_int32)), g.ctypes.data_as(C.POINTER(C.c_int32)))
        return nb, g

    def trace(self, rays, surface: bool = False) -> np.ndarray:
        L = self.L
        r = np.asarray(rays, dtype=np.int32).reshape(-1, 5)
        out = np.zeros(len(r), L.RAY_DTYPE)
        rec = L.Ray()
        fn = self.lib.orc_trace_surface if surface else self.lib.orc_trace_fluid
        for q, (a, j, k, v, h) in enumerate(r):
            fn(self.ctx, int(a), int(j), int(k), int(v), int(h), C.byref(rec))
            out[q] = tuple(getattr(rec, n) for n in L.RAY_DTYPE.names)
        return out

    def solve(self, cells, surface: bool = False, precision: int = 0, nthreads: int = 0):
        """Per (cell, run) tallies of the listed cells ((i,j,k) or (bound,j,k)
        rows): returns (de32 float32, de32d float64, de64 float64, steps)."""
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 3))
        n = len(c)
        nv = self.p.nVar
        de32 = np.zeros(n * nv, np.float32)
        de32d = np.zeros(n * nv, np.float64)
        de64 = np.zeros(n * nv, np.float64)
        steps = np.zeros(n, np.int64)
        fn = self.lib.orc_solve_surface if surface else self.lib.orc_solve_fluid
        rc = fn(self.ctx, precision, c.ctypes.data_as(C.POINTER(C.c_int32)), n, _f(de32),
                de32d.ctypes.data_as(C.POINTER(C.c_double)), de64.ctypes.data_as(C.POINTER(C.c_double)),
                steps.ctypes.data_as(C.POINTER(C.c_int64)), nthreads)
        assert rc == 0
        return de32.reshape(n, nv), de32d.reshape(n, nv), de64.reshape(n, nv), steps


    def solve_groups(self, cells, group_of_rank, precision: int = 0, nthreads: int = 0) -> np.ndarray:
        """Per (cell, run, band group) FP64 sums of the per-ray tallies
        (FP32 tallies at precision 0, full FP64 at precision 1)."""
        c = np.ascontiguousarray(np.asarray(cells, dtype=np.int32).reshape(-1, 3))
        gr = np.ascontiguousarray(np.asarray(group_of_rank, dtype=np.int32))
        ng = int(gr.max()) + 1
        out = np.zeros(len(c) * self.p.nVar * ng, np.float64)
        rc = self.lib.orc_solve_fluid_groups(self.ctx, precision, c.ctypes.data_as(C.POINTER(C.c_int32)), len(c),
                                             gr.ctypes.data_as(C.POINTER(C.c_int32)), ng,
                                             out.ctypes.data_as(C.POINTER(C.c_double)), nthreads)
        assert rc == 0
        return out.reshape(len(c), self.p.nVar, ng)


def ckdist(off, wvn, kvn, wvl, wvr, gq, n_pwrk: int = 20000, pwr: float = 0.1, nthreads: int = 0):
    """Correlated-k distributions of line-by-line spectra (the restatement of
    correlatedKT): returns (kq [nspec][nq], kavg, kmin, kmax)."""
    d = lambda a: np.ascontiguousarray(a, np.float64)
    off = np.ascontiguousarray(off, np.int64)
    wvn, kvn, wvl, wvr, gq = d(wvn), d(kvn), d(wvl), d(wvr), d(gq)
    ns, nq = len(off) - 1, len(gq)
    kq = np.zeros((ns, nq))
    kavg, kmin, kmax = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = load().orc_ckdist(ns, off.ctypes.data_as(C.POINTER(C.c_int64)), dp(wvn), dp(kvn), dp(wvl), dp(wvr),
                           n_pwrk, pwr, nq, dp(gq), dp(kq), dp(kavg), dp(kmin), dp(kmax), nthreads)
    assert rc == 0
    return kq, kavg, kmin, kmax


def ck_tables(spec: dict, gq, wq, n_pwrk: int = 20000, pwr: float = 0.1, nthreads: int = 0):
    """calcProbT after correlatedKT on spectra in synthetic_lbl's layout:
    returns dict kq [nT][nB][nq], kavg [nT][nB], kP [nT], cuprob [nT][nB],
    cuprob_g [nT][nB][nq] (FP64)."""
    d = lambda a: np.ascontiguousarray(a, np.float64)
    nB, nT, nq = spec["nB"], len(spec["temps"]), len(gq)
    off = np.ascontiguousarray(spec["off"], np.int64)
    wvn, kvn, wvl, wvr, T, gq, wq = (d(spec["wvn"]), d(spec["kvn"]), d(spec["wvl"]), d(spec["wvr"]),
                                     d(spec["temps"]), d(gq), d(wq))
    out = dict(kq=np.zeros((nT, nB, nq)), kavg=np.zeros((nT, nB)), kP=np.zeros(nT), cuprob=np.zeros((nT, nB)),
               cuprob_g=np.zeros((nT, nB, nq)))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = load().orc_ck_tables(nB, nT, off.ctypes.data_as(C.POINTER(C.c_int64)), dp(wvn), dp(kvn), dp(wvl), dp(wvr),
                              dp(T), n_pwrk, pwr, nq, dp(gq), dp(wq), dp(out["kq"]), dp(out["kavg"]), dp(out["kP"]),
                              dp(out["cuprob"]), dp(out["cuprob_g"]), nthreads)
    assert rc == 0
    return out


def all_cells(p, zstart: int = 1) -> np.ndarray:
    """Emitting cells (i, j, k global) in the library's order (i fastest)."""
    kloc = p.kmax // p.p_col
    k, j, i = np.meshgrid(np.arange(zstart, zstart + kloc), np.arange(1, p.jmax + 1), np.arange(1, p.imax + 1),
                          indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.int32)


def all_wall_cells(p, zstart: int = 1) -> np.ndarray:
    """Wall cells (bound, j, k global) in the library's order."""
    kloc = p.kmax // p.p_col
    b, k, j = np.meshgrid(np.arange(2), np.arange(zstart, zstart + kloc), np.arange(1, p.jmax + 1), indexing="ij")
    return np.stack([b.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.int32)
