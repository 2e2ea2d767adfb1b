"""Host-side session over the HIP library (the rc_* extension API).

``Solver`` keeps parameters, grids and tables resident on the GPU and solves
one or more temperature snapshots per launch.  ``mc_gpu``/``get_results``
mirror the reference's Fortran-facing entry points (mc_gpu_noint.cu:660,
:876) on numpy arrays.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import lib as L
from .tables import Tables


class Solver:
    def __init__(self, params: L.Params, tables: Tables | None = None):
        self.lib = L.load()
        self.p = params.copy()
        self.ctx = self.lib.rc_create(C.byref(self.p))
        if not self.ctx:
            raise RuntimeError("rc_create failed: " + L.last_error())
        if self.p.grey:
            self.p.srt = 0
        self.kloc = self.p.kmax // self.p.p_col
        self.ncell = self.p.imax * self.p.jmax * self.kloc
        self.nsurf = self.p.jmax * self.kloc
        self.nsnap = 0
        self.tables = None
        if tables is not None:
            self.load_tables(tables)

    def close(self) -> None:
        if self.ctx:
            self.lib.rc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_tables(self, tables: Tables) -> None:
        self.tables = tables
        L.check(self.lib.rc_load_tables(self.ctx, C.byref(tables.cview())), "rc_load_tables")

    def set_fields(self, fields: Sequence[np.ndarray], zstart: int = 1) -> None:
        arrs = [np.ascontiguousarray(f, dtype=np.float32).reshape(-1) for f in fields]
        ptrs = (C.POINTER(C.c_float) * len(arrs))(*[L.fptr(a) for a in arrs])
        L.check(self.lib.rc_set_fields(self.ctx, len(arrs), ptrs, zstart), "rc_set_fields")
        self.nsnap = len(arrs)
        self._keep = arrs

    def run(self, variant: int = L.VARIANT_BATCHED) -> None:
        L.check(self.lib.rc_run(self.ctx, variant), "rc_run")

    def sync(self) -> dict:
        st = L.Stats()
        L.check(self.lib.rc_sync(self.ctx, C.byref(st)), "rc_sync")
        return {k: getattr(st, k) for k, _ in L.Stats._fields_}

    def solve(self, variant: int = L.VARIANT_BATCHED) -> dict:
        self.run(variant)
        return self.sync()

    def results(self, s: int = 0):
        """(mean, sdev) per emitting cell, shaped [k][j][i]."""
        m = np.zeros(self.ncell, np.float32)
        sd = np.zeros(self.ncell, np.float32)
        L.check(self.lib.rc_results(self.ctx, s, L.fptr(m), L.fptr(sd)), "rc_results")
        shp = (self.kloc, self.p.jmax, self.p.imax)
        return m.reshape(shp), sd.reshape(shp)

    def tallies(self, s: int = 0) -> np.ndarray:
        """Per (cell, run) FP64 tallies, shaped [ncell][nVar]."""
        d = np.zeros(self.ncell * self.p.nVar, np.float64)
        L.check(self.lib.rc_tallies(self.ctx, s, L.dptr(d)), "rc_tallies")
        return d.reshape(self.ncell, self.p.nVar)

    def surface(self, s: int = 0):
        """Wall (mean, sdev, tallies): mean/sdev [2][k][j], tallies [2][k][j][nVar]."""
        n = 2 * self.nsurf
        m = np.zeros(n, np.float32)
        sd = np.zeros(n, np.float32)
        d = np.zeros(n * self.p.nVar, np.float64)
        L.check(self.lib.rc_surface_results(self.ctx, s, L.fptr(m), L.fptr(sd), L.dptr(d)), "rc_surface_results")
        shp = (2, self.kloc, self.p.jmax)
        return m.reshape(shp), sd.reshape(shp), d.reshape(shp + (self.p.nVar,))

    def trace(self, rays, surface: bool = False) -> np.ndarray:
        """Trace rays of snapshot 0: rows (i|bound, j, k, v, h), k global."""
        r = np.ascontiguousarray(np.asarray(rays, dtype=np.int32).reshape(-1, 5))
        out = np.zeros(len(r), L.RAY_DTYPE)
        L.check(self.lib.rc_trace(self.ctx, int(surface), L.iptr(r), len(r),
                                  out.ctypes.data_as(C.POINTER(L.Ray))), "rc_trace")
        return out

    def set_schedule(self, wg_per_cu: int = 0, threshold: int = 0, steps_per_iter: int = 0,
                     refill: int = -1) -> None:
        L.check(self.lib.rc_set_schedule(self.ctx, wg_per_cu, threshold, steps_per_iter), "rc_set_schedule")
        L.check(self.lib.rc_set_refill(self.ctx, refill), "rc_set_refill")

    def set_table_precision(self, bf16: bool) -> None:
        L.check(self.lib.rc_set_table_precision(self.ctx, int(bf16)), "rc_set_table_precision")

    def set_restriction(self, mode: int) -> None:
        """Coarse-level temperatures: 0 texture mean (default), 1 cubic
        splines (interp3D); applies to fields set afterwards."""
        L.check(self.lib.rc_set_restriction(self.ctx, mode), "rc_set_restriction")

    def set_estimator(self, mode: int) -> None:
        """Band/g pdf: 0 the reference's nearest-row CDFs (default), 1 exact
        at Tmax; applies to fields set afterwards."""
        L.check(self.lib.rc_set_estimator(self.ctx, mode), "rc_set_estimator")

    def set_reduction(self, mode: int) -> None:
        """Batched tallies: 0 FP64 atomics (default), 1 deterministic
        (per-ray tallies summed in a fixed order with compensation)."""
        L.check(self.lib.rc_set_reduction(self.ctx, mode), "rc_set_reduction")

    def set_graph(self, on: bool = True) -> None:
        L.check(self.lib.rc_set_graph(self.ctx, int(on)), "rc_set_graph")

    def autotune(self) -> dict:
        wg, th, sp, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        L.check(self.lib.rc_autotune(self.ctx, C.byref(wg), C.byref(th), C.byref(sp), C.byref(ms)), "rc_autotune")
        rf = C.c_int32()
        L.check(self.lib.rc_get_refill(self.ctx, C.byref(rf)), "rc_get_refill")
        return {"wg_per_cu": wg.value, "threshold": th.value, "steps_per_iter": sp.value,
                "refill": rf.value, "march_ms": ms.value}

    def info(self, s: int = 0) -> dict:
        f = np.zeros(5, np.float32)
        i = np.zeros(2, np.int32)
        L.check(self.lib.rc_info(self.ctx, s, L.fptr(f), L.iptr(i)), "rc_info")
        return dict(Tmax=float(f[0]), kappamax=float(f[1]), Ibmax=float(f[2]), Ibmax_w=float(f[3]),
                    Ibmax_e=float(f[4]), tm=int(i[0]), tstar=int(i[1]))


def spline_restrict(p: L.Params, level: int, T: np.ndarray, walls: bool = True):
    """interp3D on the GPU (rc_spline_restrict): (Tc [km][jm][im], Tw
    [2][km][jm] or None) of grid level `level` from the field T."""
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1)
    im, jm, km = p.maxi[level], p.maxj[level], p.maxk[level]
    tc = np.zeros(im * jm * km, np.float32)
    tw = np.zeros(2 * jm * km, np.float32) if walls else None
    L.check(L.load().rc_spline_restrict(C.byref(p), level, L.fptr(T), L.fptr(tc),
                                        L.fptr(tw) if walls else None), "rc_spline_restrict")
    return tc.reshape(km, jm, im), (tw.reshape(2, km, jm) if walls else None)


# ---- the reference's Fortran-facing entry points ---------------------------

def set_params(p: L.Params) -> None:
    L.check(L.load().rc_set_params(C.byref(p)), "rc_set_params")


def set_table_dir(d: str) -> None:
    L.check(L.load().rc_set_table_dir(d.encode()), "rc_set_table_dir")


def set_dropin_modes(estimator: int = 0, restriction: int = 0) -> None:
    """Estimator and restriction modes of the mc_gpu_ session."""
    L.check(L.load().rc_set_dropin_modes(estimator, restriction), "rc_set_dropin_modes")


def mc_gpu(T: np.ndarray, zstart: int = 1, rank: int = 0) -> None:
    """MC_GPU(TMC, xstart(3), nrank): launch the solve of field T."""
    T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1)
    z = C.c_int(zstart)
    r = C.c_int(rank)
    L.load().mc_gpu_(L.fptr(T), C.byref(z), C.byref(r))


def get_results(res: np.ndarray, var: np.ndarray, rank: int = 0) -> None:
    """GET_RESULTS(qr, vr, nrank): wait and scatter into (imax+2)(jmax+2)(kloc+2) arrays."""
    r = C.c_int(rank)
    L.load().get_results_(L.fptr(res), L.fptr(var), C.byref(r))


class Group:
    """Snapshots of several sessions (grids of different sizes) solved by
    one batched launch (rc_group_*); the results land in the sessions."""

    def __init__(self, solvers: Sequence[Solver]):
        self.lib = L.load()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.ctx for s in self.solvers])
        self.g = self.lib.rc_group_create(arr, len(self.solvers))
        if not self.g:
            raise RuntimeError("rc_group_create failed: " + L.last_error())

    def close(self) -> None:
        if self.g:
            self.lib.rc_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, wg_per_cu: int = 0, threshold: int = 0, steps_per_iter: int = 0,
                     refill: int = -1) -> None:
        L.check(self.lib.rc_group_set_schedule(self.g, wg_per_cu, threshold, steps_per_iter), "rc_group_set_schedule")
        L.check(self.lib.rc_group_set_refill(self.g, refill), "rc_group_set_refill")

    def run(self) -> None:
        L.check(self.lib.rc_group_run(self.g), "rc_group_run")

    def sync(self) -> dict:
        st = L.Stats()
        L.check(self.lib.rc_group_sync(self.g, C.byref(st)), "rc_group_sync")
        return {k: getattr(st, k) for k, _ in L.Stats._fields_}

    def solve(self) -> dict:
        self.run()
        return self.sync()

