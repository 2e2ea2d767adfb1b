"""ctypes binding of the HIP library ``libradchan.so`` (include/radchan.h).

The library is the product: every solve runs in its HIP kernels.  Loading
fails loudly if the shared object is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("RADCHAN_LIB") or os.path.join(HERE, "libradchan.so")

MAXLEV = 9
FACES = ("W", "E", "S", "N", "B", "T")

RAY_ABSORBED, RAY_WALL, RAY_LOST, RAY_CAPPED, RAY_STOPPED = range(5)
VARIANT_PLAIN, VARIANT_BATCHED, VARIANT_PLAIN_BANDS = 0, 1, 2


class Params(C.Structure):
    """rc_params: runtime form of src/param/param.h + NarrowBand.h."""

    _fields_ = [
        ("imax", C.c_int32), ("jmax", C.c_int32), ("kmax", C.c_int32), ("p_col", C.c_int32),
        ("Lx", C.c_float), ("Ly", C.c_float), ("Lz", C.c_float), ("nVar", C.c_int32),
        ("fact_mesh", C.c_double), ("chanHeight", C.c_double),
        ("photmax", C.c_int32), ("grid_num", C.c_int32),
        ("srt", C.c_int32), ("grey", C.c_int32), ("calcwalls", C.c_int32), ("grey_true", C.c_int32),
        ("maxi", C.c_int32 * MAXLEV), ("maxj", C.c_int32 * MAXLEV), ("maxk", C.c_int32 * MAXLEV),
        ("maxs", C.c_int32 * MAXLEV),
        ("bd", C.c_int32 * 6), ("adiabw", C.c_int32), ("adiabe", C.c_int32),
        ("eps", C.c_float * 6), ("Tw", C.c_float * 6),
        ("toll", C.c_float), ("tollS", C.c_float),
        ("nT", C.c_int32), ("nQ", C.c_int32), ("nB", C.c_int32),
        ("abscoeff", C.c_float), ("Tmins", C.c_float), ("wvmax", C.c_float),
        ("seed", C.c_uint64),
        ("sampler", C.c_int32), ("fp64", C.c_int32),
    ]

    def copy(self) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        return q

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class TablesC(C.Structure):
    """rc_tables: pointers into numpy arrays of a Tables object."""

    _fields_ = [
        ("nB", C.c_int32), ("nT", C.c_int32), ("nQ", C.c_int32), ("pad", C.c_int32),
        ("Tnb", C.POINTER(C.c_float)), ("kP", C.POINTER(C.c_float)),
        ("wvl", C.POINTER(C.c_float)), ("wvr", C.POINTER(C.c_float)),
        ("wvc", C.POINTER(C.c_float)), ("kavg", C.POINTER(C.c_float)),
        ("kq", C.POINTER(C.c_float)), ("prob", C.POINTER(C.c_float)),
        ("probg", C.POINTER(C.c_float)), ("wq", C.POINTER(C.c_float)),
    ]


class Ray(C.Structure):
    _fields_ = [
        ("nb", C.c_int32), ("g", C.c_int32), ("ic", C.c_int32), ("jc", C.c_int32), ("kc", C.c_int32),
        ("level", C.c_int32), ("steps", C.c_int32), ("status", C.c_int32), ("path_hash", C.c_uint32),
        ("sx", C.c_float), ("sy", C.c_float), ("sz", C.c_float), ("tra", C.c_float), ("de", C.c_float),
        ("de64", C.c_double), ("sumabs", C.c_double),
    ]


RAY_DTYPE = np.dtype([
    ("nb", "<i4"), ("g", "<i4"), ("ic", "<i4"), ("jc", "<i4"), ("kc", "<i4"), ("level", "<i4"),
    ("steps", "<i4"), ("status", "<i4"), ("path_hash", "<u4"), ("sx", "<f4"), ("sy", "<f4"),
    ("sz", "<f4"), ("tra", "<f4"), ("de", "<f4"), ("de64", "<f8"), ("sumabs", "<f8"),
])
assert RAY_DTYPE.itemsize == C.sizeof(Ray) == 72


class Stats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("march_ms", C.c_double), ("rays", C.c_int64), ("steps", C.c_int64),
        ("capped", C.c_int64), ("launches", C.c_int32), ("variant", C.c_int32),
        ("bulk_slots", C.c_int64), ("bulk_steps", C.c_int64),
    ]


_lib = None


def load() -> C.CDLL:
    """Load libradchan.so (built by ``radchan_amd.build``); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m radchan_amd.build` "
            "(the solver has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "mc_gpu_": (None, [P(C.c_float), P(C.c_int), P(C.c_int)]),
        "get_results_": (None, [P(C.c_float), P(C.c_float), P(C.c_int)]),
        "rc_default_params": (None, [P(Params)]),
        "rc_set_params": (C.c_int, [P(Params)]),
        "rc_get_params": (C.c_int, [P(Params)]),
        "rc_set_table_dir": (C.c_int, [C.c_char_p]),
        "rc_set_variant": (C.c_int, [C.c_int]),
        "rc_last_error": (C.c_char_p, []),
        "rc_version": (C.c_char_p, []),
        "rc_read_tables": (C.c_int, [C.c_char_p, C.c_double, P(TablesC)]),
        "rc_write_tables": (C.c_int, [C.c_char_p, P(TablesC), C.c_double]),
        "rc_grey_tables": (C.c_int, [P(Params), C.c_float, P(TablesC)]),
        "rc_sort_idx": (C.c_int, [P(C.c_float), C.c_int, P(C.c_int32)]),
        "rc_grid": (C.c_int, [P(Params), C.c_int] + [P(C.c_float)] * 6),
        "rc_spline_restrict": (C.c_int, [P(Params), C.c_int, P(C.c_float), P(C.c_float), P(C.c_float)]),
        "rc_spline_restrict_noalloc": (C.c_int, [P(Params), C.c_int, vp, vp, vp, vp]),
        "rc_set_restriction": (C.c_int, [C.c_void_p, C.c_int]),
        "rc_set_estimator": (C.c_int, [C.c_void_p, C.c_int]),
        "rc_set_reduction": (C.c_int, [C.c_void_p, C.c_int]),
        "rc_set_dropin_modes": (C.c_int, [C.c_int, C.c_int]),
        "rc_ckdist_device": (C.c_int, [C.c_int, C.c_void_p] + [C.c_void_p] * 4 + [C.c_int, C.c_double, C.c_int]
                             + [C.c_void_p] * 5 + [C.c_int, C.c_void_p]),
        "rc_ck_tables": (C.c_int, [C.c_int, C.c_int, C.c_int] + [P(C.c_double)] * 6 + [C.c_double, P(TablesC)]),
        "rc_lbl_bands": (C.c_int, [C.c_int64, P(C.c_double), P(C.c_double), C.c_double, C.c_double, C.c_double,
                                   P(C.c_double), C.c_int, P(C.c_int64), P(C.c_int64), P(C.c_double),
                                   P(C.c_double), P(C.c_double), P(C.c_int32)]),
        "rc_lbl_fill": (C.c_int64, [C.c_int64, P(C.c_double), P(C.c_double), C.c_double, C.c_double, C.c_int64,
                                    P(C.c_double), P(C.c_double), P(C.c_double)]),
        "rc_ckdist": (C.c_int, [C.c_int, P(C.c_int64), P(C.c_double), P(C.c_double), P(C.c_double),
                                P(C.c_double), C.c_int, C.c_double, C.c_int, P(C.c_double), P(C.c_double),
                                P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]),
        "rc_create": (vp, [P(Params)]),
        "rc_destroy": (None, [vp]),
        "rc_load_tables": (C.c_int, [vp, P(TablesC)]),
        "rc_set_fields": (C.c_int, [vp, C.c_int, P(P(C.c_float)), C.c_int]),
        "rc_run": (C.c_int, [vp, C.c_int]),
        "rc_sync": (C.c_int, [vp, P(Stats)]),
        "rc_results": (C.c_int, [vp, C.c_int, P(C.c_float), P(C.c_float)]),
        "rc_tallies": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "rc_surface_results": (C.c_int, [vp, C.c_int, P(C.c_float), P(C.c_float), P(C.c_double)]),
        "rc_trace": (C.c_int, [vp, C.c_int, P(C.c_int32), C.c_int, P(Ray)]),
        "rc_info": (C.c_int, [vp, C.c_int, P(C.c_float), P(C.c_int32)]),
        "rc_set_schedule": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "rc_get_schedule": (C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        "rc_set_refill": (C.c_int, [vp, C.c_int]),
        "rc_get_refill": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "rc_group_set_refill": (C.c_int, [vp, C.c_int]),
        "rc_set_graph": (C.c_int, [vp, C.c_int]),
        "rc_set_table_precision": (C.c_int, [vp, C.c_int]),
        "rc_autotune": (C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_float)]),
        "rc_group_create": (vp, [P(vp), C.c_int]),
        "rc_group_destroy": (None, [vp]),
        "rc_group_set_schedule": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "rc_group_run": (C.c_int, [vp]),
        "rc_group_sync": (C.c_int, [vp, P(Stats)]),
        "rc_device_synchronize": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().rc_last_error().decode()


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))
