"""Synthetic temperature fields in the solver's input layout.

A field is a float32 array of (imax+2)(jmax+2)(kmax+2) temperatures, i
fastest (idx_T, src/param/param.h:132), including ghost planes.  The x ghost
planes are set so that the wall midpoint, which the reference reads as the
adiabatic wall temperature (mc_gpu_noint.cu:391-399), equals the wall
temperature; y and z ghosts repeat the opposite interior plane (periodic).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def centres_x(p: L.Params) -> np.ndarray:
    """Level-0 x cell centres of the stretched grid (from the library)."""
    im, jm, km = p.maxi[0], p.maxj[0], p.maxk[0]
    a = [np.zeros(n, np.float32) for n in (im + 2, im + 1, jm + 2, jm + 1, km + 2, km + 1)]
    L.check(L.load().rc_grid(C.byref(p), 0, *[L.fptr(x) for x in a]), "rc_grid")
    return a[0]


def _box3(a: np.ndarray) -> np.ndarray:
    """3x3x3 box average, periodic in y and z (axes 1, 0 of [k][j][i]),
    clamped in x."""
    out = np.zeros_like(a)
    pad = np.concatenate([a[:, :, :1], a, a[:, :, -1:]], axis=2)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            sh = np.roll(np.roll(pad, dk, axis=0), dj, axis=1)
            out += sh[:, :, 0:-2] + sh[:, :, 1:-1] + sh[:, :, 2:]
    return out / 27.0


def channel_field(p: L.Params, T_wall: float, T_centre: float, noise: float = 0.0, seed: int = 0) -> np.ndarray:
    """T(x) = Tw + (Tc - Tw) (1 - (2x/Lx - 1)^2) plus Gaussian noise of
    relative size `noise`, box-smoothed over 3x3x3 cells."""
    im, jm, km = p.imax, p.jmax, p.kmax
    x = centres_x(p)[1:im + 1].astype(np.float64)
    prof = T_wall + (T_centre - T_wall) * (1.0 - (2.0 * x / p.Lx - 1.0) ** 2)
    inner = np.broadcast_to(prof[None, None, :], (km, jm, im)).copy()
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        inner = inner * (1.0 + _box3(rng.normal(0.0, noise, size=inner.shape)))
    return pad_field(p, inner, T_wall)


def uniform_field(p: L.Params, T_gas: float, T_wall: float) -> np.ndarray:
    inner = np.full((p.kmax, p.jmax, p.imax), T_gas, np.float64)
    return pad_field(p, inner, T_wall)


def pad_field(p: L.Params, inner: np.ndarray, T_wall: float) -> np.ndarray:
    """Wrap an interior [k][j][i] array with ghost planes; returns the flat
    float32 input array."""
    im, jm, km = p.imax, p.jmax, p.kmax
    T = np.zeros((km + 2, jm + 2, im + 2), np.float32)
    T[1:km + 1, 1:jm + 1, 1:im + 1] = inner.astype(np.float32)
    T[:, :, 0] = (2.0 * T_wall - T[:, :, 1].astype(np.float64)).astype(np.float32)
    T[:, :, im + 1] = (2.0 * T_wall - T[:, :, im].astype(np.float64)).astype(np.float32)
    T[:, 0, :] = T[:, jm, :]
    T[:, jm + 1, :] = T[:, 1, :]
    T[0, :, :] = T[km, :, :]
    T[km + 1, :, :] = T[1, :, :]
    return np.ascontiguousarray(T.reshape(-1))


def wavy_channel_field(p: L.Params, T_wall: float, T_centre: float, amp: float = 0.03) -> np.ndarray:
    """T(x, y, z) = [Tw + (Tc - Tw) (1 - (2x/Lx - 1)^2)] (1 + amp sin(2 pi j/jmax)
    cos(2 pi k/kmax)), built 64 z planes at a time in FP32 (for fields of
    hundreds of millions of cells, where the noisy channel_field's FP64
    temporaries do not fit comfortably); ghost planes as pad_field."""
    im, jm, km = p.imax, p.jmax, p.kmax
    x = centres_x(p)[1:im + 1].astype(np.float64)
    prof = (T_wall + (T_centre - T_wall) * (1.0 - (2.0 * x / p.Lx - 1.0) ** 2)).astype(np.float32)
    sy = (amp * np.sin(2.0 * np.pi * np.arange(1, jm + 1) / jm)).astype(np.float32)
    cz = np.cos(2.0 * np.pi * np.arange(1, km + 1) / km).astype(np.float32)
    T = np.empty((km + 2, jm + 2, im + 2), np.float32)
    tw2 = np.float32(2.0 * T_wall)
    for k0 in range(0, km, 64):
        k1 = min(km, k0 + 64)
        blk = T[k0 + 1:k1 + 1]
        np.multiply(prof[None, None, :], (np.float32(1.0) + sy[None, :] * cz[k0:k1, None])[:, :, None],
                    out=blk[:, 1:jm + 1, 1:im + 1])
        blk[:, 1:jm + 1, 0] = tw2 - blk[:, 1:jm + 1, 1]
        blk[:, 1:jm + 1, im + 1] = tw2 - blk[:, 1:jm + 1, im]
        blk[:, 0] = blk[:, jm]
        blk[:, jm + 1] = blk[:, 1]
    T[0] = T[km]
    T[km + 1] = T[1]
    return T.reshape(-1)
