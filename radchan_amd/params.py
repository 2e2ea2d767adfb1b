"""Parameter sets: the reference defaults and the benchmark/test workloads.

``reference_params()`` returns the values of src/param/param.h and
src/param/NarrowBand.h.  The named workloads follow the five configurations
the library is measured on (grey 32^3; narrow band 64^3 and 128^3; a
192x128x192 multigrid field with wall fluxes and specular z walls; a batch
of snapshots).
"""
from __future__ import annotations

import ctypes as C

from . import lib as L

TWO_PI = 6.283185307179586


def reference_params() -> L.Params:
    p = L.Params()
    L.load().rc_default_params(C.byref(p))
    return p


def _levels(p: L.Params, sizes, maxs) -> None:
    p.grid_num = len(sizes)
    for g in range(L.MAXLEV):
        if g < len(sizes):
            p.maxi[g], p.maxj[g], p.maxk[g] = sizes[g]
            p.maxs[g] = maxs[g]
        else:
            p.maxi[g] = p.maxj[g] = p.maxk[g] = 1
            p.maxs[g] = 1


def base_params(imax: int, jmax: int, kmax: int, *, photmax: int = 100, nVar: int = 1, fact_mesh: float = 0.0,
                grey: bool = False, grey_true: bool = False, srt: int = 1, seed: int = 0,
                levels=None, maxs=None) -> L.Params:
    """A one-GPU parameter set (p_col = 1, the whole domain) on a 2 x 2pi x 2pi
    channel with black x walls and periodic y and z."""
    p = reference_params()
    p.imax, p.jmax, p.kmax, p.p_col = imax, jmax, kmax, 1
    p.Lx, p.Ly, p.Lz = 2.0, TWO_PI, TWO_PI
    p.fact_mesh = fact_mesh
    p.nVar, p.photmax = nVar, photmax
    p.grey, p.grey_true = int(grey), int(grey_true)
    p.srt = 0 if grey else srt
    p.calcwalls = 0
    if levels is None:
        levels, maxs = [(imax, jmax, kmax)], [100000]
    _levels(p, levels, maxs)
    for f in range(6):
        p.eps[f], p.Tw[f] = 1.0, 0.0
    p.bd[:] = [2, 2, 1, 1, 1, 1]
    p.adiabw = p.adiabe = 1
    if grey:
        p.nB, p.nT, p.nQ = (1, 65, 1) if grey_true else (1000, 65, 1)
    else:
        p.nB, p.nT, p.nQ = 360, 20, 16
    p.seed = seed
    p.sampler = 0                 # Philox directions (the workloads' stated streams)
    return p


def config(name: str) -> L.Params:
    """Parameters of a named workload."""
    if name == "grey32":          # plain grey OERM, uniform x grid, black 500 K walls
        p = base_params(32, 32, 32, photmax=100, grey=True, fact_mesh=0.0)
        p.adiabw = p.adiabe = 0
        p.Tw[0] = p.Tw[1] = 500.0
        return p
    if name == "slab32":          # true grey slab check: uniform T, black walls
        p = base_params(32, 32, 32, photmax=100, grey=True, grey_true=True, fact_mesh=0.0)
        p.adiabw = p.adiabe = 0
        p.Tw[0] = p.Tw[1] = 500.0
        return p
    if name == "nb64":            # narrow band, stretched x, sorted bands
        return base_params(64, 64, 64, photmax=200, fact_mesh=1.5)
    if name == "nb128":
        return base_params(128, 128, 128, photmax=200, fact_mesh=1.5)
    if name == "mg192":           # multigrid, wall fluxes, specular z walls
        p = base_params(192, 128, 192, photmax=100, fact_mesh=1.5,
                        levels=[(192, 128, 192), (96, 64, 96), (48, 32, 48)], maxs=[4, 8, 64])
        p.calcwalls = 1
        p.bd[4] = p.bd[5] = 3
        return p
    if name == "big256":          # the 256 x 128 x 256 narrow-band field
        return base_params(256, 128, 256, photmax=100, fact_mesh=1.5)
    if name == "prod28":          # param.h's solver settings on 28x96x128 (its coarsening pattern scaled)
        r = reference_params()
        sizes = [(28 // a, 96 // b, 128 // c) for a, b, c in
                 zip((1, 1, 1, 1, 1, 7, 7, 7, 14), (1, 3, 6, 12, 24, 48, 96, 96, 96),
                     (1, 2, 4, 8, 16, 32, 64, 128, 128))]
        p = base_params(28, 96, 128, photmax=r.photmax, fact_mesh=r.fact_mesh, levels=sizes,
                        maxs=[r.maxs[g] for g in range(9)])
        p.bd[:] = list(r.bd)
        p.Lx, p.Ly, p.Lz = r.Lx, r.Ly, r.Lz
        p.nB, p.nT, p.nQ = r.nB, r.nT, r.nQ   # NarrowBand.h's table sizes
        p.sampler = r.sampler                 # param.h random 1: Sobol directions
        return p
    if name == "slab406":         # param.h as it stands: 406x576x2304 in 48 z slabs, one GPU's slab solved
        return reference_params()
    if name == "ens64x96":        # the small snapshots of the ensemble
        return base_params(64, 96, 64, photmax=100, fact_mesh=1.5)
    raise KeyError(name)
