"""Small solver cases shared by the parity tests (sizes the CPU oracle
finishes in seconds)."""
from radchan_amd import fields as F
from radchan_amd import params as P


def grey_case(n=16, photmax=64, nVar=2, seed=3):
    p = P.base_params(n, n, n, photmax=photmax, nVar=nVar, grey=True, fact_mesh=0.0, seed=seed)
    p.adiabw = p.adiabe = 0
    p.Tw[0] = p.Tw[1] = 500.0
    T = F.channel_field(p, 500.0, 1000.0, noise=0.02, seed=seed)
    return p, None, T


def nb_case(tables, n=16, photmax=64, nVar=2, seed=5, srt=1):
    p = P.base_params(n, n, n, photmax=photmax, nVar=nVar, fact_mesh=1.5, seed=seed, srt=srt)
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=seed)
    return p, tables, T


def mg_case(tables, photmax=48, nVar=2, seed=11):
    """three grid levels, tight step caps, wall fluxes, specular z walls,
    adiabatic black x walls"""
    p = P.base_params(24, 16, 24, photmax=photmax, nVar=nVar, fact_mesh=1.5, seed=seed,
                      levels=[(24, 16, 24), (12, 8, 12), (6, 4, 6)], maxs=[4, 8, 64])
    p.calcwalls = 1
    p.bd[4] = p.bd[5] = 3
    T = F.channel_field(p, 300.0, 1500.0, noise=0.05, seed=seed)
    return p, tables, T


def bc_case(tables, photmax=32, nVar=2, seed=13):
    """black (fixed 380 K) south / specular north y walls, bd=4 bottom and
    bd=5 top z walls, fixed-T black x walls, two levels"""
    p = P.base_params(16, 12, 20, photmax=photmax, nVar=nVar, fact_mesh=1.2, seed=seed,
                      levels=[(16, 12, 20), (8, 6, 10)], maxs=[6, 400])
    p.adiabw = p.adiabe = 0
    p.Tw[0], p.Tw[1] = 350.0, 420.0
    p.bd[2], p.bd[3] = 2, 3
    p.Tw[2] = 380.0
    p.bd[4], p.bd[5] = 4, 5
    T = F.channel_field(p, 350.0, 1400.0, noise=0.05, seed=seed)
    return p, tables, T
