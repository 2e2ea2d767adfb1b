"""Pins of the CPU oracle (no GPU): published known-answer vectors, closed
forms, brute force on small inputs, and the reference's own code where it
compiles (oracle/_ref)."""
import os
import subprocess

import math

import numpy as np
import pytest
from scipy.special import expn

from oracle import oracle as O
from radchan_amd import fields as F
from radchan_amd import params as P
from radchan_amd import lib as L
from radchan_amd.tables import GQ16, WQ16, Tables, synthetic_nb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
SIGMA = 5.670373e-08


# ---- Philox4x32-10 ---------------------------------------------------------

def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32 with 10 rounds
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_range():
    assert O.u01(0) == 2.0 ** -24
    assert O.u01(0xFFFFFFFF) == 1.0
    assert O.u01(0x80000000) == (0x800000 + 1) * 2.0 ** -24


# ---- project math ----------------------------------------------------------

def test_sincos2pi_accuracy_and_quadrants():
    us = np.concatenate([np.arange(1, 4097) / 4096.0, np.random.default_rng(0).random(2000)]).astype(np.float32)
    us = us[us > 0]
    err = 0.0
    for u in us:
        s, c = O.sincos2pi(float(u))
        ref = 2.0 * np.pi * float(u)
        err = max(err, abs(s - np.sin(ref)), abs(c - np.cos(ref)))
    assert err < 3e-7
    assert O.sincos2pi(0.25) == (1.0, 0.0) or O.sincos2pi(0.25) == (1.0, -0.0)
    assert O.sincos2pi(0.5)[1] == -1.0 and O.sincos2pi(0.75)[0] == -1.0 and O.sincos2pi(1.0) == (0.0, 1.0)


def test_expm1_exp_accuracy():
    xs = np.concatenate([-np.logspace(-12, np.log10(86.9), 3000), np.linspace(-1.0, 0.0, 1001)]).astype(np.float32)
    worst_m, worst_e = 0.0, 0.0
    for x in xs:
        em1, e = O.expm1_exp(float(x))
        rm, re = np.expm1(np.float64(x)), np.exp(np.float64(x))
        worst_m = max(worst_m, abs(em1 - rm) / abs(rm) if rm != 0 else abs(em1))
        worst_e = max(worst_e, abs(e - re) / re)
    assert worst_m < 4e-7 and worst_e < 4e-7
    assert O.expm1_exp(0.0) == (0.0, 1.0)
    assert O.expm1_exp(-100.0) == (-1.0, 0.0)


def test_fp64_sincos2pi_accuracy_and_quadrants():
    """The full-FP64 path's sin/cos of 2 pi u, against long-double libm (the
    quadrant split is exact, so the error stays within 2 ulp of 1)."""
    rng = np.random.default_rng(3)
    us = np.concatenate([np.arange(0, 4097) / 4096.0, rng.integers(0, 2 ** 24, 4000) / 2.0 ** 24]).astype(np.float32)
    two_pi = np.longdouble(2) * np.arccos(np.longdouble(-1))
    err = 0.0
    for u in us:
        s, c = O.sincos2pi_d(float(u))
        ref = two_pi * np.longdouble(u)
        err = max(err, float(abs(np.longdouble(s) - np.sin(ref))), float(abs(np.longdouble(c) - np.cos(ref))))
    assert err < 2.5e-16
    assert O.sincos2pi_d(0.0) == (0.0, 1.0) and O.sincos2pi_d(1.0) == (0.0, 1.0)
    assert O.sincos2pi_d(0.25)[0] == 1.0 and O.sincos2pi_d(0.25)[1] == 0.0
    assert O.sincos2pi_d(0.5) == (0.0, -1.0) and O.sincos2pi_d(0.75)[0] == -1.0


def test_fp64_expm1_exp_accuracy():
    """The full-FP64 path's expm1/exp on x <= 0, against libm: within one
    ulp over [-745, 0] and down to 1e-20, exact special values, subnormal
    results at the bottom of the range."""
    rng = np.random.default_rng(4)
    xs = np.concatenate([-rng.uniform(0.0, 745.0, 20000), -10.0 ** rng.uniform(-20.0, 0.0, 20000),
                         [-0.34657359027997264, -0.3465735902799727, -1.0, -708.3, -709.0]])
    wm, we = 0.0, 0.0
    for x in xs:
        em1, e = O.expm1_exp_d(float(x))
        rm, re = math.expm1(x), math.exp(x)
        wm = max(wm, abs(em1 - rm) / abs(rm) / 2.0 ** -52)
        if re > 2.3e-308:
            we = max(we, abs(e - re) / re / 2.0 ** -52)
    assert wm <= 1.0 and we <= 1.0
    assert O.expm1_exp_d(0.0) == (0.0, 1.0) and math.copysign(1.0, O.expm1_exp_d(-0.0)[0]) == -1.0
    assert O.expm1_exp_d(-746.0) == (-1.0, 0.0)
    em1, e = O.expm1_exp_d(-740.0)
    assert em1 == -1.0 and abs(e - math.exp(-740.0)) <= 2 * 5e-324


def test_planck_closed_form_and_integral():
    # I_b(T, nu) = (1/pi) C1 (100 nu)^3 / (exp(C2 100 nu / T) - 1)
    c1, c2 = 3.741771790075259e-16, 0.014387741858429
    for T in (300.0, 1000.0, 2000.0):
        for nu in (10.0, 500.0, 3000.0, 9300.0):
            w = 100.0 * nu
            ref = c1 * w ** 3 / np.pi / np.expm1(c2 * w / T)
            assert abs(O.planck(T, nu) - ref) <= 1e-7 * ref
    # pi * integral over nu of I_b = sigma T^4 (the Stefan-Boltzmann law)
    nu = np.linspace(0.5, 40000.0, 80000)
    for T in (500.0, 1500.0):
        vals = np.array([O.planck(T, float(n)) for n in nu[::40]], dtype=np.float64)
        tot = np.pi * np.trapezoid(vals, nu[::40] * 100.0)
        assert abs(tot / (SIGMA * T ** 4) - 1.0) < 2e-3


def test_upper_bound_brute_force():
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 7, 16, 360, 1000):
        cdf = np.cumsum(rng.random(n) * (rng.random(n) > 0.2)).astype(np.float32)
        if cdf[-1] == 0:
            cdf[-1] = 1.0
        cdf = (cdf / cdf[-1]).astype(np.float32)
        Rs = np.concatenate([rng.random(300), cdf, [2.0 ** -24, 1.0]]).astype(np.float32)
        for R in Rs:
            want = int(np.searchsorted(cdf, R, side="right"))
            want = min(want, n - 1)
            if R == 1.0:
                want = n - 1
            assert O.upper_bound(cdf, float(R)) == want


def test_sort_idx_stable_permutation():
    rng = np.random.default_rng(2)
    kavg = rng.integers(0, 50, 999).astype(np.float32)      # many ties
    idx = O.sort_idx(kavg)
    assert np.array_equal(idx, np.argsort(kavg, kind="stable"))
    assert np.all(np.diff(kavg[idx]) >= 0)
    assert np.array_equal(np.sort(idx), np.arange(999))
    # the library's sort gives the same permutation
    lib_idx = np.zeros(999, np.int32)
    L.check(L.load().rc_sort_idx(L.fptr(kavg), 999, L.iptr(lib_idx)), "rc_sort_idx")
    assert np.array_equal(lib_idx, idx)


def test_direction_moments():
    n = 200
    u = (np.arange(n) + 0.5) / n
    sx, sy, sz, norm = [], [], [], []
    for a in u:
        for b in u:
            s = O.iso_dir(float(np.float32(a)), float(np.float32(b)))
            sx.append(s[0]); sy.append(s[1]); sz.append(s[2])
            norm.append(s[0] ** 2 + s[1] ** 2 + s[2] ** 2)
            assert s[0] == np.float32(1.0) - np.float32(2.0) * np.float32(b)
    sx, sy, sz = map(np.array, (sx, sy, sz))
    assert np.abs(np.array(norm) - 1).max() < 1e-6
    assert abs(sx.mean()) < 1e-6 and abs(sy.mean()) < 1e-6 and abs(sz.mean()) < 1e-6
    assert abs((sx ** 2).mean() - 1 / 3) < 1e-4 and abs((sy ** 2).mean() - 1 / 3) < 1e-4
    # Lambertian: pdf of |s_x| is 2|s_x|, E|s_x| = 2/3, and |s_x| = sqrt(1-u) (west) or sqrt(u) (east)
    for bound in (0, 1):
        ax = []
        for a in u[::4]:
            for b in u:
                s = O.lamb_dir(float(np.float32(a)), float(np.float32(b)), bound)
                ax.append(abs(s[0]))
                assert (s[0] > 0) == (bound == 0)
                assert abs(s[0] ** 2 + s[1] ** 2 + s[2] ** 2 - 1) < 1e-6
        assert abs(np.mean(ax) - 2 / 3) < 1e-3


# ---- grid --------------------------------------------------------------------

def _ref_params(im, jm, km, im0):
    """level 1 of a two-level set whose level 0 has im0 x-cells"""
    p = P.reference_params()
    p.imax, p.jmax, p.kmax = im0, jm, km
    p.grid_num = 2
    p.maxi[0], p.maxj[0], p.maxk[0] = im0, jm, km
    p.maxi[1], p.maxj[1], p.maxk[1] = im, jm, km
    return p


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "ref_grid")), reason="reference grid driver not built")
@pytest.mark.parametrize("dims", [(406, 576, 2304, 406), (14, 48, 72, 406), (2, 3, 9, 406), (64, 64, 64, 64),
                                  (32, 16, 8, 64), (7, 5, 3, 7), (128, 1, 1, 128)])
def test_grid_matches_reference_build(dims):
    im, jm, km, im0 = dims
    out = subprocess.run([os.path.join(REF_DIR, "ref_grid"), str(im), str(jm), str(km), str(im0)],
                         capture_output=True, text=True, check=True).stdout.splitlines()
    ref = {}
    for line in out[:6]:
        parts = line.split()
        ref[parts[0]] = np.array([int(h, 16) for h in parts[2:]], np.uint32).view(np.float32)
    p = _ref_params(im, jm, km, im0)
    # the reference's compile-time domain: Lx=2, Ly=2pi, Lz=12pi, fact_mesh=1.6
    p.Lx, p.Ly, p.Lz, p.fact_mesh = 2.0, 6.283185307179586, 37.69911184307751, 1.6
    g = O.grid(p, 1)
    a = [np.zeros(n, np.float32) for n in (im + 2, im + 1, jm + 2, jm + 1, km + 2, km + 1)]
    L.check(L.load().rc_grid(p, 1, *[L.fptr(x) for x in a]), "rc_grid")
    for n, arr in zip(("x", "xu", "y", "yu", "z", "zu"), a):
        assert np.array_equal(g[n].view(np.uint32), ref[n].view(np.uint32)), f"oracle {n}"
        assert np.array_equal(arr.view(np.uint32), ref[n].view(np.uint32)), f"library {n}"


def test_grid_closed_form_properties():
    p = P.base_params(64, 8, 8, fact_mesh=1.5, levels=[(64, 8, 8), (16, 4, 2)], maxs=[5, 100])
    g0, g1 = O.grid(p, 0), O.grid(p, 1)
    xu = g0["xu"].astype(np.float64)
    assert xu[0] == 0.0 and abs(xu[-1] - 2.0) < 1e-6
    assert np.allclose(xu, 2.0 - xu[::-1], atol=1e-6)               # mirror symmetry
    dx = np.diff(xu)
    assert dx[0] < dx[len(dx) // 2]                                 # refined at the walls
    assert set(g1["xu"].tolist()) <= set(g0["xu"].tolist())         # coarse faces are fine faces
    # cell-size law dx(r) = 0.5 - f (r - 1/2)^2, normalised to Lx/2
    r = np.arange(1, 33) / 64.0
    law = 0.5 - 1.5 * (r - 0.5) ** 2
    assert np.allclose(dx[:32], law / law.sum(), rtol=1e-5)
    pu = P.base_params(32, 8, 8, fact_mesh=0.0)
    assert np.allclose(np.diff(O.grid(pu, 0)["xu"]), 2.0 / 32, rtol=1e-5)


# ---- tables ------------------------------------------------------------------

@pytest.mark.skipif(not os.path.exists(os.path.join(REF_DIR, "ref_readT")), reason="reference reader not built")
def test_table_reader_matches_reference_build(tmp_path):
    """readT of the reference (read_tables.cpp:21-95) built from its source,
    at its compile-time sizes nB=999, nT=135, nQ=16."""
    nB, nT, nQ = 999, 135, 16
    tab = synthetic_nb(nB=nB, nT=nT, nQ=nQ, T_lo=300.0, T_hi=2980.0, seed=3)
    props = tmp_path / "properties"
    props.mkdir()
    tab.write(str(props), 0.5)
    Tmax = 1234.5
    subprocess.run([os.path.join(REF_DIR, "ref_readT"), str(Tmax), str(tmp_path / "ref.bin")], cwd=tmp_path,
                   check=True)
    raw = np.fromfile(tmp_path / "ref.bin", dtype=np.float32)
    o = 0

    def take(n):
        nonlocal o
        a = raw[o:o + n]
        o += n
        return a

    Tnb, kP, kmax = take(nT), take(nT), take(1)
    per_band = take(nB * (4 + nT * nQ)).reshape(nB, 4 + nT * nQ)
    prob, probg = take(nT * nB), take(nT * nB * nQ)
    assert o == raw.size
    mine = O.read_tables(str(props), nB, nT, nQ, 0.5)
    lib = Tables.read(str(props), nB, nT, nQ, 0.5)
    for t in (mine, lib):
        assert np.array_equal(t.Tnb, Tnb) and np.array_equal(t.kP, kP)
        assert np.array_equal(t.wvl, per_band[:, 0]) and np.array_equal(t.wvr, per_band[:, 1])
        assert np.array_equal(t.wvc, per_band[:, 2]) and np.array_equal(t.kavg, per_band[:, 3])
        assert np.array_equal(t.kq.reshape(nB, -1), per_band[:, 4:])
        assert np.array_equal(t.prob.reshape(-1), prob) and np.array_equal(t.probg.reshape(-1), probg)
    # kappamax: linear interpolation of kP at Tmax (read_tables.cpp:35-38)
    p = P.base_params(4, 4, 4, photmax=4)
    p.nB, p.nT, p.nQ = nB, nT, nQ
    T = F.uniform_field(p, Tmax, Tmax)
    info = O.Oracle(p, mine, T).info()
    assert info["Tmax"] == np.float32(Tmax)
    assert np.float32(info["kappamax"]) == kmax[0]


def test_table_write_read_roundtrip(tmp_path, nb_tables):
    nb_tables.write(str(tmp_path), 0.5)
    for t in (Tables.read(str(tmp_path), 360, 20, 16, 0.5), O.read_tables(str(tmp_path), 360, 20, 16, 0.5)):
        for name in ("Tnb", "kP", "wvl", "wvr", "wvc", "kavg", "kq", "prob", "probg"):
            assert np.array_equal(getattr(t, name), getattr(nb_tables, name)), name


def test_grey_tables_library_equals_oracle():
    for true_grey in (False, True):
        p = P.base_params(8, 8, 8, grey=True, grey_true=true_grey)
        a = O.grey_tables(p, 1234.0)
        b = Tables(p.nB, p.nT, p.nQ)
        L.check(L.load().rc_grey_tables(p, 1234.0, b.cview()), "rc_grey_tables")
        for name in ("Tnb", "kP", "wvl", "wvr", "wvc", "kq", "prob"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), name
        assert np.all(np.diff(a.prob, axis=1) >= 0) and np.all(a.prob[:, -1] == 1.0)
        assert a.Tnb[0] == np.float32(490.0 * 0.9)


def test_synthetic_tables_consistent(nb_tables):
    t = nb_tables
    assert np.all(np.diff(t.prob, axis=1) >= 0) and np.allclose(t.prob[:, -1], 1.0)
    assert np.all(np.diff(t.probg, axis=2) >= -1e-7) and np.allclose(t.probg[:, :, -1], 1.0)
    assert np.all(np.diff(t.kq, axis=2) > 0)                          # k monotone in g
    k = t.kq.astype(np.float64)
    assert k.min() > 1e-5 and k.max() < 1e4


# ---- band selection and the srt enumeration ----------------------------------

def test_band_sequence_is_counting_sort_of_draws(nb_tables):
    p = P.base_params(8, 8, 8, photmax=257, nVar=2, fact_mesh=1.5, seed=77)
    T = F.channel_field(p, 300.0, 1500.0, noise=0.02, seed=1)
    o = O.Oracle(p, nb_tables, T)
    ts = o.info()["tstar"]
    cdf_b = nb_tables.prob[ts]
    idx = O.sort_idx(nb_tables.kavg)
    rank_of = np.empty_like(idx)
    rank_of[idx] = np.arange(len(idx))
    for (i, j, k, v) in [(1, 1, 1, 0), (8, 3, 5, 1), (4, 8, 8, 0)]:
        cid = (i - 1) + p.imax * ((j - 1) + p.jmax * (k - 1))
        keys = []
        for h in range(p.photmax):
            out = O.philox([h, v, cid, 0], [p.seed & 0xFFFFFFFF, p.seed >> 32])
            u2, u3 = O.u01(out[2]), O.u01(out[3])
            nb = min(int(np.searchsorted(cdf_b, np.float32(u2), side="right")), p.nB - 1)
            g = min(int(np.searchsorted(nb_tables.probg[ts, nb], np.float32(u3), side="right")), p.nQ - 1)
            keys.append(rank_of[nb] * p.nQ + g)
        keys = np.sort(np.array(keys))
        nbs, gs = o.band_sequence(i, j, k, v)
        assert np.array_equal(nbs * p.nQ + gs, keys)
        assert np.all(np.diff(nb_tables.kavg[idx][nbs]) >= 0)     # ascending kavg


# ---- estimator closed forms --------------------------------------------------

def test_isothermal_enclosure_is_exactly_zero(nb_tables):
    """T identical everywhere: every reciprocity term (B - Bi)/Bi is exactly
    zero when all emission goes through the same table lookup (adiabatic x
    walls at the wall-midpoint T, bd=4 z walls at the adjacent cell T,
    specular y walls), two grid levels."""
    p = P.base_params(12, 8, 10, photmax=40, nVar=2, fact_mesh=1.5, seed=9,
                      levels=[(12, 8, 10), (6, 4, 5)], maxs=[3, 1000])
    p.bd[2] = p.bd[3] = 3
    p.bd[4] = p.bd[5] = 4
    Tg = 1111.0
    T = F.uniform_field(p, Tg, Tg)
    o = O.Oracle(p, nb_tables, T)
    de32, de32d, de64, steps = o.solve(O.all_cells(p)[::7])
    assert np.all(de32 == 0) and np.all(de64 == 0) and steps.sum() > 0
    p.calcwalls = 1
    o = O.Oracle(p, nb_tables, T)
    de32, de32d, de64, steps = o.solve(O.all_wall_cells(p)[::5], surface=True)
    assert np.all(de32 == 0) and np.all(de64 == 0) and steps.sum() > 0


def _path_length_check(p, tab, T, kappa_of):
    """Uniform gas: a ray that reaches a black x wall carries
    tra = exp(-kappa d), d = distance from the emission point to that wall."""
    o = O.Oracle(p, tab, T)
    rng = np.random.default_rng(5)
    n = 400
    rays = np.stack([rng.integers(1, p.imax + 1, n), rng.integers(1, p.jmax + 1, n), rng.integers(1, p.kmax + 1, n),
                     np.zeros(n, int), rng.integers(0, p.photmax, n)], axis=1)
    r = o.trace(rays)
    x = O.grid(p, 0)["x"].astype(np.float64)
    walls = r["status"] == 1
    assert walls.sum() > n // 4
    for q in np.nonzero(walls)[0]:
        i = rays[q, 0]
        sx = float(r["sx"][q])
        d = ((p.Lx if sx > 0 else 0.0) - x[i]) / sx
        kap = kappa_of(r["nb"][q], r["g"][q])
        # the ray's wall term: -W tra (Bw - Bi)/Bi; recover tra from it
        assert d > 0
        yield q, d, kap, r[q]


def test_grey_path_length_periodic_specular_multigrid():
    """Exercises periodic and (corrected) specular walls and three grid
    levels: with uniform T only the wall term survives and its
    transmissivity must be exp(-kappa * geometric distance)."""
    for bdz, levels, maxs in [(1, [(16, 12, 10)], [100000]), (3, [(16, 12, 10), (8, 6, 5), (4, 3, 5)], [3, 5, 5000])]:
        p = P.base_params(16, 12, 10, photmax=64, grey=True, grey_true=True, fact_mesh=1.2, seed=21,
                          levels=levels, maxs=maxs)
        p.Ly, p.Lz = 1.5, 0.8
        p.adiabw = p.adiabe = 0
        p.Tw[0] = p.Tw[1] = 500.0
        p.bd[2] = p.bd[3] = 3
        p.bd[4] = p.bd[5] = bdz
        p.toll = p.tollS = 1e-12
        Tg = 1000.0
        T = F.uniform_field(p, Tg, 1000.0)
        o = O.Oracle(p, None, T)
        info = o.info()
        Bi = SIGMA * Tg ** 4 / np.pi
        Bw = SIGMA * 500.0 ** 4 / np.pi
        W = info["Ibmax64"]
        for q, d, kap, rec in _path_length_check(p, None, T, lambda nb, g: 1.0):
            tra = -rec["de64"] / (W * (Bw - Bi) / Bi)
            # FP32 positions: the summed path carries ~1 ulp per step
            tol = np.exp(-d) * (2e-5 + 1e-6 * rec["steps"] * max(d, 1.0))
            assert abs(tra - np.exp(-d)) <= tol, (d, tra, rec["steps"])


def test_grey_slab_analytic():
    """div q = 2 kappa sigma (Tg^4 - Tw^4) [E2(kappa x) + E2(kappa (L - x))]
    at cell centres of a uniform grey slab between black walls."""
    p = P.config("slab32")
    p.photmax, p.nVar, p.seed = 4000, 16, 3
    T = F.uniform_field(p, 1000.0, 500.0)
    o = O.Oracle(p, None, T)
    cells = np.array([[i, 7, 9] for i in (1, 2, 5, 9, 16, 17, 24, 31, 32)], np.int32)
    de32, de32d, de64, steps = o.solve(cells)
    x = O.grid(p, 0)["x"][cells[:, 0]].astype(np.float64)
    ana = 2 * SIGMA * (1000.0 ** 4 - 500.0 ** 4) * (expn(2, x) + expn(2, 2.0 - x))
    m = de64.mean(axis=1)
    se = de64.std(axis=1, ddof=1) / np.sqrt(p.nVar)
    z = (m - ana) / se
    assert np.all(np.abs(z) < 4.5), z
    assert abs(np.mean(m / ana) - 1) < 0.01
    # FP32 geometry and energy agree with the FP64 energy path
    assert np.abs(de32 - de64).max() <= 1e-5 * np.abs(de64).max()


def test_narrow_band_slab_analytic(nb_tables):
    """Uniform gas at a tabulated temperature between cold black walls: the
    importance-sampled estimator must reproduce
    sum_bg 2 pi k wq Ib dnu [E2(k x) + E2(k (L - x))] (per band and g)."""
    t = nb_tables
    Tg = float(t.Tnb[12])
    p = P.base_params(16, 4, 4, photmax=3000, nVar=12, fact_mesh=1.5, seed=31)
    p.adiabw = p.adiabe = 0
    p.Tw[0] = p.Tw[1] = 0.0
    T = F.uniform_field(p, Tg, Tg)
    o = O.Oracle(p, t, T)
    assert o.info()["tstar"] == 12
    cells = np.array([[i, 2, 3] for i in (1, 4, 8, 13, 16)], np.int32)
    de32, de32d, de64, steps = o.solve(cells)
    x = O.grid(p, 0)["x"][cells[:, 0]].astype(np.float64)
    k = t.kq[:, 12, :].astype(np.float64)                       # [nB][nQ]
    ib = np.array([O.planck(Tg, float(w)) for w in t.wvc], np.float64)
    dnu = (t.wvr - t.wvl).astype(np.float64) * 100.0
    wq = WQ16.astype(np.float64)
    ana = []
    for xc in x:
        e = expn(2, k * xc) + expn(2, k * (2.0 - xc))
        ana.append(np.sum(2 * np.pi * k * wq[None, :] * (ib * dnu)[:, None] * e))
    ana = np.array(ana)
    m = de64.mean(axis=1)
    se = de64.std(axis=1, ddof=1) / np.sqrt(p.nVar)
    assert np.all(np.abs((m - ana) / se) < 4.5), ((m - ana) / se, m / ana)


def test_wall_flux_grey_slab():
    """Net flux of an adiabatic black wall at Tw facing a uniform grey gas at
    Tg and an identical opposite wall: q = sigma (Tw^4 - Tg^4)(1 - 2 E3(kappa L))."""
    p = P.base_params(16, 8, 8, photmax=4000, nVar=12, grey=True, grey_true=True, fact_mesh=0.8, seed=41)
    p.calcwalls = 1
    Tg, Tw = 1000.0, 600.0
    T = F.uniform_field(p, Tg, Tw)
    o = O.Oracle(p, None, T)
    cells = np.array([[0, 3, 5], [1, 3, 5], [0, 8, 1]], np.int32)
    de32, de32d, de64, steps = o.solve(cells, surface=True)
    ana = SIGMA * (Tw ** 4 - Tg ** 4) * (1 - 2 * expn(3, 2.0))
    m = de64.mean(axis=1)
    se = de64.std(axis=1, ddof=1) / np.sqrt(p.nVar)
    assert np.all(np.abs((m - ana) / se) < 4.5), ((m - ana) / se, m / ana)


def test_step_cap_drops_energy():
    """A tiny maxs on the last level ends rays with energy left (the
    reference drops it, mc_gpu_noint.cu:252)."""
    p = P.base_params(16, 8, 8, photmax=64, grey=True, grey_true=True, fact_mesh=0.0, seed=1,
                      levels=[(16, 8, 8)], maxs=[3])
    p.adiabw = p.adiabe = 0
    p.Tw[0] = p.Tw[1] = 500.0
    T = F.channel_field(p, 500.0, 1000.0)
    o = O.Oracle(p, None, T)
    rays = np.array([[8, 4, 4, 0, h] for h in range(64)], np.int32)
    r = o.trace(rays)
    assert (r["status"] == 3).any() and np.all(r["steps"] <= 3)


def test_band_group_split_sums_to_the_cell_tallies(nb_tables):
    """The band-group split of the precision study: per (cell, run), the
    groups' FP32 tallies add up to the cell's FP64 sum of the same per-ray
    tallies, and the full-FP64 groups to the full-FP64 solve (to FP64
    reordering); a group that holds no band receives nothing."""
    p = P.base_params(10, 8, 12, photmax=48, nVar=2, fact_mesh=1.5, seed=9)
    T = F.channel_field(p, 300.0, 1500.0, noise=0.03, seed=9)
    o = O.Oracle(p, nb_tables, T)
    cells = O.all_cells(p)[::7]
    gor = np.minimum(4 * np.arange(p.nB) // p.nB, 3)
    gor[gor == 2] = 3                     # group 2 left empty
    g0 = o.solve_groups(cells, gor, precision=0)
    _, de32d, _, _ = o.solve(cells)
    assert np.abs(g0.sum(axis=2) - de32d).max() <= 1e-12 * np.abs(de32d).max()
    assert np.all(g0[:, :, 2] == 0.0)
    g1 = o.solve_groups(cells, gor, precision=1)
    _, _, de64, _ = o.solve(cells, precision=1)
    assert np.abs(g1.sum(axis=2) - de64).max() <= 1e-12 * np.abs(de64).max()
