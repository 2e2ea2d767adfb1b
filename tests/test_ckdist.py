"""Correlated-k generator (correlatedKT, nbcorrk/correlatedK.cpp:144-225, and
the k_avg of calcProbT, calc_prob.cpp:52-56).

CPU: the oracle's restatement equal bit for bit to the reference's own
correlatedKT and calcProbT built from nbcorrk/ (oracle/_ref/ref_ckdist),
and pinned by what the mathematics fixes -- the
inverse-CDF property of kq on a linear spectrum (closed form) and on random
spectra (brute-force CDF of the point pairs), k_avg as the band mean of the
pair averages, exact scaling by 2^10 (k^0.1 scales by 2 exactly), the flat
and one-point spectra.  GPU: rc_ckdist against the oracle on the same
spectra, ragged sizes and the edge cases.  The reference's own tests hold no
vectors for this path (nbcorrk/tests holds plots only)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from radchan_amd.spectra import synthetic_lbl
from radchan_amd.tables import GQ16_F64

N_PWRK, PWR = 20000, 0.1


def _one(wv, kv, gq=GQ16_F64, wl=None, wr=None, **kw):
    wv, kv = np.asarray(wv, float), np.asarray(kv, float)
    wl = wv[0] if wl is None else wl
    wr = wv[-1] if wr is None else wr
    kq, ka, lo, hi = O.ckdist([0, len(wv)], wv, kv, [wl], [wr], gq, **kw)
    return kq[0], ka[0], lo[0], hi[0]


def _rel_bin(kmin, kmax, k):
    """Relative width of the k bin holding k: (p + step)^10 / p^10 - 1."""
    step = (kmax ** PWR - kmin ** PWR) / (N_PWRK - 1)
    return (1.0 + step / k ** PWR) ** (1.0 / PWR) - 1.0


def test_linear_spectrum_inverse_cdf():
    """k rising linearly across the band has a uniform k-distribution:
    k(g) = k0 + g (k1 - k0), up to a bin width (the half-bin shift) and the
    point spacing."""
    n, k0, k1 = 20001, 1.0, 100.0
    wv = 1000.0 + 0.01 * np.arange(n)
    kv = np.linspace(k0, k1, n)
    kq, ka, lo, hi = _one(wv, kv)
    assert lo == k0 and hi == k1
    exact = k0 + GQ16_F64 * (k1 - k0)
    tol = 2.0 * _rel_bin(k0, k1, exact) * exact + 2.0 * (k1 - k0) / n
    assert np.all(np.abs(kq - exact) <= tol)
    assert abs(ka - 0.5 * (k0 + k1)) <= 2e-4 * 0.5 * (k0 + k1)
    # a planted mistake (no half-bin shift, i.e. kk = K(h)) would sit half a
    # bin low: the shift is visible above the point-spacing error
    assert np.all(kq > exact - 2.0 * (k1 - k0) / n)


def test_random_spectra_inverse_cdf_and_band_mean():
    """On seeded line spectra: the weighted fraction of pair averages below
    kq[r] brackets gq[r] within the weight of the bins next to kq, and k_avg
    is the band mean of the pair averages to a bin width."""
    d = synthetic_lbl(temps=[300.0, 1500.0])
    sel = np.arange(0, len(d["off"]) - 1, 37)
    for s in sel:
        b, e = d["off"][s], d["off"][s + 1]
        wv, kv = d["wvn"][b:e], d["kvn"][b:e]
        kq, ka, lo, hi = _one(wv, kv, wl=d["wvl"][s], wr=d["wvr"][s])
        ki = 0.5 * (kv[1:] + kv[:-1])
        w = np.diff(wv) / (d["wvr"][s] - d["wvl"][s])
        rb = _rel_bin(lo, hi, np.maximum(kq, lo))
        for r, g in enumerate(GQ16_F64):
            below = w[ki < kq[r] * (1 - 2 * rb[r])].sum()
            upto = w[ki <= kq[r] * (1 + 2 * rb[r])].sum()
            assert below <= g + 1e-12 and upto >= g - 1e-12, (s, r, below, g, upto)
        assert np.all(np.diff(kq) >= 0.0) and lo <= kq[0] and kq[-1] <= hi
        mean = (ki * w).sum()
        assert abs(ka - mean) <= _rel_bin(lo, hi, lo) * mean


def test_single_pair_lands_in_its_bin_centre():
    """Two points give one pair, a delta at ki = (k0 + k1)/2.  The bins are
    [K(b-1), K(b)) with K(h) = (pmin + h step)^10 and the generator shifts
    every k value half a bin up, so the k-distribution rises from the centre
    of ki's bin (g -> 0) to the centre of the next one (g -> 1): kq(0+) is
    within half a bin of ki and kq(1-) - kq(0+) is one bin.  A missing
    shift, a bin index off by one or a missing pair average all fail this."""
    rng = np.random.default_rng(11)
    gq = np.array([1e-9, 1.0 - 1e-9])
    for _ in range(300):
        k0, k1 = np.sort(10.0 ** rng.uniform(-3.0, 3.0, 2))
        kq, ka, lo, hi = _one([10.0, 10.01], [k0, k1], gq=gq)
        ki = 0.5 * (k0 + k1)
        step = (k1 ** PWR - k0 ** PWR) / (N_PWRK - 1)
        w = step / PWR * ki ** (1.0 - PWR)                 # dK/dh at ki: one bin width
        assert abs(kq[0] - ki) <= 0.5 * w * (1 + 1e-3), (k0, k1)
        assert abs((kq[1] - kq[0]) - w) <= 1e-3 * w
        assert abs(ka - kq[1]) <= 1e-6 * w                 # k_avg = the shifted value of ki's bin


def test_scaling_by_two_to_the_tenth_is_exact():
    """k -> 1024 k scales k^0.1 by exactly 2, so every bin stays the same and
    kq, k_avg, kmin, kmax scale by 1024 (to the libm's pow rounding)."""
    d = synthetic_lbl(temps=[800.0])
    for s in (0, 100, 400):
        b, e = d["off"][s], d["off"][s + 1]
        a = _one(d["wvn"][b:e], d["kvn"][b:e])
        c = _one(d["wvn"][b:e], 1024.0 * d["kvn"][b:e])
        np.testing.assert_allclose(c[0], 1024.0 * a[0], rtol=1e-13)
        assert abs(c[1] - 1024.0 * a[1]) <= 1e-13 * c[1]
        assert c[2] == 1024.0 * a[2] and c[3] == 1024.0 * a[3]


def test_flat_and_one_point_spectra():
    """A flat spectrum (kmin == kmax; the reference divides 0 by 0 there) and
    a single point both give kq = k; k_avg is k for the flat band and 0 for
    the single point (no pair)."""
    kq, ka, lo, hi = _one([5.0, 5.01, 5.02, 5.03], [2.5] * 4)
    np.testing.assert_allclose(kq, 2.5, rtol=1e-14)
    assert abs(ka - 2.5) <= 1e-14 * 2.5
    kq, ka, lo, hi = _one([5.0], [7.0], wl=5.0, wr=5.01)
    np.testing.assert_allclose(kq, 7.0, rtol=1e-14)
    assert ka == 0.0 and lo == hi == 7.0


def _gpu_vs_oracle(off, wvn, kvn, wvl, wvr, gq=GQ16_F64, **kw):
    from radchan_amd.spectra import ckdist
    g = ckdist(off, wvn, kvn, wvl, wvr, gq, **kw)
    kq, ka, lo, hi = O.ckdist(off, wvn, kvn, wvl, wvr, gq, **kw)
    assert np.array_equal(g["kmin"], lo) and np.array_equal(g["kmax"], hi)
    # FP64 pow of the device and of the host libm may differ in the last bit
    np.testing.assert_allclose(g["kq"], kq, rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["kavg"], ka, rtol=1e-12, atol=1e-300)
    return g, kq


@pytest.mark.gpu
def test_gpu_ckdist_matches_oracle_synthetic():
    d = synthetic_lbl(temps=[300.0, 1000.0, 2000.0, 2850.0])
    print(f"rc_ckdist: {len(d['off']) - 1} spectra of {d['off'][1]} points", flush=True)
    g, kq = _gpu_vs_oracle(d["off"], d["wvn"], d["kvn"], d["wvl"], d["wvr"])
    assert g["kernel_ms"] > 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("pwr,n_pwrk", [(0.3, 5000), (0.25, 65535)])
def test_gpu_ckdist_other_bin_laws(pwr, n_pwrk):
    """Bin laws other than the generator's k^0.1 in 20000 bins: 1/pwr not an
    integer (the kk values through pow) and an integer one, at the largest
    bin count."""
    d = synthetic_lbl(temps=[700.0], deltawv=25.0)
    _gpu_vs_oracle(d["off"], d["wvn"], d["kvn"], d["wvl"], d["wvr"], n_pwrk=n_pwrk, pwr=pwr)


@pytest.mark.gpu
@pytest.mark.parametrize("longest", [257, 513, 1025, 2049, 4097, 8192])
def test_gpu_ckdist_ragged_and_edge_cases(longest):
    """Spectra of 1, 2, 3, 257 and up to `longest` points in one launch (the
    longest picks the kernel instance: 256, 512, .., 8192 sort slots), a flat
    one, one with a plateau at kmax (pairs there fall in bin n_pwrk and
    leave gg and k_avg, as in the reference), 17 quadrature points."""
    rng = np.random.default_rng(longest)
    sizes = [1, 2, 3, min(257, longest), longest // 2 + 1, longest, 50, 64]
    wv, kv, off, wl, wr = [], [], [0], [], []
    for i, n in enumerate(sizes):
        nu = 2000.0 + 0.01 * np.arange(n)
        if i == 6:
            k = np.full(n, 3.25)
        elif i == 7:
            k = np.concatenate([rng.uniform(0.1, 1.0, 40), np.full(24, 2.0)])
        else:
            k = 10.0 ** rng.uniform(-2.0, 3.0, n)
        wv.append(nu)
        kv.append(k)
        off.append(off[-1] + n)
        wl.append(nu[0])
        wr.append(nu[-1] if n > 1 else nu[0] + 0.01)
    gq = np.sort(np.concatenate([GQ16_F64, [0.5]]))
    print(f"rc_ckdist: ragged sizes {sizes}", flush=True)
    g, kq = _gpu_vs_oracle(np.array(off), np.concatenate(wv), np.concatenate(kv), np.array(wl), np.array(wr), gq)
    np.testing.assert_allclose(g["kq"][6], 3.25, rtol=1e-14)


@pytest.mark.gpu
def test_gpu_ckdist_rejects_bad_input():
    from radchan_amd.spectra import ckdist
    nu = 100.0 + 0.01 * np.arange(8193)
    with pytest.raises(RuntimeError, match="8193 points"):
        ckdist([0, 8193], nu, np.ones(8193), [nu[0]], [nu[-1]], GQ16_F64)
    with pytest.raises(RuntimeError, match="nondecreasing"):
        ckdist([0, 10], nu[:10], np.ones(10), [nu[0]], [nu[9]], GQ16_F64[::-1].copy())


# ---- the table step (calcProbT) ---------------------------------------------

def _planck_exact(T, nu):
    """Blackbody intensity per cm^-1 (times 100: per m^-1), exact exponential."""
    C1, C2 = 3.741771790075259e-16, 0.014387741858429
    w = 100.0 * nu
    return C1 / np.pi * w ** 3 / np.expm1(C2 * w / T)


def test_ck_tables_grey_closed_form():
    """A band-independent, almost constant k = kappa (1 + 1e-7 sin): the
    correlated-k Planck mean is kappa times the fraction of sigma T^4 / pi
    that the bands' range holds (midpoint rule over 25 cm^-1 bands against
    the integral; the generator's single-precision exponential), the band
    CDF is the cumulated blackbody intensity of the bands, and every g CDF
    is the cumulated quadrature weights."""
    from scipy.integrate import quad
    from radchan_amd.tables import WQ16_F64
    kappa = 3.0
    sp = synthetic_lbl(wv0=100.0, wv1=12100.0, deltawv=25.0, temps=[500.0, 1000.0, 2000.0])
    x = np.arange(len(sp["kvn"]))
    sp["kvn"] = kappa * (1.0 + 1e-7 * np.sin(0.37 * x))
    o = O.ck_tables(sp, GQ16_F64, WQ16_F64)
    nB = sp["nB"]
    wl, wr = sp["wvl"][:nB], sp["wvr"][:nB]
    sigma = 5.670373e-08
    for i, T in enumerate(sp["temps"]):
        frac = sum(quad(lambda nu: _planck_exact(T, nu) * 100.0, a, b)[0] for a, b in zip(wl, wr)) * np.pi / (sigma * T ** 4)
        assert abs(o["kP"][i] - kappa * frac) <= 2e-4 * kappa, (T, o["kP"][i], kappa * frac)
        band = _planck_exact(T, 0.5 * (wl + wr)) * (wr - wl)
        np.testing.assert_allclose(o["cuprob"][i], np.cumsum(band) / band.sum(), atol=1e-6)
    cw = np.cumsum(WQ16_F64) / WQ16_F64.sum()
    np.testing.assert_allclose(o["cuprob_g"], np.broadcast_to(cw, o["cuprob_g"].shape), atol=1e-6)
    np.testing.assert_allclose(o["kq"], kappa, rtol=1e-6)


def test_ck_tables_library_matches_oracle():
    """rc_ck_tables (host code of the library, no GPU) on the oracle's
    distributions against the oracle's calcProbT: equal to FP32 rounding (the
    library sums k_avg per band, the reference kk ff per bin)."""
    from radchan_amd.spectra import tables_from_distributions
    from radchan_amd.tables import WQ16_F64
    sp = synthetic_lbl(wv0=150.0, wv1=2150.0, deltawv=25.0, temps=[400.0, 800.0, 1600.0])
    o = O.ck_tables(sp, GQ16_F64, WQ16_F64)
    nB = sp["nB"]
    ch = 0.5
    t = tables_from_distributions(sp["temps"], sp["wvl"][:nB], sp["wvr"][:nB], o["kq"], o["kavg"], WQ16_F64, ch)
    np.testing.assert_allclose(t.kP, o["kP"] * ch, rtol=3e-7)
    np.testing.assert_allclose(t.prob, o["cuprob"], atol=6e-8)
    np.testing.assert_allclose(t.probg, o["cuprob_g"], atol=6e-8)
    np.testing.assert_allclose(t.kq, o["kq"].transpose(1, 0, 2) * ch, rtol=3e-7)
    np.testing.assert_allclose(t.kavg, o["kavg"][-1], rtol=1e-7)
    assert np.array_equal(t.Tnb, sp["temps"].astype(np.float32))
    np.testing.assert_allclose(t.wvc, 0.5 * (sp["wvl"][:nB] + sp["wvr"][:nB]), rtol=1e-7)


@pytest.mark.gpu
def test_gpu_generator_end_to_end():
    """Spectra -> rc_ckdist -> rc_ck_tables -> a batched solve with those
    tables: the tables equal the oracle's calcProbT to FP32 rounding, and
    the solve matches the oracle's on the same tables (FP64 tallies to
    1e-12 of the FP32 per-ray sums)."""
    from radchan_amd import VARIANT_BATCHED, Solver
    from radchan_amd import fields as Fd
    from radchan_amd import params as Pm
    from radchan_amd.spectra import generate_tables
    from radchan_amd.tables import WQ16_F64
    sp = synthetic_lbl(wv0=150.0, wv1=2150.0, deltawv=25.0, temps=np.linspace(300.0, 2400.0, 8))
    p = Pm.base_params(12, 10, 8, photmax=48, nVar=2, fact_mesh=1.5, seed=3)
    p.nB, p.nT, p.nQ = sp["nB"], len(sp["temps"]), 16
    print(f"generator end to end: {p.nB} bands x {p.nT} T, field {p.imax}x{p.jmax}x{p.kmax}", flush=True)
    tab, g = generate_tables(sp, GQ16_F64, WQ16_F64, chanHeight=p.chanHeight)
    o = O.ck_tables(sp, GQ16_F64, WQ16_F64)
    np.testing.assert_allclose(tab.kP, o["kP"] * p.chanHeight, rtol=3e-7)
    np.testing.assert_allclose(tab.prob, o["cuprob"], atol=6e-8)
    np.testing.assert_allclose(tab.probg, o["cuprob_g"], atol=6e-8)
    np.testing.assert_allclose(tab.kq, o["kq"].transpose(1, 0, 2) * p.chanHeight, rtol=3e-7)
    T = Fd.channel_field(p, 500.0, 1500.0, noise=0.02, seed=3)
    s = Solver(p, tab)
    s.set_fields([T])
    s.solve(VARIANT_BATCHED)
    got = s.tallies()
    s.close()
    _, de32d, _, _ = O.Oracle(p, tab, T).solve(O.all_cells(p))
    assert np.abs(got - de32d).max() <= 1e-12 * np.abs(de32d).max()


REF_CK = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_ckdist")


@pytest.mark.skipif(not os.path.exists(REF_CK), reason="reference generator not built (make -C oracle ref)")
def test_oracle_equals_the_reference_generator(tmp_path):
    """The reference's own correlatedKT and calcProbT, compiled from
    nbcorrk/ with its compile-time sizes (2 temperatures, 20000 bins, 16
    g-points), against the oracle on the same spectra: kq, k_avg, kP, the
    band CDF and the g CDFs equal bit for bit."""
    from radchan_amd.tables import WQ16_F64
    sp = synthetic_lbl(wv0=150.0, wv1=750.0, deltawv=10.0, temps=[700.0, 1900.0], seed=3)
    nB, nT = sp["nB"], 2
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(struct.pack("<i", nB))
        f.write(np.asarray(sp["temps"], "<f8").tobytes())
        for s in range(nT * nB):
            b, e = sp["off"][s], sp["off"][s + 1]
            f.write(struct.pack("<idd", int(e - b), sp["wvl"][s], sp["wvr"][s]))
            f.write(np.asarray(sp["wvn"][b:e], "<f8").tobytes())
            f.write(np.asarray(sp["kvn"][b:e], "<f8").tobytes())
    (tmp_path / "errors").mkdir()
    subprocess.run([REF_CK, "in.bin", "out.bin"], cwd=tmp_path, check=True, capture_output=True)
    raw = np.fromfile(tmp_path / "out.bin", "<f8")
    per = 16 + 1 + 1 + 16
    kP = np.zeros(nT)
    kq, ka, cu, cg = (np.zeros((nT, nB, 16)), np.zeros((nT, nB)), np.zeros((nT, nB)), np.zeros((nT, nB, 16)))
    pos = 0
    for i in range(nT):
        kP[i] = raw[pos]
        pos += 1
        rec = raw[pos:pos + nB * per].reshape(nB, per)
        pos += nB * per
        kq[i], ka[i], cu[i], cg[i] = rec[:, :16], rec[:, 16], rec[:, 17], rec[:, 18:]
    o = O.ck_tables(sp, GQ16_F64, WQ16_F64)
    assert np.array_equal(o["kq"], kq)
    assert np.array_equal(o["kavg"], ka)
    assert np.array_equal(o["kP"], kP)
    assert np.array_equal(o["cuprob"], cu)
    assert np.array_equal(o["cuprob_g"], cg)


def test_library_rejects_bad_generator_input():
    """Input checks of rc_ckdist and rc_ck_tables, before any device work
    (so they run without a GPU)."""
    from radchan_amd.spectra import ckdist, tables_from_distributions
    from radchan_amd.tables import WQ16_F64
    nu = 100.0 + 0.01 * np.arange(8193)
    with pytest.raises(RuntimeError, match="8193 points"):
        ckdist([0, 8193], nu, np.ones(8193), [nu[0]], [nu[-1]], GQ16_F64)
    with pytest.raises(RuntimeError, match="nondecreasing"):
        ckdist([0, 10], nu[:10], np.ones(10), [nu[0]], [nu[9]], GQ16_F64[::-1].copy())
    with pytest.raises(RuntimeError, match="wv_right <= wv_left"):
        ckdist([0, 10], nu[:10], np.ones(10), [nu[9]], [nu[0]], GQ16_F64)
    with pytest.raises(RuntimeError, match="n_pwrk"):
        ckdist([0, 10], nu[:10], np.ones(10), [nu[0]], [nu[9]], GQ16_F64, n_pwrk=70000)
    kq = np.ones((2, 3, 16))
    kavg = np.ones((2, 3))
    wl, wr = np.array([1.0, 2.0, 3.0]), np.array([2.0, 3.0, 4.0])
    with pytest.raises(RuntimeError, match="increasing"):
        tables_from_distributions([900.0, 800.0], wl, wr, kq, kavg, WQ16_F64)
    with pytest.raises(RuntimeError, match="no absorption"):
        tables_from_distributions([800.0, 900.0], wl, wr, kq, np.zeros((2, 3)), WQ16_F64)
    with pytest.raises(RuntimeError, match="wv_right <= wv_left"):
        tables_from_distributions([800.0, 900.0], wr, wl, kq, kavg, WQ16_F64)


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (libamdhip64.so.7), so that its pointers are that runtime's."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # H2D
        return p

    def empty(self, n):
        return self.put(np.zeros(n))

    def get(self, p, n):
        C = self.C
        out = np.zeros(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0   # D2H
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _ckdist_device(h, off, wv, kv, wl, wr, gq, max_points):
    from radchan_amd import lib as L
    ns, nq = len(off) - 1, len(gq)
    ptr = [h.put(np.asarray(off, np.int64))] + [h.put(np.asarray(a, np.float64)) for a in (wv, kv, wl, wr, gq)]
    outs = [h.empty(ns * nq), h.empty(ns), h.empty(ns), h.empty(ns)]
    L.check(L.load().rc_ckdist_device(ns, *ptr[:5], 20000, 0.1, nq, ptr[5], *outs, max_points, None),
            "rc_ckdist_device")
    return h.get(outs[0], ns * nq).reshape(ns, nq), h.get(outs[1], ns), h.get(outs[2], ns), h.get(outs[3], ns)


@pytest.mark.gpu
def test_gpu_ckdist_device_entry_point():
    """rc_ckdist_device on spectra already in HBM: the same kernel, so the
    same bits as the host entry point; a spectrum longer than the stated
    max_points gets NaNs, the others are unaffected."""
    from radchan_amd.spectra import ckdist
    d = synthetic_lbl(temps=[400.0, 1400.0], wv0=150.0, wv1=2150.0)
    ref = ckdist(d["off"], d["wvn"], d["kvn"], d["wvl"], d["wvr"], GQ16_F64)
    h = _Hip()
    try:
        print(f"rc_ckdist_device: {len(d['off']) - 1} spectra", flush=True)
        kq, kavg, kmin, kmax = _ckdist_device(h, d["off"], d["wvn"], d["kvn"], d["wvl"], d["wvr"], GQ16_F64,
                                              int(np.diff(d["off"]).max()))
        assert np.array_equal(kq, ref["kq"]) and np.array_equal(kavg, ref["kavg"])
        assert np.array_equal(kmin, ref["kmin"]) and np.array_equal(kmax, ref["kmax"])
        rng = np.random.default_rng(1)
        sizes = [100, 600, 50]
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        wv = np.concatenate([3000.0 + 0.01 * np.arange(n) for n in sizes])
        kv = 10.0 ** rng.uniform(-2.0, 2.0, off[-1])
        wl = np.array([wv[off[i]] for i in range(3)])
        wr = np.array([wv[off[i + 1] - 1] for i in range(3)])
        r2 = ckdist(off, wv, kv, wl, wr, GQ16_F64)
        print("rc_ckdist_device: max_points below the longest spectrum", flush=True)
        kq2, kavg2, _, _ = _ckdist_device(h, off, wv, kv, wl, wr, GQ16_F64, 300)
        assert np.all(np.isnan(kq2[1])) and np.isnan(kavg2[1])
        assert np.array_equal(kq2[[0, 2]], r2["kq"][[0, 2]]) and np.array_equal(kavg2[[0, 2]], r2["kavg"][[0, 2]])
    finally:
        h.free()


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ckdist_ref.npz")


def _golden():
    g = np.load(GOLDEN)
    sp = {k: g[k] for k in ("off", "wvn", "kvn", "wvl", "wvr", "temps")}
    sp["nB"] = int(g["nB"])
    return sp, g


def test_oracle_equals_the_reference_generator_golden():
    """The oracle against the committed outputs of the reference's own
    correlatedKT + calcProbT (tests/golden/make_ckdist_golden.py): bit for
    bit, without needing the reference tree."""
    from radchan_amd.tables import WQ16_F64
    sp, g = _golden()
    o = O.ck_tables(sp, GQ16_F64, WQ16_F64)
    for k in ("kq", "kavg", "kP", "cuprob", "cuprob_g"):
        assert np.array_equal(o[k], g[k]), k


@pytest.mark.gpu
def test_gpu_generator_against_the_reference_golden():
    """The GPU generator (rc_ckdist, then rc_ck_tables) against the
    reference generator's committed outputs: kq and k_avg to FP64 rounding,
    the tables to FP32 rounding."""
    from radchan_amd.spectra import generate_tables
    from radchan_amd.tables import WQ16_F64
    sp, g = _golden()
    tab, d = generate_tables(sp, GQ16_F64, WQ16_F64, chanHeight=1.0)
    nT, nB = 2, sp["nB"]
    np.testing.assert_allclose(d["kq"].reshape(nT, nB, 16), g["kq"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(d["kavg"].reshape(nT, nB), g["kavg"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(tab.kP, g["kP"], rtol=3e-7)
    np.testing.assert_allclose(tab.prob, g["cuprob"], atol=6e-8)
    np.testing.assert_allclose(tab.probg, g["cuprob_g"], atol=6e-8)


@pytest.mark.gpu
def test_gpu_ckdist_uneven_grids_and_zero_k():
    """Spectra on uneven wavenumber grids (weights differ per pair), with
    k = 0 points (k^0.1 = 0 at the bottom of the range) and a range of ten
    decades, against the oracle."""
    rng = np.random.default_rng(21)
    wv, kv, off, wl, wr = [], [], [0], [], []
    for i, n in enumerate([40, 333, 1000, 1500]):
        nu = 4000.0 + np.cumsum(rng.uniform(0.002, 0.03, n))
        k = 10.0 ** rng.uniform(-5.0, 5.0, n)
        if i % 2 == 0:
            k[rng.integers(0, n, n // 10)] = 0.0
        wv.append(nu)
        kv.append(k)
        off.append(off[-1] + n)
        wl.append(nu[0])
        wr.append(nu[-1])
    print("rc_ckdist: uneven grids, zero k", flush=True)
    _gpu_vs_oracle(np.array(off), np.concatenate(wv), np.concatenate(kv), np.array(wl), np.array(wr))
