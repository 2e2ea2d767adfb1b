"""The generator's front end (rc_lbl_bands, rc_lbl_fill): the wide bands of
absWideBand (nbcorrk/read_files.cpp:22-73), the narrow split of the
generator's main (ckdist_hitran_part.cpp:55-114) and readFillT
(read_files.cpp:76-114).

Pinned by known answers worked out by hand from the rules, and by the
reference's own absWideBand, allocT and readFillT compiled from nbcorrk/
(oracle/_ref/ref_lbl) on files written here.  Host code only: no GPU.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from radchan_amd import spectra as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LBL = os.path.join(ROOT, "oracle", "_ref", "ref_lbl")


def _grid(w0, n):
    """n wavenumbers from w0 by 0.01, as a spectrum file prints them"""
    return np.round(w0 + 0.01 * np.arange(n), 2)


def test_narrow_split_known_answers():
    # 100 cm^-1 (10001 points) at deltawv 10: ten bands, the first holding
    # comp + 1 = 1001 points, the others 1000
    wv = _grid(2000.0, 10001)
    b = S.lbl_bands(wv, np.full(len(wv), 1e-3))
    assert b["nwide"] == 1
    assert list(b["count"]) == [1001] + [1000] * 9
    assert list(b["first"]) == [0] + [1001 + 1000 * h for h in range(9)]
    assert b["wvl"][1] == wv[1001] and b["wvr"][0] == wv[1000] and b["wvr"][-1] == wv[-1]
    assert np.array_equal(b["wvc"], (b["wvl"] + b["wvr"]) / 2)
    # 25.5 cm^-1 (2551 points): int(2.55) = 2 bands, comp = 1275: 1276 + 1275
    b = S.lbl_bands(_grid(500.0, 2551), np.full(2551, 1e-3))
    assert list(b["count"]) == [1276, 1275]
    # narrower than deltawv: one band of every point
    b = S.lbl_bands(_grid(500.0, 700), np.full(700, 1e-3))
    assert list(b["count"]) == [700] and b["first"][0] == 0
    # deltawv per wide band: the fourth wide band uses deltawv[1], the
    # seventh deltawv[2]
    kv = np.full(7 * 3000, 1e-9)
    for w in range(7):
        kv[w * 3000:w * 3000 + 2001] = 1e-3
    b = S.lbl_bands(_grid(100.0, len(kv)), kv, wvtoll=0.045, deltawv=(10.0, 5.0, 4.0))
    assert b["nwide"] == 7
    per = [2, 2, 2, 4, 4, 4, 5]                      # 2006 points, 20.05 cm^-1: int(20.05 / 10, / 5, / 4)
    assert len(b["count"]) == sum(per)


def test_wide_band_known_answers():
    """toll 1e-6, wvtoll 0.045 (4.5 weak points of 0.01): strong points
    0-99, weak 100-199, strong 200-299, weak 300-302, strong 303-399, weak
    to the end.  Band 0 takes points 0-104 (five weak points while the weak
    count is 0, 0.01, .., 0.04 < 0.045); point 200 opens band 1 (weak count
    0.05 >= 0.045), which runs through the three-point gap to 399 and the
    first five weak points after it."""
    n = 500
    kv = np.full(n, 1e-8)
    kv[0:100] = 1e-3
    kv[200:300] = 2e-3
    kv[303:400] = 3e-3
    wv = _grid(1000.0, n)
    b = S.lbl_bands(wv, kv, wvtoll=0.045, deltawv=(0.5, 0.5, 0.5))
    assert b["nwide"] == 2
    first, count = b["first"], b["count"]
    ends = first + count
    w0 = first < 200
    assert first[w0][0] == 0 and ends[w0][-1] == 105
    assert first[~w0][0] == 200 and ends[~w0][-1] == 405
    # leading weak points never open a band
    kv2 = kv.copy()
    kv2[:10] = 1e-9
    b2 = S.lbl_bands(wv, kv2, wvtoll=0.045, deltawv=(0.5, 0.5, 0.5))
    assert b2["first"][0] == 10


def test_fill_known_answers():
    """Points are taken where the wavenumber equals the next one awaited;
    kPr sums pi k 100 I_b dwv 100 / (sigma T^4) over the points read until
    the bands are full; a missing wavenumber stops the filling there."""
    wvn = np.array([10.0, 10.01, 10.02, 20.0])
    wv = np.array([9.99, 10.0, 10.005, 10.01, 10.02, 15.0, 20.0, 25.0])
    kv = np.arange(1, 9) * 1e-3
    kvn, kpr, q = S.lbl_fill(wv, kv, 1000.0, wvn)
    assert q == 4
    assert np.array_equal(kvn, np.array([2e-3, 4e-3, 5e-3, 7e-3]) * 100)
    pi, st = 3.14159265358979323846, 5.670373e-08
    C1, C2 = 3.741771790075259e-16, 0.014387741858429
    ib = lambda T, nu: 1.0 / pi * C1 * (nu * 100) ** 3 / (np.exp(np.float32(C2 * nu * 100 / T)) - 1)
    want = sum(pi * kv[i] * 100 * ib(1000.0, wv[i]) * 0.01 * 100 / (st * 1000.0 ** 4) for i in range(7))
    assert kpr == pytest.approx(want, rel=1e-6)                   # the eighth point is never read
    kvn, kpr2, q = S.lbl_fill(np.delete(wv, 3), np.delete(kv, 3), 1000.0, wvn)
    assert q == 1 and kvn[0] == 0.2 and kvn[1] == 0.0


def _write(path, wv, kv):
    with open(path, "w") as f:
        f.write("".join(f"{a:.2f} {b:.9e}\n" for a, b in zip(wv, kv)))


def _lines(wv, seed, T):
    """a seeded line spectrum k(nu) at temperature T (cm^-1)"""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(wv[0], wv[-1], 40)
    wid = 10.0 ** rng.uniform(-2, 0, 40)
    s0 = 10.0 ** rng.uniform(-4, -1, 40)
    ex = rng.uniform(-1.5, 1.0, 40)
    k = 1e-6 * np.ones_like(wv)
    for p, w, s, e in zip(pos, wid, s0, ex):
        k += s * (T / 1000.0) ** e * w * w / ((wv - p) ** 2 + w * w)
    return k


@pytest.mark.skipif(not os.path.exists(REF_LBL), reason="reference generator not built (make -C oracle ref)")
def test_front_end_equals_the_reference_generator(tmp_path):
    """The reference's absWideBand on a 2500 K spectrum with weak stretches
    at both ends, and allocT + readFillT on 2500 K and 2510 K spectra that
    hold extra wavenumbers between and around the bands' points: the wide
    bands, every filled coefficient and kPr equal the library's bit for
    bit; the narrow bands follow the split rule inside the reference's
    wide band."""
    n = 4501
    wv = _grid(3000.0, n)
    k0 = _lines(wv, 1, 2500.0)
    k0[:150] = 5e-7                       # weak head: never opens a band
    k0[-120:] = 2e-7                      # weak tail: kept (well inside wvtoll 25000)
    (tmp_path / "CO2-data").mkdir()
    _write(tmp_path / "ref.txt", wv, k0)
    wv_r, kv_r = S.read_lbl(str(tmp_path / "ref.txt"))
    b = S.lbl_bands(wv_r, kv_r)
    nb = len(b["first"])
    with open(tmp_path / "bands.bin", "wb") as f:
        f.write(struct.pack("<i", nb))
        for fi, c in zip(b["first"], b["count"]):
            f.write(struct.pack("<i", int(c)))
            f.write(np.asarray(wv_r[fi:fi + c], "<f8").tobytes())
    # temperature files: the band points plus extra wavenumbers in between
    fill = {}
    for T in (2500, 2510):
        extra = np.round(wv[::7] + 0.004, 3)
        wv_t = np.sort(np.concatenate([wv, extra]))
        kv_t = _lines(wv_t, 2, float(T))
        _write(tmp_path / "CO2-data" / f"CO2_{T}K_HITRAN.txt", wv_t, kv_t)
        w, k = S.read_lbl(str(tmp_path / "CO2-data" / f"CO2_{T}K_HITRAN.txt"))
        idx = np.concatenate([np.arange(f0, f0 + c) for f0, c in zip(b["first"], b["count"])])
        fill[T] = S.lbl_fill(w, k, float(T), wv_r[idx])
        assert fill[T][2] == len(idx)
    subprocess.run([REF_LBL, "ref.txt", "bands.bin", "out.bin"], cwd=tmp_path, check=True, capture_output=True)
    raw = (tmp_path / "out.bin").read_bytes()
    nw = struct.unpack_from("<i", raw, 0)[0]
    pos = 4
    wide = []
    for _ in range(nw):
        wide.append(struct.unpack_from("<idd", raw, pos))
        pos += 20
    assert nw == b["nwide"] == 1
    size, wfirst, wlast = wide[0]
    assert b["first"][0] == 150 and wv_r[b["first"][0]] == wfirst
    assert b["first"][-1] + b["count"][-1] == 150 + size and wv_r[b["first"][-1] + b["count"][-1] - 1] == wlast
    nexp = max(1, int((wlast - wfirst) / 10.0))
    comp = size // nexp
    assert list(b["count"]) == [comp + 1] + [comp] * (nexp - 2) + [size - (nexp - 1) * comp - 1]
    for T in (2500, 2510):
        Tr, kpr = struct.unpack_from("<dd", raw, pos)
        pos += 16
        kvn = np.frombuffer(raw, "<f8", int(b["count"].sum()), pos)
        pos += 8 * int(b["count"].sum())
        assert Tr == T
        assert np.array_equal(kvn, fill[T][0])
        assert kpr == fill[T][1]


@pytest.mark.skipif(not os.path.exists(REF_LBL), reason="reference generator not built (make -C oracle ref)")
def test_wide_bands_equal_the_reference_across_a_long_gap(tmp_path):
    """With the reference's wvtoll of 25000 (2.5 M weak points of 0.01), a
    weak stretch longer than that splits the spectrum into two wide bands:
    the reference's absWideBand and the library agree on both (first and
    last wavenumber and size)."""
    nweak = 2_500_100
    kv = np.concatenate([np.full(800, 1e-3), np.full(nweak, 1e-8), np.full(900, 2e-3), np.full(50, 1e-8)])
    wv = _grid(100.0, len(kv))
    _write(tmp_path / "ref.txt", wv, kv)
    (tmp_path / "CO2-data").mkdir()
    for T in (2500, 2510):
        _write(tmp_path / "CO2-data" / f"CO2_{T}K_HITRAN.txt", wv[:10], kv[:10])
    with open(tmp_path / "bands.bin", "wb") as f:
        f.write(struct.pack("<ii", 1, 1))
        f.write(np.asarray(wv[:1], "<f8").tobytes())
    wv_r, kv_r = S.read_lbl(str(tmp_path / "ref.txt"))
    b = S.lbl_bands(wv_r, kv_r)
    subprocess.run([REF_LBL, "ref.txt", "bands.bin", "out.bin"], cwd=tmp_path, check=True, capture_output=True)
    raw = (tmp_path / "out.bin").read_bytes()
    nw = struct.unpack_from("<i", raw, 0)[0]
    wide = [struct.unpack_from("<idd", raw, 4 + 20 * i) for i in range(nw)]
    assert nw == b["nwide"] == 2
    ends = b["first"] + b["count"]
    second = b["first"] >= 800 + nweak
    for (size, w0, w1), sel in zip(wide, (~second, second)):
        lo, hi = b["first"][sel][0], ends[sel][-1]
        assert (hi - lo, wv_r[lo], wv_r[hi - 1]) == (size, w0, w1)


@pytest.mark.gpu
def test_lbl_files_to_tables_on_the_gpu(tmp_path):
    """Spectrum files -> narrow bands and fills (host) -> correlated-k
    distributions (rc_ckdist on the GPU) -> tables: the distributions equal
    the oracle's correlatedKT on the same filled bands (to FP64 rounding,
    as in the generator tests)."""
    from radchan_amd.spectra import generate_tables, spectra_from_lbl
    from radchan_amd.tables import GQ16_F64, WQ16_F64
    from oracle import oracle as O
    print("lbl files -> tables", flush=True)
    wv = _grid(1200.0, 3001)
    _write(tmp_path / "ref.txt", wv, _lines(wv, 5, 1000.0))
    temps = [600.0, 1000.0, 1400.0]
    files = []
    for T in temps:
        _write(tmp_path / f"T{int(T)}.txt", wv, _lines(wv, 6, T))
        files.append(S.read_lbl(str(tmp_path / f"T{int(T)}.txt")))
    spec = spectra_from_lbl(S.read_lbl(str(tmp_path / "ref.txt")), temps, files)
    assert spec["nB"] == 3 and len(spec["kPr"]) == 3
    tab, d = generate_tables(spec, GQ16_F64, WQ16_F64, chanHeight=1.0)
    o = O.ck_tables(spec, GQ16_F64, WQ16_F64)
    np.testing.assert_allclose(d["kq"].reshape(o["kq"].shape), o["kq"], rtol=5e-15)
    np.testing.assert_allclose(d["kavg"].reshape(o["kavg"].shape), o["kavg"], rtol=5e-15)
    assert tab.nB == 3 and tab.nT == 3
