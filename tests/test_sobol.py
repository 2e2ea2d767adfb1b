"""The Sobol direction sampler (param.h:62-63 random 1; the reference's
emiss_angSobol, mc_gpu_noint.cu:1085-1092, on cuRAND Sobol32 generators
with Joe and Kuo's direction numbers, :84-88, :813-817).

CPU: the oracle's points equal scipy's unscrambled Sobol sequence (an
independent implementation of the same direction numbers; scipy runs in
Gray-code order, point i of it is point i xor (i >> 1) here); a cell and
run's shifted points keep the (0, m, 2)-net property (every elementary box
holds exactly one of 2^m points); the oracle's traced directions are those
of the shifted points; the grey slab closed form holds with Sobol
directions.  GPU: traces and tallies of both variants equal the oracle's
under the Sobol sampler."""
import numpy as np
import pytest
from scipy.special import expn
from scipy.stats import qmc

from radchan_amd import VARIANT_BATCHED, VARIANT_PLAIN, Solver, config
from radchan_amd import fields as F
from oracle import oracle as O
from tests import cases

SIGMA = 5.670373e-08


def test_points_equal_scipy_sobol():
    m = 12
    ref = qmc.Sobol(d=2, scramble=False).random_base2(m)
    for i in range(2 ** m):
        g = i ^ (i >> 1)
        x0, x1 = O.sobol2(g)
        assert (x0 / 2.0 ** 32, x1 / 2.0 ** 32) == (ref[i, 0], ref[i, 1]), i


@pytest.mark.parametrize("v,cell", [(0, 0), (3, 12345)])
def test_shifted_points_form_a_net(v, cell):
    m = 8
    u = np.array([O.sobol_uniforms(77, h, v, cell) for h in range(2 ** m)])
    x = np.round(u * 2.0 ** 24).astype(np.int64) - 1          # the 24-bit words behind the uniforms
    assert x.min() >= 0 and x.max() < 2 ** 24
    for k in range(m + 1):
        box = (x[:, 0] >> (24 - k)) * 2 ** (m - k) + (x[:, 1] >> (24 - (m - k)))
        assert len(np.unique(box)) == 2 ** m, k


def test_oracle_traces_use_the_shifted_points(nb_tables):
    p, tab, T = cases.nb_case(nb_tables, n=6, photmax=16, nVar=2)
    p.sampler = 1
    o = O.Oracle(p, tab, T)
    rays = np.array([[i, j, k, v, h] for (i, j, k) in ((1, 2, 3), (6, 6, 6)) for v in (0, 1) for h in range(16)],
                    np.int32)
    r = o.trace(rays)
    for q, (i, j, k, v, h) in enumerate(rays):
        cid = (i - 1) + p.imax * ((j - 1) + p.jmax * (k - 1))
        a, b = O.sobol_uniforms(p.seed, int(h), int(v), cid)
        s = O.iso_dir(a, b)
        want = [c if c != 0.0 else 1e-10 for c in s]
        assert np.array_equal(np.array([r["sx"][q], r["sy"][q], r["sz"][q]], np.float32),
                              np.array(want, np.float32)), q


def test_grey_slab_with_sobol_directions():
    """Unbiased (the closed form within 4.5 standard errors) and far less
    noisy than Philox directions on the same cells.  toll is 1e-6 here: at
    the default 1e-3 the estimator's cut-off (a ray below toll deposits its
    rest where it is, nothing in uniform gas, instead of reaching the
    wall) biases the exchange by about 1e-4 of the closed form, which the
    Sobol error bars (1e-5 to 2e-4) resolve."""
    se_rel = {}
    for sampler in (1, 0):
        p = config("slab32")
        p.photmax, p.nVar, p.seed, p.sampler, p.toll = 1024, 12, 5, sampler, 1e-6
        T = F.uniform_field(p, 1000.0, 500.0)
        A = T.reshape(p.kmax + 2, p.jmax + 2, p.imax + 2)
        A[:, :, 0] = A[:, :, -1] = 500.0
        o = O.Oracle(p, None, T)
        cells = np.array([[i, 7, 9] for i in (1, 3, 8, 16, 25, 32)], np.int32)
        _, _, de64, _ = o.solve(cells)
        x = O.grid(p, 0)["x"][cells[:, 0]].astype(np.float64)
        ana = 2 * SIGMA * (1000.0 ** 4 - 500.0 ** 4) * (expn(2, x) + expn(2, 2.0 - x))
        m = de64.mean(axis=1)
        se = de64.std(axis=1, ddof=1) / np.sqrt(p.nVar)
        assert np.all(np.abs((m - ana) / se) < 4.5), (sampler, (m - ana) / se, m / ana)
        se_rel[sampler] = float(np.mean(se / np.abs(ana)))
    assert se_rel[1] < 0.1 * se_rel[0], se_rel


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["nb", "mg"])
def test_gpu_sobol_parity(nb_tables, which):
    from tests.test_gpu_parity import INT_FIELDS, _rays
    p, tab, T = cases.nb_case(nb_tables) if which == "nb" else cases.mg_case(nb_tables)
    p.sampler = 1
    print(f"sobol sampler parity: {which}", flush=True)
    s = Solver(p, tab)
    s.set_fields([T])
    o = O.Oracle(p, tab, T)
    rays = _rays(p, 1024, 7)
    g, c = s.trace(rays), o.trace(rays)
    for f in INT_FIELDS:
        assert np.array_equal(g[f], c[f]), f
    for f in ("sx", "sy", "sz", "tra", "de"):
        assert np.array_equal(g[f].view(np.uint32), c[f].view(np.uint32)), f
    de32, de32d, _, _ = o.solve(O.all_cells(p))
    s.solve(VARIANT_PLAIN)
    assert np.array_equal(s.tallies(), de32.astype(np.float64))
    s.solve(VARIANT_BATCHED)
    assert np.abs(s.tallies() - de32d).max() <= 1e-12 * np.abs(de32d).max()
    s.close()
