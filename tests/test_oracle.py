"""Pins the C oracle (oracle/fgt_oracle.c) before it is trusted as the checker.

* against the reference's own outputs (golden fixtures written by the
  compiled reference, tests/golden/make_golden.py),
* against the reference's 1-D tables and scalars (tables.npz),
* against what the mathematics fixes: closed-form Gauss transforms of
  Gaussian sources, the Bessel integral identity, Gauss-Legendre exactness.
"""
from __future__ import annotations

import os

import numpy as np
import pytest
from scipy.special import spherical_jn

from conftest import GOLDEN, exact_gauss, golden_names, leaf_points, load_case, rel_l2
from oracle import oracle

TABLES = np.load(os.path.join(GOLDEN, "tables.npz"))


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    c = load_case(name)
    pot, info, _ = oracle.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                                 c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize,
                                 c.fvals, d1_mode=0)
    if np.linalg.norm(c.pot_ref) == 0:
        # the reference's npwlevel <= 0 / periodic small-level path: zero output
        assert info.npwlevel == 0
        assert np.all(pot == 0)
    else:
        assert rel_l2(pot, c.pot_ref) < 1e-12, name


def test_pwterms_and_nodes_match_reference():
    for bsize, delta, eps, iper, pmax_ref, npw_ref in TABLES["pwterms"]:
        pmax, npw = oracle.pwterms(bsize, delta, eps, int(iper))
        assert npw == int(npw_ref)
        if iper == 0:
            assert pmax == pytest.approx(pmax_ref, rel=1e-15)


@pytest.mark.parametrize("ipoly", [0, 1])
@pytest.mark.parametrize("n", [4, 6, 8, 12, 16])
def test_polyexps_match_reference(ipoly, n):
    x, w, u = oracle.polyexps(ipoly, n)
    np.testing.assert_allclose(x, TABLES[f"exps_{ipoly}_{n}_x"], rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, TABLES[f"exps_{ipoly}_{n}_w"], rtol=1e-13, atol=1e-15)
    uref = TABLES[f"exps_{ipoly}_{n}_u"].T    # -> [i, j]
    np.testing.assert_allclose(u, uref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("ipoly", [0, 1])
def test_polyexps_math(ipoly):
    # u interpolates: u applied to p_k(x_j) gives e_k, for k < n
    n = 9
    x, w, u = oracle.polyexps(ipoly, n)
    if ipoly == 0:
        V = np.polynomial.legendre.legvander(x, n - 1)
        xg, wg = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x, xg, atol=1e-15)
        np.testing.assert_allclose(w, wg, atol=1e-15)
    else:
        V = np.polynomial.chebyshev.chebvander(x, n - 1)
        np.testing.assert_allclose(np.sort(x), np.sort(np.cos((2 * np.arange(1, n + 1) - 1) * np.pi / (2 * n))),
                                   atol=1e-15)
        # Fejer weights integrate polynomials up to degree n-1 exactly
        for k in range(n):
            exact = (1 - (-1) ** (k + 1)) / (k + 1)
            assert np.dot(w, x ** k) == pytest.approx(exact, abs=1e-14)
    np.testing.assert_allclose(u @ V, np.eye(n), atol=1e-13)


def test_sph_bessel_matches_reference_and_scipy():
    for z, row in zip(TABLES["bessel_z"], TABLES["bessel_j"]):
        f = oracle.sph_bessel(16, z)
        np.testing.assert_allclose(f, row, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(f, spherical_jn(np.arange(17), z), rtol=1e-10, atol=1e-280)


def _table_keys(prefix):
    return sorted(k[len(prefix):] for k in TABLES.files if k.startswith(prefix))


@pytest.mark.parametrize("key", _table_keys("p2pw_"))
def test_pw_tables_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, boxdim = int(ipoly), int(n), float(delta), float(boxdim)
    ws, ts = TABLES[f"ws_{key}"], TABLES[f"ts_{key}"]
    npw = len(ts)
    t1 = oracle.poly2pw(ipoly, n, npw, ws, ts, delta, boxdim)        # [m, k]
    ref1 = TABLES[f"p2pw_{key}"].view(np.complex128).reshape(npw, n).T
    np.testing.assert_allclose(t1, ref1, rtol=0, atol=1e-13 * np.abs(ref1).max())
    t2 = oracle.pw2pot(ipoly, n, npw, ts, delta, boxdim)              # [k, i]
    ref2 = TABLES[f"pw2pot_{key}"].view(np.complex128).reshape(n, npw).T
    np.testing.assert_allclose(t2, ref2, rtol=0, atol=1e-14)
    if iper == "0":
        pmax = float(TABLES[f"pmax_{key}"])
        ws2, ts2 = oracle.pwnodes(pmax, npw)
        np.testing.assert_allclose(ts2, ts, rtol=1e-15)
        np.testing.assert_allclose(ws2, ws, rtol=1e-14)


def test_poly2pw_bessel_identity():
    # tab(m,k) = ws_k D/2 int_{-1}^1 P_m-interpolant e^{-i ts_k D x/(2 sqrt delta)}:
    # check against direct high-order quadrature of the interpolating polynomials
    n, npw, delta, boxdim = 8, 14, 1e-3, 0.25
    pmax, npw = oracle.pwterms(2 * boxdim, delta, 1e-6, 0)
    ws, ts = oracle.pwnodes(pmax, npw)
    tab = oracle.poly2pw(0, n, npw, ws, ts, delta, boxdim)
    x, _, u = oracle.polyexps(0, n)
    xq, wq = np.polynomial.legendre.leggauss(80)
    P = np.polynomial.legendre.legvander(xq, n - 1)        # [q, c]
    L = P @ u                                              # Lagrange basis at xq: [q, m]
    a = ts * boxdim / (2 * np.sqrt(delta))
    ref = (wq[:, None, None] * L[:, :, None] * np.exp(-1j * np.outer(xq, a))[:, None, :]).sum(0)
    ref = ref * ws[None, :] * boxdim / 2
    np.testing.assert_allclose(tab, ref, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("key", _table_keys("loctab_"))
def test_loctab_matches_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, eps, boxdim = int(ipoly), int(n), float(delta), float(eps), float(boxdim)
    tab, ind = oracle.loctab(eps, ipoly, n, delta, boxdim)
    rtab = TABLES[f"loctab_{key}"].reshape(13, n, n)
    rind = TABLES[f"locind_{key}"].reshape(13, n + 1, 2)
    for k in range(-6, 7):
        if abs(k) == 5:
            continue     # never referenced by a level-restricted tree
        s = k + 6
        np.testing.assert_array_equal(ind[s], rind[s], err_msg=f"band k={k}")
        scale = max(np.abs(rtab[s]).max(), 1e-300)
        np.testing.assert_allclose(tab[s], rtab[s], rtol=0, atol=1e-13 * scale, err_msg=f"k={k}")


def test_loctab_brute_force():
    # tab(j,k) (after u^T) integrates the Lagrange basis function of source
    # node j against the Gaussian centred at target node k, for the same-size
    # neighbour to the left (k = -4) and the self box (k = 0)
    n, delta, boxdim = 8, 1e-3, 0.125
    tab, _ = oracle.loctab(1e-12, 0, n, delta, boxdim)
    x, _, u = oracle.polyexps(0, n)
    xq, wq = np.polynomial.legendre.leggauss(200)
    L = np.polynomial.legendre.legvander(xq, n - 1) @ u           # [q, j]
    ys = xq * boxdim / 2                                           # source box at origin
    for k, shift in ((0, 0.0), (-4, -boxdim), (4, boxdim)):
        tx = x * boxdim / 2 + shift                                # target nodes
        G = np.exp(-(tx[None, :] - ys[:, None]) ** 2 / delta)      # [q, t]
        ref = (wq[:, None, None] * L[:, :, None] * G[:, None, :]).sum(0) * boxdim / 2   # [j, t]
        got = tab[k + 6]                                           # Fortran (j, t) -> numpy [t, j]
        np.testing.assert_allclose(got.T, ref, atol=1e-12 * np.abs(ref).max(), err_msg=f"k={k}")


def test_periodic_pwnodes_match_reference():
    keys = [k for k in TABLES.files if k.startswith("ts_") and k.endswith("_1")]
    assert keys
    for k in keys:
        ts_ref, ws_ref = TABLES[k], TABLES["ws_" + k[3:]]
        ws, ts = oracle.periodic_pwnodes(1.0, 1e-3, len(ts_ref))
        np.testing.assert_allclose(ts, ts_ref, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(ws, ws_ref, rtol=1e-14)


@pytest.mark.parametrize("name", ["f2_n8", "p2_cheb_n12", "f1_n8", "d1_free2", "d1_per2"])
def test_oracle_closed_form(name):
    """Oracle vs the exact transform of the Gaussian source (leaf nodes).

    The fixture trees resolve the source to tree_eps; eps <= 1e-6 here.  On the
    npwlevel <= 0 path the oracle is run with d1_mode=1 (PW for every box).
    """
    c = load_case(name)
    pot, info, _ = oracle.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                                 c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize,
                                 c.fvals, d1_mode=1)
    level = c.itree[c.iptr[1] - 1: c.iptr[1] - 1 + c.nboxes]
    pts = leaf_points(c.ndim, c.norder, c.ipoly, c.centers, c.boxsize, level, c.leaves)
    ex = exact_gauss(pts, c.gauss, c.delta, c.iperiod)
    err = rel_l2(pot[c.leaves], ex)
    assert err < 10 * max(c.eps, 1e-6), (name, err)


def test_oracle_plants_detected():
    """A dropped term must fail: zeroing one leaf's fvals changes the output."""
    c = load_case("c1_small")
    f = c.fvals.copy()
    big = c.leaves[np.argmax(np.abs(c.fvals[c.leaves]).max(axis=1))]
    f[big] = 0.0
    pot, _, _ = oracle.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                              c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize, f)
    assert rel_l2(pot, c.pot_ref) > 1e-3
