"""Pins the oracle's gradients and Hessians (ifpgh = 2, 3).

* against the reference's own outputs with ifpgh = 2/3 (golden fixtures
  written by the compiled reference, tests/golden/make_golden.py);
* against the reference's derivative tables (mk_pw2pgh_tables,
  mk_loctab_all's tabx_loc / tabxx_loc);
* against what the mathematics fixes: derivatives of the closed-form Gauss
  transform of Gaussian sources, and brute-force quadrature of the
  derivative near-field integrals.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import (GOLDEN, exact_gauss_deriv, golden_names, leaf_points, load_case, pgh_orders,
                      rel_l2)
from oracle import oracle

TABLES = np.load(os.path.join(GOLDEN, "tables.npz"))
PGH = [n for n in golden_names() if load_case(n).ifpgh >= 2 and load_case(n).grad_ref is not None]


def _oracle_pgh(c):
    return oracle.boxfgt_pgh(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.nboxes,
                             c.nlevels, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, c.ifpgh)


def test_pgh_fixtures_present():
    assert len(PGH) >= 5
    assert {load_case(n).ndim for n in PGH} == {1, 2, 3}
    assert any(load_case(n).ifpgh == 2 for n in PGH)


@pytest.mark.parametrize("name", PGH)
def test_oracle_pgh_matches_reference(name):
    c = load_case(name)
    pot, grad, hess, info = _oracle_pgh(c)
    assert info.npwlevel < c.nlevels, "fixture must exercise the plane-wave path"
    assert rel_l2(pot, c.pot_ref) < 1e-12
    for k in range(c.ndim):
        assert rel_l2(grad[:, :, k], c.grad_ref[:, :, k]) < 1e-12, f"grad {k}"
    if c.ifpgh >= 3:
        for k in range(hess.shape[2]):
            assert rel_l2(hess[:, :, k], c.hess_ref[:, :, k]) < 1e-11, f"hess {k}"
    else:
        assert hess is None


def _check_closed_form(ndim, norder, ipoly, iperiod, delta, eps, tree, leaves, level, gauss, grad, hess):
    # each component against the same partial derivative of the exact
    # transform; a swapped component order or a dropped sign fails here.
    # Derivatives of an eps-accurate smooth field lose accuracy with order
    # (measured: grad <= 300 eps, hess <= 1.2e4 eps on these cases).
    pts = leaf_points(ndim, norder, ipoly, tree.centers, tree.boxsize, level, leaves)
    go, ho = pgh_orders(ndim)
    for k, o in enumerate(go):
        ex = exact_gauss_deriv(pts, gauss, delta, iperiod, o)
        assert rel_l2(grad[leaves][:, :, k], ex) < 1e3 * eps, f"grad {o}"
    for k, o in enumerate(ho):
        ex = exact_gauss_deriv(pts, gauss, delta, iperiod, o)
        assert rel_l2(hess[leaves][:, :, k], ex) < 3e4 * eps, f"hess {o}"


@pytest.mark.parametrize("name", ["pgh2_n12", "pgh1_n8", "pgh2_per_n8"])
def test_oracle_pgh_closed_form(name):
    c = load_case(name)
    pot, grad, hess, _ = _oracle_pgh(c)
    level = c.itree[c.iptr[1] - 1: c.iptr[1] - 1 + c.nboxes]
    _check_closed_form(c.ndim, c.norder, c.ipoly, c.iperiod, c.delta, c.eps, c, c.leaves, level, c.gauss,
                       grad, hess)


@pytest.mark.parametrize("iperiod,delta", [(1, 1e-3), (0, 2e-3)])
def test_oracle_pgh_closed_form_3d(iperiod, delta):
    # a 3-D tree resolved to 1e-6 (the golden 3-D trees are coarse)
    import fgt_amd
    from fgt_amd.tree import gauss_source
    gauss = (np.array([[0.1, -0.2, 0.3]]), np.array([[2e-2, 1.5e-2, 3e-2]]), np.array([1.0]))
    tr = fgt_amd.vol_tree_build(3, 8, 0, iperiod, 1e-6, gauss_source(*gauss))
    pot, grad, hess, info = oracle.boxfgt_pgh(1, 3, delta, 1e-6, 0, iperiod, 8, tr.nboxes, tr.nlevels,
                                              tr.itree, tr.iptr, tr.centers, tr.boxsize, tr.fvals, 3)
    assert 1 <= info.npwlevel < tr.nlevels
    _check_closed_form(3, 8, 0, iperiod, delta, 1e-6, tr, tr.leaves, tr.level, gauss, grad, hess)


def _table_keys(prefix):
    return sorted(k[len(prefix):] for k in TABLES.files if k.startswith(prefix))


@pytest.mark.parametrize("key", _table_keys("pw2der_"))
def test_pw2pgh_tables_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, boxdim = int(ipoly), int(n), float(delta), float(boxdim)
    ts = TABLES[f"ts_{key}"]
    npw = len(ts)
    t0, t1, t2 = oracle.pw2pgh(ipoly, n, npw, ts, delta, boxdim)
    for got, nm in ((t0, "pw2pot"), (t1, "pw2der"), (t2, "pw2dxx")):
        ref = TABLES[f"{nm}_{key}"].view(np.complex128).reshape(n, npw).T
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()), err_msg=nm)
    # closed form: d/dX exp(i t X / sqrt(delta)) at X = x_i D / 2
    x, _, _ = oracle.polyexps(ipoly, n)
    ph = np.exp(1j * np.outer(ts, x) * boxdim / (2 * np.sqrt(delta)))
    d = 1j * ts[:, None] / np.sqrt(delta)
    np.testing.assert_allclose(t1, d * ph, atol=1e-12 * np.abs(t1).max())
    np.testing.assert_allclose(t2, d * d * ph, atol=1e-12 * np.abs(t2).max())


@pytest.mark.parametrize("key", _table_keys("loctabx_"))
def test_loctab_derivatives_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, eps, boxdim = int(ipoly), int(n), float(delta), float(eps), float(boxdim)
    tab, tabx, tabxx, ind = oracle.loctab3(eps, ipoly, n, delta, boxdim)
    rx = TABLES[f"loctabx_{key}"].reshape(13, n, n)
    rxx = TABLES[f"loctabxx_{key}"].reshape(13, n, n)
    for k in range(-6, 7):
        if abs(k) == 5:
            continue
        s = k + 6
        for got, ref, nm in ((tabx[s], rx[s], "x"), (tabxx[s], rxx[s], "xx")):
            scale = max(np.abs(ref).max(), 1e-300)
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale, err_msg=f"{nm} k={k}")


def test_loctab_derivatives_brute_force():
    # d/dx and d2/dx2 (target coordinate) of the near-field integral
    n, delta, boxdim = 8, 1e-3, 0.125
    tab, tabx, tabxx, _ = oracle.loctab3(1e-12, 0, n, delta, boxdim)
    x, _, u = oracle.polyexps(0, n)
    xq, wq = np.polynomial.legendre.leggauss(200)
    L = np.polynomial.legendre.legvander(xq, n - 1) @ u
    ys = xq * boxdim / 2
    for k, shift in ((0, 0.0), (-4, -boxdim), (4, boxdim)):
        tx = x * boxdim / 2 + shift
        d = tx[None, :] - ys[:, None]
        G = np.exp(-d ** 2 / delta)
        for got, ker in ((tabx, -2 * d / delta * G), (tabxx, (4 * d ** 2 / delta ** 2 - 2 / delta) * G)):
            ref = (wq[:, None, None] * L[:, :, None] * ker[:, None, :]).sum(0) * boxdim / 2
            np.testing.assert_allclose(got[k + 6].T, ref, atol=1e-11 * np.abs(ref).max(), err_msg=f"k={k}")
