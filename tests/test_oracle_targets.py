"""Pins the oracle's extra-target evaluation (STEP 11, ifpghtarg >= 1).

* against the reference's own pote / grade / hesse (golden fixtures written
  by the compiled reference with targets, tests/golden/make_golden.py);
* against what the mathematics fixes: the expansion reproduces the node
  values at the nodes, a polynomial of degree < norder exactly everywhere,
  and the leaf found for each target contains it (ties to the + side);
* against the closed-form transform at the targets.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import exact_gauss, exact_gauss_deriv, golden_names, leaf_points, load_case, pgh_orders, rel_l2
from oracle import oracle

TGT = [n for n in golden_names() if load_case(n).ifpghtarg >= 1]


def _eval(c, vals, ncomp, nd=None):
    nd = c.nd if nd is None else nd
    v = vals.reshape(c.nboxes, c.npbox, -1)
    out = oracle.evalt(c.ndim, nd * ncomp, c.ipoly, c.norder, c.nboxes, c.nlevels, c.itree, c.iptr, c.centers,
                       c.boxsize, v, c.targs)
    return out


def test_target_fixtures_present():
    assert len(TGT) >= 4
    assert {load_case(n).ndim for n in TGT} == {1, 2, 3}
    assert {load_case(n).ifpghtarg for n in TGT} == {1, 2, 3}
    assert any(load_case(n).ifpghtarg > load_case(n).ifpgh for n in TGT)


@pytest.mark.parametrize("name", TGT)
def test_oracle_targets_match_reference(name):
    c = load_case(name)
    ifp = max(c.ifpgh, c.ifpghtarg)
    pot, grad, hess, info = oracle.boxfgt_pgh(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                                              c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize,
                                              c.fvals, ifp)
    assert info.npwlevel < c.nlevels
    assert rel_l2(_eval(c, pot, 1).reshape(c.pote_ref.shape), c.pote_ref) < 1e-12
    if c.ifpghtarg >= 2:
        assert rel_l2(_eval(c, grad, c.ndim).reshape(c.grade_ref.shape), c.grade_ref) < 1e-12
    if c.ifpghtarg >= 3:
        nh = c.ndim * (c.ndim + 1) // 2
        assert rel_l2(_eval(c, hess, nh).reshape(c.hesse_ref.shape), c.hesse_ref) < 1e-12


def _leaf_of(c, t):
    """Brute force: among the leaves whose closed box holds t, the one on the +
    side of every tie (largest lower corner), as the root-to-leaf descent
    with 'child bit set when t - c >= 0' picks."""
    level = c.itree[c.iptr[1] - 1: c.iptr[1] - 1 + c.nboxes]
    best, best_lo = None, None
    for b in c.leaves:
        h = c.boxsize[level[b]] / 2
        lo, hi = c.centers[b] - h, c.centers[b] + h
        if np.all(t >= lo) and np.all(t <= hi):
            if best is None or np.all(lo >= best_lo):
                best, best_lo = b, lo
    return best


@pytest.mark.parametrize("name", ["tgt3_n8", "tgt2_up_n8", "tgt1_n8"])
def test_evalt_is_interpolation(name):
    c = load_case(name)
    level = c.itree[c.iptr[1] - 1: c.iptr[1] - 1 + c.nboxes]
    rng = np.random.default_rng(5)
    # (i) at the nodes of a few leaves the expansion returns the node values
    leaves = rng.choice(c.leaves, size=min(6, len(c.leaves)), replace=False)
    pts = leaf_points(c.ndim, c.norder, c.ipoly, c.centers, c.boxsize, level, leaves).reshape(-1, c.ndim)
    vals = rng.standard_normal((c.nboxes, c.npbox))
    c.targs = pts
    got = _eval(c, vals, 1, nd=1)
    np.testing.assert_allclose(got, vals[leaves].ravel(), rtol=0, atol=1e-11 * np.abs(vals).max())
    # (ii) a polynomial of degree < norder per axis is reproduced anywhere
    allpts = leaf_points(c.ndim, c.norder, c.ipoly, c.centers, c.boxsize, level, np.arange(c.nboxes))
    poly = lambda p: np.prod([1 + 0.3 * p[..., a] - 2.0 * p[..., a] ** 3 for a in range(c.ndim)], axis=0)  # noqa
    vals = poly(allpts)
    c.targs = rng.uniform(-0.5, 0.5, (300, c.ndim))
    got = _eval(c, vals, 1, nd=1)
    np.testing.assert_allclose(got, poly(c.targs), rtol=0, atol=1e-12)


def test_target_leaf_assignment_brute_force():
    c = load_case("tgt2_up_n8")
    level = c.itree[c.iptr[1] - 1: c.iptr[1] - 1 + c.nboxes]
    # a field equal to the box id on each leaf (constant per leaf) reveals the leaf
    vals = np.repeat(np.arange(c.nboxes, dtype=float)[:, None], c.npbox, axis=1)
    targs = load_case("tgt2_up_n8").targs
    c.targs = targs
    got = _eval(c, vals, 1, nd=1)
    for t, g in zip(targs, got):
        b = _leaf_of(c, t)
        assert b is not None and abs(g - b) < 1e-9, (t, g, b)
    del level


@pytest.mark.parametrize("name", ["tgt2_up_n8", "tgt1_n8"])
def test_oracle_targets_closed_form(name):
    c = load_case(name)
    ifp = max(c.ifpgh, c.ifpghtarg)
    pot, grad, hess, _ = oracle.boxfgt_pgh(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                                           c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize,
                                           c.fvals, ifp)
    p = _eval(c, pot, 1)
    if c.nd > 1:
        p = p[:, 0]
    assert rel_l2(p, exact_gauss(c.targs, c.gauss, c.delta, c.iperiod)) < 100 * c.eps
    go, _ = pgh_orders(c.ndim)
    if c.ifpghtarg >= 2:
        g = _eval(c, grad, c.ndim).reshape((len(c.targs), c.ndim) + ((c.nd,) if c.nd > 1 else ()))
        for k, o in enumerate(go):
            gk = g[:, k, 0] if c.nd > 1 else g[:, k]
            assert rel_l2(gk, exact_gauss_deriv(c.targs, c.gauss, c.delta, c.iperiod, o)) < 1e3 * c.eps
