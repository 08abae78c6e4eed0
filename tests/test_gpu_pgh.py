"""GPU parity of gradients and Hessians (ifpgh = 2, 3) through the C ABI.

* every ifpgh fixture: the drop-in ``boxfgt_`` against the reference's own
  pot/grad/hess (relative l2 per component <= 1e-12 / 1e-11);
* a 3-D tree resolved to 1e-6 (register-blocked 3-D kernels, n = 8) against
  the oracle (<= 1e-12) and each component against the closed-form
  derivative of the exact transform;
* the Solver API: pot is the same whatever ifpgh is, and grad/hess are
  overwritten on each execute.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import exact_gauss_deriv, golden_names, leaf_points, load_case, pgh_orders, rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source
from oracle import oracle

pytestmark = pytest.mark.gpu
PGH = [n for n in golden_names() if load_case(n).ifpgh >= 2 and load_case(n).grad_ref is not None]


def _arrays(c, ifpgh):
    tail = () if c.nd == 1 else (c.nd,)
    grad = np.zeros((c.nboxes, c.npbox, c.ndim) + tail) if ifpgh >= 2 else None
    hess = np.zeros((c.nboxes, c.npbox, c.ndim * (c.ndim + 1) // 2) + tail) if ifpgh >= 3 else None
    return np.zeros_like(c.fvals), grad, hess


@pytest.mark.parametrize("name", PGH)
def test_dropin_pgh_matches_reference(name):
    print(f"[gpu] pgh drop-in vs reference: {name}", flush=True)
    c = load_case(name)
    pot, grad, hess = _arrays(c, c.ifpgh)
    fgt_amd.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels,
                   c.ltree, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, c.ifpgh, pot, grad, hess)
    assert rel_l2(pot, c.pot_ref) < 1e-12
    for k in range(c.ndim):
        err = rel_l2(grad[:, :, k], c.grad_ref[:, :, k])
        print(f"[gpu]   grad {k}: {err:.3e}", flush=True)
        assert err < 1e-12
    if c.ifpgh >= 3:
        for k in range(hess.shape[2]):
            err = rel_l2(hess[:, :, k], c.hess_ref[:, :, k])
            print(f"[gpu]   hess {k}: {err:.3e}", flush=True)
            assert err < 1e-11
    # accumulated, never overwritten
    g0 = np.full_like(grad, 0.25)
    pot2, grad2, hess2 = _arrays(c, c.ifpgh)
    grad2 += g0
    fgt_amd.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels,
                   c.ltree, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, c.ifpgh, pot2, grad2, hess2)
    assert np.array_equal(grad2, g0 + grad)


@pytest.mark.parametrize("n,iperiod,delta", [(8, 1, 1e-3), (8, 0, 2e-3), (12, 0, 5e-3), (16, 0, 5e-3)])
def test_pgh_3d_vs_oracle_and_closed_form(n, iperiod, delta):
    """STEP 6 forms the components that share x and y table kinds in one
    pass (three z kinds at norder 8 / 12, two plus one at 16)."""
    print(f"[gpu] pgh 3-D n={n} iperiod={iperiod} delta={delta}", flush=True)
    gauss = (np.array([[0.1, -0.2, 0.3]]), np.array([[2e-2, 1.5e-2, 3e-2]]), np.array([1.0]))
    tr = fgt_amd.vol_tree_build(3, n, 0, iperiod, 1e-6, gauss_source(*gauss))
    s = fgt_amd.Solver.from_tree(tr, delta, 1e-6)
    s.set_ifpgh(3)
    s.set_fvals(tr.fvals)
    s.execute()
    pot, grad, hess = s.get_pot(), s.get_grad(), s.get_hess()
    rp, rg, rh, info = oracle.boxfgt_pgh(1, 3, delta, 1e-6, 0, iperiod, n, tr.nboxes, tr.nlevels, tr.itree,
                                         tr.iptr, tr.centers, tr.boxsize, tr.fvals, 3)
    assert 1 <= info.npwlevel < tr.nlevels
    assert rel_l2(pot, rp) < 1e-12
    for k in range(3):
        assert rel_l2(grad[:, :, k], rg[:, :, k]) < 1e-12, f"grad {k}"
    for k in range(6):
        assert rel_l2(hess[:, :, k], rh[:, :, k]) < 1e-11, f"hess {k}"
    pts = leaf_points(3, n, 0, tr.centers, tr.boxsize, tr.level, tr.leaves)
    go, ho = pgh_orders(3)
    for k, o in enumerate(go):
        assert rel_l2(grad[tr.leaves][:, :, k], exact_gauss_deriv(pts, gauss, delta, iperiod, o)) < 1e-3
    for k, o in enumerate(ho):
        assert rel_l2(hess[tr.leaves][:, :, k], exact_gauss_deriv(pts, gauss, delta, iperiod, o)) < 3e-2
    # pot does not depend on ifpgh; grad/hess are overwritten per execute
    s.execute()
    assert np.array_equal(s.get_grad(), grad)
    s.set_ifpgh(1)
    s.execute()
    assert np.array_equal(s.get_pot(), pot)
    s.close()


@pytest.mark.parametrize("norder,ipoly,iperiod,delta", [(12, 0, 0, 5e-4), (16, 1, 1, 1e-3)])
def test_pgh_2d_three_densities_vs_oracle(norder, ipoly, iperiod, delta):
    """2-D kernels (STEPs 1, 5, 6) with nd = 3 independent densities, all
    components, against the oracle (no reference fixture has 2-D nd > 1)."""
    print(f"[gpu] pgh 2-D nd=3 n={norder} ipoly={ipoly} iperiod={iperiod}", flush=True)
    gauss = (np.array([[0.1, -0.2], [-0.3, 0.25]]), np.array([[2e-2, 1e-2], [8e-3, 1.5e-2]]),
             np.array([1.0, 0.7]))
    eps = 1e-9
    tr = fgt_amd.vol_tree_build(2, norder, ipoly, iperiod, eps, gauss_source(*gauss))
    rng = np.random.default_rng(7)
    fv = np.stack([tr.fvals, -0.5 * tr.fvals, rng.uniform(-1.0, 1.0, tr.fvals.shape)], axis=-1)
    s = fgt_amd.Solver.from_tree(tr, delta, eps, nd=3)
    s.set_ifpgh(3)
    s.set_fvals(fv)
    s.execute()
    pot, grad, hess = s.get_pot(), s.get_grad(), s.get_hess()
    s.close()
    rp, rg, rh, info = oracle.boxfgt_pgh(3, 2, delta, eps, ipoly, iperiod, norder, tr.nboxes, tr.nlevels,
                                         tr.itree, tr.iptr, tr.centers, tr.boxsize, fv, 3)
    assert 1 <= info.npwlevel < tr.nlevels
    for d in range(3):
        assert rel_l2(pot[..., d], rp[..., d]) < 1e-12, f"pot {d}"
        for k in range(2):
            assert rel_l2(grad[:, :, k, d], rg[:, :, k, d]) < 1e-12, f"grad {k} density {d}"
        for k in range(3):
            assert rel_l2(hess[:, :, k, d], rh[:, :, k, d]) < 1e-11, f"hess {k} density {d}"
    # density 1 is -0.5 x density 0: the transform is linear per density
    assert np.allclose(pot[..., 1], -0.5 * pot[..., 0], rtol=0, atol=1e-13 * np.abs(pot[..., 0]).max())
