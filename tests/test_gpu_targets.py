"""GPU parity of the extra-target evaluation (ifpghtarg >= 1) through the C ABI.

* every target fixture: the drop-in ``boxfgt_`` against the reference's own
  pote / grade / hesse (relative l2 <= 1e-12), including ifpghtarg > ifpgh
  (gradients/Hessians formed only for the targets) and targets on box faces;
* the Solver API on a 3-D tree (n = 8): eval_targets against the oracle's
  evaluation of the same device outputs.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_names, load_case, rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source
from oracle import oracle

pytestmark = pytest.mark.gpu
TGT = [n for n in golden_names() if load_case(n).ifpghtarg >= 1]


@pytest.mark.parametrize("name", TGT)
def test_dropin_targets_match_reference(name):
    print(f"[gpu] targets drop-in vs reference: {name}", flush=True)
    c = load_case(name)
    tail = () if c.nd == 1 else (c.nd,)
    nh = c.ndim * (c.ndim + 1) // 2
    nt = len(c.targs)
    pot = np.zeros_like(c.fvals)
    grad = np.zeros((c.nboxes, c.npbox, c.ndim) + tail) if c.ifpgh >= 2 else None
    hess = np.zeros((c.nboxes, c.npbox, nh) + tail) if c.ifpgh >= 3 else None
    pote = np.zeros((nt,) + tail)
    grade = np.zeros((nt, c.ndim) + tail) if c.ifpghtarg >= 2 else None
    hesse = np.zeros((nt, nh) + tail) if c.ifpghtarg >= 3 else None
    fgt_amd.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels,
                   c.ltree, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, c.ifpgh, pot, grad, hess,
                   ntarg=nt, targs=c.targs, ifpghtarg=c.ifpghtarg, pote=pote, grade=grade, hesse=hesse)
    assert rel_l2(pot, c.pot_ref) < 1e-12
    err = rel_l2(pote, c.pote_ref)
    print(f"[gpu]   pote {err:.3e}", flush=True)
    assert err < 1e-12
    if c.ifpghtarg >= 2:
        assert rel_l2(grade, c.grade_ref) < 1e-12
    if c.ifpghtarg >= 3:
        assert rel_l2(hesse, c.hesse_ref) < 1e-12


def test_solver_targets_vs_oracle():
    print("[gpu] solver eval_targets vs oracle (3-D n=8)", flush=True)
    gauss = (np.array([[0.1, -0.2, 0.3]]), np.array([[2e-2, 1.5e-2, 3e-2]]), np.array([1.0]))
    tr = fgt_amd.vol_tree_build(3, 8, 0, 0, 1e-6, gauss_source(*gauss))
    s = fgt_amd.Solver.from_tree(tr, 2e-3, 1e-6)
    s.set_ifpgh(3)
    s.set_fvals(tr.fvals)
    s.execute()
    pot, grad, hess = s.get_pot(), s.get_grad(), s.get_hess()
    targs = np.random.default_rng(3).uniform(-0.5, 0.5, (2000, 3))
    pote, grade, hesse = s.eval_targets(targs, 3)
    args = (8, tr.nboxes, tr.nlevels, tr.itree, tr.iptr, tr.centers, tr.boxsize)
    ref_p = oracle.evalt(3, 1, 0, *args, pot, targs)
    ref_g = oracle.evalt(3, 3, 0, *args, grad.reshape(tr.nboxes, -1, 3), targs).reshape(grade.shape)
    ref_h = oracle.evalt(3, 6, 0, *args, hess.reshape(tr.nboxes, -1, 6), targs).reshape(hesse.shape)
    assert rel_l2(pote, ref_p) < 1e-13
    assert rel_l2(grade, ref_g) < 1e-13
    assert rel_l2(hesse, ref_h) < 1e-13
    s.close()
