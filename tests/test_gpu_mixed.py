"""Mixed precision on the GPU: FP32 (eps >= 1e-6) and BF16 hi/lo (eps >=
1e-3) matrix-core contractions with FP32 expansions, FP64 phases, weights
and output accumulation.  Every result states its error against the FP64
oracle on the same tree; the bar is relative l2 and max error <= eps.

The six (eps, delta) points are the mixed-precision configuration: the
3-Gaussian generator (seed 0, centres U[-0.4,0.4]^3, isotropic widths
log-U[1e-3,1e-2], amplitudes U[0.5,1.5]) at norder 12, trees refined to eps,
eps in {1e-3, 1e-6} x delta in {1e-2, 1e-3, 1e-4}.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source
from oracle import oracle

pytestmark = pytest.mark.gpu

C3_POINTS = [(1e-3, 1e-2), (1e-3, 1e-3), (1e-3, 1e-4), (1e-6, 1e-2), (1e-6, 1e-3), (1e-6, 1e-4)]


def c3_tree(eps, n=12):
    rng = np.random.default_rng(0)
    c = rng.uniform(-0.4, 0.4, (3, 3))
    s = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), 3))
    a = rng.uniform(0.5, 1.5, 3)
    return fgt_amd.vol_tree_build(3, n, 0, 0, eps, gauss_source(c, np.repeat(s[:, None], 3, 1), a))


def errors(pot, ref, leaves):
    p, r = pot[leaves], ref[leaves]
    return rel_l2(p, r), float(np.abs(p - r).max() / np.abs(r).max())


@pytest.mark.parametrize("eps,delta", C3_POINTS)
def test_c3_point_within_eps(eps, delta):
    """FP32 at every point; BF16 hi/lo at the eps = 1e-3 points; both within
    eps of the FP64 oracle (relative l2 and max error).  FP64 after the
    reduced modes is unchanged bit for bit."""
    print(f"[gpu] mixed precision C3 point eps={eps} delta={delta}", flush=True)
    tr = c3_tree(eps)
    ref, info, _ = oracle.boxfgt(1, 3, delta, eps, 0, 0, 12, tr.nboxes, tr.nlevels, tr.itree, tr.iptr,
                                 tr.centers, tr.boxsize, tr.fvals)
    s = fgt_amd.Solver.from_tree(tr, delta, eps)
    s.set_fvals(tr.fvals)
    s.execute()
    p64 = s.get_pot()
    assert rel_l2(p64, ref) < 1e-12
    modes = [1, 2] if eps >= 1e-3 else [1]
    if info.npwlevel > tr.nlevels:
        # cutoff below the finest level: near field only, nothing to run in
        # reduced precision (the FP64 result above is the answer)
        for m in modes:
            with pytest.raises(fgt_amd.FgtError, match="near field only"):
                s.set_precision(m)
        print(f"[gpu]   leaves={len(tr.leaves)}: near field only, FP64", flush=True)
        s.close()
        return
    for m in modes:
        s.set_precision(m)
        s.execute()
        l2, mx = errors(s.get_pot(), ref, tr.leaves)
        print(f"[gpu]   leaves={len(tr.leaves)} npw={s.stats()['npw']} mode {m}: rel l2 {l2:.3e}, "
              f"max {mx:.3e} (eps {eps})", flush=True)
        assert 0 < l2 <= eps
        assert mx <= eps
    s.set_precision(0)
    s.execute()
    assert np.array_equal(s.get_pot(), p64)
    s.close()


def test_mode_floors_are_enforced():
    """FP32 below eps 1e-6 and BF16 hi/lo below eps 1e-3 are refused (status
    2, solver unchanged)."""
    tr = c3_tree(1e-6, n=8)
    s = fgt_amd.Solver.from_tree(tr, 1e-3, 1e-6)
    with pytest.raises(fgt_amd.FgtError, match="1e-3"):
        s.set_precision(2)
    s.close()
    s = fgt_amd.Solver.from_tree(tr, 1e-3, 1e-9)
    with pytest.raises(fgt_amd.FgtError, match="1e-6"):
        s.set_precision(1)
    s.close()


@pytest.mark.parametrize("n,eps,delta,iperiod,mode", [(8, 1e-6, 1e-3, 0, 1), (12, 1e-6, 1e-4, 1, 1),
                                                      (16, 1e-3, 1e-2, 0, 2), (16, 1e-6, 3e-4, 1, 1),
                                                      (8, 1e-3, 2e-4, 1, 2)])
def test_mixed_orders_and_periodic(n, eps, delta, iperiod, mode):
    """Each norder and both boundary conditions, against the FP64 result."""
    print(f"[gpu] mixed precision n={n} eps={eps} delta={delta} iperiod={iperiod} mode={mode}", flush=True)
    rng = np.random.default_rng(n)
    g = (rng.uniform(-0.4, 0.4, (3, 3)), np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), (3, 3))),
         rng.uniform(0.5, 1.5, 3))
    tr = fgt_amd.vol_tree_build(3, n, 0, iperiod, eps, gauss_source(*g))
    s = fgt_amd.Solver.from_tree(tr, delta, eps)
    s.set_fvals(tr.fvals)
    s.execute()
    ref = s.get_pot()
    s.set_precision(mode)
    s.execute()
    l2, mx = errors(s.get_pot(), ref, tr.leaves)
    print(f"[gpu]   rel l2 vs FP64 = {l2:.3e}, max {mx:.3e}", flush=True)
    assert 0 < l2 < eps and mx < eps
    s.close()


def test_mixed_rejected_off_the_3d_path():
    tr = fgt_amd.vol_tree_build(2, 8, 0, 0, 1e-6, gauss_source(np.array([[0.1, 0.2]]), np.array([[5e-3, 5e-3]]),
                                                                np.array([1.0])))
    s = fgt_amd.Solver.from_tree(tr, 1e-3, 1e-6)
    with pytest.raises(fgt_amd.FgtError):
        s.set_precision(1)
    s.close()
