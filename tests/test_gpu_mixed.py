"""Mixed precision (FP32 plane-wave expansions, FP64 tables, phases,
arithmetic and outputs): error against the FP64 result below eps."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,eps,delta,iperiod", [(8, 1e-6, 1e-3, 0), (12, 1e-6, 1e-4, 1), (16, 1e-3, 1e-2, 0),
                                                 (16, 1e-6, 3e-4, 1)])
def test_mixed_within_eps(n, eps, delta, iperiod):
    print(f"[gpu] mixed precision n={n} eps={eps} delta={delta} iperiod={iperiod}", flush=True)
    rng = np.random.default_rng(n)
    g = (rng.uniform(-0.4, 0.4, (3, 3)), np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), (3, 3))),
         rng.uniform(0.5, 1.5, 3))
    tr = fgt_amd.vol_tree_build(3, n, 0, iperiod, eps, gauss_source(*g))
    s = fgt_amd.Solver.from_tree(tr, delta, eps)
    s.set_fvals(tr.fvals)
    s.execute()
    ref = s.get_pot()
    s.set_precision(1)
    s.execute()
    mix = s.get_pot()
    err = rel_l2(mix, ref)
    print(f"[gpu]   rel l2 vs FP64 = {err:.3e}", flush=True)
    assert 0 < err < eps
    s.set_precision(0)
    s.execute()
    assert np.array_equal(s.get_pot(), ref)
    s.close()


def test_mixed_rejected_off_the_3d_path():
    tr = fgt_amd.vol_tree_build(2, 8, 0, 0, 1e-6, gauss_source(np.array([[0.1, 0.2]]), np.array([[5e-3, 5e-3]]),
                                                                np.array([1.0])))
    s = fgt_amd.Solver.from_tree(tr, 1e-3, 1e-6)
    with pytest.raises(fgt_amd.FgtError):
        s.set_precision(1)
    s.close()
