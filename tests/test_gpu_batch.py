"""Many independent problems on several streams (fgt_batch): every problem's
potential equals its one-at-a-time Solver result bit for bit, and the oracle
within 1e-12, whatever the mix of dimensions, orders and sizes."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source
from oracle import oracle

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    for i, (ndim, n, eps, delta, iper) in enumerate([(3, 8, 1e-6, 1e-3, 0), (2, 12, 1e-9, 1e-4, 1),
                                                    (3, 16, 1e-3, 3e-3, 0), (3, 12, 1e-6, 2e-2, 1),
                                                    (2, 8, 1e-6, 5e-3, 0), (3, 8, 1e-9, 1e-2, 0)]):
        rng = np.random.default_rng(40 + i)
        ng = int(rng.integers(1, 4))
        g = (rng.uniform(-0.4, 0.4, (ng, ndim)), np.exp(rng.uniform(np.log(2e-3), np.log(1e-2), (ng, ndim))),
             rng.uniform(0.5, 1.5, ng))
        tr = fgt_amd.vol_tree_build(ndim, n, 0, iper, max(eps, 1e-6), gauss_source(*g))
        out.append((tr, delta, eps))
    return out


def test_batch_matches_single_solvers():
    print("[gpu] batch of 6 mixed problems on 3 streams", flush=True)
    probs = _problems()
    b = fgt_amd.Batch(probs, nstreams=3)
    for i, (tr, _, _) in enumerate(probs):
        b.set_fvals(i, tr.fvals)
    b.execute()
    b.execute()   # repeatable
    for i, (tr, delta, eps) in enumerate(probs):
        s = fgt_amd.Solver.from_tree(tr, delta, eps)
        s.set_fvals(tr.fvals)
        s.execute()
        ref = s.get_pot()
        s.close()
        got = b.get_pot(i)
        assert np.array_equal(got, ref), i
        if i in (0, 1):
            o, _, _ = oracle.boxfgt(1, tr.ndim, delta, eps, 0, tr.iperiod, tr.norder, tr.nboxes, tr.nlevels,
                                    tr.itree, tr.iptr, tr.centers, tr.boxsize, tr.fvals)
            assert rel_l2(got, o) < 1e-12
    st = b.stats()
    assert st["nout"] == sum(len(t.leaves) * t.norder ** t.ndim for t, _, _ in probs)
    assert st["last_ms"] > 0
    b.close()
