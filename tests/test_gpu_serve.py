"""Serving loop: requests arriving while others run overlap on the device,
and every problem's result in the mix equals its result alone."""
from __future__ import annotations

import numpy as np
import pytest

import fgt_amd
from fgt_amd.serve import Server, poisson_arrivals
from fgt_amd.tree import gauss_source

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    print("[gpu] serving loop: 5 problems, 60 requests", flush=True)
    probs = []
    for i, (ndim, n, eps, delta, iper) in enumerate([(3, 8, 1e-6, 1e-3, 0), (2, 12, 1e-9, 1e-4, 1),
                                                    (3, 12, 1e-6, 2e-2, 1), (2, 8, 1e-6, 5e-3, 0),
                                                    (3, 16, 1e-3, 3e-3, 0)]):
        rng = np.random.default_rng(70 + i)
        g = (rng.uniform(-0.4, 0.4, (2, ndim)), np.exp(rng.uniform(np.log(2e-3), np.log(1e-2), (2, ndim))),
             rng.uniform(0.5, 1.5, 2))
        probs.append((fgt_amd.vol_tree_build(ndim, n, 0, iper, max(eps, 1e-6), gauss_source(*g)), delta, eps))
    alone = []
    for tr, delta, eps in probs:
        s = fgt_amd.Solver.from_tree(tr, delta, eps)
        s.set_fvals(tr.fvals)
        s.execute()
        alone.append(s.get_pot())
        s.close()
    srv = Server(nstreams=3)
    for k, (tr, delta, eps) in enumerate(probs):
        srv.register(k, tr, delta, eps)
    arr = poisson_arrivals(400.0, 0.15, len(probs), seed=3)
    assert len(arr) > 20
    lat, wall = srv.run(arr, 0.15)
    assert len(lat) == len(arr) and all(v > 0 for v in lat.values())
    for k in range(len(probs)):
        if k in srv.resident:
            assert np.array_equal(srv.result(k), alone[k]), k
    srv.close()
