"""Many independent problems in one batch (fgt_batch): 3-D problems of norder
8/12/16 grouped into shape classes, every stage of a class one launch over
all its problems, the whole transform replayed from a captured HIP graph,
per-class tile choices from the autotuner.

* the uneven-batch generator at its full 512 problems (25 % 2-D, norder
  8/12/16, eps 1e-3/1e-6/1e-9, delta log-uniform in [1e-5, 1e-1], draws
  above 8000 leaves replaced and counted): every problem equals its
  one-at-a-time Solver result (FP64, relative l2 <= 1e-13), a sample
  matches the oracle (<= 1e-12);
* graph replays equal the eager enqueue bit for bit, with other device work
  between the replays;
* after autotuning, FP64 results stay within 1e-13 of the untuned ones and
  the cheapest precision per eps stays within eps of FP64;
* a periodic mix.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2

import fgt_amd
from fgt_amd.tree import gauss_source
from oracle import oracle

sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload generators)

pytestmark = pytest.mark.gpu


def _mixed_problems():
    out = []
    for i, (ndim, n, eps, delta, iper) in enumerate([(3, 8, 1e-6, 1e-3, 0), (2, 12, 1e-9, 1e-4, 1),
                                                    (3, 16, 1e-3, 3e-3, 0), (3, 12, 1e-6, 2e-2, 1),
                                                    (2, 8, 1e-6, 5e-3, 0), (3, 8, 1e-9, 1e-2, 0),
                                                    (3, 8, 1e-6, 4e-4, 1), (3, 12, 1e-3, 1e-4, 0),
                                                    (3, 16, 1e-6, 2e-4, 0), (3, 8, 1e-3, 8e-2, 0)]):
        rng = np.random.default_rng(40 + i)
        ng = int(rng.integers(1, 4))
        g = (rng.uniform(-0.4, 0.4, (ng, ndim)), np.exp(rng.uniform(np.log(2e-3), np.log(1e-2), (ng, ndim))),
             rng.uniform(0.5, 1.5, ng))
        tr = fgt_amd.vol_tree_build(ndim, n, 0, iper, max(eps, 1e-6), gauss_source(*g))
        out.append((tr, delta, eps))
    return out


def _single(tr, delta, eps):
    s = fgt_amd.Solver.from_tree(tr, delta, eps)
    s.set_fvals(tr.fvals)
    s.execute()
    ref = s.get_pot()
    s.close()
    return ref


def _batch(probs, nstreams=4):
    b = fgt_amd.Batch(probs, nstreams=nstreams)
    for i, (tr, _, _) in enumerate(probs):
        b.set_fvals(i, tr.fvals)
    return b


def test_batch_matches_single_solvers():
    print("[gpu] batch of 10 mixed problems (2-D side streams, 3-D classes)", flush=True)
    probs = _mixed_problems()
    b = _batch(probs, nstreams=2)
    b.execute()
    b.execute()   # graph replay
    for i, (tr, delta, eps) in enumerate(probs):
        got = b.get_pot(i)
        ref = _single(tr, delta, eps)
        assert rel_l2(got, ref) <= 1e-13, i
        if i in (0, 1):
            o, _, _ = oracle.boxfgt(1, tr.ndim, delta, eps, 0, tr.iperiod, tr.norder, tr.nboxes, tr.nlevels,
                                    tr.itree, tr.iptr, tr.centers, tr.boxsize, tr.fvals)
            assert rel_l2(got, o) < 1e-12
    st = b.stats()
    assert st["nout"] == sum(len(t.leaves) * t.norder ** t.ndim for t, _, _ in probs)
    assert st["last_ms"] > 0
    assert len(b.classes()) >= 3
    b.close()


def test_graph_replays_equal_eager_enqueue():
    """Replays of the captured graph, with eager batch runs and a solver's own
    work between them, each equal to the eager result bit for bit."""
    print("[gpu] batch graph replays vs eager enqueue", flush=True)
    probs = _mixed_problems()
    b = _batch(probs)
    b.set_graph(False)
    b.execute()
    eager = [b.get_pot(i) for i in range(len(probs))]
    b.set_graph(True)
    tr, delta, eps = probs[3]
    other = fgt_amd.Solver.from_tree(tr, delta, eps)
    other.set_fvals(2.0 * tr.fvals)
    for rep in range(3):
        b.execute()
        for i in range(len(probs)):
            assert np.array_equal(b.get_pot(i), eager[i]), (rep, i)
        other.execute()   # device work between replays
        other.get_pot()
    other.close()
    b.close()


def test_autotune_and_cheapest_precision():
    print("[gpu] batch autotune, then the cheapest precision per eps", flush=True)
    probs = _mixed_problems()
    b = _batch(probs)
    b.execute()
    ref = [b.get_pot(i) for i in range(len(probs))]
    b.autotune(2)
    cls = b.classes()
    print(f"[gpu]   classes after tuning: {cls}", flush=True)
    assert all(c["group"] in (2, 3, 4, 6, 8) for c in cls)
    assert all(c["xs5"] in (0, 1, 2) for c in cls if c["step5_tasks"] > 0)
    b.execute()
    for i in range(len(probs)):
        assert rel_l2(b.get_pot(i), ref[i]) <= 1e-13, i
    nmixed = b.set_precision(3)
    assert nmixed >= 4
    b.execute()
    for i, (tr, delta, eps) in enumerate(probs):
        got, r = b.get_pot(i)[tr.leaves], ref[i][tr.leaves]
        l2 = rel_l2(got, r)
        mx = float(np.abs(got - r).max() / np.abs(r).max())
        assert l2 <= eps and mx <= eps, (i, l2, mx, eps)
    b.close()


@pytest.mark.timeout(900)
def test_uneven_generator_512_problems():
    """The uneven generator's 512 problems in one batch: each equals its
    one-at-a-time solve, the smallest ones match the oracle."""
    probs, draws, replaced = bench.c4_problems(512)
    print(f"[gpu] uneven generator: 512 problems ({draws} draws, {replaced} over 8000 leaves replaced)",
          flush=True)
    b = _batch(probs)
    b.execute()
    b.execute()
    ms = b.stats()["last_ms"]
    print(f"[gpu]   batch {ms:.2f} ms, {len(b.classes())} classes", flush=True)
    exact = 0
    worst = 0.0
    for i, (tr, delta, eps) in enumerate(probs):
        got = b.get_pot(i)
        ref = _single(tr, delta, eps)
        err = rel_l2(got, ref)
        worst = max(worst, err)
        exact += int(np.array_equal(got, ref))
        assert err <= 1e-13, (i, err)
    print(f"[gpu]   worst rel l2 vs single solves {worst:.2e}; bit-identical {exact}/512", flush=True)
    small = sorted(range(len(probs)), key=lambda i: len(probs[i][0].leaves))[:8]
    for i in small:
        tr, delta, eps = probs[i]
        o, _, _ = oracle.boxfgt(1, tr.ndim, delta, eps, 0, tr.iperiod, tr.norder, tr.nboxes, tr.nlevels,
                                tr.itree, tr.iptr, tr.centers, tr.boxsize, tr.fvals)
        if np.linalg.norm(o) > 0:
            assert rel_l2(b.get_pot(i), o) < 1e-12, i
    b.close()


@pytest.mark.timeout(600)
def test_periodic_mix():
    probs, draws, replaced = bench.c4_problems(24, periodic=True)
    print(f"[gpu] periodic mix: 24 problems ({sum(t.iperiod for t, _, _ in probs)} periodic)", flush=True)
    b = _batch(probs)
    b.execute()
    for i, (tr, delta, eps) in enumerate(probs):
        assert rel_l2(b.get_pot(i), _single(tr, delta, eps)) <= 1e-13, i
    b.close()
