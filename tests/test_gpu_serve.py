"""Serving (fgt_server through fgt_amd.serve): requests drawn from the
uneven-batch generator, free space and periodic, arriving as a seeded
Poisson stream; every result equals its problem solved alone.

* FP64 stream at a rate that keeps several requests in flight: each kept
  result equals the one-at-a-time Solver result (relative l2 <= 1e-13);
* a device-memory budget smaller than the problem set: problems are
  evicted and planned again, requests wait for memory, all finish, results
  still equal the solo ones;
* the cheapest precision per eps: within eps of the FP64 result;
* a problem larger than the budget fails with a clear status.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2

import fgt_amd
from fgt_amd.serve import Server, poisson_arrivals, run_stream

sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload generators)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    probs, _, _ = bench.c4_problems(12, periodic=True, max_leaves=3000)
    solo = []
    for tr, delta, eps in probs:
        s = fgt_amd.Solver.from_tree(tr, delta, eps)
        s.set_fvals(tr.fvals)
        s.execute()
        solo.append(s.get_pot())
        s.close()
    return probs, solo


def _footprint(tr, delta, eps):
    """Upper bound of a problem's device bytes: values, potential and one
    FP64 expansion slot for every box that could hold one."""
    ps = fgt_amd.plan_summary(tr, delta, eps)
    return 2 * 8 * tr.nboxes * tr.npbox + 32 * ps["nexp"] * (ps["nslots"] + 1) + (4 << 20)


def _last_per_key(arr):
    last = {}
    for _, n, key in arr:
        last[key] = n
    return sorted(last.values())


def test_stream_results_equal_solo(pool):
    probs, solo = pool
    print(f"[gpu] serve: {len(probs)} problems ({sum(t.iperiod for t, _, _ in probs)} periodic), FP64 stream",
          flush=True)
    srv = Server(budget=200 << 30, nstreams=4, precision=0)
    arr = poisson_arrivals(400.0, 0.5, len(probs), seed=3)
    keep = [n for _, n, _ in arr]
    lat, wall, rids = run_stream(srv, probs, arr, keep_results=keep)
    st = srv.stats()
    print(f"[gpu]   {len(arr)} requests in {wall:.3f}s, p50 {np.percentile(list(lat.values()), 50) * 1e3:.2f} ms, "
          f"stats {st}", flush=True)
    assert st["completed"] == len(arr) and st["plans"] == len({k for _, _, k in arr})
    for t, n, key in arr:
        assert rel_l2(srv.result(rids[n]), solo[key]) <= 1e-13, (n, key)
    srv.close()


def test_burst_overlaps_on_the_device(pool):
    """Every problem requested at once: with four streams some requests start
    before others finish (their work overlaps), and results equal solo."""
    probs, solo = pool
    srv = Server(budget=200 << 30, nstreams=4, precision=0)
    arr = [(0.0, k, k) for k in range(len(probs))]
    lat, wall, rids = run_stream(srv, probs, arr, keep_results=range(len(probs)))
    inf = [srv.info(rids[k]) for k in range(len(probs))]
    overlap = sum(1 for a in inf for b in inf if a is not b and a["start"] < b["done"] and b["start"] < a["done"])
    print(f"[gpu] serve burst: {len(probs)} requests, overlapping pairs {overlap // 2}", flush=True)
    assert overlap > 0
    for k in range(len(probs)):
        assert rel_l2(srv.result(rids[k]), solo[k]) <= 1e-13, k
    srv.close()


def test_memory_admission_evicts_and_queues(pool):
    probs, solo = pool
    # each problem's device footprint (with one result buffer), measured
    sizes = []
    for k, (tr, delta, eps) in enumerate(probs):
        srv = Server(budget=200 << 30, nstreams=1, precision=0)
        srv.submit(k, tr, delta, eps)
        srv.drain(60.0)
        sizes.append(srv.stats()["used_bytes"])
        srv.close()
    res = [8 * t.nboxes * t.npbox for t, _, _ in probs]        # one result buffer each
    prob = [sz - r for sz, r in zip(sizes, res)]                # the resident problem alone
    # room for the largest problem and 40 % of the others, plus every kept
    # result and the ones in flight: the rest are evicted (planned again on
    # their next request) or wait for memory
    budget = max(prob) + 0.4 * (sum(prob) - max(prob)) + sum(res) + 4 * max(res)
    print(f"[gpu] serve: budget {budget / 2**20:.0f} MiB, problems {sum(prob) / 2**20:.0f} MiB + results "
          f"{sum(res) / 2**20:.0f} MiB", flush=True)
    assert budget < sum(prob) + sum(res) + 4 * max(res)
    srv = Server(budget=budget, nstreams=4, precision=0)
    arr = poisson_arrivals(600.0, 0.4, len(probs), seed=4)
    keep = _last_per_key(arr)
    lat, wall, rids = run_stream(srv, probs, arr, keep_results=keep)
    st = srv.stats()
    print(f"[gpu]   stats {st}", flush=True)
    assert st["completed"] == len(arr)
    assert st["evictions"] > 0 and st["plans"] > len({k for _, _, k in arr})
    key_of = {n: k for _, n, k in arr}
    for n in keep:
        assert rel_l2(srv.result(rids[n]), solo[key_of[n]]) <= 1e-13, n
    srv.close()


def test_cheapest_precision_within_eps(pool):
    probs, solo = pool
    srv = Server(budget=200 << 30, nstreams=4, precision=3)
    arr = [(0.0, k, k) for k in range(len(probs))]
    lat, wall, rids = run_stream(srv, probs, arr, keep_results=range(len(probs)))
    for k, (tr, delta, eps) in enumerate(probs):
        got, ref = srv.result(rids[k])[tr.leaves], solo[k][tr.leaves]
        assert rel_l2(got, ref) <= eps, k
        assert np.abs(got - ref).max() <= eps * np.abs(ref).max(), k
    srv.close()


def test_problem_larger_than_budget_fails_cleanly(pool):
    probs, _ = pool
    tr, delta, eps = max(probs, key=lambda p: 2 * 8 * p[0].nboxes * p[0].npbox)
    srv = Server(budget=float(2 * 8 * tr.nboxes * tr.npbox // 4), nstreams=2, precision=0)
    rid = srv.submit(0, tr, delta, eps)
    srv.drain(30.0)
    assert srv.info(rid)["state"] == "failed"
    with pytest.raises(fgt_amd.FgtError, match="budget"):
        srv.result(rid)
    srv.close()
