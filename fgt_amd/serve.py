"""Serving loop: continuous batching of box-FGT requests on one GPU.

Requests name a problem (a tree with delta and eps, registered with its
source values).  Problems stay resident on the device (plan, tables,
lists, values and expansion workspace) while they fit an HBM budget; least
recently used ones are evicted and planned again on their next request.  Each admitted request is launched on the next free HIP
stream as soon as one frees up, so requests that arrive while others run
overlap with them on the device.  Latency = completion - arrival.
"""
from __future__ import annotations

import collections
import time

import numpy as np

from . import Solver


class Server:
    def __init__(self, nstreams=4, hbm_budget=200 << 30):
        import torch
        self.torch = torch
        self.streams = [torch.cuda.Stream() for _ in range(nstreams)]
        self.busy = [None] * nstreams          # (request id, event) per stream
        self.budget = hbm_budget
        self.resident = collections.OrderedDict()   # problem key -> (solver, bytes)
        self.used = 0
        self.problems = {}                     # key -> (tree, delta, eps)

    def register(self, key, tree, delta, eps):
        self.problems[key] = (tree, delta, eps)

    def _footprint(self, s):
        st = s.stats()
        return st["expansion_bytes"] + 2 * 8 * s.nd * s.npbox * s.nboxes

    def _solver(self, key):
        if key in self.resident:
            self.resident.move_to_end(key)
            return self.resident[key][0]
        tree, delta, eps = self.problems[key]
        s = Solver.from_tree(tree, delta, eps)
        s.set_fvals(tree.fvals)
        nb = self._footprint(s)
        while self.used + nb > self.budget and self.resident:
            old, (os_, ob) = self.resident.popitem(last=False)
            self.torch.cuda.synchronize()
            os_.close()
            self.used -= ob
        self.resident[key] = (s, nb)
        self.used += nb
        return s

    def result(self, key):
        """Potential of the last request served for a resident problem."""
        self.torch.cuda.synchronize()
        return self.resident[key][0].get_pot()

    def run(self, arrivals, horizon):
        """arrivals: list of (time offset s, request id, problem key), sorted.
        Returns per-request latency (s) and the wall time."""
        torch = self.torch
        pending = collections.deque()
        lat = {}
        t0 = time.perf_counter()
        i = 0
        inflight_key = {}
        while i < len(arrivals) or pending or any(b is not None for b in self.busy):
            now = time.perf_counter() - t0
            while i < len(arrivals) and arrivals[i][0] <= now:
                pending.append(arrivals[i])
                i += 1
            for k, b in enumerate(self.busy):
                if b is not None and b[1].query():
                    rid, _, arr = b
                    lat[rid] = time.perf_counter() - t0 - arr
                    inflight_key.pop(rid, None)
                    self.busy[k] = None
            for k in range(len(self.streams)):
                if self.busy[k] is None and pending:
                    arr, rid, key = pending.popleft()
                    if key in inflight_key.values():
                        pending.appendleft((arr, rid, key))   # one request per problem at a time
                        break
                    s = self._solver(key)
                    st = self.streams[k]
                    s.execute(stream=st.cuda_stream)
                    ev = torch.cuda.Event()
                    ev.record(st)
                    self.busy[k] = (rid, ev, arr)
                    inflight_key[rid] = key
            if not pending and i < len(arrivals) and all(b is None for b in self.busy):
                time.sleep(max(0.0, min(arrivals[i][0] - (time.perf_counter() - t0), 1e-3)))
            if time.perf_counter() - t0 > horizon * 4 + 30:
                raise RuntimeError("serving loop did not drain")
        return lat, time.perf_counter() - t0

    def close(self):
        self.torch.cuda.synchronize()
        for s, _ in self.resident.values():
            s.close()
        self.resident.clear()
        self.used = 0


def poisson_arrivals(rate, duration, nkeys, seed=0):
    """Seeded Poisson arrivals over `duration` seconds at `rate` per second,
    each naming one of `nkeys` problems uniformly."""
    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, rid, int(rng.integers(nkeys))))
        rid += 1
