"""Serving: box-FGT requests on one GPU through the library's server
(fgt_server_* in include/fgt.h; the request loop, admission, plan cache and
streams are all in libfgt_amd.so).

A request names its problem by a key and brings its source values.  The
server keeps problems resident under a device-memory budget (plan cache by
key), plans and builds a missing one on the request path when it fits
(evicting idle problems least recently used first) and otherwise queues the
request until running ones finish; up to ``nstreams`` requests run at once,
overlapping on the device.  Results stay on the device until fetched.

``run_stream`` is a client: it submits Poisson arrivals at their times,
polls the server in between and records each request's latency (finish
minus scheduled arrival).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _check, _Problem, _p, lib


class _Info(C.Structure):
    _fields_ = [("budget_bytes", C.c_double), ("used_bytes", C.c_double), ("problems", C.c_longlong),
                ("resident", C.c_longlong), ("requests", C.c_longlong), ("completed", C.c_longlong),
                ("waiting", C.c_longlong), ("running", C.c_longlong), ("plans", C.c_longlong),
                ("evictions", C.c_longlong), ("mem_waits", C.c_longlong)]


STATES = {0: "waiting", 1: "running", 2: "done", 3: "failed"}


class Server:
    def __init__(self, budget=200 << 30, nstreams=4, precision=0):
        self._h = C.c_void_p()
        _check(lib().fgt_server_create(C.c_double(budget), C.c_int(nstreams), C.c_int(precision),
                                       C.byref(self._h)))
        self._keep = {}     # key -> arrays the server points at
        self._fv = {}       # request id -> values (alive until finished)
        self._shape = {}    # request id -> potential shape

    def submit(self, key, tree, delta, eps, fvals=None, version=0, nd=1):
        """Queue one request; returns its id.  fvals defaults to tree.fvals."""
        if key not in self._keep:
            arrs = [np.ascontiguousarray(tree.itree, np.int32), np.ascontiguousarray(tree.iptr, np.int32),
                    np.ascontiguousarray(tree.centers, np.float64),
                    np.ascontiguousarray(tree.boxsize, np.float64)]
            prob = _Problem(nd, tree.ndim, delta, eps, tree.ipoly, tree.iperiod, tree.norder, tree.nboxes,
                            tree.nlevels, tree.ltree, *(a.ctypes.data for a in arrs))
            self._keep[key] = (arrs, prob)
        _, prob = self._keep[key]
        fv = np.ascontiguousarray(tree.fvals if fvals is None else fvals, np.float64)
        rid = C.c_int()
        _check(lib().fgt_server_submit(self._h, C.byref(prob), C.c_longlong(key), _p(fv), C.c_longlong(version),
                                       C.byref(rid)))
        self._fv[rid.value] = fv
        self._shape[rid.value] = fv.shape
        return rid.value

    def poll(self, max_done=4096):
        buf = (C.c_int * max_done)()
        n = C.c_int()
        _check(lib().fgt_server_poll(self._h, buf, C.c_int(max_done), C.byref(n)))
        done = list(buf[:n.value])
        for r in done:
            self._fv.pop(r, None)
        return done

    def drain(self, timeout=120.0):
        _check(lib().fgt_server_drain(self._h, C.c_double(timeout)))
        self._fv.clear()

    def info(self, rid):
        t = (C.c_double * 3)()
        f = C.c_int()
        _check(lib().fgt_server_request_info(self._h, C.c_int(rid), t, C.byref(f)))
        fl = f.value
        return {"submit": t[0], "start": t[1], "done": t[2], "cache_hit": bool(fl & 1), "planned": bool(fl & 2),
                "waited_for_memory": bool(fl & 4), "uploaded": bool(fl & 8), "state": STATES[(fl >> 8) & 3]}

    def result(self, rid):
        out = np.zeros(self._shape[rid])
        _check(lib().fgt_server_result(self._h, C.c_int(rid), _p(out)))
        return out

    def release(self, rid):
        _check(lib().fgt_server_release(self._h, C.c_int(rid)))

    def stats(self):
        s = _Info()
        _check(lib().fgt_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _Info._fields_}

    def close(self):
        if self._h:
            lib().fgt_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rate, duration, nkeys, seed=0):
    """Seeded Poisson arrivals over `duration` seconds at `rate` per second,
    each naming one of `nkeys` problems uniformly: [(t, request no, key)]."""
    rng = np.random.default_rng(seed)
    t, out, n = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append((t, n, int(rng.integers(nkeys))))
        n += 1


def run_stream(server, problems, arrivals, keep_results=(), release=True, timeout=None):
    """Submit each arrival (t, n, key) at its time (problems[key] = (tree,
    delta, eps)), polling in between; returns per-arrival latency (s, finish
    on the server minus scheduled arrival), the wall time from the first
    arrival to the last finish, and {n: request id} for arrivals whose
    results are kept (others are released as they finish)."""
    keep = set(keep_results)
    rid_of, lat = {}, {}
    n_of = {}
    t0 = time.perf_counter()
    off = None     # server clock minus client clock
    horizon = arrivals[-1][0] if arrivals else 0.0
    limit = timeout if timeout is not None else 4 * horizon + 120.0
    i = 0
    while i < len(arrivals) or len(lat) < len(arrivals):
        now = time.perf_counter() - t0
        if now > limit:
            raise RuntimeError(f"stream did not finish within {limit:.0f}s ({len(lat)} of {len(arrivals)} done)")
        while i < len(arrivals) and arrivals[i][0] <= now:
            ta, n, key = arrivals[i]
            tr, delta, eps = problems[key]
            tc = time.perf_counter() - t0
            rid = server.submit(key, tr, delta, eps)
            if off is None:
                off = server.info(rid)["submit"] - tc
            rid_of[n] = rid
            n_of[rid] = (n, ta)
            i += 1
        for rid in server.poll():
            n, ta = n_of[rid]
            inf = server.info(rid)
            if inf["state"] != "done":
                raise RuntimeError(f"request {n} {inf['state']}")
            lat[n] = inf["done"] - off - ta
            if release and n not in keep:
                server.release(rid)
        if i < len(arrivals) and len(lat) == i:
            time.sleep(max(0.0, min(arrivals[i][0] - (time.perf_counter() - t0), 5e-4)))
    wall = max(lat[n] + arrivals[k][0] for k, (_, n, _) in enumerate(arrivals)) - arrivals[0][0] if arrivals else 0.0
    return lat, wall, {n: rid_of[n] for n in keep}
