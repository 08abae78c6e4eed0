"""Benchmark: box Gauss transform throughput (output nodes/s) on one MI355X.

Headline workload (`value`): the eps = 1e-6, 3-D slice of the first 512
draws of the uneven-batch generator (norder 8/12/16, delta log-uniform in
[1e-5, 1e-1], 1-10 seeded Gaussians, free space, trees refined to eps by the
library's builder -- the same arrays as the reference's vol_tree_build).  A
"step" is one batched, mixed-precision transform (STEPs 1-6) of every
problem with fvals resident in HBM and the potentials written to HBM; the
plans (tables, lists) are built once before the timed region.  Beside it,
in the same run: the FP64 batch, the device-resident FP64 solver one
problem at a time and the plain FP64 boxfgt_ drop-in one problem at a time,
and the error of the mixed result against FP64.

Secondary legs: the single-problem plain FP64 workload (3-D periodic,
norder 16, eps 1e-9, delta 1e-4, 20 anisotropic Gaussians) with the
roofline of its dominant kernel, and the norder-12 eps x delta
mixed-precision sweep.

Prints ONE JSON line on stdout (everything else goes to stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix), dense
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
# dense matrix peak of the contraction precision of each mode (TFLOP/s)
PEAK_MODE_TFLOPS = {0: 78.6, 1: 157.3, 2: 2500.0}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(seed=1, ng=20, n=16, eps=1e-9, delta=1e-4, iperiod=1):
    import fgt_amd
    from fgt_amd.tree import gauss_source
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.5, 0.5, (ng, 3))
    s = np.exp(rng.uniform(np.log(5e-4), np.log(5e-3), (ng, 3)))
    a = rng.uniform(-1.0, 1.0, ng)
    t0 = time.time()
    tree = fgt_amd.vol_tree_build(3, n, 0, iperiod, eps, gauss_source(c, s, a))
    log(f"tree: nboxes={tree.nboxes} nlevels={tree.nlevels} leaves={len(tree.leaves)} "
        f"build {time.time() - t0:.1f}s")
    return tree, dict(n=n, eps=eps, delta=delta, iperiod=iperiod, ng=ng, seed=seed)


# STEP 3 kernel by form: target groups, brick entry streams, separable bricks
M2L_KERNEL = {0: "k_m2l27", 1: "k_m2l_brick", 2: "k_m2l_sep"}


def stage_model(st, n, nd=1):
    """Algorithmic FLOPs and HBM bytes of one launch of the single-kernel stages.

    STEP 5 (k_pw2pot3, per PW leaf, 3-D): 4 n^3 nh + 8 n^2 npw nh + 8 n npw^2 nh
    flops (the x/y stages complex x complex, the last stage real output);
    bytes: read one local expansion (16 nexp) + read/write pot (16 n^3).
    STEP 3 (k_m2l_sep / k_m2l_brick / k_m2l27, per cutoff-level box): 8 nexp flops per list entry; compulsory
    bytes: every cutoff-level mp read once and every loc written once,
    2 x 16 nexp per box (re-reads of a shared source are not algorithmic).
    """
    npw, nexp = st["npw"], st["nexp"]
    nh = (npw + 1) // 2
    L = st["npw_leaves"]
    f15 = 4 * n ** 3 * nh + 8 * n * n * npw * nh + 8 * n * npw * npw * nh
    return {
        "pw2pot": {"kernel": "k_pw2pot3", "bound": "mfma", "flops": L * nd * f15,
                   "bytes": L * nd * (16 * nexp + 16 * n ** 3)},
        "m2l": {"kernel": M2L_KERNEL[st.get("m2l_form", 0)], "bound": "hbm",
                "flops": 8 * nd * nexp * st["n_m2l_pairs"],
                "bytes": 2 * 16 * nd * nexp * st["n_m2l_targets"]},
    }


def measured_traffic(kernel, summary="c2_boxfgt_fp64_hbm_traffic.md"):
    """HBM bytes per launch of `kernel` from a committed PMC summary under
    profiles/ (lines `traffic_bytes <kernel> = <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        with open(path) as f:
            for line in f:
                parts = line.replace("`", "").split()
                if len(parts) == 4 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                    return int(parts[3])
    except OSError:
        pass
    return None


def c4_problem(i, base=1000):
    """Uneven-batch generator: 25% 2-D / 75% 3-D, norder in {8,12,16}, eps in
    {1e-3,1e-6,1e-9}, delta log-uniform in [1e-5,1e-1], 1-10 Gaussians with
    centres U[-0.4,0.4], widths log-U[1e-3,1e-2], amplitudes U[0.5,1.5];
    seed = base + index; free space, Legendre, tree refined to eps."""
    rng = np.random.default_rng(base + i)
    ndim = 2 if rng.uniform() < 0.25 else 3
    n = int(rng.choice([8, 12, 16]))
    eps = float(rng.choice([1e-3, 1e-6, 1e-9]))
    delta = float(np.exp(rng.uniform(np.log(1e-5), np.log(1e-1))))
    ng = int(rng.integers(1, 11))
    c = rng.uniform(-0.4, 0.4, (ng, ndim))
    s = np.exp(rng.uniform(np.log(1e-3), np.log(1e-2), (ng, ndim)))
    a = rng.uniform(0.5, 1.5, ng)
    return ndim, n, eps, delta, (c, s, a)


def c4_problems(nprob, max_leaves=8000, periodic=False, threads=4):
    """nprob problems of the uneven generator within its stated leaf range:
    draws whose tree would exceed max_leaves leaves are replaced by the next
    draw (the generator's narrowest Gaussians at eps = 1e-9 give trees of up
    to ~2e5 leaves).  periodic: every odd draw is periodic (the served
    stream's mix).  Returns (problems, draws, replaced)."""
    import fgt_amd
    from fgt_amd.tree import gauss_source
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        ndim, n, eps, delta, g = c4_problem(i)
        iper = (i & 1) if periodic else 0
        try:
            tr = fgt_amd.vol_tree_build(ndim, n, 0, iper, eps, gauss_source(*g), max_leaves=max_leaves)
        except fgt_amd.TreeTooLarge:
            return None
        return (tr, delta, eps)

    t0 = time.time()
    probs, i = [], 0
    with ThreadPoolExecutor(threads) as ex:
        while len(probs) < nprob:
            chunk = list(ex.map(one, range(i, i + 2 * (nprob - len(probs)))))
            for k, p in enumerate(chunk):
                if p is not None and len(probs) < nprob:
                    probs.append(p)
                    last = i + k
            i += len(chunk)
    draws = last + 1
    log(f"uneven batch: {nprob} problems from {draws} draws ({draws - nprob} over {max_leaves} leaves replaced), "
        f"built in {time.time() - t0:.1f}s")
    return probs, draws, draws - nprob


def headline_problems(ndraw=512):
    """The eps = 1e-6, 3-D slice of the first ndraw draws of the uneven
    generator (norder 8/12/16, delta log-uniform in [1e-5, 1e-1], 1-10
    Gaussians, free space): every such draw, none skipped."""
    import fgt_amd
    from fgt_amd.tree import gauss_source
    t0 = time.time()
    probs = []
    for i in range(ndraw):
        ndim, n, eps, delta, g = c4_problem(i)
        if ndim != 3 or eps != 1e-6:
            continue
        tr = fgt_amd.vol_tree_build(ndim, n, 0, 0, eps, gauss_source(*g))
        probs.append((tr, delta, eps))
    log(f"headline: {len(probs)} problems (eps 1e-6, 3-D, of {ndraw} draws) built in {time.time() - t0:.1f}s")
    return probs


def pot_error(pot, ref):
    den = np.linalg.norm(ref)
    l2 = np.linalg.norm(pot - ref) / (den if den > 0 else 1.0)
    mx = np.abs(pot - ref).max() / max(np.abs(ref).max(), 1e-300)
    return l2, mx


def batch_roofline(probs, prof):
    """Per class and stage: device ms of the profiled execute and the
    algorithmic work of that launch group; roofline of the longest one among
    STEP 1 (k_p2pwF3 / k_p2pw3), STEP 3 (k_m2l27) and STEP 5 (k_pw2pot3).
    STEP 1 and 5 flops: the reference's per-leaf count 4 n^3 nh + 8 n^2 npw nh
    + 8 n npw^2 nh (nh = (npw+1)/2) per plane-wave leaf of every member;
    STEP 3 bytes: every cutoff-level expansion read once and every local
    expansion written once (2 x nexp complex of the storage precision per
    target)."""
    import fgt_amd
    rows, best = [], None
    for c in prof:
        n, prec = c["norder"], c["prec"]
        elem = 16 if prec == 0 else 8
        flops = byts = 0
        for i in c["member_ids"]:
            tr, delta, eps = probs[i]
            ps = fgt_amd.plan_summary(tr, delta, eps)
            npw = ps["npw"]
            nh = (npw + 1) // 2
            flops += ps["npw_leaves"] * (4 * n ** 3 * nh + 8 * n * n * npw * nh + 8 * n * npw * npw * nh)
            byts += 2 * elem * ps["nexp"] * ps["n_m2l_targets"]
        ms = c["stage_ms"]
        rows.append({"norder": n, "prec": prec, "npw_cap": c["npw_cap"], "members": c["members"],
                     "group": c["group"], "xs1": c["xs1"], "xs5": c["xs5"], "m2l_brick": c["m2l_brick"],
                     "m2l_ept": c["m2l_ept"],
                     "stage_ms": [round(v, 4) for v in ms], "step15_gflop": round(flops / 1e9, 3),
                     "step3_gb": round(byts / 1e9, 3)})
        for stg, kern, kind, work, xs in ((4, "k_pw2pot3", "flops", flops, c["xs5"]),
                                          (0, "k_p2pwF3+k_p2pw3", "flops", flops, c["xs1"]),
                                          (2, M2L_KERNEL[c["m2l_brick"]], "bytes", byts, -1)):
            if ms[stg] > 0 and work > 0 and (best is None or ms[stg] > best[0]):
                best = (ms[stg], kern, kind, work, n, prec, c["npw_cap"], xs)
    if best is None:
        return None, rows
    ms, kern, kind, work, n, prec, cap, xs = best
    arith = None
    if kind == "flops":
        # the peak of the arithmetic the launch group runs: its x stage (the
        # bulk of the flops) on the VALU in FP64 (xs 0), on the FP64 matrix
        # cores (xs 1) or on the reduced-precision matrix cores of the mode
        # (xs 2); FP64 vector and matrix peaks are equal on MI355X
        if xs in (0, 1):
            peak, arith = PEAK_FP64_TFLOPS, "FP64 x stage (" + ("VALU" if xs == 0 else "matrix cores") + ")"
        else:
            peak, arith = PEAK_MODE_TFLOPS[prec], f"x stage on the precision-mode-{prec} matrix cores"
        ach, unit, bound = work / (ms * 1e-3) / 1e12, "TFLOP/s", "mfma"
    else:
        ach, peak, unit, bound = work / (ms * 1e-3) / 1e9, PEAK_HBM_GBS, "GB/s", "hbm"
    # HBM bytes of the group per step from the committed PMC summary of
    # the headline batch (null for a group that summary does not cover)
    traffic = measured_traffic(f"{kern}/n{n}/prec{prec}/npw{cap}", "headline_batch_hbm_traffic.md")
    out = {"kernel": kern, "launch_group": f"norder {n}, precision mode {prec}, npw <= {cap}",
           "bound": bound, "achieved": round(ach, 3), "peak": peak, "unit": unit, "frac": round(ach / peak, 4),
           "ms": round(ms, 4), "algorithmic": work, "traffic": traffic}
    if arith:
        out["arithmetic"] = arith
    return out, rows


def headline(probs, steps, warmup, mode, drop_in=True):
    """North-star comparison on one problem set.

    value: the batched mixed-precision form (precision `mode` for every
    problem whose eps allows it; fvals resident in HBM, potential written to
    HBM); baselines in the same run: the plain FP64 boxfgt_ drop-in one
    problem at a time (host arrays: plan + H2D + STEPs 1-6 + D2H per call)
    and the device-resident FP64 solver one problem at a time (synchronised
    after each).  Errors are stated against the FP64 result of each problem.
    """
    import torch
    import fgt_amd
    nprob = len(probs)
    nout = sum(len(t.leaves) * t.npbox for t, _, _ in probs)
    b = fgt_amd.Batch(probs, nstreams=4)
    for k, (tr, _, _) in enumerate(probs):
        b.set_fvals(k, tr.fvals)
    # FP64 batch: reference values and the batched FP64 rate
    for _ in range(max(1, warmup)):
        b.execute()
    t0 = time.perf_counter()
    for _ in range(max(1, min(steps, 5))):
        b.execute()
    t_b64 = (time.perf_counter() - t0) / max(1, min(steps, 5))
    ref = [b.get_pot(k)[tr.leaves] for k, (tr, _, _) in enumerate(probs)]
    nmixed = b.set_precision(mode)
    t0 = time.perf_counter()
    b.autotune(3)
    t_tune = time.perf_counter() - t0
    for _ in range(warmup):
        b.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    torch.cuda.synchronize()
    t_mix = (time.perf_counter() - t0) / steps
    worst_l2 = worst_max = 0.0
    for k, (tr, _, eps) in enumerate(probs):
        l2, mx = pot_error(b.get_pot(k)[tr.leaves], ref[k])
        worst_l2 = max(worst_l2, l2 / eps)
        worst_max = max(worst_max, mx / eps)
    # per-stage device times: the fastest of three profiled executes (one
    # execute can hold a stall of a few ms that is not the kernel's own)
    prof = b.profile()
    for _ in range(2):
        for c, c2 in zip(prof, b.profile()):
            c["stage_ms"] = [min(x, y) for x, y in zip(c["stage_ms"], c2["stage_ms"])]
    roof, stage_tab = batch_roofline(probs, prof)
    b.close()
    # device-resident FP64 solver, one problem at a time
    solvers = []
    for tr, delta, eps in probs:
        sv = fgt_amd.Solver.from_tree(tr, delta, eps)
        sv.set_fvals(tr.fvals)
        solvers.append(sv)
    for sv in solvers:
        sv.execute()
    torch.cuda.synchronize()
    npass = 2
    t0 = time.perf_counter()
    for _ in range(npass):
        for sv in solvers:
            sv.execute()
            torch.cuda.synchronize()
    t_one = (time.perf_counter() - t0) / npass
    for sv in solvers:
        sv.close()
    # the plain FP64 drop-in as a Fortran caller sees it, one problem at a time
    t_drop = None
    if drop_in:
        pots = [np.zeros_like(tr.fvals) for tr, _, _ in probs]
        tr, delta, eps = probs[0]
        fgt_amd.boxfgt(1, 3, delta, eps, 0, 0, tr.norder, tr.npbox, tr.nboxes, tr.nlevels, tr.ltree, tr.itree,
                       tr.iptr, tr.centers, tr.boxsize, tr.fvals, 1, np.zeros_like(tr.fvals))
        t0 = time.perf_counter()
        for k, (tr, delta, eps) in enumerate(probs):
            fgt_amd.boxfgt(1, 3, delta, eps, 0, 0, tr.norder, tr.npbox, tr.nboxes, tr.nlevels, tr.ltree,
                           tr.itree, tr.iptr, tr.centers, tr.boxsize, tr.fvals, 1, pots[k])
        t_drop = time.perf_counter() - t0
        fgt_amd.lib().fgt_dropin_cache_clear()   # free the cached plans' device memory
        for k, (tr, _, _) in enumerate(probs):
            l2, _ = pot_error(pots[k][tr.leaves], ref[k])
            if l2 > 1e-12:
                raise AssertionError(f"drop-in and batch FP64 disagree on problem {k}: {l2:.3e}")
    lv = [len(t.leaves) for t, _, _ in probs]
    out = {
        "problems": nprob, "output_nodes": int(nout), "leaves_min": int(min(lv)), "leaves_max": int(max(lv)),
        "norders": sorted({int(t.norder) for t, _, _ in probs}),
        "ms_batched_mixed": round(t_mix * 1e3, 3), "nodes_per_s_batched_mixed": round(nout / t_mix, 1),
        "problems_mixed": int(nmixed), "precision_mode": mode, "autotune_s": round(t_tune, 2),
        "classes": stage_tab, "roofline": roof,
        "max_rel_l2_err_over_eps": float(f"{worst_l2:.3g}"), "max_err_over_eps": float(f"{worst_max:.3g}"),
        "error_reference": "FP64 batch on the same trees (equal to the reference boxfgt to ~1e-15)",
        "ms_batched_fp64": round(t_b64 * 1e3, 3), "nodes_per_s_batched_fp64": round(nout / t_b64, 1),
        "ms_device_one_at_a_time_fp64": round(t_one * 1e3, 3),
        "nodes_per_s_device_one_at_a_time_fp64": round(nout / t_one, 1),
        "speedup_vs_device_one_at_a_time_fp64": round(t_one / t_mix, 3),
    }
    if t_drop is not None:
        out.update({"ms_drop_in_one_at_a_time_fp64": round(t_drop * 1e3, 3),
                    "nodes_per_s_drop_in_one_at_a_time_fp64": round(nout / t_drop, 1),
                    "speedup_vs_drop_in_one_at_a_time_fp64": round(t_drop / t_mix, 3)})
    return out


def mixed_sweep(steps):
    """Mixed-precision form on the single-problem generator (3-D periodic,
    20 anisotropic Gaussians, seed 1) at norder 12: eps in {1e-3, 1e-6} x
    delta in {1e-2, 1e-3, 1e-4}.  Each point times the FP64 path and the
    cheapest reduced-precision path eps allows (BF16 hi/lo at 1e-3, FP32 at
    1e-6) on the same tree and states the mixed result's relative l2 and max
    error against the FP64 result (which matches the reference to ~1e-15)."""
    import torch
    import fgt_amd
    rows = []
    for eps in (1e-3, 1e-6):
        tree, _ = make_workload(n=12, eps=eps)
        for delta in (1e-2, 1e-3, 1e-4):
            sv = fgt_amd.Solver.from_tree(tree, delta, eps)
            sv.set_fvals(tree.fvals)
            nout = len(tree.leaves) * tree.npbox
            res = {"eps": eps, "delta": delta, "output_nodes": int(nout)}
            # the cheapest reduced precision eps allows: BF16 hi/lo at 1e-3, FP32 at 1e-6
            for mode in (0, 2 if eps >= 1e-3 else 1):
                try:
                    sv.set_precision(mode)
                except fgt_amd.FgtError as ex:
                    res["mixed"] = f"not available: {ex}"
                    break
                sv.execute()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    sv.execute()
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) / steps * 1e3
                pot = sv.get_pot()[tree.leaves]
                if mode == 0:
                    ref = pot
                    res["fp64"] = {"ms": round(ms, 3), "nodes_per_s": round(nout / ms * 1e3, 1)}
                else:
                    den = max(np.linalg.norm(ref), 1e-300)
                    res["mixed"] = {"mode": mode, "ms": round(ms, 3), "nodes_per_s": round(nout / ms * 1e3, 1),
                                    "speedup": round(res["fp64"]["ms"] / ms, 3),
                                    "rel_l2_vs_fp64": float(f"{np.linalg.norm(pot - ref) / den:.3g}"),
                                    "max_err_vs_fp64_rel_max": float(
                                        f"{np.abs(pot - ref).max() / max(np.abs(ref).max(), 1e-300):.3g}")}
            sv.close()
            rows.append(res)
    return rows


def batched_comparison(probs, nstreams, steps):
    """The uneven generator's problems as one batch (shape classes, one launch
    per stage and class, graph replay, autotuned) in FP64 and with the
    cheapest precision per eps, vs the device-resident FP64 solver one
    problem at a time (synchronised after each)."""
    import torch
    import fgt_amd
    nprob = len(probs)
    b = fgt_amd.Batch(probs, nstreams=nstreams)
    for k, (tr, _, _) in enumerate(probs):
        b.set_fvals(k, tr.fvals)
    b.autotune(3)
    b.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    t_b = (time.perf_counter() - t0) / steps
    nout = b.stats()["nout"]
    ref = [b.get_pot(k) for k in range(nprob)]
    nmixed = b.set_precision(3)
    b.autotune(3)
    b.execute()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    t_m = (time.perf_counter() - t0) / steps
    worst = 0.0
    for k, (tr, delta, eps) in enumerate(probs):
        den = np.linalg.norm(ref[k])
        err = np.linalg.norm(b.get_pot(k) - ref[k]) / (den if den > 0 else 1.0)
        worst = max(worst, err / eps)
    ncls = len(b.classes())
    b.close()
    t_p = 0.0
    for tr, delta, eps in probs:
        sv = fgt_amd.Solver.from_tree(tr, delta, eps)
        sv.set_fvals(tr.fvals)
        sv.execute()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sv.execute()
        torch.cuda.synchronize()
        t_p += time.perf_counter() - t0
        sv.close()
    dims = sum(1 for tr, _, _ in probs if tr.ndim == 2)
    return {"problems": nprob, "two_d": dims, "classes": ncls, "output_nodes": int(nout),
            "leaves_min": int(min(len(t.leaves) for t, _, _ in probs)),
            "leaves_max": int(max(len(t.leaves) for t, _, _ in probs)),
            "ms_batched": round(t_b * 1e3, 3), "ms_one_at_a_time": round(t_p * 1e3, 3),
            "nodes_per_s_batched": round(nout / t_b, 1), "nodes_per_s_one_at_a_time": round(nout / t_p, 1),
            "speedup": round(t_p / t_b, 3), "dtype": "f64",
            "mixed": {"problems_reduced_precision": nmixed, "ms": round(t_m * 1e3, 3),
                      "nodes_per_s": round(nout / t_m, 1), "speedup_vs_one_at_a_time_fp64": round(t_p / t_m, 3),
                      "max_rel_l2_err_over_eps": float(f"{worst:.3g}")}}


def serving(pool, rate, seconds, nstreams, mode, budget=200 << 30):
    """Seeded Poisson arrivals at `rate` per second for `seconds`, each
    naming a problem of `pool` (the uneven generator, free space and
    periodic), through the library's server: `nstreams` requests in flight
    with the cheapest precision per eps (mode) vs the plain FP64 path one
    request at a time (one stream, mode 0).  Each server first sees one
    request per problem (planning, building, first upload) outside the
    clock, so the measured stream is served from device-resident problems;
    throughput = output nodes of the stream / (last finish - first arrival),
    latency = finish - scheduled arrival."""
    from fgt_amd.serve import Server, poisson_arrivals, run_stream
    nodes = [len(t.leaves) * t.npbox for t, _, _ in pool]
    arr = poisson_arrivals(rate, seconds, len(pool), seed=7)
    out = {"rate_per_s": rate, "seconds": seconds, "requests": len(arr), "problems": len(pool)}
    for label, ns, m in (("served", nstreams, mode), ("one_at_a_time_fp64", 1, 0)):
        srv = Server(budget=budget, nstreams=ns, precision=m)
        t0 = time.perf_counter()
        run_stream(srv, pool, [(0.0, k, k) for k in range(len(pool))])
        warm = time.perf_counter() - t0
        lat, wall, _ = run_stream(srv, pool, arr)
        st = srv.stats()
        srv.close()
        v = np.array([lat[n] for _, n, _ in arr]) * 1e3
        done = sum(nodes[k] for _, _, k in arr)
        out[label] = {"nstreams": ns, "precision_mode": m, "nodes_per_s": round(done / wall, 1),
                      "p50_ms": round(float(np.percentile(v, 50)), 3), "p99_ms": round(float(np.percentile(v, 99)), 3),
                      "wall_s": round(wall, 3), "warm_up_s": round(warm, 2),
                      "resident_gb": round(st["used_bytes"] / 1e9, 2), "plans": st["plans"],
                      "evictions": st["evictions"]}
    return out


def host_cpus():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_baseline(probs, budget_s=15.0):
    """The reference boxfgt (oracle/_ref, OpenMP FP64) on the GPU box's host,
    over the leading problems of `probs` (same trees and fvals) until about
    budget_s seconds of CPU work; problems whose cutoff level is <= 0 are
    skipped (the reference returns zeros there without computing anything).
    Falls back to the C restatement when the reference build is absent."""
    import fgt_amd
    threads = int(os.environ.get("OMP_NUM_THREADS", host_cpus()))
    from oracle import refrun
    use_ref = refrun.available()
    if use_ref:
        refrun.lib().refglue_silence()
    else:
        log("cpu baseline: reference build missing; timing the C restatement")
        from oracle import oracle
    kind = "reference" if use_ref else "port"

    class T:
        pass
    saved = os.dup(1)
    os.dup2(2, 1)   # keep the Fortran runtime's stdout off the JSON line
    done_nodes, secs, used = 0, 0.0, 0
    try:
        for tr, delta, eps in probs:
            if fgt_amd.plan_summary(tr, delta, eps)["npwlevel"] <= 0:
                continue
            if use_ref:
                rt = T()
                for k in ("ndim", "norder", "ipoly", "iperiod", "nboxes", "nlevels", "ltree", "itree", "iptr",
                          "centers", "boxsize", "fvals"):
                    setattr(rt, k, getattr(tr, k))
                t0 = time.perf_counter()
                refrun.boxfgt(rt, delta, eps)
            else:
                t0 = time.perf_counter()
                oracle.boxfgt(1, tr.ndim, delta, eps, 0, tr.iperiod, tr.norder, tr.nboxes, tr.nlevels, tr.itree,
                              tr.iptr, tr.centers, tr.boxsize, tr.fvals)
            secs += time.perf_counter() - t0
            done_nodes += len(tr.leaves) * tr.npbox
            used += 1
            if secs >= budget_s:
                break
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    return {"value": round(done_nodes / secs, 1), "unit": "output nodes/s", "cores": threads, "kind": kind,
            "host_cpus": host_cpus(),
            "sample": f"the first {used} problems of the headline set with a cutoff level >= 1 "
                      f"({done_nodes} output nodes), one boxfgt call each, {secs:.1f}s; "
                      f"OpenMP threads = OMP_NUM_THREADS"}


def c2_leg(args):
    """Plain FP64 single problem (3-D periodic, norder 16, eps 1e-9,
    delta 1e-4): device-resident step time, per-stage HIP-event times and
    the roofline of its dominant kernel."""
    import torch
    import fgt_amd
    tree, cfg = make_workload()
    n = cfg["n"]
    t0 = time.time()
    s = fgt_amd.Solver(1, 3, cfg["delta"], cfg["eps"], 0, cfg["iperiod"], n, tree.nboxes, tree.nlevels,
                       tree.ltree, tree.itree, tree.iptr, tree.centers, tree.boxsize)
    s.set_fvals(tree.fvals)
    torch.cuda.synchronize()
    log(f"c2 plan+upload {time.time() - t0:.2f}s; stats {s.stats()}")
    for _ in range(args.warmup):
        s.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.execute()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    st = s.stats()
    nout = st["nout"]
    # per-stage device time (HIP events on the launch stream), one more step
    s.execute(timed=True)
    st2 = s.stats()
    names = ["p2pw", "merge", "m2l", "split", "pw2pot", "direct"]
    stage = dict(zip(names, st2["stage_ms"][:6]))
    log("c2 stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stage.items()))
    model = stage_model(st2, n)
    dom = max(model, key=lambda k: stage[k])
    m = model[dom]
    secs = stage[dom] * 1e-3
    if m["bound"] == "mfma":
        achieved, peak, unit = m["flops"] / secs / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
    else:
        achieved, peak, unit = m["bytes"] / secs / 1e9, PEAK_HBM_GBS, "GB/s"
    stages = {}
    for k, v in model.items():
        t = stage[k] * 1e-3
        stages[k] = {"kernel": v["kernel"], "ms": round(stage[k], 4),
                     "tflops": round(v["flops"] / t / 1e12, 3), "gbs": round(v["bytes"] / t / 1e9, 1)}
    roofline = {"kernel": m["kernel"], "workload": "c2_fp64 leg", "bound": m["bound"],
                "achieved": round(achieved, 3), "peak": peak, "unit": unit, "frac": round(achieved / peak, 4),
                "traffic": measured_traffic(m["kernel"]), "algorithmic_bytes": m["bytes"],
                "algorithmic_flops": m["flops"], "stage_ms": {k: round(v, 4) for k, v in stage.items()},
                "stages": stages}
    leg = {"workload": "3-D periodic boxfgt, Legendre, norder=16, eps=1e-9, delta=1e-4, 20 anisotropic "
                       "Gaussians (seed 1), plain FP64, device-resident",
           "ms_per_step": round(ms, 4), "nodes_per_s": round(nout / ms * 1e3, 1), "steps": args.steps,
           "nboxes": tree.nboxes, "nleaves": int(st["nleaves"]), "nlevels": tree.nlevels, "output_nodes": nout,
           "npwlevel": st["npwlevel"], "npw": st["npw"], "pw_leaves": int(st["npw_leaves"]),
           "direct_pairs": int(st["ndirect_pairs"])}
    if not args.no_drop_in:
        # the first boxfgt_ call on this tree plans and uploads; later calls on
        # the same tree arrays hit the library's plan cache and move only the
        # values and the potential
        fgt_amd.lib().fgt_dropin_cache_clear()
        pot = np.zeros_like(tree.fvals)
        calls = []
        for _ in range(4):
            pot[...] = 0.0
            ti = np.zeros(20)
            t0 = time.perf_counter()
            fgt_amd.boxfgt(1, 3, cfg["delta"], cfg["eps"], 0, cfg["iperiod"], n, tree.npbox, tree.nboxes,
                           tree.nlevels, tree.ltree, tree.itree, tree.iptr, tree.centers, tree.boxsize,
                           tree.fvals, 1, pot, timeinfo=ti)
            calls.append((time.perf_counter() - t0, ti[0], ti[7]))
        first, rep = calls[0], min(calls[1:])
        leg["drop_in_end_to_end"] = {
            "first_call_ms": round(first[0] * 1e3, 2), "plan_and_upload_ms": round(first[1] * 1e3, 2),
            "repeated_call_ms": round(rep[0] * 1e3, 2), "repeated_upload_ms": round(rep[1] * 1e3, 2),
            "download_and_accumulate_ms": round(rep[2] * 1e3, 2),
            "nodes_per_s_first": round(nout / first[0], 1), "nodes_per_s_repeated": round(nout / rep[0], 1),
            "includes": "one boxfgt_ call: plan (first call only) + H2D + STEPs 1-6 + D2H; repeated = best of "
                        "the 3 calls after the first on the same tree (plan cache hit)"}
        fgt_amd.lib().fgt_dropin_cache_clear()
        log(f"c2 drop-in end to end: first {first[0] * 1e3:.1f} ms, repeated {rep[0] * 1e3:.1f} ms")
    if args.copy_calibration:
        # plain device copy of a known size: the scale for the PMC byte counters
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b2 = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b2.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device copies done")
    s.close()
    return leg, roofline


def copy_bandwidth(reps=10):
    """Device-to-device copy of 1 GiB (torch copy_, HIP events around the
    timed copies): the practical HBM ceiling, counting bytes read + written."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    moved = 2 * a.numel() * 8
    del a, b
    return {"bytes_per_copy": moved, "ms": round(ms, 4), "gbs": round(moved / ms / 1e6, 1),
            "frac_of_peak": round(moved / ms / 1e6 / PEAK_HBM_GBS, 3),
            "what": "1 GiB device-to-device copy (read + write bytes), mean of 10"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", type=int, default=1, help="precision mode of the headline batch (1 FP32, 2 BF16 hi/lo)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true", help="skip the boxfgt_ drop-in timings")
    ap.add_argument("--no-headline", action="store_true", help="skip the headline batch (profiling runs)")
    ap.add_argument("--no-c2", action="store_true", help="skip the single-problem FP64 leg")
    ap.add_argument("--draws", type=int, default=512, help="generator draws the headline slice is taken from")
    ap.add_argument("--batch", type=int, default=0,
                    help="problems of the whole uneven generator for a batched-vs-one-at-a-time leg (0 = skip)")
    ap.add_argument("--serve-rates", default="100,1000,3000,5000", help="request arrivals per second (comma list)")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="arrival window of each serving run (0 = skip)")
    ap.add_argument("--serve-pool", type=int, default=48, help="problems the served requests are drawn from")
    ap.add_argument("--no-mixed-sweep", action="store_true",
                    help="skip the norder-12 eps x delta mixed-precision sweep")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the C2 leg, copy 1 GiB on the device 3 times (reference kernel for byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              f"sharding of the box FGT is design-only in this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    import fgt_amd  # noqa: F401
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.init()

    head = probs = None
    if not args.no_headline:
        probs = headline_problems(args.draws)
        head = headline(probs, args.steps, args.warmup, args.mode, drop_in=not args.no_drop_in)
        log(f"headline: {head}")
    c2 = roofline = None
    if not args.no_c2:
        c2, roofline = c2_leg(args)
        c2["roofline"] = roofline
        log(f"c2: {c2}")
    sweep = None
    if not args.no_mixed_sweep:
        sweep = mixed_sweep(max(1, min(args.steps, 5)))
        log(f"mixed sweep: {sweep}")
    batched = served = None
    if args.batch > 0:
        allp, _, _ = c4_problems(args.batch)
        batched = batched_comparison(allp, 4, max(1, min(args.steps, 5)))
        log(f"batched: {batched}")
    rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
    if args.serve_seconds > 0 and rates and args.serve_pool > 0:
        pool, draws, replaced = c4_problems(args.serve_pool, periodic=True)
        served = {"pool": {"problems": len(pool), "periodic": int(sum(t.iperiod for t, _, _ in pool)),
                           "output_nodes": int(sum(len(t.leaves) * t.npbox for t, _, _ in pool)),
                           "draws": draws, "replaced_over_8000_leaves": replaced},
                  "budget_gb": round((200 << 30) / 1e9, 1),
                  "rates": [serving(pool, r, args.serve_seconds, 4, 3) for r in rates]}
        log(f"served: {served}")
    hbm_copy = copy_bandwidth()
    log(f"hbm copy: {hbm_copy}")
    cpu = None
    if not args.no_cpu_baseline and probs is not None:
        cpu = cpu_baseline(probs)
        log(f"cpu baseline: {cpu}")

    if head is not None:
        value, ms, dtype = head["nodes_per_s_batched_mixed"], head["ms_batched_mixed"], "f32"
        if head.get("roofline"):
            roofline = dict(head["roofline"], workload="headline batch")
        workload = (f"{head['problems']} problems: the eps=1e-6, 3-D slice of the first {args.draws} draws of the "
                    f"uneven generator (norder 8/12/16, delta log-uniform in [1e-5,1e-1], 1-10 Gaussians, free "
                    f"space), batched, mixed precision (FP32 contractions and expansions, FP64 phases, weights "
                    f"and output accumulation)")
        config = {"workload": workload, "problems": head["problems"], "output_nodes": head["output_nodes"],
                  "precision_mode": args.mode}
    else:
        value, ms, dtype = c2["nodes_per_s"], c2["ms_per_step"], "f64"
        config = {"workload": c2["workload"]}
    out = {
        "metric": "output nodes/sec at eps=1e-6 3D (batched vs plain FP64); p50/p99 latency",
        "value": value,
        "unit": "output nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded Gaussian sources)",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "headline": head,
        "c2_fp64": c2,
        "mixed_precision_sweep": sweep,
        "batched": batched,
        "served": served,
        "hbm_copy": hbm_copy,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
