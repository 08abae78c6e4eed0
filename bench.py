"""Benchmark: box Gauss transform throughput (output nodes/s) on one MI355X.

Workload (default): 3-D periodic box FGT at FP64, Legendre nodes, norder=16,
eps=1e-9, delta=1e-4; source = 20 anisotropic Gaussians with seeded centres
in the unit box, per-axis widths log-uniform in [5e-4, 5e-3], amplitudes
uniform in [-1, 1]; adaptive tree refined to eps by the library's builder
(same arrays as the reference's vol_tree_build).  A "step" is one full
transform (STEPs 1-6) of the resident tree: fvals are in HBM and the
potential is written to HBM inside the timed region; the plan (tables,
lists) is built once before it.

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


def stage_model(st, n, nd=1):
    """Algorithmic FLOPs and HBM bytes of one launch of the single-kernel stages.

    STEP 5 (k_pw2pot3, per PW leaf, 3-D): 4 n^3 nh + 8 n^2 npw nh + 8 n npw^2 nh
    flops (the x/y stages complex x complex, the last stage real output);
    bytes: read one local expansion (16 nexp) + read/write pot (16 n^3).
    STEP 3 (k_m2l27, per cutoff-level box): 8 nexp flops per list entry; compulsory
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
        "m2l": {"kernel": "k_m2l27", "bound": "hbm", "flops": 8 * nd * nexp * st["n_m2l_pairs"],
                "bytes": 2 * 16 * nd * nexp * st["n_m2l_targets"]},
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (lines `traffic_bytes <kernel> = <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", "c2_boxfgt_fp64_hbm_traffic.md")
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


def c4_problems(nprob):
    """nprob problems from the uneven generator; trees above 8000 leaves are
    skipped, keeping leaf counts in the 50-8000 range."""
    import fgt_amd
    from fgt_amd.tree import gauss_source
    t0 = time.time()
    probs, i = [], 0
    while len(probs) < nprob:
        ndim, n, eps, delta, g = c4_problem(i)
        i += 1
        tr = fgt_amd.vol_tree_build(ndim, n, 0, 0, eps, gauss_source(*g))
        if len(tr.leaves) > 8000:
            continue
        probs.append((tr, delta, eps))
    log(f"batch: {nprob} problems ({i} drawn) built in {time.time() - t0:.1f}s")
    return probs


def mixed_sweep(steps):
    """Mixed-precision form on the single-problem generator (3-D periodic,
    20 anisotropic Gaussians, seed 1) at norder 12: eps in {1e-3, 1e-6} x
    delta in {1e-2, 1e-3, 1e-4}.  Each point times the FP64 path and the
    FP32-expansion path on the same tree and states the mixed result's
    relative l2 and max error against the FP64 result (which matches the
    reference to ~1e-15)."""
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
            for mode in (0, 1):
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
                    res["mixed"] = {"ms": round(ms, 3), "nodes_per_s": round(nout / ms * 1e3, 1),
                                    "speedup": round(res["fp64"]["ms"] / ms, 3),
                                    "rel_l2_vs_fp64": float(f"{np.linalg.norm(pot - ref) / den:.3g}"),
                                    "max_err_vs_fp64_rel_max": float(
                                        f"{np.abs(pot - ref).max() / max(np.abs(ref).max(), 1e-300):.3g}")}
            sv.close()
            rows.append(res)
    return rows


def batched_comparison(probs, nstreams, steps):
    """All problems executed overlapped on nstreams streams (fgt_batch) vs
    one at a time (device-resident solver per problem, synchronised after
    each)."""
    import torch
    import fgt_amd
    nprob = len(probs)
    b = fgt_amd.Batch(probs, nstreams=nstreams)
    for k, (tr, _, _) in enumerate(probs):
        b.set_fvals(k, tr.fvals)
    b.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    t_b = (time.perf_counter() - t0) / steps
    nout = b.stats()["nout"]
    ref = [b.get_pot(k) for k in range(nprob)]
    # mixed precision (FP32 expansions) where eps >= 1e-6 and the 3-D path applies
    nmixed = b.set_precision(1, 1e-6)
    b.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    t_m = (time.perf_counter() - t0) / steps
    worst = 0.0
    for k, (tr, delta, eps) in enumerate(probs):
        den = np.linalg.norm(ref[k])
        err = np.linalg.norm(b.get_pot(k) - ref[k]) / (den if den > 0 else 1.0)
        worst = max(worst, err / eps)
    b.close()
    solvers = []
    for tr, delta, eps in probs:
        sv = fgt_amd.Solver.from_tree(tr, delta, eps)
        sv.set_fvals(tr.fvals)
        solvers.append(sv)
    for sv in solvers:
        sv.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for sv in solvers:
            sv.execute()
            torch.cuda.synchronize()
    t_p = (time.perf_counter() - t0) / steps
    for sv in solvers:
        sv.close()
    dims = sum(1 for tr, _, _ in probs if tr.ndim == 2)
    return {"problems": nprob, "two_d": dims, "nstreams": nstreams, "output_nodes": int(nout),
            "leaves_min": int(min(len(t.leaves) for t, _, _ in probs)),
            "leaves_max": int(max(len(t.leaves) for t, _, _ in probs)),
            "ms_batched": round(t_b * 1e3, 3), "ms_one_at_a_time": round(t_p * 1e3, 3),
            "nodes_per_s_batched": round(nout / t_b, 1), "nodes_per_s_one_at_a_time": round(nout / t_p, 1),
            "speedup": round(t_p / t_b, 3), "dtype": "f64",
            "mixed": {"problems_fp32_expansions": nmixed, "ms": round(t_m * 1e3, 3),
                      "nodes_per_s": round(nout / t_m, 1), "speedup_vs_one_at_a_time_fp64": round(t_p / t_m, 3),
                      "max_rel_l2_err_over_eps": float(f"{worst:.3g}")}}


def serving(probs, rate, seconds, nstreams):
    """Poisson arrivals (seeded) at `rate` per second for `seconds`, each
    naming one of the problems; served with continuous admission on
    nstreams streams and, for comparison, one at a time (one stream).
    Problems are made resident before the clock starts."""
    from fgt_amd.serve import Server, poisson_arrivals
    nout = [len(t.leaves) * t.norder ** t.ndim for t, _, _ in probs]
    arr = poisson_arrivals(rate, seconds, len(probs), seed=7)
    out = {"rate_per_s": rate, "seconds": seconds, "requests": len(arr), "problems": len(probs)}
    for label, ns in (("served", nstreams), ("one_at_a_time", 1)):
        srv = Server(nstreams=ns)
        for k, (tr, delta, eps) in enumerate(probs):
            srv.register(k, tr, delta, eps)
            srv._solver(k)
        lat, wall = srv.run(arr, seconds)
        srv.close()
        v = np.array([lat[r] for _, r, _ in arr]) * 1e3
        done = sum(nout[k] for _, _, k in arr)
        out[label] = {"nstreams": ns, "nodes_per_s": round(done / wall, 1), "p50_ms": round(float(np.percentile(v, 50)), 3),
                      "p99_ms": round(float(np.percentile(v, 99)), 3), "wall_s": round(wall, 3)}
    return out


def cpu_baseline(tree, cfg):
    """Reference boxfgt (oracle/_ref, OpenMP) on the same tree and fvals."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        from oracle import refrun
        if not refrun.available():
            raise RuntimeError("reference build missing")
        refrun.lib().refglue_silence()
        kind = "reference"

        class T:
            pass
        rt = T()
        for k in ("ndim", "norder", "ipoly", "iperiod", "nboxes", "nlevels", "ltree", "itree", "iptr",
                  "centers", "boxsize", "fvals"):
            setattr(rt, k, getattr(tree, k))

        def run():
            return refrun.boxfgt(rt, cfg["delta"], cfg["eps"])
    except Exception as e:  # pragma: no cover - fallback when the reference .so is absent
        log(f"cpu baseline: reference unavailable ({e}); timing the C restatement")
        from oracle import oracle
        kind = "port"

        def run():
            return oracle.boxfgt(1, 3, cfg["delta"], cfg["eps"], 0, tree.iperiod, tree.norder, tree.nboxes,
                                 tree.nlevels, tree.itree, tree.iptr, tree.centers, tree.boxsize, tree.fvals)
    # keep the Fortran runtime's stdout off our JSON line
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            run()
            times.append(time.perf_counter() - t0)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    return min(times), kind, threads


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true", help="skip the end-to-end boxfgt_ call timing")
    ap.add_argument("--batch", type=int, default=128,
                    help="problems in the uneven-batch comparison (0 = skip)")
    ap.add_argument("--serve-rates", default="1000,3000,5000", help="request arrivals per second (comma list)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window of each serving run (0 = skip)")
    ap.add_argument("--no-mixed-sweep", action="store_true",
                    help="skip the norder-12 eps x delta mixed-precision sweep")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy 1 GiB on the device 3 times (reference kernel for byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              f"sharding of the box FGT is design-only in this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    import fgt_amd
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.init()

    tree, cfg = make_workload()
    n = cfg["n"]
    t0 = time.time()
    s = fgt_amd.Solver(1, 3, cfg["delta"], cfg["eps"], 0, cfg["iperiod"], n, tree.nboxes, tree.nlevels,
                       tree.ltree, tree.itree, tree.iptr, tree.centers, tree.boxsize)
    s.set_fvals(tree.fvals)
    torch.cuda.synchronize()
    log(f"plan+upload {time.time() - t0:.2f}s; stats {s.stats()}")

    for _ in range(args.warmup):
        s.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.execute()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    st = s.stats()
    nout = st["nout"]
    value = nout / (ms * 1e-3)

    # per-stage device time (HIP events on the launch stream), one more step
    s.execute(timed=True)
    st2 = s.stats()
    stage_ms = st2["stage_ms"]
    names = ["p2pw", "merge", "m2l", "split", "pw2pot", "direct"]
    stage = dict(zip(names, stage_ms[:6]))
    log("stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stage.items()))
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
    roofline = {"kernel": m["kernel"], "bound": m["bound"], "achieved": round(achieved, 3), "peak": peak,
                "unit": unit, "frac": round(achieved / peak, 4), "traffic": measured_traffic(m["kernel"]),
                "algorithmic_bytes": m["bytes"], "algorithmic_flops": m["flops"],
                "stage_ms": {k: round(v, 4) for k, v in stage.items()}, "stages": stages}

    if args.copy_calibration:
        # plain device copy of a known size: the scale for the PMC byte counters
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device copies done")

    # the drop-in as a host caller sees it: one boxfgt_ call with host
    # arrays = plan (host tables and lists) + H2D + STEPs 1-6 + D2H; best of 2
    e2e = None
    if not args.no_drop_in:
        pot = np.zeros_like(tree.fvals)
        best, split = None, None
        for _ in range(2):
            pot[...] = 0.0
            ti = np.zeros(20)
            t0 = time.perf_counter()
            fgt_amd.boxfgt(1, 3, cfg["delta"], cfg["eps"], 0, cfg["iperiod"], n, tree.npbox, tree.nboxes,
                           tree.nlevels, tree.ltree, tree.itree, tree.iptr, tree.centers, tree.boxsize,
                           tree.fvals, 1, pot, timeinfo=ti)
            dt = time.perf_counter() - t0
            if best is None or dt < best:
                # timeinfo: [0] plan + fvals upload, [1..6] device stages, [7] download and +=
                best, split = dt, {"plan_and_upload_ms": round(ti[0] * 1e3, 2),
                                   "download_and_accumulate_ms": round(ti[7] * 1e3, 2)}
        e2e = {"ms": round(best * 1e3, 2), "nodes_per_s": round(nout / best, 1), **split,
               "includes": "plan + H2D + STEPs 1-6 + D2H, one boxfgt_ call with host arrays, best of 2"}
        log(f"drop-in end to end: {best * 1e3:.1f} ms")

    sweep = None
    if not args.no_mixed_sweep:
        sweep = mixed_sweep(max(1, min(args.steps, 5)))
        log(f"mixed sweep: {sweep}")
    batched = served = None
    if args.batch > 0:
        probs = c4_problems(args.batch)
        batched = batched_comparison(probs, 4, max(1, min(args.steps, 5)))
        log(f"batched: {batched}")
        rates = [float(r) for r in args.serve_rates.split(",") if r.strip()]
        if args.serve_seconds > 0 and rates:
            served = [serving(probs, r, args.serve_seconds, 4) for r in rates]
            log(f"served: {served}")

    cpu = None
    if not args.no_cpu_baseline:
        secs, kind, threads = cpu_baseline(tree, cfg)
        cpu = {"value": round(nout / secs, 1), "unit": "output nodes/s", "cores": threads, "kind": kind,
               "sample": f"the full workload ({nout} output nodes), one boxfgt call, best of 2, "
                         f"{secs:.2f}s"}
        log(f"cpu baseline: {secs:.2f}s ({kind}, {threads} threads)")

    out = {
        "metric": "output nodes/sec at eps=1e-6 3D (batched vs plain FP64); p50/p99 latency",
        "value": round(value, 1),
        "unit": "output nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian sources)",
        "config": {"workload": "plain FP64 drop-in, one problem: 3-D periodic boxfgt, Legendre, norder=16, "
                               "eps=1e-9, delta=1e-4, 20 anisotropic Gaussians (seed 1)",
                   "nboxes": tree.nboxes, "nleaves": int(st["nleaves"]), "nlevels": tree.nlevels,
                   "output_nodes": nout, "npwlevel": st["npwlevel"], "npw": st["npw"],
                   "pw_leaves": int(st["npw_leaves"]), "direct_pairs": int(st["ndirect_pairs"])},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "drop_in_end_to_end": e2e,
        "mixed_precision_sweep": sweep,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)
    s.close()


if __name__ == "__main__":
    main()
