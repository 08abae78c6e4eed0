"""Writes the golden fixtures in tests/golden/ from the reference build.

Run in the build container only (it needs /root/reference, compiled by
``make -C oracle ref`` into oracle/_ref/libfgtref.so):

    python tests/golden/make_golden.py

Fixtures are data only: tree arrays, inputs and the reference's outputs.
  * boxfgt_<name>.npz : a tree from the reference's vol_tree_mem/vol_tree_build
    (src/common/tree_vol_coeffs_nd.f:83,444), its fvals, and the reference
    boxfgt potential (src/bfgt/boxfgt.f:5) for the stated (eps, delta).
  * boxfgt_pgh*/pg*.npz : the same with ifpgh = 2/3 (gradients, Hessians).
  * boxfgt_tgt*.npz : with extra targets (ifpghtarg = 1..3): targets and the
    reference's pote / grade / hesse.
  * tree_<name>.npz : a tree from vol_tree_mem / vol_tree_build for the
    compiled source function of tests/csrc/sources.c, for the refinement
    norms (iptype), box-size powers (eta), nd, zk and ifnewtree stated in
    TREE_FUN_CASES.
  * tables.npz : the reference's 1-D tables and scalars (fgtterms.f,
    legeexps.f, chebexps.f, besseljs3d.f, bfgt_volrouts.f) for a grid of
    parameters.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402

# name: (ndim, norder, ipoly, iperiod, tree_eps, eps, delta, seed, ng, width range, nd)
CASES = {
    "c1_small":      (3, 8, 0, 0, 1e-2, 1e-6, 1e-3, 0, 3, (1e-3, 1e-2), 1),
    "p3_n8":         (3, 8, 0, 1, 1e-2, 1e-9, 1e-4, 1, 4, (5e-4, 5e-3), 1),
    "p3_n16":        (3, 16, 0, 1, 1e-1, 1e-9, 1e-4, 2, 2, (2e-3, 5e-3), 1),
    "cheb3_n6":      (3, 6, 1, 0, 1e-2, 1e-6, 1e-4, 3, 3, (1e-3, 1e-2), 1),
    "c3_n12_e3_d2":  (3, 12, 0, 0, 1e-1, 1e-3, 1e-2, 0, 3, (1e-3, 1e-2), 1),
    "c3_n12_e6_d4":  (3, 12, 0, 0, 1e-1, 1e-6, 1e-4, 0, 3, (1e-3, 1e-2), 1),
    "f2_n8":         (2, 8, 0, 0, 1e-6, 1e-9, 1e-4, 4, 5, (1e-4, 1e-2), 1),
    "p2_cheb_n12":   (2, 12, 1, 1, 1e-6, 1e-6, 1e-3, 5, 4, (1e-3, 1e-2), 1),
    "f1_n8":         (1, 8, 0, 0, 1e-9, 1e-6, 1e-3, 6, 3, (1e-4, 1e-2), 1),
    "nd2_f3_n8":     (3, 8, 0, 0, 1e-2, 1e-6, 1e-3, 7, 3, (1e-3, 1e-2), 2),
    # npwlevel <= 0 in free space and the periodic npwlevel <= 1 branch:
    # the reference returns a zero potential here (it forms no leaf PW)
    "d1_free2":      (2, 8, 0, 0, 1e-6, 1e-6, 1e-1, 8, 3, (1e-3, 1e-2), 1),
    "d1_per2":       (2, 8, 0, 1, 1e-6, 1e-3, 3e-2, 9, 3, (1e-3, 1e-2), 1),
}


# gradients and Hessians (ifpgh = 2 / 3), small trees:
# name: (ndim, norder, ipoly, iperiod, tree_eps, eps, delta, seed, ng, width range, nd, ifpgh)
PGH_CASES = {
    "pgh3_n6":       (3, 6, 0, 0, 1e-1, 1e-6, 1e-2, 10, 2, (2e-3, 1e-2), 1, 3),
    "pgh3_per_n8":   (3, 8, 0, 1, 1e-1, 1e-9, 1e-3, 11, 2, (2e-3, 5e-3), 1, 3),
    "pg3_cheb_nd2":  (3, 6, 1, 0, 1e-1, 1e-6, 3e-3, 12, 2, (2e-3, 1e-2), 2, 2),
    "pgh2_per_n8":   (2, 8, 0, 1, 1e-6, 1e-6, 1e-3, 13, 4, (1e-3, 1e-2), 1, 3),
    "pgh2_n12":      (2, 12, 0, 0, 1e-6, 1e-9, 1e-4, 14, 3, (1e-3, 1e-2), 1, 3),
    "pgh1_n8":       (1, 8, 0, 0, 1e-9, 1e-6, 1e-3, 15, 3, (1e-4, 1e-2), 1, 3),
    "pgh2_cheb_nd2_n16": (2, 16, 1, 1, 1e-6, 1e-9, 1e-3, 20, 3, (1e-3, 1e-2), 2, 3),
}


# extra targets (ifpghtarg >= 1):
# name: (ndim, norder, ipoly, iperiod, tree_eps, eps, delta, seed, ng, width range, nd, ifpgh, ifpghtarg, ntarg)
TARG_CASES = {
    "tgt3_n8":       (3, 8, 0, 0, 1e-2, 1e-6, 1e-3, 16, 3, (1e-3, 1e-2), 1, 1, 1, 400),
    "tgt3_pgh_n6":   (3, 6, 0, 1, 1e-1, 1e-6, 1e-3, 17, 2, (2e-3, 1e-2), 1, 3, 3, 300),
    "tgt2_up_n8":    (2, 8, 1, 0, 1e-6, 1e-6, 1e-3, 18, 3, (1e-3, 1e-2), 2, 1, 3, 500),
    "tgt1_n8":       (1, 8, 0, 1, 1e-9, 1e-6, 1e-3, 19, 3, (1e-4, 1e-2), 1, 2, 2, 200),
}


def make_targets(ndim, nt, seed, tr):
    """Uniform points in the root box, plus leaf centres and points on box faces."""
    rng = np.random.default_rng(seed + 1000)
    t = rng.uniform(-0.5, 0.5, (nt, ndim))
    nch = tr.itree[tr.iptr[3] - 1: tr.iptr[3] - 1 + tr.nboxes]
    leaves = np.nonzero(nch == 0)[0]
    k = min(len(leaves), nt // 8)
    t[:k] = tr.centers[leaves[:k]]                       # leaf centres
    t[k:2 * k] = np.round(t[k:2 * k] * 8) / 8            # on box faces / corners
    return t


def make_source(ndim, seed, ng, wr, iperiod):
    rng = np.random.default_rng(seed)
    lo, hi = (-0.5, 0.5) if iperiod else (-0.4, 0.4)
    c = rng.uniform(lo, hi, (ng, ndim))
    s = np.exp(rng.uniform(np.log(wr[0]), np.log(wr[1]), (ng, ndim)))
    a = rng.uniform(-1.0, 1.0, ng) if iperiod else rng.uniform(0.5, 1.5, ng)
    return c, s, a


def boxfgt_case(name, spec):
    ndim, n, ipoly, iper, teps, eps, delta, seed, ng, wr, nd = spec
    c, s, a = make_source(ndim, seed, ng, wr, iper)
    dp = refrun.gauss_params(c, s, a)
    tr = refrun.build_tree(ndim, n, ipoly, iper, teps, dp, ng)
    f = tr.fvals
    if nd == 2:
        # second density: a smooth function of the first, values at the same nodes
        f = np.stack([tr.fvals, np.cos(3.0 * tr.fvals)], axis=-1)
    pot, _ = refrun.boxfgt(tr, delta, eps, nd=nd, fvals=f)
    out = os.path.join(HERE, f"boxfgt_{name}.npz")
    # boxfgt reads fvals only at leaves and writes pot only there: store leaves
    nch = tr.itree[tr.iptr[3] - 1: tr.iptr[3] - 1 + tr.nboxes]
    leaves = np.nonzero(nch == 0)[0].astype(np.int32)
    assert np.all(pot[nch != 0] == 0)
    np.savez_compressed(out, ndim=ndim, norder=n, ipoly=ipoly, iperiod=iper, nboxes=tr.nboxes,
                        nlevels=tr.nlevels, ltree=tr.ltree, itree=tr.itree, iptr=tr.iptr,
                        centers=tr.centers, boxsize=tr.boxsize, leaves=leaves,
                        fvals_leaf=f[leaves], pot_leaf=pot[leaves], eps=eps,
                        delta=delta, nd=nd, tree_eps=teps, gauss_c=c, gauss_s=s, gauss_a=a)
    print(f"{name}: nboxes={tr.nboxes} nlevels={tr.nlevels} |pot|={np.abs(pot).max():.3e} "
          f"size={os.path.getsize(out)/1e6:.2f} MB")


def pgh_case(name, spec):
    ndim, n, ipoly, iper, teps, eps, delta, seed, ng, wr, nd, ifpgh = spec
    c, s, a = make_source(ndim, seed, ng, wr, iper)
    dp = refrun.gauss_params(c, s, a)
    tr = refrun.build_tree(ndim, n, ipoly, iper, teps, dp, ng)
    f = tr.fvals
    if nd == 2:
        f = np.stack([tr.fvals, np.cos(3.0 * tr.fvals)], axis=-1)
    pot, grad, hess, _ = refrun.boxfgt(tr, delta, eps, nd=nd, fvals=f, ifpgh=ifpgh)
    out = os.path.join(HERE, f"boxfgt_{name}.npz")
    nch = tr.itree[tr.iptr[3] - 1: tr.iptr[3] - 1 + tr.nboxes]
    leaves = np.nonzero(nch == 0)[0].astype(np.int32)
    extra = {} if hess is None else {"hess_leaf": hess[leaves]}
    np.savez_compressed(out, ndim=ndim, norder=n, ipoly=ipoly, iperiod=iper, nboxes=tr.nboxes,
                        nlevels=tr.nlevels, ltree=tr.ltree, itree=tr.itree, iptr=tr.iptr,
                        centers=tr.centers, boxsize=tr.boxsize, leaves=leaves,
                        fvals_leaf=f[leaves], pot_leaf=pot[leaves], grad_leaf=grad[leaves],
                        ifpgh=ifpgh, eps=eps, delta=delta, nd=nd, tree_eps=teps, gauss_c=c, gauss_s=s,
                        gauss_a=a, **extra)
    print(f"{name}: nboxes={tr.nboxes} nlevels={tr.nlevels} leaves={len(leaves)} "
          f"|grad|={np.abs(grad).max():.3e} size={os.path.getsize(out)/1e6:.2f} MB")


def targ_case(name, spec):
    ndim, n, ipoly, iper, teps, eps, delta, seed, ng, wr, nd, ifpgh, ifpghtarg, nt = spec
    c, s, a = make_source(ndim, seed, ng, wr, iper)
    dp = refrun.gauss_params(c, s, a)
    tr = refrun.build_tree(ndim, n, ipoly, iper, teps, dp, ng)
    f = tr.fvals
    if nd == 2:
        f = np.stack([tr.fvals, np.cos(3.0 * tr.fvals)], axis=-1)
    targs = make_targets(ndim, nt, seed, tr)
    pot, grad, hess, pote, grade, hesse, _ = refrun.boxfgt(tr, delta, eps, nd=nd, fvals=f, ifpgh=ifpgh,
                                                           targs=targs, ifpghtarg=ifpghtarg)
    out = os.path.join(HERE, f"boxfgt_{name}.npz")
    nch = tr.itree[tr.iptr[3] - 1: tr.iptr[3] - 1 + tr.nboxes]
    leaves = np.nonzero(nch == 0)[0].astype(np.int32)
    # grad/hess on the leaves are covered by the pgh fixtures; keep the file small
    extra = {"pote": pote, "targs": targs, "ifpghtarg": ifpghtarg}
    if grade is not None:
        extra["grade"] = grade
    if hesse is not None:
        extra["hesse"] = hesse
    np.savez_compressed(out, ndim=ndim, norder=n, ipoly=ipoly, iperiod=iper, nboxes=tr.nboxes,
                        nlevels=tr.nlevels, ltree=tr.ltree, itree=tr.itree, iptr=tr.iptr,
                        centers=tr.centers, boxsize=tr.boxsize, leaves=leaves,
                        fvals_leaf=f[leaves], pot_leaf=pot[leaves], ifpgh=ifpgh, eps=eps, delta=delta, nd=nd,
                        tree_eps=teps, gauss_c=c, gauss_s=s, gauss_a=a, **extra)
    print(f"{name}: nboxes={tr.nboxes} nlevels={tr.nlevels} nt={nt} |pote|={np.abs(pote).max():.3e} "
          f"size={os.path.getsize(out)/1e6:.2f} MB")


def _dp(x):
    return C.byref(C.c_double(x))


def _ip(x):
    return C.byref(C.c_int(x))


def tables():
    L = refrun.lib()
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = {}
    # fgtpwterms / pw nodes
    rows = []
    for bsize in (0.125, 0.25, 0.5, 1.0, 2.0):
        for delta in (1e-5, 1e-4, 1e-3, 1e-2, 1e-1):
            for eps in (1e-3, 1e-6, 1e-9, 1e-12):
                for iper in (0, 1):
                    pmax, npw = C.c_double(0.0), C.c_int(0)
                    L.fgtpwterms_(_dp(bsize), _dp(delta), _dp(eps), _ip(iper), C.byref(pmax), C.byref(npw))
                    rows.append((bsize, delta, eps, iper, pmax.value if iper == 0 else 0.0, npw.value))
    out["pwterms"] = np.array(rows)
    # nodes/weights/u for both bases
    for ipoly, fn in ((0, L.legeexps_), (1, L.chebexps_)):
        for n in (4, 6, 8, 12, 16):
            x = np.zeros(n); w = np.zeros(n); u = np.zeros((n, n)); v = np.zeros((n, n))
            fn(_ip(2), _ip(n), P(x), P(u), P(v), P(w))
            out[f"exps_{ipoly}_{n}_x"] = x
            out[f"exps_{ipoly}_{n}_w"] = w
            out[f"exps_{ipoly}_{n}_u"] = u   # numpy [j, i] == Fortran u(i, j)
    # spherical Bessel
    zs = np.array([1e-3, 0.1, 0.5, 1.0, 3.0, 7.5, 12.0, 25.0, 40.0])
    bes = np.zeros((len(zs), 17))
    for i, z in enumerate(zs):
        zc = np.array([z, 0.0])
        fjs = np.zeros(2 * 17)
        fjd = np.zeros(2 * 17)
        L.besseljs3d_(_ip(16), P(zc), _dp(1.0), P(fjs), _ip(0), P(fjd))
        bes[i] = fjs[0::2]
    out["bessel_z"] = zs
    out["bessel_j"] = bes
    # PW tables and near-field tables
    for ipoly in (0, 1):
        for (n, delta, eps, boxdim, iper) in ((8, 1e-3, 1e-6, 0.25, 0), (12, 1e-4, 1e-9, 0.0625, 0),
                                               (16, 1e-4, 1e-9, 0.0625, 0), (8, 1.0, 1e-6, 0.5, 1)):
            key = f"{ipoly}_{n}_{delta:g}_{eps:g}_{boxdim:g}_{iper}"
            if iper == 0:
                pmax, npw = C.c_double(0.0), C.c_int(0)
                L.fgtpwterms_(_dp(2 * boxdim), _dp(delta), _dp(eps), _ip(0), C.byref(pmax), C.byref(npw))
                npw = npw.value
                ws = np.zeros(npw); ts = np.zeros(npw)
                L.get_pwnodes_(_ip(0), C.byref(pmax), _ip(npw), P(ws), P(ts))
                out[f"pmax_{key}"] = np.array(pmax.value)
            else:
                npw = 2 * int(1.0 * np.sqrt(np.log(1 / eps) / 1e-3) / np.pi + 1) + 1
                ws = np.zeros(npw); ts = np.zeros(npw)
                L.get_periodic_pwnodes_(_dp(1.0), _dp(1e-3), _dp(eps), _ip(npw), P(ws), P(ts))
            t1 = np.zeros(2 * n * npw)
            L.mk_poly2pw_tables_(_ip(n), _ip(ipoly), _ip(npw), _ip(100), P(ws), P(ts), _dp(delta),
                                 _dp(boxdim), P(t1))
            x = np.zeros(n); w = np.zeros(n); u = np.zeros((n, n)); v = np.zeros((n, n))
            (L.legeexps_ if ipoly == 0 else L.chebexps_)(_ip(2), _ip(n), P(x), P(u), P(v), P(w))
            t2 = np.zeros(2 * n * npw); t2x = np.zeros(2 * n * npw); t2xx = np.zeros(2 * n * npw)
            L.mk_pw2pgh_tables_(_ip(n), _ip(npw), P(ts), P(x), _dp(delta), _dp(boxdim), P(t2), P(t2x),
                                P(t2xx))
            out[f"ws_{key}"] = ws
            out[f"ts_{key}"] = ts
            out[f"p2pw_{key}"] = t1
            out[f"pw2pot_{key}"] = t2
            out[f"pw2der_{key}"] = t2x
            out[f"pw2dxx_{key}"] = t2xx
            if iper == 0:
                nt = 6
                tl = np.zeros(n * n * 13); tlx = np.zeros(n * n * 13); tlxx = np.zeros(n * n * 13)
                ind = np.zeros(2 * (n + 1) * 13, np.int32)
                L.mk_loctab_all_(_dp(eps), _ip(ipoly), _ip(n), _ip(50), _dp(delta), _dp(boxdim), _ip(0),
                                 _ip(nt), P(tl), P(tlx), P(tlxx), P(ind))
                out[f"loctab_{key}"] = tl
                out[f"loctabx_{key}"] = tlx
                out[f"loctabxx_{key}"] = tlxx
                out[f"locind_{key}"] = ind
    path = os.path.join(HERE, "tables.npz")
    np.savez_compressed(path, **out)
    print(f"tables: {len(out)} arrays, {os.path.getsize(path)/1e6:.2f} MB")


# boxfgt with ifnewtree = 1 (refine / coarsen / 2:1 rebalance after the
# solve).  Generate these with OMP_NUM_THREADS=1: the reference's threaded
# 2:1 rebalance shares one coefficient scratch array between threads
# (vol_tree_refine_boxes_flag_interp, tree_vol_coeffs_nd.f:3129-3227, gcoefs /
# hcoefs are not private), so with several threads the gradient and Hessian
# of the boxes it adds depend on thread timing.  Spec: (ndim, n, ipoly, iperiod, tree_eps, eps, delta, seed, ng, width
# range, nd, ifpgh, ifpghtarg, ntarg).  Trees are built with room for three
# times their boxes, as vol_tree_mem does for ifnewtree = 1; tree_eps above
# eps leaves the potential under-resolved so that boxes get refined.
NEWTREE_CASES = {
    "nt2_coarsen":   (2, 8, 0, 0, 1e-3, 1e-6, 1e-2, 7, 3, (1e-3, 1e-2), 1, 1, 0, 0),
    "nt2_per":       (2, 8, 0, 1, 1e-2, 1e-4, 1e-3, 7, 3, (1e-3, 1e-2), 1, 1, 0, 0),
    "nt2_gen2_n6":   (2, 6, 0, 0, 1e-6, 1e-9, 1e-3, 7, 3, (1e-3, 1e-2), 1, 1, 0, 0),
    "nt3_free":      (3, 8, 0, 0, 1e-2, 1e-4, 1e-3, 7, 3, (1e-3, 1e-2), 1, 1, 0, 0),
    "nt3_per":       (3, 8, 0, 1, 1e-2, 1e-4, 1e-3, 7, 3, (1e-3, 1e-2), 1, 1, 0, 0),
    "nt1_per":       (1, 8, 0, 1, 1e-6, 1e-9, 1e-3, 7, 3, (1e-4, 1e-3), 1, 1, 0, 0),
    "nt2_pgh3":      (2, 8, 0, 1, 1e-2, 1e-4, 1e-3, 7, 3, (1e-3, 1e-2), 1, 3, 0, 0),
    "nt2_nd2":       (2, 8, 0, 0, 1e-3, 1e-5, 1e-3, 7, 3, (1e-3, 1e-2), 2, 1, 0, 0),
    "nt2_cheb":      (2, 6, 1, 1, 1e-2, 1e-4, 1e-3, 7, 3, (1e-3, 1e-2), 1, 2, 0, 0),
    "nt2_tgt":       (2, 12, 0, 1, 1e-3, 1e-6, 1e-3, 7, 3, (1e-3, 1e-2), 1, 1, 2, 300),
}


def newtree_case(name, spec):
    if os.environ.get("OMP_NUM_THREADS") != "1":
        raise SystemExit("run the ifnewtree fixtures with OMP_NUM_THREADS=1")
    ndim, n, ipoly, iper, teps, eps, delta, seed, ng, wr, nd, ifpgh, ifpghtarg, nt = spec
    c, s, a = make_source(ndim, seed, ng, wr, iper)
    dp = refrun.gauss_params(c, s, a)
    tr = refrun.build_tree(ndim, n, ipoly, iper, teps, dp, ng, ifnewtree=1)
    nb0 = int(tr.itree[2 * tr.nlevels + 1])
    f = tr.fvals
    if nd == 2:
        f = np.stack([tr.fvals, np.cos(3.0 * tr.fvals)], axis=-1)
    nch = tr.itree[tr.iptr[3] - 1: tr.iptr[3] - 1 + nb0]
    leaves = np.nonzero(nch == 0)[0].astype(np.int32)
    targs = None
    if nt:
        rng = np.random.default_rng(seed + 2000)
        targs = rng.uniform(-0.5, 0.5, (nt, ndim))
    out = {}
    res = refrun.boxfgt(tr, delta, eps, nd=nd, fvals=f, ifpgh=ifpgh, targs=targs, ifpghtarg=ifpghtarg,
                        ifnewtree=1, out_tree=out)
    nl1 = out["nlevels"]
    nb1 = int(out["itree"][2 * nl1 + 1])
    extra = {}
    if nt:
        pot, grad, hess, pote, grade, hesse, _ = res
        extra.update(targs=targs, pote=pote)
        if grade is not None:
            extra["grade"] = grade
        if hesse is not None:
            extra["hesse"] = hesse
    elif ifpgh == 1:
        pot, _ = res
        grad = hess = None
    else:
        pot, grad, hess, _ = res
    if ifpgh >= 2:
        extra["grad_out"] = grad[:nb1]
    if ifpgh >= 3:
        extra["hess_out"] = hess[:nb1]
    path = os.path.join(HERE, f"boxfgt_{name}.npz")
    np.savez_compressed(path, ndim=ndim, norder=n, ipoly=ipoly, iperiod=iper, nboxes=tr.nboxes,
                        nlevels=tr.nlevels, ltree=tr.ltree, itree=tr.itree, iptr=tr.iptr,
                        centers=tr.centers[:nb0], boxsize=tr.boxsize, leaves=leaves, fvals_leaf=f[leaves],
                        eps=eps, delta=delta, nd=nd, ifpgh=ifpgh, ifpghtarg=ifpghtarg, tree_eps=teps,
                        gauss_c=c, gauss_s=s, gauss_a=a, ifnewtree=1,
                        itree_out=out["itree"], centers_out=out["centers"][:nb1], nlevels_out=nl1,
                        pot_out=pot[:nb1], **extra)
    print(f"{name}: capacity={tr.nboxes} boxes {nb0} -> {nb1}, nlevels {tr.nlevels} -> {nl1}, "
          f"size={os.path.getsize(path)/1e6:.2f} MB")


# name: (ndim, norder, ipoly, iperiod, eps, iptype, eta, nd, zk, ifnewtree, seed, ng, width range)
TREE_FUN_CASES = {
    "f2_l2":         (2, 8, 0, 0, 1e-6, 2, 0.0, 1, 0.0, 0, 31, 3, (1e-3, 1e-2)),
    "f2_l1_eta1":    (2, 8, 0, 0, 1e-6, 1, 1.0, 1, 0.0, 0, 32, 3, (1e-3, 1e-2)),
    "f2_max_nd2_per": (2, 6, 1, 1, 1e-3, 0, 0.0, 2, 0.0, 0, 33, 3, (1e-3, 1e-2)),
    "f3_l2_eta2":    (3, 6, 0, 0, 1e-3, 2, 2.0, 1, 0.0, 0, 34, 2, (3e-3, 1e-2)),
    "f3_max_per_nd2": (3, 4, 1, 1, 5e-2, 0, 0.5, 2, 0.0, 0, 35, 2, (1e-2, 3e-2)),
    "f1_l1":         (1, 8, 0, 0, 1e-9, 1, 0.0, 1, 0.0, 0, 36, 3, (1e-4, 1e-3)),
    "f2_zk":         (2, 8, 0, 0, 1e-4, 2, 0.0, 1, 60.0, 0, 37, 2, (1e-2, 3e-2)),
    "f2_newtree":    (2, 8, 0, 0, 1e-5, 2, 0.0, 1, 0.0, 1, 38, 3, (1e-3, 1e-2)),
}


def test_source_lib(outdir):
    """tests/csrc/sources.c compiled into outdir (the source function handed
    to both tree builders)."""
    import subprocess
    so = os.path.join(outdir, "libtestsource.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-o", so, os.path.join(ROOT, "tests", "csrc", "sources.c"),
                    "-lm"], check=True)
    L = C.CDLL(so)
    return L, L.test_source


def source_pars(ndim, seed, ng, wr, iperiod):
    c, s, a = make_source(ndim, seed, ng, wr, iperiod)
    dp = np.ascontiguousarray(np.hstack([c, s, a.reshape(-1, 1)]).ravel())
    ip = np.array([ndim, ng, iperiod], np.int32)
    return dp, ip


def tree_fun_case(name, spec, fun):
    ndim, n, ipoly, iper, eps, iptype, eta, nd, zk, ifnt, seed, ng, wr = spec
    dp, ip = source_pars(ndim, seed, ng, wr, iper)
    tr = refrun.build_tree_fun(ndim, n, ipoly, iper, eps, fun, dp, ip, nd=nd, iptype=iptype, eta=eta, zk=zk,
                               ifnewtree=ifnt)
    nused = int(tr.itree[2 * tr.nlevels + 1])
    path = os.path.join(HERE, f"tree_{name}.npz")
    np.savez_compressed(path, ndim=ndim, norder=n, ipoly=ipoly, iperiod=iper, eps=eps, iptype=iptype, eta=eta,
                        nd=nd, zk=zk, ifnewtree=ifnt, dpars=dp, ipars=ip, nboxes=tr.nboxes, nlevels=tr.nlevels,
                        ltree=tr.ltree, itree=tr.itree, iptr=tr.iptr, centers=tr.centers[:nused],
                        boxsize=tr.boxsize, fvals=tr.fvals[:nused])
    print(f"tree_{name}: boxes {nused} (arrays {tr.nboxes}), nlevels {tr.nlevels}, "
          f"size={os.path.getsize(path)/1e6:.2f} MB")


def tables_mref3():
    """mk_loctab_all with mrefinelev = 3 (193 offsets per level), the tables
    boxfgt uses when it refines the tree after the solve."""
    L = refrun.lib()
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    out = {}
    for ipoly, n, delta, eps, boxdim in ((0, 8, 1e-3, 1e-6, 0.25), (1, 6, 1e-4, 1e-9, 0.0625),
                                         (0, 12, 1e-4, 1e-8, 0.125)):
        key = f"{ipoly}_{n}_{delta:g}_{eps:g}_{boxdim:g}"
        nt = 96
        ns = 2 * nt + 1
        tl = np.zeros(n * n * ns); tlx = np.zeros(n * n * ns); tlxx = np.zeros(n * n * ns)
        ind = np.zeros(2 * (n + 1) * ns, np.int32)
        L.mk_loctab_all_(_dp(eps), _ip(ipoly), _ip(n), _ip(50), _dp(delta), _dp(boxdim), _ip(3),
                         _ip(nt), P(tl), P(tlx), P(tlxx), P(ind))
        out[f"loctab_{key}"] = tl
        out[f"locind_{key}"] = ind
    path = os.path.join(HERE, "tables_mref3.npz")
    np.savez_compressed(path, **out)
    print(f"tables_mref3: {len(out)} arrays, {os.path.getsize(path)/1e6:.2f} MB")


if __name__ == "__main__":
    refrun.lib().refglue_silence()
    only = sys.argv[1:]
    if not only or "tables" in only:
        tables()
    if not only or "tables_mref3" in only:
        tables_mref3()
    for name, spec in CASES.items():
        if not only or name in only:
            boxfgt_case(name, spec)
    for name, spec in PGH_CASES.items():
        if not only or name in only:
            pgh_case(name, spec)
    for name, spec in TARG_CASES.items():
        if not only or name in only:
            targ_case(name, spec)
    for name, spec in NEWTREE_CASES.items():
        if not only or name in only or "newtree" in only:
            newtree_case(name, spec)
    if not only or "tree_fun" in only or any(k in only for k in TREE_FUN_CASES):
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            _, fun = test_source_lib(td)
            for name, spec in TREE_FUN_CASES.items():
                if not only or name in only or "tree_fun" in only:
                    tree_fun_case(name, spec, fun)
