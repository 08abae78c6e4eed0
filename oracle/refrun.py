"""ctypes driver for the reference build in ``oracle/_ref/libfgtref.so``.

Test infrastructure only: tests use it (in the build container, where the
reference sources exist) to build trees with the reference tree builder and
to run the reference ``boxfgt`` on them, producing the golden fixtures under
``tests/golden/``.  ``bench.py``'s CPU-baseline leg may time it.

Reference interfaces driven here:
  * ``vol_tree_mem`` / ``vol_tree_build`` (src/common/tree_vol_coeffs_nd.f:83,444)
  * ``boxfgt`` (src/bfgt/boxfgt.f:5-8), 28 by-reference arguments.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(HERE, "_ref", "libfgtref.so")

_lib = None


def available() -> bool:
    return os.path.exists(REF_SO)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"reference build missing: {REF_SO} (make -C oracle ref)")
        _lib = C.CDLL(REF_SO)
    return _lib


@dataclass
class Tree:
    """Reference tree arrays (1-based box ids inside ``itree``)."""
    ndim: int
    norder: int
    ipoly: int
    iperiod: int
    nboxes: int
    nlevels: int
    ltree: int
    itree: np.ndarray      # int32 (ltree,)
    iptr: np.ndarray       # int32 (8,)
    centers: np.ndarray    # f64 (nboxes, ndim)  == Fortran centers(ndim,nboxes)
    boxsize: np.ndarray    # f64 (nlevels+1,)
    fvals: np.ndarray      # f64 (nboxes, npbox) == Fortran fvals(1,npbox,nboxes)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def gauss_params(centers, widths, amps) -> np.ndarray:
    """Pack Gaussians (c[ndim], s[ndim], a) per source into the dpars layout."""
    centers = np.atleast_2d(np.asarray(centers, np.float64))
    widths = np.atleast_2d(np.asarray(widths, np.float64))
    amps = np.asarray(amps, np.float64).reshape(-1, 1)
    return np.ascontiguousarray(np.hstack([centers, widths, amps]).ravel())


def build_tree(ndim, norder, ipoly, iperiod, eps, dpars, ng, ifnewtree=0) -> Tree:
    """ifnewtree=1: arrays sized for three times the boxes (vol_tree_mem's
    room for boxfgt's refinement); boxes past the tree's own stay zero."""
    L = lib()
    dp = np.ascontiguousarray(dpars, np.float64)
    nb, nl, lt = C.c_int(), C.c_int(), C.c_int()
    L.refglue_tree_mem_nt(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod),
                          C.c_double(eps), C.c_int(norder), C.c_int(ng), _p(dp), C.c_int(ifnewtree),
                          C.byref(nb), C.byref(nl), C.byref(lt))
    nboxes, nlevels, ltree = nb.value, nl.value, lt.value
    npbox = norder ** ndim
    itree = np.zeros(ltree, np.int32)
    iptr = np.zeros(8, np.int32)
    centers = np.zeros((nboxes, ndim), np.float64)
    boxsize = np.zeros(nlevels + 1, np.float64)
    fvals = np.zeros((nboxes, npbox), np.float64)
    L.refglue_tree_build(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod),
                         C.c_double(eps), C.c_int(norder), C.c_int(ng), _p(dp),
                         C.c_int(nboxes), C.c_int(nlevels), C.c_int(ltree),
                         _p(itree), _p(iptr), _p(centers), _p(boxsize), _p(fvals))
    return Tree(ndim, norder, ipoly, iperiod, nboxes, nlevels, ltree, itree,
                iptr, centers, boxsize, fvals)


def build_tree_fun(ndim, norder, ipoly, iperiod, eps, fun, dpars, ipars, nd=1, iptype=2, eta=0.0, zk=0.0,
                   ifnewtree=0) -> Tree:
    """The reference's vol_tree_mem / vol_tree_build for a compiled source
    function ``fun`` (a ctypes function pointer with the reference's
    by-reference ``fun(nd, xyz, dpars, zpars, ipars, f)`` convention).
    fvals is (nboxes, npbox) for nd = 1, else (nboxes, npbox, nd)."""
    L = lib()
    dp = np.ascontiguousarray(dpars, np.float64)
    ip = np.ascontiguousarray(ipars, np.int32)
    fp = C.cast(fun, C.c_void_p)
    nb, nl, lt = C.c_int(), C.c_int(), C.c_int()
    L.refglue_tree_mem_fun(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod), C.c_double(eps), C.c_double(zk),
                           C.c_int(norder), C.c_int(iptype), C.c_double(eta), fp, C.c_int(nd), _p(dp), _p(ip),
                           C.c_int(ifnewtree), C.byref(nb), C.byref(nl), C.byref(lt))
    nboxes, nlevels, ltree = nb.value, nl.value, lt.value
    npbox = norder ** ndim
    itree = np.zeros(ltree, np.int32)
    iptr = np.zeros(8, np.int32)
    centers = np.zeros((nboxes, ndim), np.float64)
    boxsize = np.zeros(nlevels + 1, np.float64)
    fvals = np.zeros((nboxes, npbox) if nd == 1 else (nboxes, npbox, nd), np.float64)
    L.refglue_tree_build_fun(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod), C.c_double(eps), C.c_double(zk),
                             C.c_int(norder), C.c_int(iptype), C.c_double(eta), fp, C.c_int(nd), _p(dp), _p(ip),
                             C.c_int(nboxes), C.c_int(nlevels), C.c_int(ltree), _p(itree), _p(iptr), _p(centers),
                             _p(boxsize), _p(fvals))
    return Tree(ndim, norder, ipoly, iperiod, nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals)


def boxfgt(tree: Tree, delta: float, eps: float, nd: int = 1, fvals=None, ifpgh: int = 1, targs=None,
           ifpghtarg: int = 0, ifnewtree: int = 0, out_tree=None):
    """Run the reference boxfgt.

    ifnewtree=1 rewrites the tree (refinement, coarsening, 2:1 rebalance);
    out_tree (a dict) then receives the rewritten itree, centers and nlevels.

    Returns (pot, timeinfo[20]) for ifpgh=1, (pot, grad, hess, timeinfo) for
    ifpgh=2/3 (hess None for ifpgh=2).  pot is shaped like fvals; grad is
    Fortran grad(nd,ndim,npbox,nboxes) as [nboxes, npbox, ndim(, nd)], hess
    likewise with ndim*(ndim+1)/2 components.  With targs (nt, ndim) and
    ifpghtarg >= 1 the result is (pot, grad, hess, pote, grade, hesse, timeinfo)
    with pote [nt(, nd)], grade [nt, ndim(, nd)], hesse [nt, nhess(, nd)].
    """
    L = lib()
    npbox = tree.norder ** tree.ndim
    f = np.ascontiguousarray(tree.fvals if fvals is None else fvals, np.float64)
    pot = np.zeros_like(f)
    ndim = tree.ndim
    nhess = ndim * (ndim + 1) // 2
    tail = () if nd == 1 else (nd,)
    grad = np.zeros((tree.nboxes, npbox, ndim) + tail if ifpgh >= 2 else 1, np.float64)
    hess = np.zeros((tree.nboxes, npbox, nhess) + tail if ifpgh >= 3 else 1, np.float64)
    nt = 0 if targs is None else len(targs)
    tail2 = () if nd == 1 else (nd,)
    targs = np.zeros((1, tree.ndim)) if targs is None else np.ascontiguousarray(targs, np.float64)
    pote = np.zeros((max(nt, 1),) + tail2, np.float64)
    grade = np.zeros((max(nt, 1), ndim) + tail2, np.float64)
    hesse = np.zeros((max(nt, 1), nhess) + tail2, np.float64)
    dum = np.zeros(1, np.float64)
    timeinfo = np.zeros(100, np.float64)
    itree = tree.itree.copy()   # the reference may write colleagues back (periodic, npwlevel<=1)
    centers = tree.centers.copy()
    nlev = C.c_int(tree.nlevels)
    ints = [C.c_int(v) for v in (nd, tree.ndim)]
    L.boxfgt_(C.byref(ints[0]), C.byref(ints[1]), C.byref(C.c_double(delta)),
              C.byref(C.c_double(eps)), C.byref(C.c_int(tree.ipoly)),
              C.byref(C.c_int(tree.iperiod)), C.byref(C.c_int(tree.norder)),
              C.byref(C.c_int(npbox)), C.byref(C.c_int(tree.nboxes)),
              C.byref(nlev), C.byref(C.c_int(tree.ltree)),
              _p(itree), _p(tree.iptr), _p(centers), _p(tree.boxsize),
              _p(f), C.byref(C.c_int(ifpgh)), _p(pot), _p(grad), _p(hess),
              C.byref(C.c_int(ifnewtree)), C.byref(C.c_int(nt)), _p(targs),
              C.byref(C.c_int(ifpghtarg if nt else 0)), _p(pote), _p(grade), _p(hesse), _p(timeinfo))
    if out_tree is not None:
        out_tree.update(itree=itree, centers=centers, nlevels=nlev.value)
    if nt and ifpghtarg >= 1:
        return (pot, grad if ifpgh >= 2 else None, hess if ifpgh >= 3 else None, pote,
                grade if ifpghtarg >= 2 else None, hesse if ifpghtarg >= 3 else None, timeinfo[:20].copy())
    if ifpgh == 1:
        return pot, timeinfo[:20].copy()
    return pot, grad, (hess if ifpgh >= 3 else None), timeinfo[:20].copy()
