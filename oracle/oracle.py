"""ctypes wrapper for the C restatement ``oracle/_build/liboracle.so``.

Test infrastructure only (the checker and the CPU-baseline leg of bench.py).
See fgt_oracle.h for what each entry point restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


class OracleInfo(C.Structure):
    _fields_ = [("npwlevel", C.c_int), ("npw", C.c_int), ("nlevstart", C.c_int),
                ("nexp", C.c_int), ("periodic_small", C.c_int), ("pmax", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = C.CDLL(SO)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels,
           itree, iptr, centers, boxsize, fvals, d1_mode=0):
    """Potential (nboxes, npbox[, nd]) from the oracle; returns (pot, info, timeinfo)."""
    L = lib()
    itree = np.ascontiguousarray(itree, np.int32)
    iptr = np.ascontiguousarray(iptr, np.int32)
    centers = np.ascontiguousarray(centers, np.float64)
    boxsize = np.ascontiguousarray(boxsize, np.float64)
    fvals = np.ascontiguousarray(fvals, np.float64)
    pot = np.zeros_like(fvals)
    info = OracleInfo()
    ti = np.zeros(8, np.float64)
    rc = L.oracle_boxfgt(C.c_int(nd), C.c_int(ndim), C.c_double(delta), C.c_double(eps),
                         C.c_int(ipoly), C.c_int(iperiod), C.c_int(norder),
                         C.c_int(nboxes), C.c_int(nlevels), _p(itree), _p(iptr),
                         _p(centers), _p(boxsize), _p(fvals), _p(pot),
                         C.c_int(d1_mode), _p(ti), C.byref(info))
    if rc != 0:
        raise RuntimeError(f"oracle_boxfgt failed with code {rc}")
    return pot, info, ti


def boxfgt_pgh(nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels,
               itree, iptr, centers, boxsize, fvals, ifpgh, d1_mode=0):
    """(pot, grad, hess, info): grad [nboxes, npbox, ndim(, nd)], hess likewise
    with ndim*(ndim+1)/2 components (None for ifpgh < 3)."""
    L = lib()
    itree = np.ascontiguousarray(itree, np.int32)
    iptr = np.ascontiguousarray(iptr, np.int32)
    centers = np.ascontiguousarray(centers, np.float64)
    boxsize = np.ascontiguousarray(boxsize, np.float64)
    fvals = np.ascontiguousarray(fvals, np.float64)
    npbox = norder ** ndim
    tail = () if nd == 1 else (nd,)
    pot = np.zeros_like(fvals)
    grad = np.zeros((nboxes, npbox, ndim) + tail if ifpgh >= 2 else 1)
    hess = np.zeros((nboxes, npbox, ndim * (ndim + 1) // 2) + tail if ifpgh >= 3 else 1)
    info = OracleInfo()
    rc = L.oracle_boxfgt_pgh(C.c_int(nd), C.c_int(ndim), C.c_double(delta), C.c_double(eps),
                             C.c_int(ipoly), C.c_int(iperiod), C.c_int(norder),
                             C.c_int(nboxes), C.c_int(nlevels), _p(itree), _p(iptr),
                             _p(centers), _p(boxsize), _p(fvals), C.c_int(ifpgh), _p(pot), _p(grad),
                             _p(hess), C.c_int(d1_mode), None, C.byref(info))
    if rc != 0:
        raise RuntimeError(f"oracle_boxfgt_pgh failed with code {rc}")
    return pot, (grad if ifpgh >= 2 else None), (hess if ifpgh >= 3 else None), info


def evalt(ndim, nd, ipoly, norder, nboxes, nlevels, itree, iptr, centers, boxsize, vals, targs):
    """Values (nboxes, npbox[, nd]) on the leaves -> values at targs (nt, ndim): (nt[, nd])."""
    targs = np.ascontiguousarray(targs, np.float64)
    nt = targs.shape[0]
    out = np.zeros((nt,) + ((nd,) if nd > 1 else ()))
    rc = lib().oracle_evalt(C.c_int(ndim), C.c_int(nd), C.c_int(ipoly), C.c_int(norder), C.c_int(nboxes),
                            C.c_int(nlevels), _p(np.ascontiguousarray(itree, np.int32)),
                            _p(np.ascontiguousarray(iptr, np.int32)), _p(np.ascontiguousarray(centers)),
                            _p(np.ascontiguousarray(boxsize)), _p(np.ascontiguousarray(vals, np.float64)),
                            C.c_int(nt), _p(targs), _p(out))
    if rc != 0:
        raise RuntimeError(f"oracle_evalt failed with code {rc}")
    return out


def pw2pgh(ipoly, n, npw, ts, delta, boxdim):
    """(tab_pw2pot, tab_pw2der, tab_pw2dxx), each [j, i] like the reference (npw, n)."""
    t = [np.zeros(2 * n * npw) for _ in range(3)]
    lib().oracle_pw2pgh(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(np.ascontiguousarray(ts)),
                        C.c_double(delta), C.c_double(boxdim), _p(t[0]), _p(t[1]), _p(t[2]))
    return tuple(a.view(np.complex128).reshape(n, npw).T for a in t)


def loctab3(eps, ipoly, n, delta, boxdim):
    """(tab, tabx, tabxx, ind) as loctab, with the derivative tables."""
    t = [np.zeros(n * n * 13) for _ in range(3)]
    ind = np.zeros(2 * (n + 1) * 13, np.int32)
    lib().oracle_loctab3(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                         C.c_double(boxdim), _p(t[0]), _p(t[1]), _p(t[2]), _p(ind))
    return tuple(a.reshape(13, n, n) for a in t) + (ind.reshape(13, n + 1, 2),)


def pwterms(bsize, delta, eps, iperiod):
    pmax = C.c_double()
    npw = C.c_int()
    lib().oracle_pwterms(C.c_double(bsize), C.c_double(delta), C.c_double(eps),
                         C.c_int(iperiod), C.byref(pmax), C.byref(npw))
    return pmax.value, npw.value


def pwnodes(pmax, npw):
    ws = np.zeros(npw); ts = np.zeros(npw)
    lib().oracle_pwnodes(C.c_double(pmax), C.c_int(npw), _p(ws), _p(ts))
    return ws, ts


def periodic_pwnodes(bs0, delta, npw):
    ws = np.zeros(npw); ts = np.zeros(npw)
    lib().oracle_periodic_pwnodes(C.c_double(bs0), C.c_double(delta), C.c_int(npw), _p(ws), _p(ts))
    return ws, ts


def polyexps(ipoly, n):
    x = np.zeros(n); w = np.zeros(n); u = np.zeros(n * n)
    lib().oracle_polyexps(C.c_int(ipoly), C.c_int(n), _p(x), _p(w), _p(u))
    return x, w, u.reshape(n, n).T   # u[i, j]: coefficient i from value j


def sph_bessel(nterms, z):
    f = np.zeros(nterms + 1)
    lib().oracle_sph_bessel(C.c_int(nterms), C.c_double(z), _p(f))
    return f


def poly2pw(ipoly, n, npw, ws, ts, delta, boxdim):
    t = np.zeros(2 * n * npw)
    lib().oracle_poly2pw(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(np.ascontiguousarray(ws)),
                         _p(np.ascontiguousarray(ts)), C.c_double(delta), C.c_double(boxdim), _p(t))
    t = t.view(np.complex128)
    return t.reshape(npw, n).T      # [m, j] like tab_poly2pw(n, npw)


def pw2pot(ipoly, n, npw, ts, delta, boxdim):
    t = np.zeros(2 * n * npw)
    lib().oracle_pw2pot(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(np.ascontiguousarray(ts)),
                        C.c_double(delta), C.c_double(boxdim), _p(t))
    return t.view(np.complex128).reshape(n, npw).T   # [j, i] like tab_pw2pot(npw, n)


def loctab(eps, ipoly, n, delta, boxdim):
    tab = np.zeros(n * n * 13)
    ind = np.zeros(2 * (n + 1) * 13, np.int32)
    lib().oracle_loctab(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                        C.c_double(boxdim), _p(tab), _p(ind))
    # Fortran tab(i,j,k) -> numpy [k+6, j, i]; ind(2,n+1,k) -> [k+6, j, 2]
    return tab.reshape(13, n, n), ind.reshape(13, n + 1, 2)
