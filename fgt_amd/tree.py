"""Adaptive volume tree (host) in the reference's array layout.

``Tree`` holds the arrays ``vol_tree_build`` produces
(src/common/tree_vol_coeffs_nd.f:444-687): ``itree`` (1-based box ids,
sections located by ``iptr``), ``centers``, ``boxsize`` and ``fvals``.
``vol_tree_build`` builds one with the library's C++ builder (libfgt_amd.so,
``fgt_vol_tree_*`` in include/fgt_tree.h) for a source made of anisotropic
Gaussians, ``vol_tree_build_fun`` for a caller's compiled source function
(the reference's ``fun`` convention) with any refinement norm.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np


@dataclass
class Tree:
    ndim: int
    norder: int
    ipoly: int
    iperiod: int
    nboxes: int
    nlevels: int
    ltree: int
    itree: np.ndarray      # int32 (ltree,)
    iptr: np.ndarray       # int32 (8,)
    centers: np.ndarray    # float64 (nboxes, ndim)
    boxsize: np.ndarray    # float64 (nlevels+1,)
    fvals: np.ndarray      # float64 (nboxes, npbox) or (nboxes, npbox, nd)

    @property
    def npbox(self) -> int:
        return self.norder ** self.ndim

    @property
    def nused(self) -> int:
        """Boxes of the tree (1..laddr(2,nlevels)); nboxes may be larger
        (room for boxfgt's refinement)."""
        return int(self.itree[2 * self.nlevels + 1])

    def section(self, k: int, width: int = 1) -> np.ndarray:
        """itree section k (1-based iptr index) as a (nboxes, width) view."""
        start = int(self.iptr[k - 1]) - 1
        return self.itree[start:start + width * self.nboxes].reshape(self.nboxes, width)

    @property
    def nchild(self) -> np.ndarray:
        return self.section(4)[:, 0]

    @property
    def level(self) -> np.ndarray:
        return self.section(2)[:, 0]

    @property
    def leaves(self) -> np.ndarray:
        return np.nonzero(self.nchild[:self.nused] == 0)[0]

    @classmethod
    def from_npz(cls, z) -> "Tree":
        return cls(int(z["ndim"]), int(z["norder"]), int(z["ipoly"]), int(z["iperiod"]),
                   int(z["nboxes"]), int(z["nlevels"]), int(z["ltree"]), z["itree"].astype(np.int32),
                   z["iptr"].astype(np.int32), z["centers"].astype(np.float64),
                   z["boxsize"].astype(np.float64), z["fvals"].astype(np.float64))


def gauss_source(centers, widths, amps) -> np.ndarray:
    """Pack Gaussians as rows (c[ndim], s[ndim], a): f = sum a exp(-sum (y-c)^2/s)."""
    centers = np.atleast_2d(np.asarray(centers, np.float64))
    widths = np.atleast_2d(np.asarray(widths, np.float64))
    amps = np.asarray(amps, np.float64).reshape(-1, 1)
    return np.ascontiguousarray(np.hstack([centers, widths, amps]))


class TreeTooLarge(RuntimeError):
    """The tree would exceed the requested leaf limit."""


def _copy_out(L, h, ndim, norder, ipoly, iperiod, nb, nl, lt, nd=1, newtree=False):
    from . import _check, _p
    nboxes, nlevels, ltree = nb.value, nl.value, lt.value
    if newtree:
        _check(L.fgt_vol_tree_newtree_room(h, C.byref(nb), C.byref(lt)))
        nboxes, ltree = nb.value, lt.value
    itree = np.zeros(ltree, np.int32)
    iptr = np.zeros(8, np.int32)
    centers = np.zeros((nboxes, ndim), np.float64)
    boxsize = np.zeros(nlevels + 1, np.float64)
    shape = (nboxes, norder ** ndim) if nd == 1 else (nboxes, norder ** ndim, nd)
    fvals = np.zeros(shape, np.float64)
    _check(L.fgt_vol_tree_build(h, _p(itree), _p(iptr), _p(centers), _p(boxsize), _p(fvals)))
    return Tree(ndim, norder, ipoly, iperiod, nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals)


def vol_tree_build_fun(ndim, norder, ipoly, iperiod, eps, fun, dpars, ipars, nd=1, iptype=2, eta=0.0, zk=0.0,
                       ifnewtree=0, boxlen=1.0, zpars=None) -> Tree:
    """vol_tree_mem + vol_tree_build for a caller's source
    (src/common/tree_vol_coeffs_nd.f:83-687): ``fun`` is a compiled function
    (a ctypes function pointer or its address) with the reference's
    convention ``fun(nd, xyz, dpars, zpars, ipars, f)``, every argument by
    reference, safe to call from several threads; ``dpars``/``ipars`` (and
    ``zpars``, complex) are passed through.  iptype 0/1/2 = max / L1 / L2
    refinement norm, eta = the box-size power on the tail error, zk = the
    Helmholtz parameter's real part.  ifnewtree = 1 sizes the arrays for
    three times the boxes (room for boxfgt's refinement)."""
    from . import lib, _check, _p
    L = lib()
    dp = np.ascontiguousarray(dpars, np.float64)
    ip = np.ascontiguousarray(ipars, np.int32)
    zp = np.zeros(2) if zpars is None else np.ascontiguousarray(np.asarray(zpars, np.complex128).view(np.float64))
    fptr = C.cast(fun, C.c_void_p) if not isinstance(fun, int) else C.c_void_p(fun)
    h = C.c_void_p()
    nb, nl, lt = C.c_int(), C.c_int(), C.c_int()
    _check(L.fgt_vol_tree_mem_fun(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod), C.c_double(eps), C.c_double(zk),
                                  C.c_double(boxlen), C.c_int(norder), C.c_int(iptype), C.c_double(eta), fptr,
                                  C.c_int(nd), _p(dp), _p(zp), _p(ip), C.c_int(ifnewtree), C.byref(h), C.byref(nb),
                                  C.byref(nl), C.byref(lt)))
    return _copy_out(L, h, ndim, norder, ipoly, iperiod, nb, nl, lt, nd=nd)


def vol_tree_build(ndim, norder, ipoly, iperiod, eps, gauss, boxlen=1.0, max_leaves=0, ifnewtree=0) -> Tree:
    """Adaptive level-restricted tree resolving ``gauss`` to ``eps``.

    Same refinement rule, level restriction and array layout as the
    reference's vol_tree_mem + vol_tree_build with iptype = 2 (L2), eta = 0
    (src/common/tree_vol_coeffs_nd.f:83-779,1242-1733).  With iperiod = 1 the
    Gaussians are periodised over the 3^ndim nearest images.  max_leaves > 0:
    raise TreeTooLarge as soon as the tree would have more leaves.
    ifnewtree = 1 sizes the arrays for three times the boxes (room for
    boxfgt's refinement, as vol_tree_mem does).
    """
    from . import lib, _check, _p  # local import: avoids a cycle at import time
    L = lib()
    g = np.ascontiguousarray(gauss, np.float64)
    ng = g.shape[0]
    h = C.c_void_p()
    nb, nl, lt = C.c_int(), C.c_int(), C.c_int()
    rc = L.fgt_vol_tree_mem_capped(C.c_int(ndim), C.c_int(ipoly), C.c_int(iperiod), C.c_double(eps),
                                   C.c_int(norder), C.c_double(boxlen), C.c_int(ng), _p(g),
                                   C.c_longlong(max_leaves), C.byref(h), C.byref(nb), C.byref(nl), C.byref(lt))
    if rc == 3:
        raise TreeTooLarge(f"tree exceeds {max_leaves} leaves")
    _check(rc)
    return _copy_out(L, h, ndim, norder, ipoly, iperiod, nb, nl, lt, newtree=bool(ifnewtree))
