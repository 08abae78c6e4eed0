"""fgt_amd -- MI355X box Gauss transform behind the reference's ``boxfgt`` interface.

The compute path is ``libfgt_amd.so`` (C++ host planner + HIP kernels for
gfx950), reached through its C ABI (``include/fgt.h``).  This module mirrors
the reference's Fortran interface for Python callers:

* :func:`boxfgt` -- same arguments and meaning as ``boxfgt``
  (src/bfgt/boxfgt.f:5-8): accumulates the potential into ``pot``.
* :class:`Solver` -- device-resident plan for repeated solves on one tree.
* :func:`vol_tree_build` -- host tree builder producing the arrays of the
  reference's ``vol_tree_mem``/``vol_tree_build``.

There is no CPU fallback: if the library (or a GPU) is missing, calls fail.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfgt_amd.so")

_lib = None


class FgtError(RuntimeError):
    """Raised when the library reports a nonzero status."""


class _SolverStats(C.Structure):
    _fields_ = [("npwlevel", C.c_int), ("npw", C.c_int), ("nexp", C.c_int),
                ("nlevstart", C.c_int), ("nslots", C.c_int),
                ("periodic_small", C.c_int), ("all_pw", C.c_int),
                ("nleaves", C.c_longlong), ("npw_leaves", C.c_longlong),
                ("ndirect_pairs", C.c_longlong), ("nout", C.c_longlong),
                ("expansion_bytes", C.c_longlong), ("stage_ms", C.c_double * 8),
                ("n_m2l_targets", C.c_longlong), ("n_m2l_pairs", C.c_longlong),
                ("n_parent_child", C.c_longlong), ("m2l_bricks", C.c_longlong), ("m2l_form", C.c_int)]


class _PlanInfo(C.Structure):
    _fields_ = [("npwlevel", C.c_int), ("npw", C.c_int), ("nexp", C.c_int),
                ("nlevstart", C.c_int), ("nslots", C.c_int),
                ("periodic_small", C.c_int), ("all_pw", C.c_int),
                ("nleaves", C.c_longlong), ("npw_leaves", C.c_longlong),
                ("ndirect_pairs", C.c_longlong), ("n_m2l_pairs", C.c_longlong),
                ("n_parent_child", C.c_longlong), ("n_m2l_targets", C.c_longlong)]


def lib():
    """Load libfgt_amd.so (built in-tree by ``python -m fgt_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FgtError(f"{LIB_PATH} is missing: build it with `python -m fgt_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.fgt_solver_create.restype = C.c_int
        L.fgt_solver_fvals_device.restype = C.c_void_p
        L.fgt_solver_pot_device.restype = C.c_void_p
        L.fgt_solver_grad_device.restype = C.c_void_p
        L.fgt_solver_hess_device.restype = C.c_void_p
        L.fgt_plan_direct_pairs.restype = C.c_longlong
        _lib = L
    return _lib


def last_status():
    buf = C.create_string_buffer(512)
    st = lib().fgt_last_status(buf, 512)
    return st, buf.value.decode(errors="replace")


def _check(rc):
    if rc != 0:
        st, msg = last_status()
        raise FgtError(f"fgt status {rc}: {msg}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i(v):
    return C.byref(C.c_int(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels,
           ltree, itree, iptr, centers, boxsize, fvals, ifpgh, pot, grad=None,
           hess=None, ifnewtree=0, ntarg=0, targs=None, ifpghtarg=0, pote=None,
           grade=None, hesse=None, timeinfo=None, newtree=None):
    """Box Gauss transform, argument for argument as the reference ``boxfgt``.

    Arrays use the reference's Fortran memory order; with C-ordered numpy
    arrays that means ``centers[nboxes, ndim]``, ``fvals[nboxes, npbox(, nd)]``
    and ``pot`` of the same shape as ``fvals`` (accumulated, ``+=``);
    ``grad[nboxes, npbox, ndim(, nd)]`` for ifpgh >= 2 and
    ``hess[nboxes, npbox, ndim*(ndim+1)/2(, nd)]`` for ifpgh = 3, also
    accumulated (Hessian order 1-D xx; 2-D xx, xy, yy; 3-D xx, yy, zz, xy,
    xz, yz).  With ifpghtarg >= 1, ``targs[ntarg, ndim]`` and
    ``pote[ntarg(, nd)]``, ``grade[ntarg, ndim(, nd)]``,
    ``hesse[ntarg, nhess(, nd)]`` receive the values at the targets
    (overwritten).  Raises :class:`FgtError` for invalid input instead of
    silently returning.

    ifnewtree = 1 refines, coarsens and 2:1-rebalances the tree after the
    solve, as the reference does: ``nboxes`` is then the capacity of the
    arrays (vol_tree_mem's three times the tree's boxes), ``itree``,
    ``centers``, ``pot`` (and ``grad``/``hess``) are rewritten in place (so
    they must be C-contiguous int32 / float64 arrays) and the new level count
    is stored in ``newtree["nlevels"]`` when a dict is passed.
    """
    if ifnewtree == 1:
        for arr, dt, nm in ((itree, np.int32, "itree"), (centers, np.float64, "centers")):
            if not (isinstance(arr, np.ndarray) and arr.dtype == dt and arr.flags.c_contiguous):
                raise TypeError(f"with ifnewtree = 1, {nm} must be a C-contiguous {dt.__name__} array "
                                "(it is rewritten in place)")
    itree = np.ascontiguousarray(itree, np.int32)
    iptr = np.ascontiguousarray(iptr, np.int32)
    centers = np.ascontiguousarray(centers, np.float64)
    boxsize = np.ascontiguousarray(boxsize, np.float64)
    fvals = np.ascontiguousarray(fvals, np.float64)
    if not (isinstance(pot, np.ndarray) and pot.dtype == np.float64 and pot.flags.c_contiguous):
        raise TypeError("pot must be a C-contiguous float64 array (it is accumulated in place)")
    if pot.size != nd * npbox * nboxes or fvals.size != nd * npbox * nboxes:
        raise ValueError("fvals/pot must hold nd*npbox*nboxes values")
    for arr, need, nm in ((grad, 2, "grad"), (hess, 3, "hess")):
        if ifpgh >= need:
            ncomp = ndim if nm == "grad" else ndim * (ndim + 1) // 2
            if not (isinstance(arr, np.ndarray) and arr.dtype == np.float64 and arr.flags.c_contiguous
                    and arr.size == nd * ncomp * npbox * nboxes):
                raise TypeError(f"{nm} must be a C-contiguous float64 array of nd*{ncomp}*npbox*nboxes values")
    if ifpghtarg >= 1 and ntarg > 0:
        targs = np.ascontiguousarray(targs, np.float64)
        if targs.size != ndim * ntarg:
            raise ValueError("targs must hold ndim*ntarg values")
        nhess = ndim * (ndim + 1) // 2
        for arr, need, ncomp, nm in ((pote, 1, 1, "pote"), (grade, 2, ndim, "grade"), (hesse, 3, nhess, "hesse")):
            if ifpghtarg >= need and not (isinstance(arr, np.ndarray) and arr.dtype == np.float64
                                          and arr.flags.c_contiguous and arr.size == nd * ncomp * ntarg):
                raise TypeError(f"{nm} must be a C-contiguous float64 array of nd*{ncomp}*ntarg values")
    dummy = np.zeros(1, np.float64)
    ti = np.zeros(20, np.float64) if timeinfo is None else timeinfo
    nl = C.c_int(nlevels)
    rc = lib().fgt_boxfgt(_i(nd), _i(ndim), _d(delta), _d(eps), _i(ipoly), _i(iperiod),
                          _i(norder), _i(npbox), _i(nboxes), C.byref(nl), _i(ltree),
                          _p(itree), _p(iptr), _p(centers), _p(boxsize), _p(fvals),
                          _i(ifpgh), _p(pot), _p(dummy if grad is None else grad),
                          _p(dummy if hess is None else hess), _i(ifnewtree), _i(ntarg),
                          _p(dummy if targs is None else targs), _i(ifpghtarg),
                          _p(dummy if pote is None else pote), _p(dummy if grade is None else grade),
                          _p(dummy if hesse is None else hesse), _p(ti))
    _check(rc)
    if newtree is not None:
        newtree["nlevels"] = nl.value
    return pot


class Solver:
    """Device-resident box FGT on one tree (plan once, solve many times)."""

    def __init__(self, nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels,
                 ltree, itree, iptr, centers, boxsize):
        self._h = C.c_void_p()
        self.nd, self.ndim, self.norder, self.nboxes = nd, ndim, norder, nboxes
        self.npbox = norder ** ndim
        self._keep = [np.ascontiguousarray(itree, np.int32), np.ascontiguousarray(iptr, np.int32),
                      np.ascontiguousarray(centers, np.float64), np.ascontiguousarray(boxsize, np.float64)]
        it, ip, ce, bs = self._keep
        _check(lib().fgt_solver_create(C.c_int(nd), C.c_int(ndim), C.c_double(delta), C.c_double(eps),
                                       C.c_int(ipoly), C.c_int(iperiod), C.c_int(norder),
                                       C.c_int(nboxes), C.c_int(nlevels), C.c_int(ltree),
                                       _p(it), _p(ip), _p(ce), _p(bs), C.byref(self._h)))

    @classmethod
    def from_tree(cls, tree, delta, eps, nd=1):
        return cls(nd, tree.ndim, delta, eps, tree.ipoly, tree.iperiod, tree.norder,
                   tree.nboxes, tree.nlevels, tree.ltree, tree.itree, tree.iptr,
                   tree.centers, tree.boxsize)

    def set_fvals(self, fvals):
        f = np.ascontiguousarray(fvals, np.float64)
        if f.size != self.nd * self.npbox * self.nboxes:
            raise ValueError("fvals size mismatch")
        _check(lib().fgt_solver_set_fvals(self._h, _p(f)))

    def execute(self, stream=None, timed=False):
        _check(lib().fgt_solver_execute(self._h, C.c_void_p(stream or 0), C.c_int(1 if timed else 0)))

    def set_precision(self, mode):
        """0 FP64; 1 FP32 matrix-core contractions and FP32 expansions (eps >=
        1e-6); 2 BF16 hi/lo contractions (eps >= 1e-3).  Phases, weights and
        the output accumulation stay FP64."""
        _check(lib().fgt_solver_set_precision(self._h, C.c_int(mode)))

    def set_ifpgh(self, ifpgh):
        """1 potential, 2 + gradient, 3 + Hessian for later executes."""
        _check(lib().fgt_solver_set_ifpgh(self._h, C.c_int(ifpgh)))
        self.ifpgh = ifpgh

    def _shape(self, ncomp):
        tail = () if self.nd == 1 else (self.nd,)
        return (self.nboxes, self.npbox, ncomp) + tail

    def get_grad(self, out=None, accumulate=False):
        if out is None:
            out = np.zeros(self._shape(self.ndim))
        _check(lib().fgt_solver_get_grad(self._h, _p(out), C.c_int(1 if accumulate else 0)))
        return out

    def get_hess(self, out=None, accumulate=False):
        if out is None:
            out = np.zeros(self._shape(self.ndim * (self.ndim + 1) // 2))
        _check(lib().fgt_solver_get_hess(self._h, _p(out), C.c_int(1 if accumulate else 0)))
        return out

    def eval_targets(self, targs, ifpghtarg=1):
        """Values of the last execute at targets [nt, ndim]: (pote, grade, hesse)."""
        t = np.ascontiguousarray(targs, np.float64).reshape(-1, self.ndim)
        nt = t.shape[0]
        tail = () if self.nd == 1 else (self.nd,)
        nh = self.ndim * (self.ndim + 1) // 2
        pote = np.zeros((nt,) + tail)
        grade = np.zeros((nt, self.ndim) + tail) if ifpghtarg >= 2 else None
        hesse = np.zeros((nt, nh) + tail) if ifpghtarg >= 3 else None
        d = np.zeros(1)
        _check(lib().fgt_solver_eval_targets(self._h, C.c_int(nt), _p(t), C.c_int(ifpghtarg), _p(pote),
                                             _p(d if grade is None else grade), _p(d if hesse is None else hesse)))
        return pote, grade, hesse

    def get_pot(self, out=None, accumulate=False):
        if out is None:
            out = np.zeros((self.nboxes, self.npbox) if self.nd == 1 else (self.nboxes, self.npbox, self.nd))
        _check(lib().fgt_solver_get_pot(self._h, _p(out), C.c_int(1 if accumulate else 0)))
        return out

    def stats(self):
        s = _SolverStats()
        _check(lib().fgt_solver_stats_get(self._h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in _SolverStats._fields_ if k != "stage_ms"}
        d["stage_ms"] = list(s.stage_ms)
        return d

    @property
    def fvals_device_ptr(self):
        return lib().fgt_solver_fvals_device(self._h)

    @property
    def pot_device_ptr(self):
        return lib().fgt_solver_pot_device(self._h)

    def close(self):
        if self._h:
            lib().fgt_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_summary(tree, delta, eps, nd=1):
    """Host-only plan scalars (cutoff level, npw, list sizes) for a tree."""
    info = _PlanInfo()
    t = tree
    rc = lib().fgt_plan_summary(C.c_int(nd), C.c_int(t.ndim), C.c_double(delta), C.c_double(eps),
                                C.c_int(t.ipoly), C.c_int(t.iperiod), C.c_int(t.norder),
                                C.c_int(t.nboxes), C.c_int(t.nlevels), C.c_int(t.ltree),
                                _p(np.ascontiguousarray(t.itree, np.int32)),
                                _p(np.ascontiguousarray(t.iptr, np.int32)),
                                _p(np.ascontiguousarray(t.centers, np.float64)),
                                _p(np.ascontiguousarray(t.boxsize, np.float64)), C.byref(info))
    if rc != 0:
        raise FgtError(f"plan failed with status {rc}")
    return {k: getattr(info, k) for k, _ in _PlanInfo._fields_}


class _Problem(C.Structure):
    _fields_ = [("nd", C.c_int), ("ndim", C.c_int), ("delta", C.c_double), ("eps", C.c_double),
                ("ipoly", C.c_int), ("iperiod", C.c_int), ("norder", C.c_int), ("nboxes", C.c_int),
                ("nlevels", C.c_int), ("ltree", C.c_int), ("itree", C.c_void_p), ("iptr", C.c_void_p),
                ("centers", C.c_void_p), ("boxsize", C.c_void_p)]


class Batch:
    """Many independent box FGT problems planned and executed together: 3-D
    problems of norder 8/12/16 in shape classes, one launch per stage and
    class; the others on ``nstreams`` side streams; the whole transform
    replayed from a captured HIP graph.

    ``problems``: sequence of (tree, delta, eps) with ``tree`` a
    :class:`fgt_amd.tree.Tree` (or any object with the same attributes).
    """

    def __init__(self, problems, nstreams=4, nd=1):
        self._h = C.c_void_p()
        self._keep = []
        arr = (_Problem * len(problems))()
        self.shapes = []
        for i, (t, delta, eps) in enumerate(problems):
            it = np.ascontiguousarray(t.itree, np.int32)
            ip = np.ascontiguousarray(t.iptr, np.int32)
            ce = np.ascontiguousarray(t.centers, np.float64)
            bs = np.ascontiguousarray(t.boxsize, np.float64)
            self._keep += [it, ip, ce, bs]
            arr[i] = _Problem(nd, t.ndim, delta, eps, t.ipoly, t.iperiod, t.norder, t.nboxes, t.nlevels, t.ltree,
                              it.ctypes.data, ip.ctypes.data, ce.ctypes.data, bs.ctypes.data)
            npbox = t.norder ** t.ndim
            self.shapes.append((t.nboxes, npbox) if nd == 1 else (t.nboxes, npbox, nd))
        _check(lib().fgt_batch_create(C.c_int(len(problems)), arr, C.c_int(nstreams), C.byref(self._h)))

    def set_precision(self, mode, eps_min=0.0):
        """Per-problem precision: 0 FP64; 1 FP32 where eps >= 1e-6; 2 BF16
        hi/lo where eps >= 1e-3; 3 the cheapest mode each eps allows.
        Problems below eps_min or off the 3-D path stay FP64; returns how
        many are not FP64."""
        n = C.c_int()
        _check(lib().fgt_batch_set_precision(self._h, C.c_int(mode), C.c_double(eps_min), C.byref(n)))
        return n.value

    def set_fvals(self, i, fvals):
        f = np.ascontiguousarray(fvals, np.float64)
        if f.size != int(np.prod(self.shapes[i])):
            raise ValueError("fvals size mismatch")
        _check(lib().fgt_batch_set_fvals(self._h, C.c_int(i), _p(f)))

    def execute(self):
        _check(lib().fgt_batch_execute(self._h))

    def set_graph(self, on):
        """Replay the batch from a captured HIP graph (default) or enqueue
        every launch eagerly."""
        _check(lib().fgt_batch_set_graph(self._h, C.c_int(1 if on else 0)))

    def autotune(self, reps=3):
        """Pick, per shape class, the STEP 3 targets per block and the STEP 1 /
        STEP 5 x-stage variants by timing them on this batch's data."""
        _check(lib().fgt_batch_autotune(self._h, C.c_int(reps)))

    def profile(self):
        """One eager execute with per-stage events; per class: STEP 1..6 ms
        and the member problem indices."""
        _check(lib().fgt_batch_profile(self._h))
        out = []
        for k, c in enumerate(self.classes()):
            ms = (C.c_double * 6)()
            _check(lib().fgt_batch_class_stage_ms(self._h, C.c_int(k), ms))
            mem = (C.c_int * max(1, c["members"]))()
            _check(lib().fgt_batch_class_members(self._h, C.c_int(k), mem))
            out.append(dict(c, stage_ms=list(ms), member_ids=list(mem)[:c["members"]]))
        return out

    def classes(self):
        """Shape classes: dicts of norder, nd, precision, plane-wave cap,
        member count and the chosen group / x-stage variants."""
        out = []
        for k in range(lib().fgt_batch_nclasses(self._h)):
            v = (C.c_int * 12)()
            _check(lib().fgt_batch_class_info(self._h, C.c_int(k), v))
            out.append(dict(zip(("norder", "nd", "prec", "npw_cap", "members", "group", "xs1", "xs5",
                                 "step5_tasks", "near_targets", "m2l_ept", "m2l_brick"), list(v))))
        return out

    def get_pot(self, i):
        out = np.zeros(self.shapes[i])
        _check(lib().fgt_batch_get_pot(self._h, C.c_int(i), _p(out), C.c_int(0)))
        return out

    def stats(self):
        n, ms = C.c_longlong(), C.c_double()
        _check(lib().fgt_batch_stats(self._h, C.byref(n), C.byref(ms)))
        return {"nout": n.value, "last_ms": ms.value}

    def close(self):
        if self._h:
            lib().fgt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .tree import Tree, TreeTooLarge, vol_tree_build, vol_tree_build_fun  # noqa: E402  (host tree builder)

__all__ = ["boxfgt", "Solver", "Batch", "FgtError", "plan_summary", "Tree", "TreeTooLarge", "vol_tree_build", "vol_tree_build_fun", "lib",
           "last_status"]
