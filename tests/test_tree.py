"""Host tree builder (fgt_vol_tree_*) vs trees built by the reference.

Every golden fixture holds a tree built by the reference's vol_tree_mem /
vol_tree_build for a stated Gaussian source and tolerance; the library's
builder must reproduce the arrays exactly (ids, levels, children,
colleagues) and the node values to rounding.  The tree_*.npz fixtures do
the same for a compiled source function (tests/csrc/sources.c) under each
refinement norm, eta, several densities, zk and ifnewtree.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import golden_names, load_case

import fgt_amd
from fgt_amd.tree import gauss_source

NAMES = [n for n in golden_names()]


@pytest.mark.parametrize("name", NAMES)
def test_tree_matches_reference(name):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"boxfgt_{name}.npz"))
    c = load_case(name)
    g = gauss_source(z["gauss_c"], z["gauss_s"], z["gauss_a"])
    t = fgt_amd.vol_tree_build(c.ndim, c.norder, c.ipoly, c.iperiod, float(z["tree_eps"]), g)
    assert (t.nboxes, t.nlevels, t.ltree) == (c.nboxes, c.nlevels, c.ltree)
    np.testing.assert_array_equal(t.iptr, c.iptr)
    np.testing.assert_array_equal(t.itree, c.itree)
    np.testing.assert_array_equal(t.centers, c.centers)
    np.testing.assert_array_equal(t.boxsize, c.boxsize)
    fl = c.fvals[c.leaves] if c.nd == 1 else c.fvals[c.leaves][..., 0]
    np.testing.assert_allclose(t.fvals[c.leaves], fl, rtol=1e-13, atol=1e-13 * np.abs(fl).max())


def test_tree_level_restriction_and_colleagues():
    """2:1 balance and symmetric colleague lists on a sharper source."""
    rng = np.random.default_rng(11)
    g = gauss_source(rng.uniform(-0.4, 0.4, (3, 3)), np.full((3, 3), 1e-3), rng.uniform(0.5, 1, 3))
    t = fgt_amd.vol_tree_build(3, 6, 0, 0, 1e-5, g)
    assert t.nlevels >= 4
    lev, nch = t.level, t.nchild
    ncoll = t.section(6)[:, 0]
    coll = t.section(7, 27)
    # colleagues are symmetric and at the same level
    for b in range(t.nboxes):
        for j in coll[b, :ncoll[b]] - 1:
            assert lev[j] == lev[b]
            assert b + 1 in coll[j, :ncoll[j]]
    # leaves touching each other differ by at most one level
    leaves = np.nonzero(nch == 0)[0]
    cen = t.centers[leaves]
    half = t.boxsize[lev[leaves]] / 2
    d = np.abs(cen[:, None, :] - cen[None, :, :]).max(axis=2)
    touch = d <= (half[:, None] + half[None, :]) * 1.0001
    dl = np.abs(lev[leaves][:, None].astype(int) - lev[leaves][None, :].astype(int))
    assert np.all(dl[touch] <= 1)


# ---- trees for a caller's source function (vol_tree_mem's fun, iptype, eta, nd, zk, ifnewtree) ----

FUN_NAMES = sorted(os.path.basename(p)[len("tree_"):-4]
                   for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "tree_*.npz")))


@pytest.fixture(scope="module")
def test_source(tmp_path_factory):
    """tests/csrc/sources.c compiled the way the fixtures' generator compiles
    it, so this library's builder sees the values the reference's saw."""
    d = tmp_path_factory.mktemp("src")
    so = str(d / "libtestsource.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-o", so,
                    os.path.join(os.path.dirname(__file__), "csrc", "sources.c"), "-lm"], check=True)
    L = ctypes.CDLL(so)
    return L, L.test_source


def _build_fun(z, fun, ifnewtree=None):
    return fgt_amd.vol_tree_build_fun(int(z["ndim"]), int(z["norder"]), int(z["ipoly"]), int(z["iperiod"]),
                                      float(z["eps"]), fun, z["dpars"], z["ipars"], nd=int(z["nd"]),
                                      iptype=int(z["iptype"]), eta=float(z["eta"]), zk=float(z["zk"]),
                                      ifnewtree=int(z["ifnewtree"]) if ifnewtree is None else ifnewtree)


@pytest.mark.parametrize("name", FUN_NAMES)
def test_tree_for_source_function_matches_reference(name, test_source):
    """Every array of the reference's tree for the same compiled source: the
    refinement decisions under each norm (iptype 0 max, 1 L1, 2 L2), eta, nd
    densities, zk's forced refinement and ifnewtree's room."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", f"tree_{name}.npz"))
    _, fun = test_source
    t = _build_fun(z, fun)
    assert (t.nboxes, t.nlevels, t.ltree) == (int(z["nboxes"]), int(z["nlevels"]), int(z["ltree"]))
    np.testing.assert_array_equal(t.iptr, z["iptr"])
    np.testing.assert_array_equal(t.itree, z["itree"])
    nused = len(z["centers"])
    assert t.nused == nused
    np.testing.assert_array_equal(t.centers[:nused], z["centers"])
    np.testing.assert_array_equal(t.centers[nused:], 0.0)
    np.testing.assert_array_equal(t.boxsize, z["boxsize"])
    ref = z["fvals"]
    np.testing.assert_allclose(t.fvals[:nused], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_array_equal(t.fvals[nused:], 0.0)


def test_source_function_tree_feeds_the_solver_layout(test_source):
    """nd = 2 source: fvals (nboxes, npbox, nd), density 0 equal to the nd = 1
    tree's values on the same boxes."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tree_f2_max_nd2_per.npz"))
    _, fun = test_source
    t2 = _build_fun(z, fun)
    assert t2.fvals.shape == (t2.nboxes, t2.npbox, 2)
    ip = np.array(z["ipars"])
    f = np.zeros(2)
    x = np.array([0.1, -0.2])
    ndv = ctypes.c_int(2)
    fun(ctypes.byref(ndv), x.ctypes.data_as(ctypes.c_void_p), np.asarray(z["dpars"]).ctypes.data_as(ctypes.c_void_p),
        None, ip.astype(np.int32).ctypes.data_as(ctypes.c_void_p), f.ctypes.data_as(ctypes.c_void_p))
    assert f[1] == pytest.approx(f[0] * (1 + np.cos(3 * 0.1 + 1)))


def test_source_function_rejects_bad_arguments(test_source):
    _, fun = test_source
    dp = np.array([0.0, 0.0, 0.01, 0.01, 1.0])
    ip = np.array([2, 1, 0], np.int32)
    for kw in ({"iptype": 3}, {"nd": 0}, {"ifnewtree": 2}):
        args = dict(nd=1, iptype=2, ifnewtree=0)
        args.update(kw)
        with pytest.raises(fgt_amd.FgtError):
            fgt_amd.vol_tree_build_fun(2, 8, 0, 0, 1e-6, fun, dp, ip, **args)


def test_python_source_function_builds_the_same_tree(test_source):
    """A Python function as the source (a ctypes callback with the
    reference's by-reference convention, called from the builder's threads)
    gives the tree of the compiled source of the same values."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tree_f2_l1_eta1.npz"))
    dp = np.array(z["dpars"])
    ip = np.array(z["ipars"], np.int32)
    ndim, ng = int(ip[0]), int(ip[1])
    c = dp.reshape(ng, 2 * ndim + 1)

    proto = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_double))

    def py_source(nd, xyz, dpars, zpars, ipars, f):
        x = [xyz[k] for k in range(ndim)]
        acc = 0.0
        for g in range(ng):
            rr = 0.0
            for k in range(ndim):
                d = x[k] - c[g, k]
                rr += d * d / c[g, ndim + k]
            acc += c[g, 2 * ndim] * np.exp(-rr)
        for q in range(nd[0]):
            f[q] = acc * (1.0 + q * np.cos(3.0 * x[0] + q))

    cb = proto(py_source)
    t = fgt_amd.vol_tree_build_fun(ndim, int(z["norder"]), int(z["ipoly"]), int(z["iperiod"]), float(z["eps"]), cb,
                                   dp, ip, nd=1, iptype=int(z["iptype"]), eta=float(z["eta"]))
    np.testing.assert_array_equal(t.itree, z["itree"])
    np.testing.assert_allclose(t.fvals[:t.nused], z["fvals"], rtol=1e-13, atol=1e-13 * np.abs(z["fvals"]).max())
