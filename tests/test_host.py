"""CPU tests of the product library: exports, host planner tables, error paths.

No kernels run here (no GPU in the build container); the planner's FP64
tables and lists are compared with the reference's (tables.npz) and with the
oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden_names, load_case

import fgt_amd
from oracle import oracle

TABLES = np.load(os.path.join(GOLDEN, "tables.npz"))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def declared_symbols():
    names = []
    for h in ("fgt.h", "fgt_host.h", "fgt_tree.h"):
        path = os.path.join(ROOT, "include", h)
        if not os.path.exists(path):
            continue
        txt = open(path).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(fgt_\w+|boxfgt_)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = fgt_amd.lib()
    syms = declared_symbols()
    assert "boxfgt_" in syms and "fgt_solver_execute" in syms
    for s in syms:
        assert hasattr(L, s), s


def test_pwterms_match_reference():
    L = fgt_amd.lib()
    for bsize, delta, eps, iper, pmax_ref, npw_ref in TABLES["pwterms"]:
        pmax, npw = C.c_double(), C.c_int()
        L.fgt_host_pwterms(C.c_double(bsize), C.c_double(delta), C.c_double(eps), C.c_int(int(iper)),
                           C.byref(pmax), C.byref(npw))
        assert npw.value == int(npw_ref)
        if iper == 0:
            assert pmax.value == pytest.approx(pmax_ref, rel=1e-15)


@pytest.mark.parametrize("ipoly", [0, 1])
@pytest.mark.parametrize("n", [4, 8, 12, 16])
def test_polyexps_match_reference(ipoly, n):
    x, w, u = np.zeros(n), np.zeros(n), np.zeros(n * n)
    fgt_amd.lib().fgt_host_polyexps(C.c_int(ipoly), C.c_int(n), _p(x), _p(w), _p(u))
    np.testing.assert_allclose(x, TABLES[f"exps_{ipoly}_{n}_x"], atol=2e-15)
    np.testing.assert_allclose(u.reshape(n, n), TABLES[f"exps_{ipoly}_{n}_u"], rtol=1e-12, atol=1e-13)


def test_bessel_match_reference():
    for z, row in zip(TABLES["bessel_z"], TABLES["bessel_j"]):
        f = np.zeros(17)
        fgt_amd.lib().fgt_host_sph_bessel(C.c_int(16), C.c_double(z), _p(f))
        np.testing.assert_allclose(f, row, rtol=1e-12, atol=1e-300)


def _keys(prefix):
    return sorted(k[len(prefix):] for k in TABLES.files if k.startswith(prefix))


@pytest.mark.parametrize("key", _keys("p2pw_"))
def test_pw_tables_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, boxdim = int(ipoly), int(n), float(delta), float(boxdim)
    ws, ts = TABLES[f"ws_{key}"], TABLES[f"ts_{key}"]
    npw = len(ts)
    t1 = np.zeros(2 * n * npw)
    fgt_amd.lib().fgt_host_poly2pw(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(ws), _p(ts),
                                   C.c_double(delta), C.c_double(boxdim), _p(t1))
    ref1 = TABLES[f"p2pw_{key}"]
    np.testing.assert_allclose(t1, ref1, rtol=0, atol=1e-13 * np.abs(ref1).max())
    t2 = np.zeros(2 * n * npw)
    fgt_amd.lib().fgt_host_pw2pot(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(ts), C.c_double(delta),
                                  C.c_double(boxdim), _p(t2))
    np.testing.assert_allclose(t2, TABLES[f"pw2pot_{key}"], rtol=0, atol=1e-14)


@pytest.mark.parametrize("key", _keys("loctab_"))
def test_loctab_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, eps, boxdim = int(ipoly), int(n), float(delta), float(eps), float(boxdim)
    tab = np.zeros(n * n * 13)
    ind = np.zeros(2 * (n + 1) * 13, np.int32)
    fgt_amd.lib().fgt_host_loctab(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                                  C.c_double(boxdim), _p(tab), _p(ind))
    tab = tab.reshape(13, n, n)
    ind = ind.reshape(13, n + 1, 2)
    rtab = TABLES[f"loctab_{key}"].reshape(13, n, n)
    rind = TABLES[f"locind_{key}"].reshape(13, n + 1, 2)
    for k in range(-6, 7):
        if abs(k) == 5:
            continue
        s = k + 6
        np.testing.assert_array_equal(ind[s], rind[s], err_msg=f"band k={k}")
        # the product zeroes entries outside the band; compare inside it
        mask = np.zeros((n, n), bool)
        lo_c, hi_c = rind[s, n]
        for col in range(n):
            lo, hi = rind[s, col]
            if lo_c <= col + 1 <= hi_c and lo >= 1:
                mask[col, lo - 1:hi] = True
        scale = max(np.abs(rtab[s]).max(), 1e-300)
        np.testing.assert_allclose(tab[s][mask], rtab[s][mask], rtol=0, atol=1e-13 * scale)
        assert np.all(tab[s][~mask] == 0)


@pytest.mark.parametrize("key", _keys("pw2der_"))
def test_pw2pgh_tables_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, boxdim = int(ipoly), int(n), float(delta), float(boxdim)
    ts = TABLES[f"ts_{key}"]
    npw = len(ts)
    t = [np.zeros(2 * n * npw) for _ in range(3)]
    fgt_amd.lib().fgt_host_pw2pgh(C.c_int(ipoly), C.c_int(n), C.c_int(npw), _p(ts), C.c_double(delta),
                                  C.c_double(boxdim), _p(t[0]), _p(t[1]), _p(t[2]))
    for got, nm in zip(t, ("pw2pot", "pw2der", "pw2dxx")):
        ref = TABLES[f"{nm}_{key}"]
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * max(1.0, np.abs(ref).max()), err_msg=nm)


@pytest.mark.parametrize("key", _keys("loctabx_"))
def test_loctab_derivatives_match_reference(key):
    ipoly, n, delta, eps, boxdim, iper = key.split("_")
    ipoly, n, delta, eps, boxdim = int(ipoly), int(n), float(delta), float(eps), float(boxdim)
    tx, txx = np.zeros(n * n * 13), np.zeros(n * n * 13)
    fgt_amd.lib().fgt_host_loctab_der(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                                      C.c_double(boxdim), _p(tx), _p(txx))
    rind = TABLES[f"locind_{key}"].reshape(13, n + 1, 2)
    for got, nm in ((tx.reshape(13, n, n), "loctabx"), (txx.reshape(13, n, n), "loctabxx")):
        rtab = TABLES[f"{nm}_{key}"].reshape(13, n, n)
        for k in range(-6, 7):
            if abs(k) == 5:
                continue
            s = k + 6
            mask = np.zeros((n, n), bool)
            lo_c, hi_c = rind[s, n]
            for col in range(n):
                lo, hi = rind[s, col]
                if lo_c <= col + 1 <= hi_c and lo >= 1:
                    mask[col, lo - 1:hi] = True
            scale = max(np.abs(rtab[s]).max(), 1e-300)
            np.testing.assert_allclose(got[s][mask], rtab[s][mask], rtol=0, atol=1e-12 * scale,
                                       err_msg=f"{nm} k={k}")
            assert np.all(got[s][~mask] == 0)


@pytest.mark.parametrize("name", golden_names())
def test_plan_summary_matches_oracle(name):
    c = load_case(name)
    L = fgt_amd.lib()

    class T:  # minimal tree view for plan_summary
        pass
    t = T()
    t.ndim, t.ipoly, t.iperiod, t.norder = c.ndim, c.ipoly, c.iperiod, c.norder
    t.nboxes, t.nlevels, t.ltree, t.itree, t.iptr = c.nboxes, c.nlevels, c.ltree, c.itree, c.iptr
    t.centers, t.boxsize = c.centers, c.boxsize
    info = fgt_amd.plan_summary(t, c.delta, c.eps, nd=c.nd)
    _, oinfo, _ = oracle.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder,
                                c.nboxes, c.nlevels, c.itree, c.iptr, c.centers, c.boxsize, c.fvals)
    assert info["npwlevel"] == oinfo.npwlevel
    assert info["npw"] == oinfo.npw
    assert info["nexp"] == oinfo.nexp
    assert info["nleaves"] == len(c.leaves)
    del L


def _call_boxfgt(c, ifpgh=1, ifpghtarg=0, ifnewtree=0, ipoly=None, norder=None):
    L = fgt_amd.lib()
    pot = np.zeros_like(c.fvals)
    ti = np.zeros(20)
    d = np.zeros(1)
    I = lambda v: C.byref(C.c_int(v))  # noqa: E731
    D = lambda v: C.byref(C.c_double(v))  # noqa: E731
    rc = L.fgt_boxfgt(I(c.nd), I(c.ndim), D(c.delta), D(c.eps), I(c.ipoly if ipoly is None else ipoly),
                      I(c.iperiod), I(c.norder if norder is None else norder), I(c.npbox), I(c.nboxes),
                      I(c.nlevels), I(c.ltree), _p(c.itree), _p(c.iptr), _p(c.centers), _p(c.boxsize),
                      _p(c.fvals), I(ifpgh), _p(pot), _p(d), _p(d), I(ifnewtree), I(0), _p(d), I(ifpghtarg),
                      _p(d), _p(d), _p(d), _p(ti))
    return rc, pot


@pytest.mark.parametrize("kw", [dict(ifpgh=4), dict(ifpgh=0), dict(ifpghtarg=5), dict(ifpghtarg=-1),
                                dict(ifnewtree=2), dict(ipoly=3), dict(norder=40)])
def test_unsupported_options_report_status_and_leave_outputs(kw):
    c = load_case("c1_small")
    rc, pot = _call_boxfgt(c, **kw)
    assert rc != 0
    st, msg = fgt_amd.last_status()
    assert st == rc and msg
    assert np.all(pot == 0)


def test_corrupt_tree_is_rejected():
    c = load_case("c1_small")
    c.iptr = c.iptr.copy()
    c.iptr[3] += 1
    rc, _ = _call_boxfgt(c)
    assert rc == 2
    assert "iptr" in fgt_amd.last_status()[1]


@pytest.mark.parametrize("what", ["parent_out_of_range", "parent_wrong_level", "level_vs_laddr"])
def test_corrupt_parent_or_level_is_rejected(what):
    """A leaf's parent id out of range (or one level off), or a level entry
    that disagrees with laddr, is refused with status 2 before any list is
    built (an out-of-range parent used to be read as an index)."""
    c = load_case("c1_small")
    it = c.itree.copy()
    nlev = c.nlevels
    ilevel = c.iptr[1] - 1
    iparent = c.iptr[2] - 1
    nchild = c.iptr[3] - 1
    lev = it[ilevel:ilevel + c.nboxes]
    leaves = [b for b in range(c.nboxes) if it[nchild + b] == 0 and lev[b] >= 2]
    b = leaves[len(leaves) // 2]
    if what == "parent_out_of_range":
        it[iparent + b] = 100000000
    elif what == "parent_wrong_level":
        it[iparent + b] = 1          # the root, two or more levels up
    else:
        it[ilevel + b] = (lev[b] + 1) % (nlev + 1)
    c.itree = it
    with pytest.raises(fgt_amd.FgtError):
        fgt_amd.plan_summary(c, c.delta, c.eps)
    rc, _ = _call_boxfgt(c)
    assert rc == 2
    msg = fgt_amd.last_status()[1]
    assert "iparent" in msg or "level" in msg


MREF3 = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tables_mref3.npz"))


@pytest.mark.parametrize("key", sorted(k[len("loctab_"):] for k in MREF3.files if k.startswith("loctab_")))
def test_loctab_mrefinelev3_match_reference(key):
    """The 193-offset near-field tables of a refined tree (mk_loctab_all with
    mrefinelev = 3, reference values) - bands bit for bit, values to 1e-13."""
    ipoly, n, delta, eps, boxdim = key.split("_")
    ipoly, n, delta, eps, boxdim = int(ipoly), int(n), float(delta), float(eps), float(boxdim)
    ns = 193
    tab = np.zeros(n * n * ns)
    ind = np.zeros(2 * (n + 1) * ns, np.int32)
    fgt_amd.lib().fgt_host_loctab_mref(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                                       C.c_double(boxdim), C.c_int(3), _p(tab), _p(ind))
    tab, ind = tab.reshape(ns, n, n), ind.reshape(ns, n + 1, 2)
    rtab, rind = MREF3[f"loctab_{key}"].reshape(ns, n, n), MREF3[f"locind_{key}"].reshape(ns, n + 1, 2)
    made = 0
    for s in range(ns):
        if not rind[s].any() and not rtab[s].any():
            continue   # an offset no geometric case produces (left unset by the reference)
        made += 1
        np.testing.assert_array_equal(ind[s], rind[s], err_msg=f"band k={s - 96}")
        mask = np.zeros((n, n), bool)
        lo_c, hi_c = rind[s, n]
        for col in range(n):
            lo, hi = rind[s, col]
            if lo_c <= col + 1 <= hi_c and lo >= 1:
                mask[col, lo - 1:hi] = True
        scale = max(np.abs(rtab[s]).max(), 1e-300)
        np.testing.assert_allclose(tab[s][mask], rtab[s][mask], rtol=0, atol=1e-13 * scale)
        assert np.all(tab[s][~mask] == 0)
    assert made > 100
    # the mrefinelev = 0 offsets are the same tables at 8x the slot index
    t0 = np.zeros(n * n * 13)
    i0 = np.zeros(2 * (n + 1) * 13, np.int32)
    fgt_amd.lib().fgt_host_loctab(C.c_double(eps), C.c_int(ipoly), C.c_int(n), C.c_double(delta),
                                  C.c_double(boxdim), _p(t0), _p(i0))
    t0, i0 = t0.reshape(13, n, n), i0.reshape(13, n + 1, 2)
    for k in range(-6, 7):
        if abs(k) != 5:
            np.testing.assert_array_equal(t0[k + 6], tab[8 * k + 96])
            np.testing.assert_array_equal(i0[k + 6], ind[8 * k + 96])
