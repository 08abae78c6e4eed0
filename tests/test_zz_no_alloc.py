"""The no-allocation entry point fgt_boxfgt_noalloc (device fvals / pot,
every other device buffer carved from the library's static per-GPU
workspace) against the boxfgt drop-in on the same inputs, through the C ABI.

Two calls in a row, the second on a smaller problem (the workspace keeps the
first call's contents), a problem larger than the workspace (refused with
status 2 before anything is launched), and a further call after the refusal.
Bar: relative l2 <= 1e-13 against the drop-in (both FP64; the summation
order of the STEP 5 and STEP 6 contributions differs)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import load_case, rel_l2

import fgt_amd

pytestmark = pytest.mark.gpu


def _noalloc(c, fvals_dev, pot_dev, stream):
    L = fgt_amd.lib()
    L.fgt_boxfgt_noalloc.restype = C.c_int
    it = np.ascontiguousarray(c.itree, np.int32)
    ip = np.ascontiguousarray(c.iptr, np.int32)
    ce = np.ascontiguousarray(c.centers, np.float64)
    bs = np.ascontiguousarray(c.boxsize, np.float64)
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    return L.fgt_boxfgt_noalloc(C.c_int(c.nd), C.c_int(c.ndim), C.c_double(c.delta), C.c_double(c.eps),
                                C.c_int(c.ipoly), C.c_int(c.iperiod), C.c_int(c.norder), C.c_int(c.nboxes),
                                C.c_int(c.nlevels), C.c_int(c.ltree), P(it), P(ip), P(ce), P(bs),
                                C.c_void_p(fvals_dev), C.c_void_p(pot_dev), C.c_void_p(stream))


def _dropin(c, p0):
    pot = p0.copy()
    fgt_amd.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels,
                   c.ltree, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, 1, pot)
    return pot


def _run(c, p0, stream):
    import torch
    fv = torch.from_numpy(np.ascontiguousarray(c.fvals)).cuda()
    pot = torch.from_numpy(p0.copy()).cuda()
    torch.cuda.synchronize()
    rc = _noalloc(c, fv.data_ptr(), pot.data_ptr(), stream)
    out = pot.cpu().numpy()
    return rc, out


def test_noalloc_matches_dropin_twice_then_refuses_oversize():
    import torch
    st = torch.cuda.Stream()
    rng = np.random.default_rng(5)
    for name in ("c1_small", "c3_n12_e3_d2"):   # the second problem is smaller
        print(f"[gpu] fgt_boxfgt_noalloc vs boxfgt drop-in: {name}", flush=True)
        c = load_case(name)
        ref = _dropin(c, np.zeros_like(c.fvals))
        rc, out = _run(c, np.zeros_like(c.fvals), st.cuda_stream)
        assert rc == 0, fgt_amd.last_status()
        err = rel_l2(out, ref)
        print(f"[gpu]   rel l2 vs drop-in = {err:.3e}", flush=True)
        assert err <= 1e-13
        # pot is accumulated, not overwritten: same as the drop-in adding
        # onto the same values
        p0 = rng.uniform(-1.0, 1.0, c.fvals.shape)
        rc, out2 = _run(c, p0, st.cuda_stream)
        assert rc == 0
        assert rel_l2(out2, _dropin(c, p0)) <= 1e-13
    # larger than the workspace: status 2, the caller's pot untouched
    c = load_case("pgh3_per_n8")
    print("[gpu] fgt_boxfgt_noalloc: oversize problem is refused", flush=True)
    p0 = rng.uniform(-1.0, 1.0, c.fvals.shape)
    rc, out = _run(c, p0, st.cuda_stream)
    assert rc == 2
    assert "no-allocation scratch" in fgt_amd.last_status()[1]
    assert np.array_equal(out, p0)
    # and a valid call after the refusal
    c = load_case("tgt3_n8")
    ref = _dropin(c, np.zeros_like(c.fvals))
    rc, out = _run(c, np.zeros_like(c.fvals), 0)
    assert rc == 0
    assert rel_l2(out, ref) <= 1e-13
