"""Host tree builder (fgt_vol_tree_*) vs trees built by the reference.

Every golden fixture holds a tree built by the reference's vol_tree_mem /
vol_tree_build for a stated Gaussian source and tolerance; the library's
builder must reproduce the arrays exactly (ids, levels, children,
colleagues) and the node values to rounding.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_names, load_case

import fgt_amd
from fgt_amd.tree import gauss_source

NAMES = [n for n in golden_names()]


@pytest.mark.parametrize("name", NAMES)
def test_tree_matches_reference(name):
    import os
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
