"""Shared test fixtures: golden-fixture loading, GPU marker, exact solutions."""
from __future__ import annotations

import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden_names():
    return sorted(os.path.basename(p)[7:-4] for p in glob.glob(os.path.join(GOLDEN, "boxfgt_*.npz")))


class Case:
    """A golden boxfgt case: reference tree arrays, inputs, reference output."""

    def __init__(self, name):
        z = np.load(os.path.join(GOLDEN, f"boxfgt_{name}.npz"))
        self.name = name
        for k in ("ndim", "norder", "ipoly", "iperiod", "nboxes", "nlevels", "ltree", "nd"):
            setattr(self, k, int(z[k]))
        self.eps = float(z["eps"])
        self.delta = float(z["delta"])
        self.itree = z["itree"].astype(np.int32)
        self.iptr = z["iptr"].astype(np.int32)
        self.centers = z["centers"]
        self.boxsize = z["boxsize"]
        self.leaves = z["leaves"]
        self.npbox = self.norder ** self.ndim
        shape = (self.nboxes, self.npbox) if self.nd == 1 else (self.nboxes, self.npbox, self.nd)
        self.fvals = np.zeros(shape)
        self.fvals[self.leaves] = z["fvals_leaf"]
        self.pot_ref = np.zeros(shape)
        self.pot_ref[self.leaves] = z["pot_leaf"]
        self.gauss = (z["gauss_c"], z["gauss_s"], z["gauss_a"])

    def tree_args(self):
        return (self.nboxes, self.nlevels, self.ltree, self.itree, self.iptr, self.centers, self.boxsize)


def load_case(name) -> Case:
    return Case(name)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def leaf_points(ndim, norder, ipoly, centers, boxsize, level, leaves):
    """Physical coordinates of the tensor nodes of each leaf (x fastest)."""
    from oracle import oracle
    x, _, _ = oracle.polyexps(ipoly, norder)
    grids = np.meshgrid(*([x] * ndim), indexing="ij")
    # node index m1 + n*m2 + n^2*m3 with m1 (x) fastest
    ref = np.stack([g.transpose(*range(ndim)[::-1]).ravel() for g in grids], axis=1)
    pts = []
    for b in leaves:
        bs = boxsize[level[b]] / 2
        pts.append(centers[b] + ref * bs)
    return np.stack(pts)    # (nleaf, npbox, ndim)


def exact_gauss(points, gauss, delta, iperiod, images=2):
    """Exact transform of sum_g a exp(-sum_k (y_k-c_k)^2/s_k) over the unit box.

    Free space: per axis int_{-1/2}^{1/2} e^{-(y-c)^2/s} e^{-(x-y)^2/delta} dy
    in closed form with erf; periodic: the full-line convolution of the
    periodised source, summed over lattice images.
    """
    from scipy.special import erf
    c, s, a = gauss
    out = np.zeros(points.shape[:-1])
    ndim = points.shape[-1]
    for g in range(len(a)):
        term = np.full(points.shape[:-1], a[g])
        for k in range(ndim):
            x = points[..., k]
            sk, ck = s[g, k], c[g, k]
            if iperiod == 0:
                tau = sk * delta / (sk + delta)
                mu = (ck * delta + x * sk) / (sk + delta)
                fac = np.exp(-(x - ck) ** 2 / (sk + delta)) * np.sqrt(np.pi * tau) / 2
                term = term * fac * (erf((0.5 - mu) / np.sqrt(tau)) - erf((-0.5 - mu) / np.sqrt(tau)))
            else:
                acc = 0.0
                for m in range(-images, images + 1):
                    acc = acc + np.exp(-(x - ck - m) ** 2 / (sk + delta))
                term = term * acc * np.sqrt(np.pi * sk * delta / (sk + delta))
        out += term
    return out


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


requires_gpu = pytest.mark.skipif(not gpu_available(), reason="no GPU")
