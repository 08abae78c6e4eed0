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
    """Fixtures of the solve on a fixed tree (the ifnewtree = 1 fixtures,
    boxfgt_nt*, hold rewritten trees and are read by their own test)."""
    return sorted(os.path.basename(p)[7:-4] for p in glob.glob(os.path.join(GOLDEN, "boxfgt_*.npz"))
                  if not os.path.basename(p).startswith("boxfgt_nt"))


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
        self.ifpgh = int(z["ifpgh"]) if "ifpgh" in z else 1
        tail = () if self.nd == 1 else (self.nd,)
        self.grad_ref = self.hess_ref = None
        if self.ifpgh >= 2 and "grad_leaf" in z:
            self.grad_ref = np.zeros((self.nboxes, self.npbox, self.ndim) + tail)
            self.grad_ref[self.leaves] = z["grad_leaf"]
        if self.ifpgh >= 3 and "hess_leaf" in z:
            self.hess_ref = np.zeros((self.nboxes, self.npbox, self.ndim * (self.ndim + 1) // 2) + tail)
            self.hess_ref[self.leaves] = z["hess_leaf"]
        # extra targets (ifpghtarg >= 1): targets and the reference's values there
        self.ifpghtarg = int(z["ifpghtarg"]) if "ifpghtarg" in z else 0
        self.targs = z["targs"] if "targs" in z else None
        self.pote_ref = z["pote"] if "pote" in z else None
        self.grade_ref = z["grade"] if "grade" in z else None
        self.hesse_ref = z["hesse"] if "hesse" in z else None

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


def _factor_deriv(x, ck, sk, delta, iperiod, order, images=2):
    """d^order/dx^order of the 1-D factor of exact_gauss (order 0, 1, 2).

    Periodic: the lattice-image sum of Gaussians of variance S = s + delta.
    Free space: the product of the two Gaussians is A(x) exp(-(y-mu)^2/tau)
    with A = exp(-(x-c)^2/S), mu = (c delta + x s)/S, tau = s delta/S, so the
    factor is A(x) B(mu(x)) with B the erf difference over [-1/2, 1/2];
    B' and B'' are Gaussians at the interval ends.
    """
    from scipy.special import erf
    S = sk + delta
    if iperiod:
        acc = np.zeros_like(x)
        for m in range(-images, images + 1):
            d = x - ck - m
            e = np.exp(-d * d / S)
            acc = acc + (e if order == 0 else (-2 * d / S) * e if order == 1 else (4 * d * d / S ** 2 - 2 / S) * e)
        return acc * np.sqrt(np.pi * sk * delta / S)
    tau = sk * delta / S
    r = sk / S
    mu = (ck * delta + x * sk) / S
    A = np.exp(-(x - ck) ** 2 / S)
    A1 = -2 * (x - ck) / S * A
    A2 = (4 * (x - ck) ** 2 / S ** 2 - 2 / S) * A
    B = np.sqrt(np.pi * tau) / 2 * (erf((0.5 - mu) / np.sqrt(tau)) - erf((-0.5 - mu) / np.sqrt(tau)))
    e1, e2 = np.exp(-(0.5 - mu) ** 2 / tau), np.exp(-(0.5 + mu) ** 2 / tau)
    B1 = -e1 + e2
    B2 = -2 * (0.5 - mu) / tau * e1 - 2 * (0.5 + mu) / tau * e2
    if order == 0:
        return A * B
    if order == 1:
        return A1 * B + A * B1 * r
    return A2 * B + 2 * A1 * B1 * r + A * B2 * r * r


def exact_gauss_deriv(points, gauss, delta, iperiod, orders):
    """Partial derivative of exact_gauss of the per-axis orders (tuple of 0/1/2)."""
    c, s, a = gauss
    out = np.zeros(points.shape[:-1])
    ndim = points.shape[-1]
    for g in range(len(a)):
        term = np.full(points.shape[:-1], a[g])
        for k in range(ndim):
            term = term * _factor_deriv(points[..., k], c[g, k], s[g, k], delta, iperiod, orders[k])
        out += term
    return out


def pgh_orders(ndim):
    """Per-axis derivative orders of grad and Hessian components, in the
    reference's component order (Hessian: 1-D xx; 2-D xx,xy,yy; 3-D
    xx,yy,zz,xy,xz,yz)."""
    grad = [tuple(1 if a == k else 0 for a in range(ndim)) for k in range(ndim)]
    hess = {1: [(2,)], 2: [(2, 0), (1, 1), (0, 2)],
            3: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]}[ndim]
    return grad, hess


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


requires_gpu = pytest.mark.skipif(not gpu_available(), reason="no GPU")
