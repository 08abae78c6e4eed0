// retree.h -- rewriting the tree after a solve (boxfgt with ifnewtree = 1).
//
// After STEPs 1-6 the reference checks every leaf's potential for
// resolution, refines the leaves that are under-resolved (evaluating the
// potential on the new boxes from the parent's local expansion and
// near-field list), coarsens sibling leaves whose merged potential is
// resolved, and restores the 2:1 level restriction with interpolated values
// (src/bfgt/boxfgt.f:1026-1337, src/common/tree_vol_coeffs_nd.f:1975-3232).
// This file holds the host side of that: the tree edits, the polynomial
// transforms on the outputs and the refinement decisions.  The potential of
// new boxes comes from a caller-supplied evaluator (the device path).
#pragma once

#include <functional>
#include <vector>

namespace fgt {

struct NewTree {
    int ndim = 3, mc = 8, mnb = 27, nd = 1, n = 8, npbox = 0, ipoly = 0, iperiod = 0, ifpgh = 1;
    int cap = 0;       // boxes the caller's arrays hold
    int nlevels = 0;
    std::vector<double> boxsize;   // levels 0..nlevels of the input tree
    // tree, 0-based ids, one entry per box of the current tree
    std::vector<int> level, parent, nchild, child, ncoll, coll;
    std::vector<double> centers;
    std::vector<int> lfirst, llast;
    bool coll_valid = true;        // ncoll/coll describe the current boxes
    // outputs (the caller's arrays, cap boxes each; grad/hess null below
    // their ifpgh) and the coefficients of the potential (cap boxes)
    double* pot = nullptr;
    double* grad = nullptr;
    double* hess = nullptr;
    std::vector<double> coefs;

    int count() const { return static_cast<int>(level.size()); }
    int nhess() const { return ndim * (ndim + 1) / 2; }
    size_t wp() const { return static_cast<size_t>(nd) * npbox; }
    size_t wg() const { return wp() * ndim; }
    size_t wh() const { return wp() * nhess(); }
};

// A box created by the refinement: its id, parent, level and the original
// leaf it descends from (whose near-field list and plane-wave flag it takes).
struct NewBox {
    int id, parent, level, origin;
};

// Potential (and gradient / Hessian when ifpgh >= 2 / 3) of each new box,
// nd*npbox (nd*ndim*npbox, nd*nhess*npbox) values per box in the order of
// `boxes`; the new boxes' centres are in T.centers.
using NewBoxEval = std::function<void(const NewTree& T, const std::vector<NewBox>& boxes, std::vector<double>& pot,
                                      std::vector<double>& grad, std::vector<double>& hess)>;

// Reads the caller's tree (1-based reference arrays; nboxes = capacity).
void nt_read(NewTree& T, const int* itree, const int* iptr, const double* centers, const double* boxsize);
// Writes the current tree back (laddr of its levels, the per-box sections
// for its boxes, colleagues when they were recomputed) and sets *nlevels.
void nt_write(const NewTree& T, int* itree, const int* iptr, double* centers, int* nlevels);

struct NewTreeStats {
    double rnorm = 0, reps = 0;
    int nboxes0 = 0, nrefine = 0, maxrefinelev = 0, nnew = 0, ndeleted = 0, nrebalance = 0;
    int skipped_finest = 0;   // flagged boxes at the finest level, left unrefined
    int generations = 0;
};

// STEPs 8-10 of boxfgt with ifnewtree = 1.  Throws std::length_error when
// the caller's arrays cannot hold the refined tree.
void nt_rewrite(NewTree& T, double eps, const std::vector<char>& ifpwexp, const NewBoxEval& eval,
                NewTreeStats* st);

}  // namespace fgt
