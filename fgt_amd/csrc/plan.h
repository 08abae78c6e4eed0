// plan.h -- host-side plan for one box FGT solve on one adaptive tree.
//
// The plan turns the reference tree arrays (1-based itree/iptr/centers/
// boxsize, as produced by vol_tree_build, src/common/tree_vol_coeffs_nd.f:444)
// into everything the device needs:
//   * the cutoff level / plane-wave length rules of boxfgt
//     (src/bfgt/boxfgt.f:132-162, src/common/fgtterms.f:72-211),
//   * the FP64 1-D tables (bfgt_volrouts.f, bfgt_pwrouts.f:1076-1543),
//   * the interaction lists (tree_routs_nd.f:864-1628),
//   * flat, level-ordered work lists for the kernels of each stage.
// All of it is computed in FP64 on the host; nothing here touches the GPU.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace fgt {

using cplx = std::complex<double>;

// Number of 1-D near-field table slots: offsets k = -6..6 (mrefinelev = 0).
constexpr int kLocSlots = 13;

struct TreeView {
    int ndim = 0, nboxes = 0, nlevels = 0, mc = 0, mnbors = 0;
    // 0-based box ids throughout; -1 = none.
    std::vector<int> lfirst, llast;        // per level, inclusive box range
    std::vector<int> level, parent, nchild;
    std::vector<int> child;                // mc per box
    std::vector<int> ncoll, coll;          // mnbors per box
    std::vector<double> centers;           // ndim per box
    std::vector<double> boxsize;           // nlevels+1
};

struct PlanInput {
    int nd, ndim;
    double delta, eps;
    int ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree;
    const int* itree;
    const int* iptr;
    const double* centers;
    const double* boxsize;
};

// Work lists; every "slot" indexes a pair (mp, loc) of plane-wave expansions.
struct Plan {
    // problem
    int nd = 1, ndim = 3, n = 8, ipoly = 0, iperiod = 0, npbox = 0;
    int nboxes = 0, nlevels = 0;
    double delta = 0, eps = 0, delta0 = 0, bs0 = 1;
    // cutoff / plane waves
    int npwlevel = 0, npw = 0, nh = 0, nexp = 0, nlevstart = 0, nmax = 0;
    int iper_eff = 0;            // periodicity used by the list/shift rules
    bool periodic_small = false; // periodic Fourier-series path (npwlevel <= 1)
    bool all_pw = false;         // every box carries PW expansions
    bool do_pw = false;          // some PW work exists
    double pmax = 0;
    TreeView tree;
    std::vector<char> ifpwexp;

    // 1-D tables (FP64)
    std::vector<double> ws, ts;             // npw
    std::vector<double> xnodes;             // n, on [-1,1]
    // p2pw[l][k][m]: value m -> frequency k (tab_poly2pw(m,k) at level l)
    std::vector<cplx> p2pw;
    // pw2pot[kind][l][i][k]: frequency k -> node i at level l; kind 0 =
    // tab_pw2pot, 1 = tab_pw2der, 2 = tab_pw2dxx (first/second derivative)
    std::vector<cplx> pw2pot;
    size_t pw2pot_ks = 0;   // kind stride (elements)
    // merge/split 1-D phases wms[klev-1][j] = exp(i ts_j xmin2)^(2^(klev-1))
    std::vector<cplx> wms;
    // M2L 1-D phases wm2l[o][j], o = offset+1 in {-1,0,1}
    std::vector<cplx> wm2l;
    // near-field tables per target level: loc[l][slot][k][j] (source value j
    // -> target node k), zero outside the band; band[l][slot][k] = (lo, hi)
    // 0-based inclusive row range (lo > hi: empty); colr[l][slot] = column range
    // loctab[kind][l][slot][k][j], kind as for pw2pot (tab_loc, tabx_loc,
    // tabxx_loc), all three masked to the same band
    std::vector<double> loctab;
    size_t loctab_ks = 0;   // kind stride (elements)
    std::vector<int> locband;   // 2 per (l, slot, k)
    std::vector<int> loccols;   // 2 per (l, slot)

    // expansion slots
    int nslots = 0;
    std::vector<int> slot_of_box;   // nboxes, -1 if no expansion

    // S1 / S5: PW leaves (box, level, slot), ordered by level
    std::vector<int> leaf_box, leaf_lev, leaf_slot;
    // S2 / S4 per level: for parents at level l (l = nlevels-1 .. nlevstart)
    // pc_off[l]..pc_off[l+1]: entries of (parent slot, child slot, sign k)
    std::vector<int> pc_level_off;          // nlevels+2
    std::vector<int> pc_parent, pc_child, pc_sign;
    // S2 grouping: one task per parent: par_off per level, par_slot, first entry
    std::vector<int> par_level_off;         // nlevels+2
    std::vector<int> par_slot, par_first, par_count;
    // S3: per target slot at nlevstart: m2l_off..., (src slot, 3^d index)
    std::vector<int> m2l_tgt, m2l_off, m2l_src, m2l_idx;
    // S6: per target leaf: src list with table slots per axis
    std::vector<int> dir_tgt, dir_lev, dir_off, dir_src, dir_code;

    // bookkeeping
    int64_t nleaves = 0;      // all leaves (the output-node count uses these)
    int64_t npw_leaves = 0;
    int64_t ndirect_pairs = 0;
};

// Modified list 1 of box b (raw, before the empty-band filter of the work
// lists); needs the plan's tree, cutoff level and plane-wave flags.
void modified_list1(const Plan& P, int b, std::vector<int>& out);

// Builds the plan.  Returns 0 on success, otherwise a nonzero status with a
// message in *err.
int make_plan(const PlanInput& in, Plan& plan, std::string* err);

// Pieces exposed for tests (restated from the reference, see plan.cpp).
void pw_terms(double bsize, double delta, double eps, int iperiod,
              double* pmax, int* npw);
void poly_exps(int ipoly, int n, double* x, double* w, double* u);
void sph_bessel(int nterms, double z, double* fjs);
// One level's value->PW table out[k*n + m] and PW->value table out[i*npw + k].
void make_p2pw(int ipoly, int n, int npw, const double* ws, const double* ts,
               double delta, double boxdim, cplx* out);
void make_pw2pot(int n, int npw, const double* ts, const double* xs, double delta,
                 double boxdim, cplx* out);
// The same table with its first/second derivative versions (out1/out2
// nullable), each out[i*npw + k].
void make_pw2pgh(int n, int npw, const double* ts, const double* xs, double delta,
                 double boxdim, cplx* out0, cplx* out1, cplx* out2);
// One target level's near-field tables: tab[slot][k][j] (masked to the band),
// band[slot][k] = (lo, hi) 0-based rows, cols[slot] = (first, last) columns;
// tabx/tabxx (nullable): the first/second derivative tables, same mask.
// mref: refinement depth the tables cover (mrefinelev of boxfgt.f:601-607):
// 0 for the solve (13 slots, offsets -6..6), 3 when the tree is refined
// afterwards (193 slots, offsets -96..96); slot = offset + loc_nloctab(mref).
void make_loctab_level(double eps, int ipoly, int n, double delta, double boxdim,
                       double* tab, int* band, int* cols, double* tabx = nullptr,
                       double* tabxx = nullptr, int mref = 0);
int loc_nloctab(int mref);

}  // namespace fgt
