// retree.cpp -- host side of boxfgt with ifnewtree = 1 (see retree.h).
//
// STEP 8  refine:  src/bfgt/boxfgt.f:1026-1259 (decisions, box creation,
//                  fgt_vol_tree_reorg tree_vol_coeffs_nd.f:1975-2155)
// STEP 9  coarsen: boxfgt.f:1261-1307, vol_tree_coarsen 2160-2381,
//                  fgt_vol_tree_reorg_after_coarsen 2386-2567
// STEP 10 2:1 rebalance with interpolation: boxfgt.f:1309-1337,
//                  vol_tree_fix_lr_interp 2573-2914, vol_tree_reorg_pgh
//                  2918-3123, vol_tree_refine_boxes_flag_interp 3129-3227,
//                  vol_updateflags 1735-1853, computecoll
//                  (src/common/tree_routs_nd.f:106-242)
// Helpers: tens_prod_get_rmask (polytens.f:1579-1616), fun_err
// (tree_vol_coeffs_nd.f:860-943), treedata_lpnorm / treedata_trans_nd
// (tree_data_routs_nd.f:1003-1098, 16-95), ortho_trans_nd / ortho_eval_nd
// (tensor_prod_routs_nd.f:21-500), the child / parent interpolation
// matrices (tree_data_routs_nd.f:1477-1620) and tree_data_c2p_copy_nd
// (1100-1300).  Box numbering, flag passes and the order of every edit
// follow the reference so that the rewritten tree is the same array for
// array.
#include "retree.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "plan.h"

namespace fgt {

namespace {

// sign of axis k of child j: + when bit k of j is set (get_child_box_sign,
// src/common/pts_tree_nd.f:1418-1464)
inline double child_sign(int j, int k) { return ((j >> k) & 1) ? 1.0 : -1.0; }

void poly_vals(int ipoly, double x, int m, double* p) {
    p[0] = 1.0;
    if (m == 0) return;
    p[1] = x;
    for (int k = 1; k < m; ++k)
        p[k + 1] = ipoly == 0 ? ((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1) : 2 * x * p[k] - p[k - 1];
}

// out(nd, n2^d) from in(nd, n1^d): along x, then y, then z,
// out_k = sum_{k1} A[a][k * n1 + k1] in_{k1}, k1 ascending.
void tensor_apply(int ndim, int nd, int n1, int n2, const double* in, double* out, const double* const* A) {
    std::vector<double> cur(in, in + static_cast<size_t>(nd) * static_cast<size_t>(std::pow(n1, ndim) + 0.5));
    int dims[3] = {n1, n1, n1};
    for (int a = 0; a < ndim; ++a) {
        // cur layout (nd, dims[0], dims[1], dims[2]) with ndim axes
        size_t inner = nd, outer = 1;
        for (int b = 0; b < a; ++b) inner *= dims[b];
        for (int b = a + 1; b < ndim; ++b) outer *= dims[b];
        std::vector<double> nxt(inner * n2 * outer);
        const double* M = A[a];
        for (size_t o = 0; o < outer; ++o)
            for (int k = 0; k < n2; ++k)
                for (size_t i = 0; i < inner; ++i) {
                    double s = 0;
                    for (int k1 = 0; k1 < dims[a]; ++k1) s += M[k * dims[a] + k1] * cur[(o * dims[a] + k1) * inner + i];
                    nxt[(o * n2 + k) * inner + i] = s;
                }
        dims[a] = n2;
        cur.swap(nxt);
    }
    std::copy(cur.begin(), cur.end(), out);
}

struct Math {
    int ndim, n, nd, ipoly, npbox;
    std::vector<double> xs, ws, ut;    // nodes, weights, values->coefficients as [k*n + k1]
    std::vector<double> rmask;
    double rsum = 0;
    std::vector<double> wts;           // tensor weights on [-1,1]^d
    // child->parent (n2 points of the parent per axis) and parent->child
    // (n points) interpolation, per sign (0 = -, 1 = +): [k * n + k1]
    std::vector<double> c2p[2], p2c[2];
    int n2;

    Math(int ndim_, int n_, int nd_, int ipoly_) : ndim(ndim_), n(n_), nd(nd_), ipoly(ipoly_) {
        npbox = 1;
        for (int a = 0; a < ndim; ++a) npbox *= n;
        xs.resize(n);
        ws.resize(n);
        std::vector<double> u(static_cast<size_t>(n) * n);
        poly_exps(ipoly, n, xs.data(), ws.data(), u.data());
        ut.resize(u.size());
        for (int k = 0; k < n; ++k)
            for (int k1 = 0; k1 < n; ++k1) ut[k * n + k1] = u[k + k1 * n];   // umat(k, k1)
        // tail mask, iptype = 2: coefficient powers with sum (p+1)^2 >= m^2
        const int morder = ndim == 1 ? n - 1 : n;
        rmask.assign(npbox, 0.0);
        int cnt = 0;
        for (int i = 0; i < npbox; ++i) {
            int r = i, s = 0;
            for (int a = 0; a < ndim; ++a) {
                const int p = r % n;
                r /= n;
                s += (p + 1) * (p + 1);
            }
            if (s >= morder * morder) {
                rmask[i] = 1.0;
                ++cnt;
            }
        }
        rsum = std::sqrt(static_cast<double>(cnt));
        wts.assign(npbox, 1.0);
        for (int i = 0; i < npbox; ++i) {
            int r = i;
            double w = 1.0;
            for (int a = 0; a < ndim; ++a) {
                w *= ws[r % n];
                r /= n;
            }
            wts[i] = w;
        }
        n2 = n / 2;
        std::vector<double> p(n + 1);
        for (int sg = 0; sg < 2; ++sg) {
            c2p[sg].assign(static_cast<size_t>(n2) * n, 0.0);
            for (int i = 0; i < n2; ++i) {
                const double x = sg ? xs[n2 + i] * 2 - 1.0 : xs[i] * 2 + 1.0;
                poly_vals(ipoly, x, n - 1, p.data());
                for (int k1 = 0; k1 < n; ++k1) c2p[sg][i * n + k1] = p[k1];
            }
            p2c[sg].assign(static_cast<size_t>(n) * n, 0.0);
            for (int i = 0; i < n; ++i) {
                const double x = sg ? xs[i] / 2 + 0.5 : xs[i] / 2 - 0.5;
                poly_vals(ipoly, x, n - 1, p.data());
                for (int k1 = 0; k1 < n; ++k1) p2c[sg][i * n + k1] = p[k1];
            }
        }
    }

    // values -> coefficients for m interleaved functions
    void trans(int m, const double* v, double* c) const {
        const double* A[3] = {ut.data(), ut.data(), ut.data()};
        tensor_apply(ndim, m, n, n, v, c, A);
    }
    // coefficients -> values at the nodes of child j (parent->child)
    void eval_child(int m, int j, const double* c, double* v) const {
        const double* A[3];
        for (int a = 0; a < ndim; ++a) A[a] = p2c[(j >> a) & 1].data();
        tensor_apply(ndim, m, n, n, c, v, A);
    }
    // coefficients of child j -> values at the parent's nodes inside it
    // (n2 per axis)
    void eval_c2p(int m, int j, const double* c, double* v) const {
        const double* A[3];
        for (int a = 0; a < ndim; ++a) A[a] = c2p[(j >> a) & 1].data();
        tensor_apply(ndim, m, n, n2, c, v, A);
    }
    // copy child j's n2^d block into the parent's grid (m functions)
    void c2p_copy(int m, int j, const double* v2, double* v) const {
        const int ox = (j & 1) ? n2 : 0, oy = (j & 2) ? n2 : 0, oz = (j & 4) ? n2 : 0;
        const int ny2 = ndim >= 2 ? n2 : 1, nz2 = ndim >= 3 ? n2 : 1;
        for (int i = 0; i < nz2; ++i)
            for (int jj = 0; jj < ny2; ++jj)
                for (int k = 0; k < n2; ++k) {
                    const size_t src = (static_cast<size_t>(i) * ny2 + jj) * n2 + k;
                    size_t dst = static_cast<size_t>(k + ox);
                    if (ndim >= 2) dst += static_cast<size_t>(n) * (jj + oy);
                    if (ndim >= 3) dst += static_cast<size_t>(n) * n * (i + oz);
                    for (int q = 0; q < m; ++q) v[dst * m + q] = v2[src * m + q];
                }
    }
    // fun_err, iptype = 2: max over densities of the masked coefficient l2 norm
    double tail(const double* c, double rscale) const {
        double e = 0;
        for (int q = 0; q < nd; ++q) {
            double s = 0;
            for (int i = 0; i < npbox; ++i) s += c[static_cast<size_t>(i) * nd + q] * c[static_cast<size_t>(i) * nd + q] * rmask[i];
            e = std::max(e, std::sqrt(s));
        }
        return e * rscale;
    }
};

double wrap_d(double dis, int iperiod, double bs0) {
    if (iperiod == 1) {
        const double dp = dis + bs0, dm = dis - bs0;
        if (std::fabs(dis) > std::fabs(dp)) dis = dp;
        if (std::fabs(dis) > std::fabs(dm)) dis = dm;
    }
    return dis;
}

// the reference multiplies box sizes by the default-real literal 1.05
constexpr double kCollFac = static_cast<double>(1.05f);

struct Editor {
    NewTree& T;
    const Math& M;
    std::vector<int> iflag;

    Editor(NewTree& t, const Math& m) : T(t), M(m) {}

    double* prow(int b) const { return T.pot + static_cast<size_t>(b) * T.wp(); }
    double* grow(int b) const { return T.grad + static_cast<size_t>(b) * T.wg(); }
    double* hrow(int b) const { return T.hess + static_cast<size_t>(b) * T.wh(); }
    double* crow(int b) const { return const_cast<double*>(T.coefs.data()) + static_cast<size_t>(b) * T.wp(); }

    bool touching(int a, int b, double distest) const {
        for (int k = 0; k < T.ndim; ++k) {
            const double dis = wrap_d(T.centers[T.ndim * a + k] - T.centers[T.ndim * b + k], T.iperiod, T.boxsize[0]);
            if (!(std::fabs(dis) <= distest)) return false;
        }
        return true;
    }

    int add_box(int lev, int par) {
        const int b = T.count();
        if (b >= T.cap) throw std::length_error("the refined tree needs more than the " + std::to_string(T.cap) +
                                                " boxes the caller's arrays hold");
        T.level.push_back(lev);
        T.parent.push_back(par);
        T.nchild.push_back(0);
        for (int j = 0; j < T.mc; ++j) T.child.push_back(-1);
        for (int k = 0; k < T.ndim; ++k) T.centers.push_back(0.0);
        T.ncoll.push_back(0);
        for (int j = 0; j < T.mnb; ++j) T.coll.push_back(-1);
        if (static_cast<int>(iflag.size()) == b) iflag.push_back(0);   // flags in use (STEP 10)
        return b;
    }

    // New numbering: order[i] = old id of new box i (i < order.size()); the
    // rows of every per-box array are moved, rows past the new count of the
    // caller's output arrays are left as they are.
    void renumber(const std::vector<int>& order, std::vector<int>* extra = nullptr) {
        const int m = static_cast<int>(order.size()), n0 = T.count();
        std::vector<int> newid(n0, -1);
        for (int i = 0; i < m; ++i) newid[order[i]] = i;
        auto perm_i = [&](std::vector<int>& v, int w) {
            std::vector<int> o(static_cast<size_t>(m) * w);
            for (int i = 0; i < m; ++i)
                for (int k = 0; k < w; ++k) o[static_cast<size_t>(w) * i + k] = v[static_cast<size_t>(w) * order[i] + k];
            v.swap(o);
        };
        auto perm_rows = [&](double* a, size_t w) {
            if (!a) return;
            std::vector<double> tmp(static_cast<size_t>(m) * w);
            for (int i = 0; i < m; ++i)
                std::copy(a + static_cast<size_t>(order[i]) * w, a + static_cast<size_t>(order[i] + 1) * w,
                          tmp.begin() + static_cast<size_t>(i) * w);
            std::copy(tmp.begin(), tmp.end(), a);
        };
        perm_i(T.level, 1);
        perm_i(T.parent, 1);
        perm_i(T.nchild, 1);
        perm_i(T.child, T.mc);
        perm_i(T.ncoll, 1);
        perm_i(T.coll, T.mnb);
        if (static_cast<int>(iflag.size()) == n0) perm_i(iflag, 1);
        if (extra) perm_i(*extra, 1);
        std::vector<double> c(static_cast<size_t>(m) * T.ndim);
        for (int i = 0; i < m; ++i)
            for (int k = 0; k < T.ndim; ++k) c[static_cast<size_t>(T.ndim) * i + k] = T.centers[static_cast<size_t>(T.ndim) * order[i] + k];
        T.centers.swap(c);
        perm_rows(T.pot, T.wp());
        perm_rows(T.coefs.data(), T.wp());
        if (T.ifpgh >= 2) perm_rows(T.grad, T.wg());
        if (T.ifpgh >= 3) perm_rows(T.hess, T.wh());
        for (auto& p : T.parent)
            if (p >= 0) p = newid[p];
        for (auto& c2 : T.child)
            if (c2 >= 0) c2 = newid[c2];
        for (auto& c2 : T.coll)
            if (c2 >= 0) c2 = newid[c2];
    }

    void computecoll() {
        const int n = T.count(), mnb = T.mnb, mc = T.mc;
        std::fill(T.ncoll.begin(), T.ncoll.end(), 0);
        std::fill(T.coll.begin(), T.coll.end(), -1);
        T.ncoll[0] = 1;
        T.coll[0] = 0;
        const double bs0 = T.boxsize[0];
        for (int l = 1; l <= T.nlevels; ++l)
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                const int dad = T.parent[b];
                for (int i = 0; i < T.ncoll[dad]; ++i) {
                    const int jb = T.coll[static_cast<size_t>(mnb) * dad + i];
                    for (int j = 0; j < mc; ++j) {
                        const int kb = T.child[static_cast<size_t>(mc) * jb + j];
                        if (kb < 0) continue;
                        bool ok = true;
                        for (int k = 0; k < T.ndim; ++k) {
                            double dis = std::fabs(T.centers[T.ndim * kb + k] - T.centers[T.ndim * b + k]);
                            if (T.iperiod == 1) dis = std::min(dis, bs0 - dis);
                            if (dis > kCollFac * T.boxsize[l]) { ok = false; break; }
                        }
                        if (ok) T.coll[static_cast<size_t>(mnb) * b + T.ncoll[b]++] = kb;
                    }
                }
            }
        (void)n;
        T.coll_valid = true;
    }

    // -------------------------------------------------------------- STEP 9
    int coarsen(double reps) {
        const int mc = T.mc, nd = T.nd;
        std::vector<char> del(T.count(), 0);
        int ndel = 0;
        std::vector<double> v2(static_cast<size_t>(T.wh() + T.wg() + T.wp()));
        std::vector<double> gc(T.wg()), hc(T.wh());
        for (int l = T.nlevels - 1; l >= 0; --l) {
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                if (T.nchild[b] != mc) continue;
                bool leafkids = true;
                for (int j = 0; j < mc; ++j)
                    if (T.nchild[T.child[static_cast<size_t>(mc) * b + j]] != 0) leafkids = false;
                if (!leafkids) continue;
                // the children's potential at the parent's nodes (written
                // whether or not the merge is accepted, as the reference does)
                for (int j = 0; j < mc; ++j) {
                    const int c = T.child[static_cast<size_t>(mc) * b + j];
                    M.eval_c2p(nd, j, crow(c), v2.data());
                    M.c2p_copy(nd, j, v2.data(), prow(b));
                }
                M.trans(nd, prow(b), crow(b));
                const double erra = M.tail(crow(b), 1.0) / M.rsum;   // eta = 0
                if (!(erra <= reps)) continue;
                for (int j = 0; j < mc; ++j) {
                    const int c = T.child[static_cast<size_t>(mc) * b + j];
                    if (T.ifpgh >= 2) {
                        M.trans(nd * T.ndim, grow(c), gc.data());
                        M.eval_c2p(nd * T.ndim, j, gc.data(), v2.data());
                        M.c2p_copy(nd * T.ndim, j, v2.data(), grow(b));
                    }
                    if (T.ifpgh >= 3) {
                        M.trans(nd * T.nhess(), hrow(c), hc.data());
                        M.eval_c2p(nd * T.nhess(), j, hc.data(), v2.data());
                        M.c2p_copy(nd * T.nhess(), j, v2.data(), hrow(b));
                    }
                }
                for (int j = 0; j < mc; ++j) {
                    const int c = T.child[static_cast<size_t>(mc) * b + j];
                    ++ndel;
                    del[c] = 1;
                    T.parent[c] = -1;
                    T.nchild[c] = 0;
                    for (int q = 0; q < mc; ++q) T.child[static_cast<size_t>(mc) * c + q] = -1;
                }
                T.nchild[b] = 0;
                for (int j = 0; j < mc; ++j) T.child[static_cast<size_t>(mc) * b + j] = -1;
            }
        }
        if (ndel == 0) return 0;
        // fgt_vol_tree_reorg_after_coarsen: drop the deleted boxes, keep the
        // level order; levels left empty leave the tree
        std::vector<int> order;
        order.reserve(T.count() - ndel);
        const int nl = T.nlevels;
        std::vector<int> nf(T.lfirst.size()), nlast(T.llast.size());
        for (int l = 0; l <= nl; ++l) {
            nf[l] = static_cast<int>(order.size());
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b)
                if (!del[b]) order.push_back(b);
            nlast[l] = static_cast<int>(order.size()) - 1;
        }
        renumber(order);
        for (int l = 0; l <= nl; ++l) {
            T.lfirst[l] = nf[l];
            T.llast[l] = nlast[l];
        }
        int nnew = nl;
        for (int l = nl; l >= 0; --l)
            if (T.llast[l] < T.lfirst[l]) --nnew;
        T.nlevels = nnew;
        T.coll_valid = false;
        return ndel;
    }

    // -------------------------------------------------------------- STEP 10
    // vol_tree_refine_boxes_flag_interp over boxes [f, l] of level lev
    void refine_interp(int f, int l, int lev) {
        const int mc = T.mc, nd = T.nd;
        std::vector<double> gc(T.wg()), hc(T.wh());
        const double bsh = T.boxsize[lev + 1] / 2;
        for (int b = f; b <= l; ++b) {
            if (iflag[b] <= 0) continue;
            if (T.ifpgh >= 2) M.trans(nd * T.ndim, grow(b), gc.data());
            if (T.ifpgh >= 3) M.trans(nd * T.nhess(), hrow(b), hc.data());
            T.nchild[b] = mc;
            for (int j = 0; j < mc; ++j) {
                const int c = add_box(lev + 1, b);
                for (int k = 0; k < T.ndim; ++k) T.centers[T.ndim * c + k] = T.centers[T.ndim * b + k] + child_sign(j, k) * bsh;
                M.eval_child(nd, j, crow(b), prow(c));
                M.trans(nd, prow(c), crow(c));
                if (T.ifpgh >= 2) M.eval_child(nd * T.ndim, j, gc.data(), grow(c));
                if (T.ifpgh >= 3) M.eval_child(nd * T.nhess(), j, hc.data(), hrow(c));
                T.child[static_cast<size_t>(mc) * b + j] = c;
                if (iflag[b] == 1) iflag[c] = 3;
                if (iflag[b] == 2) iflag[c] = 0;
            }
        }
    }

    void update_flags(int lev, int f, int l) {
        const double distest = 1.05 * (T.boxsize[lev] + T.boxsize[lev + 1]) / 2.0;
        const int mc = T.mc, mnb = T.mnb;
        for (int b = f; b <= l; ++b) {
            if (iflag[b] != 3) continue;
            iflag[b] = 0;
            bool hit = false;
            for (int i = 0; i < T.ncoll[b] && !hit; ++i) {
                const int jb = T.coll[static_cast<size_t>(mnb) * b + i];
                for (int j = 0; j < mc && !hit; ++j) {
                    const int kb = T.child[static_cast<size_t>(mc) * jb + j];
                    if (kb >= 0 && T.nchild[kb] > 0 && touching(kb, b, distest)) hit = true;
                }
            }
            if (hit) iflag[b] = 1;
        }
    }

    // vol_tree_reorg_pgh: per level >= 2 the level's boxes, then its tail
    void reorg_tail(const std::vector<int>& tf, const std::vector<int>& tl) {
        std::vector<int> order;
        order.reserve(T.count());
        std::vector<int> nf(T.lfirst.size()), nlast(T.llast.size());
        for (int l = 0; l <= T.nlevels; ++l) {
            nf[l] = static_cast<int>(order.size());
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) order.push_back(b);
            if (l >= 2)
                for (int b = tf[l]; b <= tl[l]; ++b) order.push_back(b);
            nlast[l] = static_cast<int>(order.size()) - 1;
        }
        renumber(order);
        for (int l = 0; l <= T.nlevels; ++l) {
            T.lfirst[l] = nf[l];
            T.llast[l] = nlast[l];
        }
    }

    int fix_lr_interp() {
        const int n0 = T.count(), mnb = T.mnb, mc = T.mc, nl = T.nlevels;
        iflag.assign(n0, 0);
        for (int lev = nl; lev >= 2; --lev) {
            const double distest = 1.05 * (T.boxsize[lev - 1] + T.boxsize[lev - 2]) / 2.0;
            for (int b = T.lfirst[lev]; b <= T.llast[lev]; ++b) {
                const int dad = T.parent[b], gd = T.parent[dad];
                for (int i = 0; i < T.ncoll[gd]; ++i) {
                    const int jb = T.coll[static_cast<size_t>(mnb) * gd + i];
                    if (T.nchild[jb] == 0 && iflag[jb] == 0 && touching(jb, dad, distest)) iflag[jb] = 1;
                }
            }
        }
        for (int lev = nl; lev >= 1; --lev) {
            const double distest = 1.05 * (T.boxsize[lev] + T.boxsize[lev - 1]) / 2.0;
            for (int b = T.lfirst[lev]; b <= T.llast[lev]; ++b) {
                if (iflag[b] != 1 && iflag[b] != 2) continue;
                const int dad = T.parent[b];
                for (int i = 0; i < T.ncoll[dad]; ++i) {
                    const int jb = T.coll[static_cast<size_t>(mnb) * dad + i];
                    if (T.nchild[jb] == 0 && iflag[jb] == 0 && touching(jb, b, distest)) iflag[jb] = 2;
                }
            }
        }
        std::vector<int> tf(nl + 1, 0), tl(nl + 1, -1);
        for (int lev = 1; lev <= nl - 2; ++lev) {
            tf[lev + 1] = T.count();
            refine_interp(T.lfirst[lev], T.llast[lev], lev);
            tl[lev + 1] = T.count() - 1;
        }
        reorg_tail(tf, tl);
        computecoll();
        for (auto& f : iflag)
            if (f != 3) f = 0;
        std::fill(tf.begin(), tf.end(), 0);
        std::fill(tl.begin(), tl.end(), -1);
        for (int lev = 2; lev <= nl - 2; ++lev) {
            update_flags(lev, T.lfirst[lev], T.llast[lev]);
            update_flags(lev, tf[lev], tl[lev]);
            tf[lev + 1] = T.count();
            refine_interp(T.lfirst[lev], T.llast[lev], lev);
            refine_interp(tf[lev], tl[lev], lev);
            tl[lev + 1] = T.count() - 1;
            for (int b = tf[lev + 1]; b <= tl[lev + 1]; ++b) {
                T.ncoll[b] = 0;
                const int dad = T.parent[b];
                for (int i = 0; i < T.ncoll[dad]; ++i) {
                    const int jb = T.coll[static_cast<size_t>(mnb) * dad + i];
                    for (int j = 0; j < mc; ++j) {
                        const int kb = T.child[static_cast<size_t>(mc) * jb + j];
                        if (kb < 0) continue;
                        bool ok = true;
                        for (int k = 0; k < T.ndim; ++k) {
                            const double dis =
                                wrap_d(T.centers[T.ndim * kb + k] - T.centers[T.ndim * b + k], T.iperiod, T.boxsize[0]);
                            if (std::fabs(dis) > kCollFac * T.boxsize[lev + 1]) {
                                ok = false;
                                break;
                            }
                        }
                        if (ok) T.coll[static_cast<size_t>(mnb) * b + T.ncoll[b]++] = kb;
                    }
                }
            }
        }
        reorg_tail(tf, tl);
        computecoll();
        iflag.clear();
        return T.count() - n0;
    }
};

}  // namespace

void nt_read(NewTree& T, const int* itree, const int* iptr, const double* centers, const double* boxsize) {
    const int nl = T.nlevels, mc = T.mc, mnb = T.mnb, d = T.ndim;
    T.boxsize.assign(boxsize, boxsize + nl + 1);
    T.lfirst.resize(nl + 1);
    T.llast.resize(nl + 1);
    for (int l = 0; l <= nl; ++l) {
        T.lfirst[l] = itree[2 * l] - 1;
        T.llast[l] = itree[2 * l + 1] - 1;
    }
    const int nb = T.llast[nl] + 1;
    if (nb < 1 || nb > T.cap) throw std::invalid_argument("laddr does not fit the box capacity");
    T.level.resize(nb);
    T.parent.resize(nb);
    T.nchild.resize(nb);
    T.child.resize(static_cast<size_t>(mc) * nb);
    T.ncoll.resize(nb);
    T.coll.resize(static_cast<size_t>(mnb) * nb);
    for (int b = 0; b < nb; ++b) {
        T.level[b] = itree[iptr[1] - 1 + b];
        const int p = itree[iptr[2] - 1 + b];
        T.parent[b] = p > 0 ? p - 1 : -1;
        T.nchild[b] = itree[iptr[3] - 1 + b];
        for (int j = 0; j < mc; ++j) {
            const int c = itree[iptr[4] - 1 + mc * b + j];
            T.child[static_cast<size_t>(mc) * b + j] = c > 0 ? c - 1 : -1;
        }
        T.ncoll[b] = itree[iptr[5] - 1 + b];
        for (int j = 0; j < mnb; ++j) {
            const int c = itree[iptr[6] - 1 + mnb * b + j];
            T.coll[static_cast<size_t>(mnb) * b + j] = c > 0 ? c - 1 : -1;
        }
    }
    T.centers.assign(centers, centers + static_cast<size_t>(d) * nb);
    T.coll_valid = true;
}

void nt_write(const NewTree& T, int* itree, const int* iptr, double* centers, int* nlevels) {
    const int mc = T.mc, mnb = T.mnb, nb = T.count();
    for (size_t l = 0; l < T.lfirst.size(); ++l) {
        itree[2 * l] = T.lfirst[l] + 1;
        itree[2 * l + 1] = T.llast[l] + 1;
    }
    for (int b = 0; b < nb; ++b) {
        itree[iptr[1] - 1 + b] = T.level[b];
        itree[iptr[2] - 1 + b] = T.parent[b] >= 0 ? T.parent[b] + 1 : -1;
        itree[iptr[3] - 1 + b] = T.nchild[b];
        for (int j = 0; j < mc; ++j) {
            const int c = T.child[static_cast<size_t>(mc) * b + j];
            itree[iptr[4] - 1 + mc * b + j] = c >= 0 ? c + 1 : -1;
        }
        if (T.coll_valid) {
            itree[iptr[5] - 1 + b] = T.ncoll[b];
            for (int j = 0; j < mnb; ++j) {
                const int c = T.coll[static_cast<size_t>(mnb) * b + j];
                itree[iptr[6] - 1 + mnb * b + j] = c >= 0 ? c + 1 : -1;
            }
        }
    }
    std::copy(T.centers.begin(), T.centers.end(), centers);
    *nlevels = T.nlevels;
}

void nt_rewrite(NewTree& T, double eps, const std::vector<char>& ifpwexp, const NewBoxEval& eval, NewTreeStats* st) {
    NewTreeStats s;
    const Math M(T.ndim, T.n, T.nd, T.ipoly);
    Editor E(T, M);
    const int nd = T.nd, mc = T.mc;
    T.coefs.assign(static_cast<size_t>(T.cap) * T.wp(), 0.0);
    // ---- STEP 8: which leaves are under-resolved (L2 tail of the
    //      potential's coefficients against eps times its L2 norm)
    double rn = 0;
    for (int l = 0; l <= T.nlevels; ++l) {
        double sc = 1.0;
        for (int a = 0; a < T.ndim; ++a) sc *= T.boxsize[l] / 2;
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
            if (T.nchild[b] != 0) continue;
            const double* v = E.prow(b);
            for (int i = 0; i < T.npbox; ++i)
                for (int q = 0; q < nd; ++q) {
                    const double f = v[static_cast<size_t>(i) * nd + q];
                    rn += f * f * M.wts[i] * sc;
                }
        }
    }
    s.rnorm = std::sqrt(rn);
    const double reps = eps * (std::sqrt(1.0 / std::pow(T.boxsize[0], T.ndim)) * s.rnorm);
    s.reps = reps;
    const int nboxes0 = T.count();
    s.nboxes0 = nboxes0;
    std::vector<int> ifrefine(nboxes0, 0), irefinelev(nboxes0, 0), imaxref(nboxes0, 0), origin(nboxes0);
    for (int b = 0; b < nboxes0; ++b) origin[b] = b;
    for (int l = 0; l <= T.nlevels; ++l)
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b)
            if (T.nchild[b] == 0) M.trans(nd, E.prow(b), E.crow(b));
    int maxref = 0;
    for (int l = 0; l <= T.nlevels - 1; ++l) {
        const double rscale = T.boxsize[l] / 2;   // eta = 1
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
            if (T.nchild[b] != 0) continue;
            const double erra = M.tail(E.crow(b), rscale) / M.rsum;
            if (erra > reps) {
                ifrefine[b] = 1;
                imaxref[b] = static_cast<int>(std::ceil(std::log(erra / reps) / std::log(2.0) / T.n));
                ++s.nrefine;
                maxref = std::max(maxref, imaxref[b]);
            }
        }
    }
    s.maxrefinelev = maxref;
    std::vector<int> cur;
    for (int b = 0; b < nboxes0; ++b)
        if (ifrefine[b] == 1 && irefinelev[b] <= maxref) cur.push_back(b);
    std::vector<double> np, ng, nh;
    while (!cur.empty()) {
        std::vector<NewBox> kids;
        for (int p : cur) {
            const int lev = T.level[p];
            if (lev >= T.nlevels) {
                // a child would lie below the finest level the tables and the
                // caller's boxsize array describe
                ++s.skipped_finest;
                continue;
            }
            irefinelev[p] += 1;
            T.nchild[p] = mc;
            const double bsh = T.boxsize[lev] / 4;
            for (int j = 0; j < mc; ++j) {
                const int c = E.add_box(lev + 1, p);
                for (int k = 0; k < T.ndim; ++k) T.centers[T.ndim * c + k] = T.centers[T.ndim * p + k] + child_sign(j, k) * bsh;
                T.child[static_cast<size_t>(mc) * p + j] = c;
                ifrefine.push_back(0);
                irefinelev.push_back(irefinelev[p] + 1);
                imaxref.push_back(0);
                origin.push_back(origin[p]);
                kids.push_back(NewBox{c, p, lev + 1, origin[p]});
            }
        }
        if (kids.empty()) break;
        ++s.generations;
        s.nnew += static_cast<int>(kids.size());
        np.assign(kids.size() * T.wp(), 0.0);
        ng.assign(T.ifpgh >= 2 ? kids.size() * T.wg() : 0, 0.0);
        nh.assign(T.ifpgh >= 3 ? kids.size() * T.wh() : 0, 0.0);
        eval(T, kids, np, ng, nh);
        cur.clear();
        for (size_t i = 0; i < kids.size(); ++i) {
            const int c = kids[i].id;
            double* pr = E.prow(c);
            for (size_t q = 0; q < T.wp(); ++q) pr[q] += np[i * T.wp() + q];
            if (T.ifpgh >= 2) {
                double* gr = E.grow(c);
                for (size_t q = 0; q < T.wg(); ++q) gr[q] += ng[i * T.wg() + q];
            }
            if (T.ifpgh >= 3) {
                double* hr = E.hrow(c);
                for (size_t q = 0; q < T.wh(); ++q) hr[q] += nh[i * T.wh() + q];
            }
            M.trans(nd, pr, E.crow(c));
            const double erra = M.tail(E.crow(c), T.boxsize[T.level[c]] / 2) / M.rsum;
            if (erra > reps) {
                ifrefine[c] = 1;
                // refine no deeper below an original leaf than its estimate
                int k = 1, kb = c, kdad = -1;
                for (; k <= maxref; ++k) {
                    kdad = T.parent[kb];
                    if (kdad < nboxes0) break;
                    kb = kdad;
                }
                if (k > imaxref[kdad]) ifrefine[c] = 0;
            }
            if (ifrefine[c] == 1 && irefinelev[c] <= maxref) cur.push_back(c);
        }
    }
    if (s.nnew > 0) {
        // fgt_vol_tree_reorg: each level's boxes, then its new boxes
        std::vector<int> order;
        order.reserve(T.count());
        std::vector<int> nf(T.nlevels + 1), nlast(T.nlevels + 1);
        for (int l = 0; l <= T.nlevels; ++l) {
            nf[l] = static_cast<int>(order.size());
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) order.push_back(b);
            if (l >= 1)
                for (int b = nboxes0; b < T.count(); ++b)
                    if (T.level[b] == l) order.push_back(b);
            nlast[l] = static_cast<int>(order.size()) - 1;
        }
        E.renumber(order);
        T.lfirst = nf;
        T.llast = nlast;
        T.coll_valid = false;
        // ---- STEP 9 (when the refinement added at least a tenth of the boxes)
        if (!(s.nnew < 0.1 * nboxes0)) s.ndeleted = E.coarsen(reps);
        // ---- STEP 10
        E.computecoll();
        s.nrebalance = E.fix_lr_interp();
    }
    if (st) *st = s;
}

}  // namespace fgt
