// tree.cpp -- host adaptive volume tree builder (include/fgt_tree.h).
//
// Builds the level-restricted adaptive tree that the reference's
// vol_tree_mem / vol_tree_build produce (src/common/tree_vol_coeffs_nd.f),
// for a source given as a sum of anisotropic Gaussians:
//   * refinement by the L2 norm of the tail coefficients of the
//     tensor-product interpolant, scaled by a running L2 norm of the source
//     (vol_tree_find_box_refine 693-779, fun_err 860-943, update_rints
//     949-1092; tail mask tens_prod_get_rmask, src/common/polytens.f:1579-1616),
//   * children numbered level by level in parent order, signs mmm, pmm, ...
//     (vol_tree_refine_boxes 787-857; get_child_box_sign,
//     src/common/pts_tree_nd.f:1418-1464),
//   * 2:1 level restriction by flag / flag+ / flag++ passes and two
//     reorganisations (vol_tree_fix_lr 1242-1557, vol_tree_reorg 1561-1733,
//     vol_updateflags 1736-1841, vol_tree_refine_boxes_flag 1848-1920),
//   * colleagues within 1.05 box sizes, with periodic wrap
//     (computecoll, src/common/tree_routs_nd.f:106-242).
// One pass suffices: the running norm seen by each level is the same value
// the reference's second pass reads back from its first pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fgt_tree.h"
#include "plan.h"

namespace {

struct Builder {
    int ndim = 3, ipoly = 0, iperiod = 0, norder = 8, npbox = 0, mc = 8, mnb = 27;
    double eps = 1e-6, boxlen = 1.0;
    std::vector<double> gauss;  // ng rows of (c[ndim], s[ndim], a)
    int ng = 0;
    // a caller's source instead of the Gaussians: fun(nd, xyz, dpars, zpars,
    // ipars, f) at one point, nd values (the reference's `fun` convention)
    fgt_source_fn fun = nullptr;
    double* dpars = nullptr;
    void* zpars = nullptr;
    int* ipars = nullptr;
    int nd = 1;          // densities (values per node)
    int iptype = 2;      // refinement norm: 0 max, 1 L1, 2 L2
    double eta = 0.0;    // tail error scaled by (box size / 2)^eta
    double zk = 0.0;     // real part of the Helmholtz parameter: boxes with zk * size > 5 always refine
    int cap = 0;         // box capacity of the output arrays (0: the tree's own count)
    // tree (0-based box ids)
    std::vector<int> level, parent, nchild, child, ncoll, coll, iflag;
    std::vector<double> centers, fvals, boxsize;
    std::vector<int> lfirst, llast;
    int nlevels = 0;
    // discretisation
    std::vector<double> grid;   // ndim per node, on [-1/2, 1/2]
    std::vector<double> wts;    // tensor weights on [-1,1]^ndim
    std::vector<double> umat;   // n x n, u[i + j*n]
    std::vector<double> rmask;
    double rsum = 0;

    int nb() const { return static_cast<int>(level.size()); }

    double source(const double* y) const {
        int nimg = 1;
        for (int k = 0; k < ndim; ++k) nimg *= iperiod ? 3 : 1;
        const int stride = 2 * ndim + 1;
        double acc = 0.0;
        for (int g = 0; g < ng; ++g) {
            const double* p = &gauss[static_cast<size_t>(g) * stride];
            for (int im = 0; im < nimg; ++im) {
                int code = im;
                double rr = 0.0;
                for (int k = 0; k < ndim; ++k) {
                    double shift = 0.0;
                    if (iperiod) {
                        shift = static_cast<double>(code % 3 - 1) * boxlen;
                        code /= 3;
                    }
                    const double d = y[k] - p[k] - shift;
                    rr += d * d / p[ndim + k];
                }
                acc += p[2 * ndim] * std::exp(-rr);
            }
        }
        return acc;
    }

    int add_box(int lev, int par, const double* cen, double bs) {
        const int b = nb();
        level.push_back(lev);
        parent.push_back(par);
        nchild.push_back(0);
        for (int j = 0; j < mc; ++j) child.push_back(-1);
        for (int k = 0; k < ndim; ++k) centers.push_back(cen[k]);
        iflag.push_back(0);
        bsize_of.push_back(bs);
        fvals.resize(static_cast<size_t>(b + 1) * npbox * nd);
        return b;
    }

    // Node values of every box created since the last call, boxes in
    // parallel.  Gaussian images whose exponent exceeds 750 on the whole box
    // are skipped: exp underflows to exactly +0 there, so the sums are
    // unchanged bit for bit.
    std::vector<double> bsize_of;
    int evaluated = 0;
    void flush() {
        if (fun) {
            flush_fun();
            return;
        }
        const int n = nb();
        const int stride = 2 * ndim + 1;
        int nimg = 1;
        for (int k = 0; k < ndim; ++k) nimg *= iperiod ? 3 : 1;
#pragma omp parallel for schedule(dynamic, 4)
        for (int b = evaluated; b < n; ++b) {
            const double bs = bsize_of[b];
            const double* cen = &centers[static_cast<size_t>(ndim) * b];
            std::vector<int> live;
            live.reserve(static_cast<size_t>(ng) * nimg);
            for (int g = 0; g < ng; ++g) {
                const double* p = &gauss[static_cast<size_t>(g) * stride];
                for (int im = 0; im < nimg; ++im) {
                    int code = im;
                    double lb = 0.0;
                    for (int k = 0; k < ndim; ++k) {
                        double shift = 0.0;
                        if (iperiod) {
                            shift = static_cast<double>(code % 3 - 1) * boxlen;
                            code /= 3;
                        }
                        const double d = std::fabs(cen[k] - p[k] - shift) - bs / 2;
                        if (d > 0) lb += d * d / p[ndim + k];
                    }
                    live.push_back(lb < 750.0 ? 1 : 0);
                }
            }
            double y[3];
            for (int l = 0; l < npbox; ++l) {
                for (int k = 0; k < ndim; ++k) y[k] = cen[k] + grid[ndim * l + k] * bs;
                double acc = 0.0;
                for (int g = 0; g < ng; ++g) {
                    const double* p = &gauss[static_cast<size_t>(g) * stride];
                    for (int im = 0; im < nimg; ++im) {
                        if (!live[static_cast<size_t>(g) * nimg + im]) continue;
                        int code = im;
                        double rr = 0.0;
                        for (int k = 0; k < ndim; ++k) {
                            double shift = 0.0;
                            if (iperiod) {
                                shift = static_cast<double>(code % 3 - 1) * boxlen;
                                code /= 3;
                            }
                            const double d = y[k] - p[k] - shift;
                            rr += d * d / p[ndim + k];
                        }
                        acc += p[2 * ndim] * std::exp(-rr);
                    }
                }
                fvals[static_cast<size_t>(b) * npbox + l] = acc;
            }
        }
        evaluated = n;
    }

    // the caller's source at every node of the boxes created since the last
    // call (xyz = centre + node * box size, as vol_tree_refine_boxes forms
    // it); boxes in parallel, so fun must be safe to call concurrently (the
    // reference calls it from OpenMP loops too)
    void flush_fun() {
        const int n = nb();
#pragma omp parallel for schedule(dynamic, 4)
        for (int b = evaluated; b < n; ++b) {
            const double bs = bsize_of[b];
            double xyz[3];
            int ndv = nd;
            for (int l = 0; l < npbox; ++l) {
                for (int k = 0; k < ndim; ++k) xyz[k] = centers[static_cast<size_t>(ndim) * b + k] + grid[ndim * l + k] * bs;
                fun(&ndv, xyz, dpars, zpars, ipars, &fvals[(static_cast<size_t>(b) * npbox + l) * nd]);
            }
        }
        evaluated = n;
    }

    // children of b at level lev+1 (box size bs); sign of axis k for child j
    // is + when bit k of j is set
    void refine(int b, double bs) {
        nchild[b] = mc;
        const double bsh = bs / 2;
        double cen[3];
        for (int j = 0; j < mc; ++j) {
            for (int k = 0; k < ndim; ++k)
                cen[k] = centers[ndim * b + k] + (((j >> k) & 1) ? bsh : -bsh);
            const int c = add_box(level[b] + 1, b, cen, bs);
            child[static_cast<size_t>(mc) * b + j] = c;
        }
    }

    // the box's part of the running source norm (update_rints,
    // tree_vol_coeffs_nd.f:949-1092): iptype 2 sum f^2 w rsc (squared L2
    // mass), 1 sum |f| w rsc, 0 the largest value (signed, as the reference
    // takes maxval)
    double box_norm(int b, double rsc) {
        flush();
        const double* v = &fvals[static_cast<size_t>(b) * npbox * nd];
        if (iptype == 0) {
            double m = v[0];
            for (size_t i = 1; i < static_cast<size_t>(npbox) * nd; ++i) m = std::max(m, v[i]);
            return m;
        }
        double s = 0;
        for (int l = 0; l < npbox; ++l)
            for (int q = 0; q < nd; ++q) {
                const double f = v[static_cast<size_t>(l) * nd + q];
                s += (iptype == 2 ? f * f : std::fabs(f)) * wts[l] * rsc;
            }
        return s;
    }

    // tail error of box b (values already evaluated): the tensor
    // coefficients of each density, masked to the tail, in the iptype norm
    // (fun_err, tree_vol_coeffs_nd.f:860-943), largest over the densities
    double tail_error_flushed(int b) const {
        const int n = norder;
        double err = 0;
        std::vector<double> a(npbox), t(npbox);
        for (int q = 0; q < nd; ++q) {
            for (int i = 0; i < npbox; ++i) a[i] = fvals[(static_cast<size_t>(b) * npbox + i) * nd + q];
            int stride = 1;
            for (int ax = 0; ax < ndim; ++ax) {
                for (int i = 0; i < npbox; ++i) {
                    const int ia = (i / stride) % n;
                    const int base = i - ia * stride;
                    double s = 0;
                    for (int j = 0; j < n; ++j) s += umat[ia + j * n] * a[base + j * stride];
                    t[i] = s;
                }
                a.swap(t);
                stride *= n;
            }
            double e = 0;
            for (int i = 0; i < npbox; ++i) {
                if (iptype == 0) {
                    if (rmask[i] > 0.5 && e < std::fabs(a[i])) e = std::fabs(a[i]);
                } else if (iptype == 1) {
                    e += std::fabs(a[i]) * rmask[i];
                } else {
                    e += a[i] * a[i] * rmask[i];
                }
            }
            if (iptype == 2) e = std::sqrt(e);
            err = std::max(err, e);
        }
        return err;
    }

    double wrap(double dis) const {
        if (iperiod == 1) {
            const double bs0 = boxsize[0];
            const double dp = dis + bs0, dm = dis - bs0;
            if (std::fabs(dis) > std::fabs(dp)) dis = dp;
            if (std::fabs(dis) > std::fabs(dm)) dis = dm;
        }
        return dis;
    }

    bool touching(int a, int b, double distest, bool strict_le = true) const {
        for (int k = 0; k < ndim; ++k) {
            const double dis = std::fabs(wrap(centers[ndim * a + k] - centers[ndim * b + k]));
            if (strict_le ? !(dis <= distest) : !(dis < distest)) return false;
        }
        return true;
    }

    void computecoll() {
        const int n = nb();
        ncoll.assign(n, 0);
        coll.assign(static_cast<size_t>(mnb) * n, -1);
        ncoll[0] = 1;
        coll[0] = 0;
        const double bs0 = boxsize[0];
        for (int l = 1; l <= nlevels; ++l)
            for (int b = lfirst[l]; b <= llast[l]; ++b) {
                const int dad = parent[b];
                for (int i = 0; i < ncoll[dad]; ++i) {
                    const int jb = coll[static_cast<size_t>(mnb) * dad + i];
                    for (int j = 0; j < mc; ++j) {
                        const int kb = child[static_cast<size_t>(mc) * jb + j];
                        if (kb < 0) continue;
                        bool ok = true;
                        for (int k = 0; k < ndim; ++k) {
                            double dis = std::fabs(centers[ndim * kb + k] - centers[ndim * b + k]);
                            if (iperiod == 1) dis = std::min(dis, bs0 - dis);
                            if (dis > 1.05 * boxsize[l]) { ok = false; break; }
                        }
                        if (ok) coll[static_cast<size_t>(mnb) * b + ncoll[b]++] = kb;
                    }
                }
            }
    }

    // Renumber so that each level's boxes are contiguous: per level the
    // boxes of [lfirst,llast] then those of the tail range (vol_tree_reorg).
    void reorg(const std::vector<int>& tfirst, const std::vector<int>& tlast) {
        flush();
        const int n = nb();
        std::vector<int> newid(n, -1);
        std::vector<int> order;
        order.reserve(n);
        std::vector<int> nf(nlevels + 1), nl(nlevels + 1);
        for (int l = 0; l <= nlevels; ++l) {
            nf[l] = static_cast<int>(order.size());
            for (int b = lfirst[l]; b <= llast[l]; ++b) order.push_back(b);
            if (l >= 2)
                for (int b = tfirst[l]; b <= tlast[l]; ++b) order.push_back(b);
            nl[l] = static_cast<int>(order.size()) - 1;
        }
        for (int i = 0; i < n; ++i) newid[order[i]] = i;
        auto perm_i = [&](std::vector<int>& v, int w) {
            std::vector<int> o(v.size());
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < w; ++k) o[static_cast<size_t>(w) * i + k] = v[static_cast<size_t>(w) * order[i] + k];
            v.swap(o);
        };
        auto perm_d = [&](std::vector<double>& v, int w) {
            std::vector<double> o(v.size());
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < w; ++k) o[static_cast<size_t>(w) * i + k] = v[static_cast<size_t>(w) * order[i] + k];
            v.swap(o);
        };
        perm_i(level, 1);
        perm_i(parent, 1);
        perm_i(nchild, 1);
        perm_i(child, mc);
        perm_i(iflag, 1);
        perm_d(centers, ndim);
        perm_d(fvals, npbox * nd);
        perm_d(bsize_of, 1);
        for (auto& p : parent)
            if (p >= 0) p = newid[p];
        for (auto& c : child)
            if (c >= 0) c = newid[c];
        lfirst = nf;
        llast = nl;
    }

    // vol_tree_refine_boxes_flag: refine boxes of [f,l] with iflag > 0
    void refine_flagged(int f, int l) {
        for (int b = f; b <= l; ++b) {
            if (iflag[b] <= 0) continue;
            const int first_child = nb();
            refine(b, boxsize[level[b] + 1]);
            for (int j = 0; j < mc; ++j) {
                const int c = first_child + j;
                if (iflag[b] == 1) iflag[c] = 3;
                if (iflag[b] == 2) iflag[c] = 0;
            }
        }
    }

    void update_flags(int lev, int f, int l) {
        const double distest = 1.05 * (boxsize[lev] + boxsize[lev + 1]) / 2.0;
        for (int b = f; b <= l; ++b) {
            if (iflag[b] != 3) continue;
            iflag[b] = 0;
            bool hit = false;
            for (int i = 0; i < ncoll[b] && !hit; ++i) {
                const int jb = coll[static_cast<size_t>(mnb) * b + i];
                for (int j = 0; j < mc && !hit; ++j) {
                    const int kb = child[static_cast<size_t>(mc) * jb + j];
                    if (kb >= 0 && nchild[kb] > 0 && touching(kb, b, distest)) hit = true;
                }
            }
            if (hit) iflag[b] = 1;
        }
    }

    void fix_lr() {
        const int n0 = nb();
        std::fill(iflag.begin(), iflag.end(), 0);
        // flag: leaf colleagues of a grandparent touching the parent
        for (int lev = nlevels; lev >= 2; --lev) {
            const double distest = 1.05 * (boxsize[lev - 1] + boxsize[lev - 2]) / 2.0;
            for (int b = lfirst[lev]; b <= llast[lev]; ++b) {
                const int dad = parent[b], gd = parent[dad];
                for (int i = 0; i < ncoll[gd]; ++i) {
                    const int jb = coll[static_cast<size_t>(mnb) * gd + i];
                    if (nchild[jb] == 0 && iflag[jb] == 0 && touching(jb, dad, distest)) iflag[jb] = 1;
                }
            }
        }
        // flag+: bigger leaves touching a flagged / flag+ box
        for (int lev = nlevels; lev >= 1; --lev) {
            const double distest = 1.05 * (boxsize[lev] + boxsize[lev - 1]) / 2.0;
            for (int b = lfirst[lev]; b <= llast[lev]; ++b) {
                if (iflag[b] != 1 && iflag[b] != 2) continue;
                const int dad = parent[b];
                for (int i = 0; i < ncoll[dad]; ++i) {
                    const int jb = coll[static_cast<size_t>(mnb) * dad + i];
                    if (nchild[jb] == 0 && iflag[jb] == 0 && touching(jb, b, distest)) iflag[jb] = 2;
                }
            }
        }
        (void)n0;
        std::vector<int> tf(nlevels + 1, 0), tl(nlevels + 1, -1);
        for (int lev = 1; lev <= nlevels - 2; ++lev) {
            tf[lev + 1] = nb();
            refine_flagged(lfirst[lev], llast[lev]);
            tl[lev + 1] = nb() - 1;
        }
        reorg(tf, tl);
        computecoll();
        for (auto& f : iflag)
            if (f != 3) f = 0;
        std::fill(tf.begin(), tf.end(), 0);
        std::fill(tl.begin(), tl.end(), -1);
        for (int lev = 2; lev <= nlevels - 2; ++lev) {
            update_flags(lev, lfirst[lev], llast[lev]);
            update_flags(lev, tf[lev], tl[lev]);
            tf[lev + 1] = nb();
            refine_flagged(lfirst[lev], llast[lev]);
            refine_flagged(tf[lev], tl[lev]);
            tl[lev + 1] = nb() - 1;
            // colleagues of the boxes just created
            ncoll.resize(nb(), 0);
            coll.resize(static_cast<size_t>(mnb) * nb(), -1);
            for (int b = tf[lev + 1]; b <= tl[lev + 1]; ++b) {
                ncoll[b] = 0;
                const int dad = parent[b];
                for (int i = 0; i < ncoll[dad]; ++i) {
                    const int jb = coll[static_cast<size_t>(mnb) * dad + i];
                    for (int j = 0; j < mc; ++j) {
                        const int kb = child[static_cast<size_t>(mc) * jb + j];
                        if (kb < 0) continue;
                        bool ok = true;
                        for (int k = 0; k < ndim; ++k)
                            if (std::fabs(wrap(centers[ndim * kb + k] - centers[ndim * b + k])) > 1.05 * boxsize[lev + 1]) {
                                ok = false;
                                break;
                            }
                        if (ok) coll[static_cast<size_t>(mnb) * b + ncoll[b]++] = kb;
                    }
                }
            }
        }
        reorg(tf, tl);
        computecoll();
    }

    // stop with TooLarge once a level leaves more than max_leaves leaves
    // (0: no limit); the count before the level restriction is a lower bound
    long long max_leaves = 0;
    struct TooLarge {};

    void build() {
        mc = 1 << ndim;
        mnb = 1;
        for (int k = 0; k < ndim; ++k) mnb *= 3;
        npbox = 1;
        for (int k = 0; k < ndim; ++k) npbox *= norder;
        const int n = norder;
        std::vector<double> x(n), w(n);
        umat.resize(static_cast<size_t>(n) * n);
        fgt::poly_exps(ipoly, n, x.data(), w.data(), umat.data());
        // grid on [-1/2,1/2]^ndim and tensor weights on [-1,1]^ndim, x fastest
        grid.resize(static_cast<size_t>(ndim) * npbox);
        wts.resize(npbox);
        for (int l = 0; l < npbox; ++l) {
            int r = l, ik[3] = {0, 0, 0};
            for (int k = 0; k < ndim; ++k) {
                ik[k] = r % n;
                r /= n;
                grid[ndim * l + k] = x[ik[k]] / 2;
            }
            // product taken from the slowest axis down, as polytens_exps_nd does
            double wt = w[ik[ndim - 1]];
            for (int k = ndim - 2; k >= 0; --k) wt *= w[ik[k]];
            wts[l] = wt;
        }
        // tail mask: sum_k (p_k+1)^e >= m^e, e = max(iptype, 1), m = n (n-1
        // in 1-D); rsum = the mask's count in the iptype norm
        rmask.assign(npbox, 0.0);
        const int morder = (ndim == 1) ? n - 1 : n;
        const int pe = std::max(iptype, 1);
        auto ipw = [](int v, int e) { int r = 1; while (e-- > 0) r *= v; return r; };
        double cnt = 0;
        for (int l = 0; l < npbox; ++l) {
            int r = l, i1 = 0;
            for (int k = 0; k < ndim; ++k) {
                const int p = r % n;
                r /= n;
                i1 += ipw(p + 1, pe);
            }
            if (i1 >= ipw(morder, pe)) {
                rmask[l] = 1.0;
                cnt += 1;
            }
        }
        rsum = iptype == 2 ? std::sqrt(cnt) : (iptype == 0 ? 1.0 : cnt);

        boxsize.assign(1, boxlen);
        double cen0[3] = {0, 0, 0};
        add_box(0, -1, cen0, boxlen);
        lfirst.assign(1, 0);
        llast.assign(1, 0);
        // the root's norm takes |f| for iptype 0 (its children take maxval)
        std::vector<double> rintbs(1, 0.0);
        if (iptype == 0) {
            flush();
            for (size_t i = 0; i < static_cast<size_t>(npbox) * nd; ++i)
                rintbs[0] = std::max(rintbs[0], std::fabs(fvals[i]));
        } else {
            rintbs[0] = box_norm(0, boxlen * boxlen / mc);
        }
        double rint = iptype == 2 ? std::sqrt(rintbs[0]) : rintbs[0];
        const int nlmax = 200;
        int lev = 0;
        for (lev = 0; lev < nlmax - 1; ++lev) {
            double rsc = 1.0;
            if (iptype == 2) rsc = std::sqrt(1.0 / std::pow(boxsize[0], ndim));
            if (iptype == 1) rsc = 1.0 / std::pow(boxsize[0], ndim);
            rsc *= rint;
            const double rscale2 = std::pow(boxsize[lev] / 2, eta);
            const int f = lfirst[lev], l = llast[lev];
            std::vector<char> ref(l - f + 1, 0);
            flush();
            if (zk * boxsize[lev] > 5) {
                std::fill(ref.begin(), ref.end(), 1);
            } else {
                // boxes of one level are tested independently
#pragma omp parallel for schedule(dynamic, 16)
                for (int b = f; b <= l; ++b)
                    ref[b - f] = (tail_error_flushed(b) * rscale2) / rsum > eps * rsc ? 1 : 0;
            }
            bool any = false;
            for (char c : ref) any = any || c;
            if (!any) break;
            boxsize.push_back(boxsize[lev] / 2);
            lfirst.push_back(nb());
            for (int b = f; b <= l; ++b)
                if (ref[b - f]) refine(b, boxsize[lev + 1]);
            llast.push_back(nb() - 1);
            if (max_leaves > 0 && nb() - (nb() - 1) / mc > max_leaves) throw TooLarge{};
            // running norm: refined boxes replaced by their children
            // (iptype 0: the largest child maximum so far)
            const double crs = std::pow(boxsize[lev + 1], ndim) / mc;
            rintbs.resize(nb(), 0.0);
            if (iptype == 0) {
                for (int b = f; b <= l; ++b)
                    if (nchild[b] > 0)
                        for (int j = 0; j < mc; ++j) {
                            const int c = child[static_cast<size_t>(mc) * b + j];
                            rintbs[c] = box_norm(c, crs);
                            if (rintbs[c] > rint) rint = rintbs[c];
                        }
            } else {
                double acc = iptype == 2 ? rint * rint : rint;
                for (int b = f; b <= l; ++b)
                    if (nchild[b] > 0) acc -= rintbs[b];
                for (int b = f; b <= l; ++b)
                    if (nchild[b] > 0)
                        for (int j = 0; j < mc; ++j) {
                            const int c = child[static_cast<size_t>(mc) * b + j];
                            rintbs[c] = box_norm(c, crs);
                            acc += rintbs[c];
                        }
                rint = iptype == 2 ? std::sqrt(acc) : acc;
            }
        }
        nlevels = lev;
        computecoll();
        if (nlevels >= 2) fix_lr();
        if (max_leaves > 0 && nb() - (nb() - 1) / mc > max_leaves) throw TooLarge{};
        flush();
    }
};

thread_local std::string g_tree_err;

}  // namespace

struct fgt_tree_builder {
    Builder b;
};

extern "C" {

int fgt_vol_tree_mem_capped(int ndim, int ipoly, int iperiod, double eps, int norder, double boxlen, int ng,
                            const double* gauss, long long max_leaves, fgt_tree_builder** out, int* nboxes,
                            int* nlevels, int* ltree) {
    if (!out || !nboxes || !nlevels || !ltree || !gauss) return 2;
    *out = nullptr;
    if (ndim < 1 || ndim > 3 || norder < 2 || norder > 16 || ng < 1 || (ipoly != 0 && ipoly != 1) ||
        (iperiod != 0 && iperiod != 1) || !(eps > 0) || !(boxlen > 0) || max_leaves < 0)
        return 2;
    try {
        auto tb = std::make_unique<fgt_tree_builder>();
        Builder& B = tb->b;
        B.ndim = ndim;
        B.ipoly = ipoly;
        B.iperiod = iperiod;
        B.eps = eps;
        B.norder = norder;
        B.boxlen = boxlen;
        B.ng = ng;
        B.max_leaves = max_leaves;
        B.gauss.assign(gauss, gauss + static_cast<size_t>(ng) * (2 * ndim + 1));
        B.build();
        *nboxes = B.nb();
        *nlevels = B.nlevels;
        *ltree = (4 + B.mc + B.mnb) * B.nb() + 2 * (B.nlevels + 1);
        *out = tb.release();
        return 0;
    } catch (const Builder::TooLarge&) {
        g_tree_err = "tree exceeds " + std::to_string(max_leaves) + " leaves";
        return 3;
    } catch (const std::exception& e) {
        g_tree_err = e.what();
        return 5;
    }
}

int fgt_vol_tree_mem(int ndim, int ipoly, int iperiod, double eps, int norder, double boxlen, int ng,
                     const double* gauss, fgt_tree_builder** out, int* nboxes, int* nlevels, int* ltree) {
    return fgt_vol_tree_mem_capped(ndim, ipoly, iperiod, eps, norder, boxlen, ng, gauss, 0, out, nboxes, nlevels,
                                   ltree);
}

int fgt_vol_tree_mem_fun(int ndim, int ipoly, int iperiod, double eps, double zk, double boxlen, int norder,
                         int iptype, double eta, fgt_source_fn fun, int nd, double* dpars, void* zpars, int* ipars,
                         int ifnewtree, fgt_tree_builder** out, int* nboxes, int* nlevels, int* ltree) {
    if (!out || !nboxes || !nlevels || !ltree || !fun) return 2;
    *out = nullptr;
    if (ndim < 1 || ndim > 3 || norder < 2 || norder > 16 || nd < 1 || (ipoly != 0 && ipoly != 1) ||
        (iperiod != 0 && iperiod != 1) || iptype < 0 || iptype > 2 || (ifnewtree != 0 && ifnewtree != 1) ||
        !(eps > 0) || !(boxlen > 0) || !(eta >= 0))
        return 2;
    try {
        auto tb = std::make_unique<fgt_tree_builder>();
        Builder& B = tb->b;
        B.ndim = ndim;
        B.ipoly = ipoly;
        B.iperiod = iperiod;
        B.eps = eps;
        B.norder = norder;
        B.boxlen = boxlen;
        B.fun = fun;
        B.dpars = dpars;
        B.zpars = zpars;
        B.ipars = ipars;
        B.nd = nd;
        B.iptype = iptype;
        B.eta = eta;
        B.zk = zk;
        B.build();
        B.cap = ifnewtree ? 3 * B.nb() : B.nb();
        *nboxes = B.cap;
        *nlevels = B.nlevels;
        *ltree = (4 + B.mc + B.mnb) * B.cap + 2 * (B.nlevels + 1);
        *out = tb.release();
        return 0;
    } catch (const std::exception& e) {
        g_tree_err = e.what();
        return 5;
    }
}

int fgt_vol_tree_newtree_room(fgt_tree_builder* tb, int* nboxes, int* ltree) {
    if (!tb || !nboxes || !ltree) return 2;
    Builder& B = tb->b;
    B.cap = 3 * B.nb();
    *nboxes = B.cap;
    *ltree = (4 + B.mc + B.mnb) * B.cap + 2 * (B.nlevels + 1);
    return 0;
}

int fgt_vol_tree_build(fgt_tree_builder* tb, int* itree, int* iptr, double* centers, double* boxsize,
                       double* fvals) {
    if (!tb || !itree || !iptr || !centers || !boxsize || !fvals) return 2;
    const Builder& B = tb->b;
    const int nb = B.nb(), nl = B.nlevels, mc = B.mc, mnb = B.mnb;
    // sections laid out for the capacity; boxes past the tree's own are zero
    const int cap = std::max(B.cap, nb);
    iptr[0] = 1;
    iptr[1] = 2 * (nl + 1) + 1;
    iptr[2] = iptr[1] + cap;
    iptr[3] = iptr[2] + cap;
    iptr[4] = iptr[3] + cap;
    iptr[5] = iptr[4] + mc * cap;
    iptr[6] = iptr[5] + cap;
    iptr[7] = iptr[6] + mnb * cap;
    std::memset(itree, 0, sizeof(int) * ((size_t)(4 + mc + mnb) * cap + 2 * (nl + 1)));
    for (int l = 0; l <= nl; ++l) {
        itree[2 * l] = B.lfirst[l] + 1;
        itree[2 * l + 1] = B.llast[l] + 1;
    }
    for (int b = 0; b < nb; ++b) {
        itree[iptr[1] - 1 + b] = B.level[b];
        itree[iptr[2] - 1 + b] = B.parent[b] >= 0 ? B.parent[b] + 1 : -1;
        itree[iptr[3] - 1 + b] = B.nchild[b];
        for (int j = 0; j < mc; ++j) {
            const int c = B.child[static_cast<size_t>(mc) * b + j];
            itree[iptr[4] - 1 + mc * b + j] = c >= 0 ? c + 1 : -1;
        }
        itree[iptr[5] - 1 + b] = B.ncoll[b];
        for (int j = 0; j < mnb; ++j) {
            const int c = B.coll[static_cast<size_t>(mnb) * b + j];
            itree[iptr[6] - 1 + mnb * b + j] = c >= 0 ? c + 1 : -1;
        }
    }
    std::memcpy(centers, B.centers.data(), sizeof(double) * B.centers.size());
    std::memset(centers + B.centers.size(), 0, sizeof(double) * (size_t)B.ndim * (cap - nb));
    std::memcpy(boxsize, B.boxsize.data(), sizeof(double) * B.boxsize.size());
    std::memcpy(fvals, B.fvals.data(), sizeof(double) * B.fvals.size());
    std::memset(fvals + B.fvals.size(), 0, sizeof(double) * (size_t)B.npbox * B.nd * (cap - nb));
    delete tb;
    return 0;
}

}  // extern "C"
