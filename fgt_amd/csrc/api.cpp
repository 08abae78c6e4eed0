// api.cpp -- C ABI (include/fgt.h): the boxfgt drop-in and the
// device-resident solver.  Host orchestration of the stage kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fgt.h"
#include "kernels.h"
#include "plan.h"

namespace {

// eps floors of the reduced-precision contraction modes (their error against
// the FP64 result stays below eps above these)
constexpr double kEpsFp32 = 1e-6 * (1.0 - 1e-12);
constexpr double kEpsBf16 = 1e-3 * (1.0 - 1e-12);

thread_local int g_status = 0;
thread_local std::string g_msg;

void set_status(int st, const std::string& m) {
    g_status = st;
    g_msg = m;
}

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Bump allocator over a caller-provided device range (the no-allocation
// entry point's scratch): while one is installed on this thread, device
// buffers are carved from it instead of hipMalloc'd.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* take(size_t b) {
        const size_t off = (used + 255) & ~size_t(255);
        if (off + b > cap)
            throw std::invalid_argument("problem needs more device workspace than the " +
                                        std::to_string(cap >> 20) + " MiB no-allocation scratch");
        used = off + b;
        return base + off;
    }
};
thread_local Arena* g_arena = nullptr;

// A device buffer owned by the solver (or carved from the installed arena).
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool from_arena = false;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p && !from_arena) (void)hipFree(p);
        p = nullptr;
        from_arena = false;
    }
    void alloc(size_t b) {
        release();
        bytes = b;
        if (!b) return;
        if (g_arena) {
            p = g_arena->take(b);
            from_arena = true;
        } else {
            hcheck(hipMalloc(&p, b), "hipMalloc");
        }
    }
    template <class T>
    void upload(const std::vector<T>& v) {
        alloc(sizeof(T) * v.size());
        if (!v.empty()) hcheck(hipMemcpy(p, v.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct fgt_solver {
    fgt::Plan plan;
    fgt::Geom geom{};
    fgt::DevTables tabs{};
    // tables
    DBuf d_p2pw, d_pw2pot, d_pw2potT, d_pw2potS, d_wms, d_wm2l, d_one, d_loctab, d_locband;
    // data
    DBuf d_fvals, d_pot, d_expn;
    int ifpgh = 1;          // 1 pot, 2 + grad, 3 + hess
    DBuf d_grad, d_hess;    // grad(nd,ndim,npbox,nboxes), hess(nd,nhess,npbox,nboxes)
    // extra targets (STEP 11): set up on first use
    bool t_ready = false;
    DBuf d_t_u, d_t_leaves, d_t_centers, d_t_bsz, d_t_coefs;
    int t_nleaf = 0;
    // lists
    DBuf d_leaf_box, d_leaf_lev, d_leaf_slot;
    DBuf d_par_slot, d_par_first, d_par_count, d_pc_parent, d_pc_child, d_pc_sign, d_pc_sign_split;
    DBuf d_m2l_tgt, d_m2l_off, d_m2l_src, d_m2l_idx;
    DBuf d_m2l_toff, d_m2l_poff, d_m2l_soff, d_m2l_so;   // STEP 3 entry streams per group
    int n_m2l_groups = 0;
    DBuf d_dir_tgt, d_dir_lev, d_dir_off, d_dir_src, d_dir_code, d_dir_flags;
    // 3-D fast path: STEP 4 folded into STEP 5 for leaves below the cutoff level
    bool fast3d = false;
    DBuf d_s5_tasks;   // fgt::PotTask list of the 3-D STEP 5 kernel
    int n_s5 = 0, n_s5_sib = 0;   // tasks, of which the first n_s5_sib are sibling tasks
    DBuf d_p2pwS, d_fz_lev, d_fz_slot, d_fz_kids, d_fz_ksg;          // fused STEP 1+2 parents
    DBuf d_s1_box, d_s1_lev, d_s1_slot;                              // standalone STEP 1 leaves
    DBuf d_mp_slot, d_mp_first, d_mp_count, d_mp_child, d_mp_sign, d_mp_acc;  // merges of non-leaf children
    std::vector<int> mp_level_off;
    int n_fz = 0, n_s1 = 0;
    DBuf d_pcs_parent, d_pcs_child, d_pcs_sign;   // splits into non-leaf children only
    std::vector<int> pcs_level_off;
    double stage_ms[8] = {0};
    // STEP 6 on a second stream, overlapping STEPs 1-4 (it reads only fvals;
    // STEP 5 waits for it, so pot gets each sum once): made on first use
    hipStream_t st_near = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    ~fgt_solver() {
        if (st_near) (void)hipStreamDestroy(st_near);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
    }
};

// Many independent problems (uneven sizes, orders, delta) solved together:
// each has its own plan and device buffers; their executes are spread over
// a few HIP streams (largest first, each onto the least-loaded stream) so
// small problems overlap on the device instead of running one after another.
struct fgt_batch {
    std::vector<std::unique_ptr<fgt_solver>> solvers;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_of;     // per problem
    std::vector<int> order;         // launch order (largest first)
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double last_ms = 0.0;
    ~fgt_batch() {
        for (auto s : streams) (void)hipStreamDestroy(s);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
    }
};

namespace {

// Map problem-axis sign bits / offsets / table codes onto the (x, y, z) axes
// of the kernels: 3-D (x,y,z); 2-D (x,_,z); 1-D (_,_,z).
int map_sign(int ndim, int s) {
    if (ndim == 3) return s;
    if (ndim == 2) return (s & 1) | ((s & 2) << 1);
    return (s & 1) << 2;
}
int real_mask(int ndim) { return ndim == 3 ? 7 : (ndim == 2 ? 5 : 4); }
int map_m2l(int ndim, int idx) {
    if (ndim == 3) return idx;
    if (ndim == 2) return (idx % 3) + 3 * 1 + 9 * (idx / 3);
    return 1 + 3 * 1 + 9 * idx;
}
int map_code(int ndim, int c) {
    const int a0 = c & 15, a1 = (c >> 4) & 15, a2 = (c >> 8) & 15;
    if (ndim == 3) return a0 | (a1 << 4) | (a2 << 8);
    if (ndim == 2) return a0 | (6 << 4) | (a1 << 8);
    return 6 | (6 << 4) | (a0 << 8);
}

std::vector<double2> to_d2(const std::vector<fgt::cplx>& v) {
    std::vector<double2> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) o[i] = make_double2(v[i].real(), v[i].imag());
    return o;
}

// Expansion workspace: nslots slots of (mp, loc), nd * nexp complex each,
// elem bytes per complex (16 FP64, 8 FP32); one extra slot past the last
// stays zero (padding entries of the grouped STEP 3 lists read it).
void alloc_expansions(fgt_solver& S, size_t elem) {
    const fgt::Plan& P = S.plan;
    const size_t slot_elems = 2 * (size_t)P.nd * P.nexp;
    S.d_expn.alloc(elem * slot_elems * (P.nslots + 1));
    // an arena's contents persist between calls: its user clears the zero
    // slot on its own stream before each execute
    if (!S.d_expn.from_arena)
        hcheck(hipMemset(static_cast<char*>(S.d_expn.p) + elem * slot_elems * P.nslots, 0, elem * slot_elems),
               "zero slot");
}

// own_io = false: the caller brings device fvals / pot (no-allocation
// entry point), so the solver keeps none of its own.
void build_solver(fgt_solver& S, const fgt::PlanInput& in, bool own_io = true) {
    std::string err;
    const int rc = fgt::make_plan(in, S.plan, &err);
    if (rc != 0) throw std::invalid_argument(err);
    const fgt::Plan& P = S.plan;
    const int ndim = P.ndim, n = P.n;
    fgt::Geom& g = S.geom;
    g.nd = P.nd;
    g.ndim = ndim;
    g.nx = ndim == 1 ? 1 : n;
    g.ny = ndim == 3 ? n : 1;
    g.nz = n;
    g.px = ndim == 1 ? 1 : P.npw;
    g.py = ndim == 3 ? P.npw : 1;
    g.pz = P.nh;
    g.npw = P.npw;
    g.npbox = P.npbox;
    g.nexp = P.nexp;

    // device limits for the LDS-staged kernels
    int dev = 0;
    hcheck(hipGetDevice(&dev), "hipGetDevice");
    int lds_max = 0;   // one attribute query (the full property struct is slow to fill)
    hcheck(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
           "hipDeviceGetAttribute");
    const size_t lds_limit = static_cast<size_t>(lds_max);
    const size_t lds_p2pw = sizeof(double2) * ((size_t)P.npw * n + (size_t)g.nx * g.ny + (size_t)g.py * g.nx);
    const size_t lds_pw2pot = sizeof(double2) * ((size_t)n * P.npw + (size_t)g.nx * g.py);
    if (P.do_pw && (lds_p2pw > lds_limit || lds_pw2pot > lds_limit))
        throw std::invalid_argument("plane-wave count too large for the LDS-staged kernels");

    if (own_io) {
        S.d_fvals.alloc(sizeof(double) * (size_t)P.nd * P.npbox * P.nboxes);
        S.d_pot.alloc(sizeof(double) * (size_t)P.nd * P.npbox * P.nboxes);
    }

    S.d_loctab.upload(P.loctab);
    S.d_locband.upload(P.locband);
    std::vector<double2> one{make_double2(1.0, 0.0)};
    S.d_one.upload(one);
    if (P.do_pw) {
        S.d_p2pw.upload(to_d2(P.p2pw));
        S.d_pw2pot.upload(to_d2(P.pw2pot));
        std::vector<fgt::cplx> tr(P.pw2pot.size());
        const size_t per = (size_t)n * P.npw;
        for (size_t l = 0; l * per < P.pw2pot.size(); ++l)
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < P.npw; ++k) tr[l * per + (size_t)k * n + i] = P.pw2pot[l * per + (size_t)i * P.npw + k];
        S.d_pw2potT.upload(to_d2(tr));
        S.d_wms.upload(to_d2(P.wms));
        S.d_wm2l.upload(to_d2(P.wm2l));
        alloc_expansions(S, sizeof(double2));
    }
    fgt::DevTables& t = S.tabs;
    t.p2pw = S.d_p2pw.as<double2>();
    t.pw2pot = S.d_pw2pot.as<double2>();
    t.pw2potT = S.d_pw2potT.as<double2>();
    t.pw2pot_ks = P.pw2pot_ks;
    t.loctab_ks = P.loctab_ks;
    t.lds_limit = lds_limit;
    t.n = n;
    t.npw = P.npw;
    S.fast3d = P.do_pw && fgt::pw2pot_3d_supported(g, t);
    if (S.fast3d) {
        const int nl = P.nlevels, np = P.npw;
        const size_t per = (size_t)n * np;
        // [kind][l][v][k][i], kind = value / first / second derivative table
        const size_t sks = 3 * per * (nl + 1);
        std::vector<fgt::cplx> ts(3 * sks, fgt::cplx(0.0));
        for (int kind = 0; kind < 3; ++kind)
            for (int l = P.nlevstart; l <= nl; ++l) {
                const int klev = nl - (l - 1);
                const fgt::cplx* w = (l > P.nlevstart) ? &P.wms[(size_t)(klev - 1) * np] : nullptr;
                const fgt::cplx* U = &P.pw2pot[kind * P.pw2pot_ks + l * per];
                fgt::cplx* o = &ts[kind * sks + 3 * (size_t)l * per];
                for (int k = 0; k < np; ++k)
                    for (int i = 0; i < n; ++i) {
                        const fgt::cplx u = U[(size_t)i * np + k];
                        o[(size_t)k * n + i] = u;
                        if (w) {
                            o[per + (size_t)k * n + i] = u * w[k];
                            o[2 * per + (size_t)k * n + i] = u * std::conj(w[k]);
                        }
                    }
            }
        S.d_pw2potS.upload(to_d2(ts));
        t.pw2potS_ks = sks;
        const fgt::TreeView& T = P.tree;
        // STEP 5 tasks: leaves below the first PW level evaluate their
        // parent's expansion through the split-phased tables; siblings with
        // the same x split sign form one task (shared x contraction).  Those
        // tasks come first (more work per task), then one task per leaf
        // evaluating its own expansion.
        std::vector<fgt::PotTask> fam, own;
        std::unordered_map<long long, int> fam_of;
        for (size_t i = 0; i < P.leaf_box.size(); ++i) {
            const int b = P.leaf_box[i], l = P.leaf_lev[i];
            if (l > P.nlevstart) {
                const int par = T.parent[b];
                int sg[3];
                for (int a = 0; a < 3; ++a) sg[a] = T.centers[3 * b + a] - T.centers[3 * par + a] < 0 ? 1 : 0;
                const long long key = 2LL * par + sg[0];
                auto it = fam_of.find(key);
                if (it == fam_of.end() || fam[it->second].box[3] >= 0) {
                    fgt::PotTask pt{};
                    pt.slot = P.slot_of_box[par];
                    pt.lev = l;
                    pt.vx = 1 + sg[0];
                    pt.nvy = 0;
                    for (int c = 0; c < 4; ++c) pt.box[c] = -1, pt.code[c] = 0;
                    fam_of[key] = (int)fam.size();
                    fam.push_back(pt);
                    it = fam_of.find(key);
                }
                fgt::PotTask& pt = fam[it->second];
                const int vy = 1 + sg[1];
                int j = 0;
                while (j < pt.nvy && pt.vy[j] != vy) ++j;
                if (j == pt.nvy) pt.vy[pt.nvy++] = vy;
                int c = 0;
                while (pt.box[c] >= 0) ++c;
                pt.box[c] = b;
                pt.code[c] = j | ((1 + sg[2]) << 1);
            } else {
                fgt::PotTask pt{};
                pt.slot = P.leaf_slot[i];
                pt.lev = l;
                pt.vx = 0;
                pt.nvy = 1;
                pt.vy[0] = pt.vy[1] = 0;
                for (int c = 0; c < 4; ++c) pt.box[c] = -1, pt.code[c] = 0;
                pt.box[0] = b;
                own.push_back(pt);
            }
        }
        S.n_s5_sib = (int)fam.size();
        fam.insert(fam.end(), own.begin(), own.end());
        S.n_s5 = (int)fam.size();
        S.d_s5_tasks.upload(fam);
        // split entries, keeping non-leaf children only
        std::vector<int> pp, pcc, psg;
        S.pcs_level_off.assign(nl + 2, 0);
        for (int l = 0; l <= nl; ++l) {
            S.pcs_level_off[l] = (int)pp.size();
            if (l < P.nlevstart || l >= nl) continue;
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                if (T.nchild[b] == 0) continue;
                for (int i = 0; i < T.nchild[b]; ++i) {
                    const int c = T.child[(size_t)T.mc * b + i];
                    if (T.nchild[c] == 0) continue;
                    int k = 0;
                    for (int a = 0; a < 3; ++a)
                        if (T.centers[3 * c + a] - T.centers[3 * b + a] < 0) k |= 1 << a;
                    pp.push_back(P.slot_of_box[b]);
                    pcc.push_back(P.slot_of_box[c]);
                    psg.push_back(k);
                }
            }
        }
        S.pcs_level_off[nl + 1] = (int)pp.size();
        S.d_pcs_parent.upload(pp);
        S.d_pcs_child.upload(pcc);
        S.d_pcs_sign.upload(psg);

        // fused STEP 1+2: value->PW tables times the child->parent merge phase
        std::vector<fgt::cplx> t1(3 * per * (nl + 1), fgt::cplx(0.0));
        for (int l = P.nlevstart; l <= nl; ++l) {
            const int klev = nl - (l - 1);
            const fgt::cplx* w = (l > P.nlevstart) ? &P.wms[(size_t)(klev - 1) * np] : nullptr;
            for (int k = 0; k < np; ++k)
                for (int m = 0; m < n; ++m) {
                    const fgt::cplx v = P.p2pw[l * per + (size_t)k * n + m];
                    t1[(3 * (size_t)l + 0) * per + (size_t)k * n + m] = v;
                    if (w) {
                        t1[(3 * (size_t)l + 1) * per + (size_t)k * n + m] = v * w[k];
                        t1[(3 * (size_t)l + 2) * per + (size_t)k * n + m] = v * std::conj(w[k]);
                    }
                }
        }
        S.d_p2pwS.upload(to_d2(t1));
        std::vector<int> fzl, fzs, fzk, fzg, s1b, s1l, s1s;
        std::vector<int> mps, mpf, mpc, mch, msg, macc;
        S.mp_level_off.assign(nl + 2, 0);
        for (int l = 0; l <= nl; ++l) {
            S.mp_level_off[l] = (int)mps.size();
            if (l < P.nlevstart) continue;
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                if (T.nchild[b] == 0) {
                    if (l == P.nlevstart && P.ifpwexp[b]) {
                        s1b.push_back(b);
                        s1l.push_back(l);
                        s1s.push_back(P.slot_of_box[b]);
                    }
                    continue;
                }
                if (l >= nl) continue;
                int kids[8], ksg[8], nk = 0, nnl = 0;
                for (int c = 0; c < 8; ++c) kids[c] = -1, ksg[c] = 0;
                for (int i = 0; i < T.nchild[b]; ++i) {
                    const int c = T.child[(size_t)T.mc * b + i];
                    int k = 0;
                    for (int a = 0; a < 3; ++a)
                        if (T.centers[3 * b + a] - T.centers[3 * c + a] < 0) k |= 1 << a;
                    if (T.nchild[c] == 0) {
                        kids[nk] = c;
                        ksg[nk] = k;
                        ++nk;
                    } else {
                        if (nnl == 0) {
                            mps.push_back(P.slot_of_box[b]);
                            mpf.push_back((int)mch.size());
                        }
                        mch.push_back(P.slot_of_box[c]);
                        msg.push_back(k);
                        ++nnl;
                    }
                }
                if (nnl > 0) {
                    mpc.push_back(nnl);
                    macc.push_back(nk > 0 ? 1 : 0);
                }
                if (nk > 0) {
                    fzl.push_back(l);
                    fzs.push_back(P.slot_of_box[b]);
                    for (int c = 0; c < 8; ++c) {
                        fzk.push_back(kids[c]);
                        fzg.push_back(ksg[c]);
                    }
                }
            }
        }
        S.mp_level_off[nl + 1] = (int)mps.size();
        S.n_fz = (int)fzl.size();
        S.n_s1 = (int)s1b.size();
        S.d_fz_lev.upload(fzl);
        S.d_fz_slot.upload(fzs);
        S.d_fz_kids.upload(fzk);
        S.d_fz_ksg.upload(fzg);
        S.d_s1_box.upload(s1b);
        S.d_s1_lev.upload(s1l);
        S.d_s1_slot.upload(s1s);
        S.d_mp_slot.upload(mps);
        S.d_mp_first.upload(mpf);
        S.d_mp_count.upload(mpc);
        S.d_mp_child.upload(mch);
        S.d_mp_sign.upload(msg);
        S.d_mp_acc.upload(macc);
    }
    t.p2pwS = S.d_p2pwS.as<double2>();
    t.pw2potS = S.d_pw2potS.as<double2>();
    t.wms = S.d_wms.as<double2>();
    t.wm2l = S.d_wm2l.as<double2>();
    t.one = S.d_one.as<double2>();
    t.loctab = S.d_loctab.as<double>();
    t.locband = S.d_locband.as<int>();
    t.n = n;
    t.npw = P.npw;

    // work lists
    S.d_leaf_box.upload(P.leaf_box);
    S.d_leaf_lev.upload(P.leaf_lev);
    S.d_leaf_slot.upload(P.leaf_slot);
    S.d_par_slot.upload(P.par_slot);
    S.d_par_first.upload(P.par_first);
    S.d_par_count.upload(P.par_count);
    S.d_pc_parent.upload(P.pc_parent);
    S.d_pc_child.upload(P.pc_child);
    std::vector<int> sg(P.pc_sign.size()), sgs(P.pc_sign.size());
    for (size_t i = 0; i < sg.size(); ++i) {
        sg[i] = map_sign(ndim, P.pc_sign[i]);
        sgs[i] = sg[i] ^ real_mask(ndim);
    }
    S.d_pc_sign.upload(sg);
    S.d_pc_sign_split.upload(sgs);
    S.d_m2l_tgt.upload(P.m2l_tgt);
    S.d_m2l_off.upload(P.m2l_off);
    S.d_m2l_src.upload(P.m2l_src);
    std::vector<int> mi(P.m2l_idx.size());
    for (size_t i = 0; i < mi.size(); ++i) mi[i] = map_m2l(ndim, P.m2l_idx[i]);
    S.d_m2l_idx.upload(mi);
    if (P.do_pw) {
        // STEP 3 entry streams: kM2lGroup consecutive targets per group; per
        // target its own expansion (unit phase, o = 13), then its list, the
        // final entry flagged; each group padded with zero-slot entries
        const long long se = 2LL * P.nd * P.nexp;
        const int nt = (int)P.m2l_tgt.size();
        const int G = fgt::m2l_group();
        const int ng = (nt + G - 1) / G;
        std::vector<long long> toff((size_t)ng * G, 0), soff;
        std::vector<int> gent(ng + 1, 0), code;
        for (int gi = 0; gi < ng; ++gi) {
            gent[gi] = (int)soff.size();
            for (int tk = gi * G; tk < std::min(nt, (gi + 1) * G); ++tk) {
                toff[tk] = se * P.m2l_tgt[tk];
                soff.push_back(toff[tk]);
                code.push_back(13);
                for (int j = P.m2l_off[tk]; j < P.m2l_off[tk + 1]; ++j) {
                    soff.push_back(se * P.m2l_src[j]);
                    code.push_back(mi[j]);
                }
                code.back() |= 1 << 5;
            }
            while ((soff.size() - gent[gi]) % fgt::kM2lBatch) {
                soff.push_back(se * P.nslots);   // the zero slot
                code.push_back(13);
            }
            if ((int)soff.size() - gent[gi] > fgt::kM2lMaxEntries)
                throw std::runtime_error("STEP 3 group entry stream longer than kM2lMaxEntries");
        }
        gent[ng] = (int)soff.size();
        S.n_m2l_groups = ng;
        S.d_m2l_toff.upload(toff);
        S.d_m2l_poff.upload(gent);
        S.d_m2l_soff.upload(soff);
        S.d_m2l_so.upload(code);
    }

    // near-field targets, longest list first (one block per target: the
    // longest-processing-time order shortens the tail of the launch); entries
    // of a target sorted by (z, y, x) table slot so the 3-D kernel can share
    // the y and z contractions within a group
    const int ndt = (int)P.dir_tgt.size();
    std::vector<int> order(ndt);
    for (int i = 0; i < ndt; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return P.dir_off[x + 1] - P.dir_off[x] > P.dir_off[y + 1] - P.dir_off[y];
    });
    std::vector<int> dt(ndt), dl(ndt), doff(ndt + 1, 0);
    std::vector<int> ds(P.dir_src.size()), dc(P.dir_code.size()), df(P.dir_code.size(), 0);
    for (int q = 0; q < ndt; ++q) {
        const int t2 = order[q];
        const int a0 = P.dir_off[t2], b0 = P.dir_off[t2 + 1];
        dt[q] = P.dir_tgt[t2];
        dl[q] = P.dir_lev[t2];
        const int a = doff[q], b = a + (b0 - a0);
        doff[q + 1] = b;
        std::vector<int> idx(b0 - a0);
        for (int i = a0; i < b0; ++i) idx[i - a0] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            const int cx = map_code(ndim, P.dir_code[x]), cy = map_code(ndim, P.dir_code[y]);
            return (cx >> 4) < (cy >> 4);
        });
        for (int i = a; i < b; ++i) {
            ds[i] = P.dir_src[idx[i - a]];
            dc[i] = map_code(ndim, P.dir_code[idx[i - a]]);
        }
        for (int i = a; i < b; ++i) {
            const bool last = (i + 1 == b);
            if (last || (dc[i] >> 4) != (dc[i + 1] >> 4)) df[i] |= 1;
            if (last || (dc[i] >> 8) != (dc[i + 1] >> 8)) df[i] |= 2;
        }
    }
    S.d_dir_tgt.upload(dt);
    S.d_dir_lev.upload(dl);
    S.d_dir_off.upload(doff);
    S.d_dir_src.upload(ds);
    S.d_dir_code.upload(dc);
    S.d_dir_flags.upload(df);
}

// Output components of ifpgh and the table kind each uses per axis
// (gnd_pw2pgh, bfgt_pwrouts.f:233-1065; gnd_tens_prod_to_pghloc,
// bfgt_local.f:14-920): pot; grad x, y, z; Hessian in the reference's order
// (1-D xx; 2-D xx, xy, yy; 3-D xx, yy, zz, xy, xz, yz).  Reference axes map
// onto the kernels' (x, y, z) like the geometry: 2-D (x, y) -> (x, z),
// 1-D x -> z.
struct Comp {
    fgt::Kinds kd;
    fgt::OutDesc od;
};

std::vector<Comp> components(fgt_solver& S, double* pot_override = nullptr) {
    const fgt::Plan& P = S.plan;
    const int ndim = P.ndim, nd = P.nd, nhess = ndim * (ndim + 1) / 2;
    auto axes = [&](int r0, int r1, int r2) {
        if (ndim == 3) return fgt::Kinds{r0, r1, r2};
        if (ndim == 2) return fgt::Kinds{r0, 0, r1};
        return fgt::Kinds{0, 0, r0};
    };
    std::vector<Comp> c;
    if (pot_override) {
        c.push_back({axes(0, 0, 0), {pot_override, nd, 0}});
        return c;
    }
    c.push_back({axes(0, 0, 0), {S.d_pot.as<double>(), nd, 0}});
    if (S.ifpgh >= 2)
        for (int a = 0; a < ndim; ++a)
            c.push_back({axes(a == 0, a == 1, a == 2), {S.d_grad.as<double>(), nd * ndim, a * nd}});
    if (S.ifpgh >= 3) {
        static const int h1[1][3] = {{2, 0, 0}};
        static const int h2[3][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}};
        static const int h3[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
        const int(*h)[3] = ndim == 1 ? h1 : (ndim == 2 ? h2 : h3);
        for (int k = 0; k < nhess; ++k)
            c.push_back({axes(h[k][0], h[k][1], h[k][2]), {S.d_hess.as<double>(), nd * nhess, k * nd}});
    }
    return c;
}

// fork_near: run STEP 6 on the solver's second stream, overlapping STEPs
// 1-4 on `st` (single-problem executes; a batch already overlaps problems).
// The outputs are the same bit for bit: each output node receives one STEP 5
// and one STEP 6 sum onto zero, and two-term addition is commutative.
// Caller-provided buffers (no-allocation entry point): device fvals, a
// device potential that is accumulated into (not zeroed) and the expansion
// workspace; the potential only.
struct ExecOverride {
    const double* fvals;
    double* pot;
    double2* expn;
};

void execute(fgt_solver& S, hipStream_t st, bool timed, bool fork_near = false,
             const ExecOverride* ov = nullptr) {
    const fgt::Plan& P = S.plan;
    const fgt::Geom& g = S.geom;
    const fgt::DevTables& t = S.tabs;
    const double* fv = ov ? ov->fvals : S.d_fvals.as<double>();
    double2* ex = ov ? ov->expn : S.d_expn.as<double2>();
    if (ov) timed = fork_near = false;
    hipEvent_t ev[8], evn[2];
    if (timed) {
        for (auto& e : ev) hcheck(hipEventCreate(&e), "hipEventCreate");
        for (auto& e : evn) hcheck(hipEventCreate(&e), "hipEventCreate");
    }
    auto mark = [&](int i) { if (timed) hcheck(hipEventRecord(ev[i], st), "hipEventRecord"); };
    if (!ov) {
        hcheck(hipMemsetAsync(S.d_pot.p, 0, S.d_pot.bytes, st), "hipMemsetAsync pot");
        if (S.ifpgh >= 2) hcheck(hipMemsetAsync(S.d_grad.p, 0, S.d_grad.bytes, st), "hipMemsetAsync grad");
        if (S.ifpgh >= 3) hcheck(hipMemsetAsync(S.d_hess.p, 0, S.d_hess.bytes, st), "hipMemsetAsync hess");
    }
    const std::vector<Comp> comps = components(S, ov ? ov->pot : nullptr);
    auto near_field = [&](hipStream_t sn) {
        for (const Comp& c : comps)
            fgt::launch_direct(g, t, (int)P.dir_tgt.size(), S.d_dir_tgt.as<int>(), S.d_dir_lev.as<int>(),
                               S.d_dir_off.as<int>(), S.d_dir_src.as<int>(), S.d_dir_code.as<int>(),
                               S.d_dir_flags.as<int>(), fv, c.kd, c.od, sn);
    };
    const bool fork = fork_near && P.do_pw && !P.dir_tgt.empty();
    if (fork) {
        if (!S.st_near) {
            hcheck(hipStreamCreateWithFlags(&S.st_near, hipStreamNonBlocking), "hipStreamCreate near");
            hcheck(hipEventCreateWithFlags(&S.ev_fork, hipEventDisableTiming), "hipEventCreate");
            hcheck(hipEventCreateWithFlags(&S.ev_join, hipEventDisableTiming), "hipEventCreate");
        }
        hcheck(hipEventRecord(S.ev_fork, st), "hipEventRecord");   // after the zeroing
        hcheck(hipStreamWaitEvent(S.st_near, S.ev_fork, 0), "hipStreamWaitEvent");
        if (timed) hcheck(hipEventRecord(evn[0], S.st_near), "hipEventRecord");
        near_field(S.st_near);
        if (timed) hcheck(hipEventRecord(evn[1], S.st_near), "hipEventRecord");
        hcheck(hipEventRecord(S.ev_join, S.st_near), "hipEventRecord");
    }
    mark(0);
    if (P.do_pw) {
        if (S.fast3d) {
            // STEP 1: leaves at the cutoff level alone; every other PW leaf is
            // folded into its parent's expansion (STEP 1 + 2 fused), all levels
            // in one launch (they read only fvals)
            fgt::launch_p2pw(g, t, S.n_s1, S.d_s1_box.as<int>(), S.d_s1_lev.as<int>(), S.d_s1_slot.as<int>(),
                             fv, ex, st);
            if (!fgt::p2pw_fused_3d(g, t, S.n_fz, S.d_fz_lev.as<int>(), S.d_fz_slot.as<int>(),
                                    S.d_fz_kids.as<int>(), S.d_fz_ksg.as<int>(), fv,
                                    ex, st))
                throw std::runtime_error("fused STEP 1+2 kernel not available for this shape");
            hcheck(hipGetLastError(), "k_p2pwF3 launch");
            mark(1);
            // STEP 2: non-leaf children, fine to coarse
            for (int l = P.nlevels - 1; l >= P.nlevstart; --l) {
                const int a = S.mp_level_off[l], b = S.mp_level_off[l + 1];
                fgt::launch_merge(g, t, b - a, S.d_mp_slot.as<int>() + a, S.d_mp_first.as<int>() + a,
                                  S.d_mp_count.as<int>() + a, S.d_mp_child.as<int>(), S.d_mp_sign.as<int>(),
                                  S.d_mp_acc.as<int>() + a, P.nlevels - l, ex, st);
            }
            mark(2);
        } else {
            // STEP 1: every PW leaf, all levels in one launch
            fgt::launch_p2pw(g, t, (int)P.leaf_box.size(), S.d_leaf_box.as<int>(), S.d_leaf_lev.as<int>(),
                             S.d_leaf_slot.as<int>(), fv, ex, st);
            mark(1);
            // STEP 2: fine to coarse
            for (int l = P.nlevels - 1; l >= P.nlevstart; --l) {
                const int a = P.par_level_off[l], b = P.par_level_off[l + 1];
                fgt::launch_merge(g, t, b - a, S.d_par_slot.as<int>() + a, S.d_par_first.as<int>() + a,
                                  S.d_par_count.as<int>() + a, S.d_pc_child.as<int>(), S.d_pc_sign.as<int>(),
                                  nullptr, P.nlevels - l, ex, st);
            }
            mark(2);
        }
        // STEP 3
        fgt::launch_m2l27(g, t, S.n_m2l_groups, fgt::m2l_group(), 2LL * P.nd * P.nexp * P.nslots,
                          S.d_m2l_toff.as<long long>(), S.d_m2l_poff.as<int>(), S.d_m2l_soff.as<long long>(),
                          S.d_m2l_so.as<int>(), ex, st);
        mark(3);
        // STEP 4: coarse to fine (fast 3-D path: only into non-leaf children;
        // leaves read their parent's expansion in STEP 5 through phased tables)
        for (int l = P.nlevstart; l <= P.nlevels - 1; ++l) {
            if (S.fast3d) {
                const int a = S.pcs_level_off[l], b = S.pcs_level_off[l + 1];
                fgt::launch_split(g, t, b - a, S.d_pcs_parent.as<int>() + a, S.d_pcs_child.as<int>() + a,
                                  S.d_pcs_sign.as<int>() + a, P.nlevels - l, ex, st);
            } else {
                const int a = P.pc_level_off[l], b = P.pc_level_off[l + 1];
                fgt::launch_split(g, t, b - a, S.d_pc_parent.as<int>() + a, S.d_pc_child.as<int>() + a,
                                  S.d_pc_sign_split.as<int>() + a, P.nlevels - l, ex, st);
            }
        }
        mark(4);
        // STEP 5, once per output component (pot, grad, hess entries)
        if (fork) {
            hcheck(hipStreamWaitEvent(st, S.ev_join, 0), "hipStreamWaitEvent");
            mark(7);   // STEP 5 starts here (after the wait for STEP 6)
        }
        for (const Comp& c : comps) {
            if (S.fast3d)
                fgt::launch_pw2pot_tasks(g, t, S.n_s5, S.n_s5_sib, S.d_s5_tasks.as<fgt::PotTask>(),
                                         ex, c.kd, c.od, st);
            else
                fgt::launch_pw2pot(g, t, (int)P.leaf_box.size(), S.d_leaf_box.as<int>(), S.d_leaf_lev.as<int>(),
                                   S.d_leaf_slot.as<int>(), ex, c.kd, c.od, st);
        }
        mark(5);
    } else {
        for (int i = 1; i <= 5; ++i) mark(i);
    }
    // STEP 6
    if (!fork) near_field(st);
    mark(6);
    if (timed) {
        hcheck(hipEventSynchronize(ev[6]), "hipEventSynchronize");
        for (int i = 0; i < 6; ++i) {
            float ms = 0;
            // with the fork: STEP 5 from the end of the wait, STEP 6 on its own stream
            const hipEvent_t a = (fork && i == 4) ? ev[7] : ev[i], b = (fork && i == 5) ? evn[1] : ev[i + 1];
            hcheck(hipEventElapsedTime(&ms, (fork && i == 5) ? evn[0] : a, b), "hipEventElapsedTime");
            S.stage_ms[i] = ms;
        }
        S.stage_ms[6] = 0.0;
        if (fork) {
            // whole execute, first to last event of either stream
            float ms = 0;
            hcheck(hipEventElapsedTime(&ms, ev[0], ev[6]), "hipEventElapsedTime");
            S.stage_ms[6] = ms;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (auto& e : evn) (void)hipEventDestroy(e);
    }
}

// STEP 11 (extra targets, src/bfgt/boxfgt.f:1343-1378): values of the last
// execute (device pot / grad / hess) -> coefficients on every leaf -> values
// at the targets.  Targets are assigned to leaves on the host by the
// reference's root-to-leaf descent (child bit a set when t_a - c_a >= 0,
// find_childbox_ind, src/common/pts_tree_nd.f:713-740) and grouped by leaf.
void eval_targets(fgt_solver& S, int nt, const double* targs, int ifpghtarg, double* pote, double* grade,
                  double* hesse, hipStream_t st) {
    const fgt::Plan& P = S.plan;
    const fgt::TreeView& T = P.tree;
    const int ndim = P.ndim, n = P.n, nd = P.nd, mc = 1 << ndim;
    if (ifpghtarg > S.ifpgh) throw std::invalid_argument("targets need ifpgh >= ifpghtarg on the solver");
    if (nt <= 0) return;
    if (!S.t_ready) {
        std::vector<double> x(n), w(n), u((size_t)n * n);
        fgt::poly_exps(P.ipoly, n, x.data(), w.data(), u.data());
        S.d_t_u.upload(u);
        std::vector<int> leaves;
        std::vector<double> bsz(P.nboxes);
        for (int b = 0; b < P.nboxes; ++b) {
            if (T.nchild[b] == 0) leaves.push_back(b);
            bsz[b] = T.boxsize[T.level[b]];
        }
        S.t_nleaf = (int)leaves.size();
        S.d_t_leaves.upload(leaves);
        S.d_t_centers.upload(T.centers);
        S.d_t_bsz.upload(bsz);
        S.t_ready = true;
    }
    std::vector<int> leaf(nt), perm(nt);
    for (int it = 0; it < nt; ++it) {
        int b = 0;
        while (T.nchild[b] > 0) {
            int k = 0;
            for (int a = 0; a < ndim; ++a)
                if (targs[(size_t)ndim * it + a] - T.centers[(size_t)ndim * b + a] >= 0) k |= 1 << a;
            const int c = T.child[(size_t)mc * b + k];
            if (c < 0) throw std::invalid_argument("tree has a box with a missing child");
            b = c;
        }
        leaf[it] = b;
        perm[it] = it;
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return leaf[a] < leaf[b]; });
    std::vector<double> ts((size_t)nt * ndim);
    std::vector<int> tbox, tfirst, tcount;
    for (int i = 0; i < nt; ++i) {
        for (int a = 0; a < ndim; ++a) ts[(size_t)ndim * i + a] = targs[(size_t)ndim * perm[i] + a];
        if (i == 0 || leaf[perm[i]] != leaf[perm[i - 1]] || tcount.back() == 256) {
            tbox.push_back(leaf[perm[i]]);
            tfirst.push_back(i);
            tcount.push_back(0);
        }
        ++tcount.back();
    }
    DBuf d_ts, d_perm, d_tbox, d_tfirst, d_tcount, d_out;
    d_ts.upload(ts);
    d_perm.upload(perm);
    d_tbox.upload(tbox);
    d_tfirst.upload(tfirst);
    d_tcount.upload(tcount);
    const int nhess = ndim * (ndim + 1) / 2;
    struct Out { const DBuf* vals; int ncomp; double* host; };
    const Out outs[3] = {{&S.d_pot, 1, pote}, {&S.d_grad, ndim, grade}, {&S.d_hess, nhess, hesse}};
    for (int L = 0; L < ifpghtarg; ++L) {
        const int ncd = nd * outs[L].ncomp;
        const size_t need = sizeof(double) * (size_t)ncd * P.npbox * P.nboxes;
        if (S.d_t_coefs.bytes < need) S.d_t_coefs.alloc(need);
        d_out.alloc(sizeof(double) * (size_t)ncd * nt);
        fgt::launch_vals2coefs(ndim, n, ncd, S.d_t_u.as<double>(), S.t_nleaf, S.d_t_leaves.as<int>(),
                               outs[L].vals->as<double>(), S.d_t_coefs.as<double>(), st);
        hcheck(hipGetLastError(), "k_vals2coefs launch");
        fgt::launch_evalt(ndim, n, P.ipoly, ncd, (int)tbox.size(), d_tbox.as<int>(), d_tfirst.as<int>(),
                          d_tcount.as<int>(), S.d_t_centers.as<double>(), S.d_t_bsz.as<double>(), d_ts.as<double>(),
                          d_perm.as<int>(), S.d_t_coefs.as<double>(), d_out.as<double>(), st);
        hcheck(hipGetLastError(), "k_evalt launch");
        hcheck(hipStreamSynchronize(st), "hipStreamSynchronize");
        hcheck(hipMemcpy(outs[L].host, d_out.p, d_out.bytes, hipMemcpyDeviceToHost), "download target values");
    }
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        set_status(0, "");
        return 0;
    } catch (const std::invalid_argument& e) {
        set_status(2, e.what());
        return 2;
    } catch (const std::bad_alloc&) {
        set_status(5, "host allocation failed");
        return 5;
    } catch (const std::exception& e) {
        set_status(6, e.what());
        return 6;
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int fgt_last_status(char* msg, int len) {
    if (msg && len > 0) {
        std::strncpy(msg, g_msg.c_str(), (size_t)len - 1);
        msg[len - 1] = '\0';
    }
    return g_status;
}

int fgt_solver_create(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder,
                      int nboxes, int nlevels, int ltree, const int* itree, const int* iptr,
                      const double* centers, const double* boxsize, fgt_solver** out) {
    if (!out) return 2;
    *out = nullptr;
    return guarded([&] {
        if (!itree || !iptr || !centers || !boxsize) throw std::invalid_argument("null tree array");
        int npbox = 1;
        for (int i = 0; i < ndim && i < 3; ++i) npbox *= norder;
        fgt::PlanInput in{nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree,
                          itree, iptr, centers, boxsize};
        auto S = std::make_unique<fgt_solver>();
        build_solver(*S, in);
        *out = S.release();
    });
}

int fgt_solver_set_fvals(fgt_solver* s, const double* fvals) {
    return guarded([&] {
        if (!s || !fvals) throw std::invalid_argument("null argument");
        // an execute queued on any stream may still read fvals
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hcheck(hipMemcpy(s->d_fvals.p, fvals, s->d_fvals.bytes, hipMemcpyHostToDevice), "upload fvals");
    });
}

int fgt_solver_execute(fgt_solver* s, void* stream, int timed) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null solver");
        execute(*s, static_cast<hipStream_t>(stream), timed != 0, true);
    });
}

int fgt_solver_get_pot(fgt_solver* s, double* pot, int accumulate) {
    return guarded([&] {
        if (!s || !pot) throw std::invalid_argument("null argument");
        // wait for executes queued on any stream (a caller's non-blocking
        // stream is not ordered with the null-stream copy below)
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const size_t cnt = s->d_pot.bytes / sizeof(double);
        if (!accumulate) {
            hcheck(hipMemcpy(pot, s->d_pot.p, s->d_pot.bytes, hipMemcpyDeviceToHost), "download pot");
            return;
        }
        std::vector<double> tmp(cnt);
        hcheck(hipMemcpy(tmp.data(), s->d_pot.p, s->d_pot.bytes, hipMemcpyDeviceToHost), "download pot");
        for (size_t i = 0; i < cnt; ++i) pot[i] += tmp[i];
    });
}

int fgt_solver_stats_get(const fgt_solver* s, fgt_solver_stats* o) {
    if (!s || !o) return 2;
    const fgt::Plan& P = s->plan;
    std::memset(o, 0, sizeof *o);
    o->npwlevel = P.npwlevel;
    o->npw = P.npw;
    o->nexp = P.nexp;
    o->nlevstart = P.nlevstart;
    o->nslots = P.nslots;
    o->periodic_small = P.periodic_small;
    o->all_pw = P.all_pw;
    o->nleaves = P.nleaves;
    o->npw_leaves = P.npw_leaves;
    o->ndirect_pairs = P.ndirect_pairs;
    o->nout = (long long)P.nd * P.npbox * P.nleaves;
    o->expansion_bytes = (long long)s->d_expn.bytes;
    for (int i = 0; i < 8; ++i) o->stage_ms[i] = s->stage_ms[i];
    o->n_m2l_targets = (long long)P.m2l_tgt.size();
    o->n_m2l_pairs = (long long)P.m2l_src.size();
    o->n_parent_child = (long long)P.pc_child.size();
    return 0;
}

int fgt_solver_set_ifpgh(fgt_solver* s, int ifpgh) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null solver");
        if (ifpgh < 1 || ifpgh > 3) throw std::invalid_argument("ifpgh must be 1, 2 or 3");
        const fgt::Plan& P = s->plan;
        const size_t nodes = (size_t)P.nd * P.npbox * P.nboxes;
        if (ifpgh >= 2 && !s->d_grad.p) s->d_grad.alloc(sizeof(double) * nodes * P.ndim);
        if (ifpgh >= 3 && !s->d_hess.p) s->d_hess.alloc(sizeof(double) * nodes * (P.ndim * (P.ndim + 1) / 2));
        s->ifpgh = ifpgh;
    });
}

}  // extern "C"

namespace {
// Host <-> device transfers of the drop-in's host arrays go through a pinned
// staging buffer (allocated once per process, grown on demand, two halves):
// chunk i is copied by DMA while the host threads move or accumulate chunk
// i-1.  Pageable hipMemcpy plus a single-threaded += ran at ~2.4 GB/s.
constexpr size_t kStageChunk = size_t(16) << 20;   // bytes per half
struct PinnedStage {
    std::mutex m;
    char* p = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
PinnedStage& pinned_stage() {
    static PinnedStage g;
    std::lock_guard<std::mutex> lk(g.m);
    if (!g.p) {
        void* q = nullptr;
        hcheck(hipHostMalloc(&q, 2 * kStageChunk, hipHostMallocDefault), "hipHostMalloc staging");
        g.p = static_cast<char*>(q);
        hcheck(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking), "hipStreamCreate staging");
        for (auto& e : g.ev) hcheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate staging");
    }
    return g;
}

void host_move(char* dst, const char* src, size_t bytes, bool accumulate) {
    if (!accumulate) {
        const long long nb = (long long)((bytes + 65535) / 65536);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < nb; ++i) {
            const size_t off = (size_t)i * 65536, len = std::min<size_t>(65536, bytes - off);
            std::memcpy(dst + off, src + off, len);
        }
        return;
    }
    double* d = reinterpret_cast<double*>(dst);
    const double* q = reinterpret_cast<const double*>(src);
    const long long n = (long long)(bytes / sizeof(double));
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) d[i] += q[i];
}

// device -> host (out = or += device values); waits for all device work
// first, like a blocking hipMemcpy (the staging stream is non-blocking)
void download(const DBuf& b, double* out, int accumulate, const char* what) {
    hcheck(hipDeviceSynchronize(), what);
    PinnedStage& g = pinned_stage();
    std::lock_guard<std::mutex> lk(g.m);
    const size_t total = b.bytes;
    const size_t nch = (total + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t c) {
        const size_t off = c * kStageChunk, len = std::min(kStageChunk, total - off);
        char* half = g.p + (c & 1) * kStageChunk;
        hcheck(hipMemcpyAsync(half, static_cast<const char*>(b.p) + off, len, hipMemcpyDeviceToHost, g.st), what);
        hcheck(hipEventRecord(g.ev[c & 1], g.st), what);
    };
    if (nch > 0) issue(0);
    for (size_t c = 0; c < nch; ++c) {
        hcheck(hipEventSynchronize(g.ev[c & 1]), what);
        if (c + 1 < nch) issue(c + 1);   // the other half, free since chunk c-1 was consumed
        const size_t off = c * kStageChunk, len = std::min(kStageChunk, total - off);
        host_move(reinterpret_cast<char*>(out) + off, g.p + (c & 1) * kStageChunk, len, accumulate != 0);
    }
}

// host -> device; waits for all device work first
void upload_host(const DBuf& b, const double* in, const char* what) {
    hcheck(hipDeviceSynchronize(), what);
    PinnedStage& g = pinned_stage();
    std::lock_guard<std::mutex> lk(g.m);
    const size_t total = b.bytes;
    const size_t nch = (total + kStageChunk - 1) / kStageChunk;
    for (size_t c = 0; c < nch; ++c) {
        const size_t off = c * kStageChunk, len = std::min(kStageChunk, total - off);
        char* half = g.p + (c & 1) * kStageChunk;
        if (c >= 2) hcheck(hipEventSynchronize(g.ev[c & 1]), what);   // DMA of chunk c-2 is done
        host_move(half, reinterpret_cast<const char*>(in) + off, len, false);
        hcheck(hipMemcpyAsync(static_cast<char*>(b.p) + off, half, len, hipMemcpyHostToDevice, g.st), what);
        hcheck(hipEventRecord(g.ev[c & 1], g.st), what);
    }
    hcheck(hipStreamSynchronize(g.st), what);
}
}  // namespace

extern "C" {

int fgt_solver_set_precision(fgt_solver* s, int mode) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null solver");
        if (mode < 0 || mode > 2)
            throw std::invalid_argument("precision mode must be 0 (FP64), 1 (FP32) or 2 (BF16 hi/lo)");
        if (mode == 1 && s->plan.eps < kEpsFp32)
            throw std::invalid_argument("FP32 contractions need eps >= 1e-6 (this plan's eps is smaller)");
        if (mode == 2 && s->plan.eps < kEpsBf16)
            throw std::invalid_argument("BF16 hi/lo contractions need eps >= 1e-3 (this plan's eps is smaller)");
        if (mode >= 1 && !s->plan.do_pw)
            throw std::invalid_argument("mixed precision: this plan forms no plane-wave expansions (near field only)");
        if (mode >= 1 && !s->fast3d)
            throw std::invalid_argument("mixed precision needs the 3-D register-blocked path (ndim 3, norder 8/12/16)");
        const int f32 = mode >= 1 ? 1 : 0;
        if (f32 != s->tabs.expn_f32) {
            hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
            alloc_expansions(*s, f32 ? sizeof(float2) : sizeof(double2));
            s->tabs.expn_f32 = f32;
        }
        s->tabs.prec = mode;
    });
}

int fgt_solver_get_grad(fgt_solver* s, double* grad, int accumulate) {
    return guarded([&] {
        if (!s || !grad) throw std::invalid_argument("null argument");
        if (s->ifpgh < 2) throw std::invalid_argument("gradients not requested (fgt_solver_set_ifpgh)");
        download(s->d_grad, grad, accumulate, "download grad");
    });
}

int fgt_solver_get_hess(fgt_solver* s, double* hess, int accumulate) {
    return guarded([&] {
        if (!s || !hess) throw std::invalid_argument("null argument");
        if (s->ifpgh < 3) throw std::invalid_argument("Hessians not requested (fgt_solver_set_ifpgh)");
        download(s->d_hess, hess, accumulate, "download hess");
    });
}

int fgt_solver_eval_targets(fgt_solver* s, int ntarg, const double* targs, int ifpghtarg, double* pote,
                            double* grade, double* hesse) {
    return guarded([&] {
        if (!s) throw std::invalid_argument("null solver");
        if (ifpghtarg < 1 || ifpghtarg > 3) throw std::invalid_argument("ifpghtarg must be 1, 2 or 3");
        if (ntarg < 0) throw std::invalid_argument("ntarg < 0");
        if (ntarg > 0 && (!targs || !pote || (ifpghtarg >= 2 && !grade) || (ifpghtarg >= 3 && !hesse)))
            throw std::invalid_argument("null target array");
        // the values come from the last execute, which may be queued on
        // another stream
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        eval_targets(*s, ntarg, targs, ifpghtarg, pote, grade, hesse, nullptr);
    });
}

void* fgt_solver_grad_device(fgt_solver* s) { return s ? s->d_grad.p : nullptr; }
void* fgt_solver_hess_device(fgt_solver* s) { return s ? s->d_hess.p : nullptr; }

int fgt_batch_create(int nprob, const fgt_problem* probs, int nstreams, fgt_batch** out) {
    if (!out) return 2;
    *out = nullptr;
    return guarded([&] {
        if (nprob < 1 || !probs) throw std::invalid_argument("empty batch");
        if (nstreams < 1 || nstreams > 16) throw std::invalid_argument("nstreams must be 1..16");
        auto B = std::make_unique<fgt_batch>();
        std::vector<double> work(nprob);
        for (int i = 0; i < nprob; ++i) {
            const fgt_problem& q = probs[i];
            if (!q.itree || !q.iptr || !q.centers || !q.boxsize) throw std::invalid_argument("null tree array");
            int npbox = 1;
            for (int a = 0; a < q.ndim && a < 3; ++a) npbox *= q.norder;
            fgt::PlanInput in{q.nd, q.ndim, q.delta, q.eps, q.ipoly, q.iperiod, q.norder, npbox, q.nboxes,
                              q.nlevels, q.ltree, q.itree, q.iptr, q.centers, q.boxsize};
            auto S = std::make_unique<fgt_solver>();
            build_solver(*S, in);
            work[i] = (double)S->plan.nd * S->plan.npbox * S->plan.nleaves;
            B->solvers.push_back(std::move(S));
        }
        for (int k = 0; k < nstreams; ++k) {
            hipStream_t st;
            hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
            B->streams.push_back(st);
        }
        B->order.resize(nprob);
        for (int i = 0; i < nprob; ++i) B->order[i] = i;
        std::stable_sort(B->order.begin(), B->order.end(), [&](int a, int b) { return work[a] > work[b]; });
        std::vector<double> load(nstreams, 0.0);
        B->stream_of.assign(nprob, 0);
        for (int i : B->order) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            B->stream_of[i] = k;
            load[k] += work[i];
        }
        hcheck(hipEventCreate(&B->t0), "hipEventCreate");
        hcheck(hipEventCreate(&B->t1), "hipEventCreate");
        *out = B.release();
    });
}

int fgt_batch_set_precision(fgt_batch* b, int mode, double eps_min, int* nmixed) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (mode < 0 || mode > 3) throw std::invalid_argument("batch precision mode must be 0..3");
        int cnt = 0;
        for (auto& S : b->solvers) {
            const double eps = S->plan.eps;
            const bool ok = mode >= 1 && S->plan.do_pw && S->fast3d && eps >= eps_min;
            int m = 0;
            if (ok) {
                if (mode == 1) m = eps >= kEpsFp32 ? 1 : 0;
                else if (mode == 2) m = eps >= kEpsBf16 ? 2 : 0;
                else m = eps >= kEpsBf16 ? 2 : (eps >= kEpsFp32 ? 1 : 0);
            }
            if (fgt_solver_set_precision(S.get(), m) != 0) throw std::runtime_error(g_msg);
            cnt += m > 0;
        }
        if (nmixed) *nmixed = cnt;
    });
}

int fgt_batch_set_fvals(fgt_batch* b, int i, const double* fvals) {
    return guarded([&] {
        if (!b || !fvals || i < 0 || i >= (int)b->solvers.size()) throw std::invalid_argument("bad argument");
        fgt_solver& S = *b->solvers[i];
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hcheck(hipMemcpy(S.d_fvals.p, fvals, S.d_fvals.bytes, hipMemcpyHostToDevice), "upload fvals");
    });
}

int fgt_batch_execute(fgt_batch* b) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        // all streams start after t0 and t1 waits for all of them
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hcheck(hipEventRecord(b->t0, b->streams[0]), "hipEventRecord");
        for (size_t k = 1; k < b->streams.size(); ++k)
            hcheck(hipStreamWaitEvent(b->streams[k], b->t0, 0), "hipStreamWaitEvent");
        for (int i : b->order) execute(*b->solvers[i], b->streams[b->stream_of[i]], false);
        std::vector<hipEvent_t> done(b->streams.size());
        for (size_t k = 1; k < b->streams.size(); ++k) {
            hcheck(hipEventCreate(&done[k]), "hipEventCreate");
            hcheck(hipEventRecord(done[k], b->streams[k]), "hipEventRecord");
            hcheck(hipStreamWaitEvent(b->streams[0], done[k], 0), "hipStreamWaitEvent");
        }
        hcheck(hipEventRecord(b->t1, b->streams[0]), "hipEventRecord");
        hcheck(hipEventSynchronize(b->t1), "hipEventSynchronize");
        for (size_t k = 1; k < done.size(); ++k) (void)hipEventDestroy(done[k]);
        float ms = 0;
        hcheck(hipEventElapsedTime(&ms, b->t0, b->t1), "hipEventElapsedTime");
        b->last_ms = ms;
    });
}

int fgt_batch_get_pot(fgt_batch* b, int i, double* pot, int accumulate) {
    return guarded([&] {
        if (!b || !pot || i < 0 || i >= (int)b->solvers.size()) throw std::invalid_argument("bad argument");
        download(b->solvers[i]->d_pot, pot, accumulate, "download pot");
    });
}

int fgt_batch_stats(const fgt_batch* b, long long* nout, double* last_ms) {
    if (!b) return 2;
    long long n = 0;
    for (const auto& S : b->solvers) n += (long long)S->plan.nd * S->plan.npbox * S->plan.nleaves;
    if (nout) *nout = n;
    if (last_ms) *last_ms = b->last_ms;
    return 0;
}

void fgt_batch_destroy(fgt_batch* b) { delete b; }

void* fgt_solver_fvals_device(fgt_solver* s) { return s ? s->d_fvals.p : nullptr; }
void* fgt_solver_pot_device(fgt_solver* s) { return s ? s->d_pot.p : nullptr; }

void fgt_solver_destroy(fgt_solver* s) { delete s; }

int fgt_boxfgt(int* nd, int* ndim, double* delta, double* eps, int* ipoly, int* iperiod, int* norder,
               int* npbox, int* nboxes, int* nlevels, int* ltree, int* itree, int* iptr, double* centers,
               double* boxsize, double* fvals, int* ifpgh, double* pot, double* grad, double* hess,
               int* ifnewtree, int* ntarg, double* targs, int* ifpghtarg, double* pote, double* grade,
               double* hesse, double* timeinfo) {
    (void)ntarg; (void)targs; (void)pote; (void)grade; (void)hesse;
    if (timeinfo)
        for (int i = 0; i < 20; ++i) timeinfo[i] = 0.0;
    return guarded([&] {
        if (!nd || !ndim || !delta || !eps || !ipoly || !iperiod || !norder || !npbox || !nboxes ||
            !nlevels || !ltree || !itree || !iptr || !centers || !boxsize || !fvals || !ifpgh || !pot ||
            !ifnewtree || !ifpghtarg)
            throw std::invalid_argument("null argument");
        if (*ifpgh < 1 || *ifpgh > 3) throw std::invalid_argument("ifpgh must be 1, 2 or 3");
        if (*ifpgh >= 2 && !grad) throw std::invalid_argument("null grad");
        if (*ifpgh >= 3 && !hess) throw std::invalid_argument("null hess");
        if (*ifpghtarg < 0 || *ifpghtarg > 3) throw std::invalid_argument("ifpghtarg must be 0..3");
        const int nt = (*ifpghtarg >= 1 && ntarg) ? *ntarg : 0;
        if (*ifpghtarg >= 1 && (!ntarg || *ntarg < 0)) throw std::invalid_argument("bad ntarg");
        if (nt > 0 && (!targs || !pote || (*ifpghtarg >= 2 && !grade) || (*ifpghtarg >= 3 && !hesse)))
            throw std::invalid_argument("null target array");
        if (*ifnewtree != 0) throw std::invalid_argument("ifnewtree = 1 is not supported");
        const double t0 = now_s();
        fgt::PlanInput in{*nd, *ndim, *delta, *eps, *ipoly, *iperiod, *norder, *npbox, *nboxes, *nlevels,
                          *ltree, itree, iptr, centers, boxsize};
        fgt_solver S;
        build_solver(S, in);
        // gradients / Hessians are formed when the caller or the targets need
        // them (boxfgt.f:202-272 routes ifpgh < ifpghtarg the same way)
        const int ifpgh_dev = std::max(*ifpgh, nt > 0 ? *ifpghtarg : 0);
        if (ifpgh_dev > 1) {
            const int rc = fgt_solver_set_ifpgh(&S, ifpgh_dev);
            if (rc) throw std::runtime_error(g_msg);
        }
        upload_host(S.d_fvals, fvals, "upload fvals");
        const double t1 = now_s();
        execute(S, nullptr, timeinfo != nullptr, true);
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const double t2 = now_s();
        // accumulate into the caller's arrays (+=), as the reference does
        download(S.d_pot, pot, 1, "download pot");
        if (*ifpgh >= 2) download(S.d_grad, grad, 1, "download grad");
        if (*ifpgh >= 3) download(S.d_hess, hess, 1, "download hess");
        const double t3 = now_s();
        if (nt > 0) {
            // STEP 11 evaluates what the reference holds at that point: the
            // caller's accumulated arrays for the outputs it requested, fresh
            // values for the ones formed only for the targets
            upload_host(S.d_pot, pot, "upload pot");
            if (*ifpgh >= 2 && *ifpghtarg >= 2) upload_host(S.d_grad, grad, "upload grad");
            if (*ifpgh >= 3 && *ifpghtarg >= 3) upload_host(S.d_hess, hess, "upload hess");
            eval_targets(S, nt, targs, *ifpghtarg, pote, grade, hesse, nullptr);
            if (timeinfo) timeinfo[11] = now_s() - t3;
        }
        if (timeinfo) {
            timeinfo[0] = t1 - t0;
            for (int i = 0; i < 6; ++i) timeinfo[1 + i] = S.stage_ms[i] * 1e-3;
            timeinfo[7] = t3 - t2;
        }
        (void)t2;
    });
}

int fgt_boxfgt_noalloc(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int nboxes,
                       int nlevels, int ltree, const int* itree, const int* iptr, const double* centers,
                       const double* boxsize, const double* fvals, double* pot, void* stream) {
    return guarded([&] {
        if (!itree || !iptr || !centers || !boxsize || !fvals || !pot) throw std::invalid_argument("null argument");
        const hipStream_t st = static_cast<hipStream_t>(stream);
        size_t cap = 0;
        char* base = static_cast<char*>(fgt::scratch_region(&cap));
        // every call on this device shares the scratch: the previous call
        // ended with a synchronisation; wait for the caller's stream too
        hcheck(hipStreamSynchronize(st), "hipStreamSynchronize");
        Arena arena{base, cap, 0};
        struct Installed {
            explicit Installed(Arena* a) { g_arena = a; }
            ~Installed() { g_arena = nullptr; }
        } installed(&arena);
        int npbox = 1;
        for (int i = 0; i < ndim && i < 3; ++i) npbox *= norder;
        fgt::PlanInput in{nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree,
                          itree, iptr, centers, boxsize};
        fgt_solver S;   // every device buffer of S is carved from the scratch
        build_solver(S, in, false);
        ExecOverride ov{fvals, pot, S.d_expn.as<double2>()};
        if (S.plan.do_pw) {
            // the scratch keeps the last call's contents: clear the zero slot
            const size_t slot = sizeof(double2) * 2 * (size_t)S.plan.nd * S.plan.nexp;
            fgt::launch_zero(static_cast<char*>(S.d_expn.p) + slot * S.plan.nslots, slot, st);
        }
        execute(S, st, false, false, &ov);
        hcheck(hipGetLastError(), "launch");
        hcheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}

void boxfgt_(int* nd, int* ndim, double* delta, double* eps, int* ipoly, int* iperiod, int* norder,
             int* npbox, int* nboxes, int* nlevels, int* ltree, int* itree, int* iptr, double* centers,
             double* boxsize, double* fvals, int* ifpgh, double* pot, double* grad, double* hess,
             int* ifnewtree, int* ntarg, double* targs, int* ifpghtarg, double* pote, double* grade,
             double* hesse, double* timeinfo) {
    (void)fgt_boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree, itree, iptr,
                     centers, boxsize, fvals, ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, ifpghtarg, pote,
                     grade, hesse, timeinfo);
}

}  // extern "C"
