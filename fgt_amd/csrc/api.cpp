// api.cpp -- C ABI (include/fgt.h): the boxfgt drop-in and the
// device-resident solver.  Host orchestration of the stage kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/fgt.h"
#include "kernels.h"
#include "plan.h"
#include "retree.h"

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
// device bytes allocated on this thread by DBuf::alloc (the server's
// admission accounting reads it around a solver build)
thread_local size_t g_tally = 0;

// A device buffer owned by the solver (or carved from the installed arena).
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool from_arena = false;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            from_arena = o.from_arena;
            o.p = nullptr;
            o.bytes = 0;
            o.from_arena = false;
        }
        return *this;
    }
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
            g_tally += b;
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
    DBuf d_mb_toff, d_mb_gent, d_mb_soff, d_mb_code;     // STEP 3 over sibling bricks
    int n_m2l_bricks = -1;                                // -1: bricks not used
    bool m2l_sep_ok = false;                              // bricks' lists geometric: separable form
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
    // host copies of the device work lists (the batch concatenates them
    // per shape class)
    struct HostLists {
        std::vector<fgt::PotTask> s5;
        std::vector<int> pcs_parent, pcs_child, pcs_sign;
        std::vector<int> fz_lev, fz_slot, fz_kids, fz_ksg, s1_box, s1_lev, s1_slot;
        std::vector<int> mp_slot, mp_first, mp_count, mp_child, mp_sign, mp_acc;
        std::vector<long long> m2l_toff, m2l_soff;
        std::vector<int> m2l_gent, m2l_code;
        std::vector<int> dir_tgt, dir_lev, dir_off, dir_src, dir_code, dir_flags;
    } h;
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
// table slot codes, 8 bits per axis; unit axes take the centre slot
// (same-box table), which the kernels read but do not apply
int map_code(int ndim, int c, int centre = 6) {
    const int a0 = c & 255, a1 = (c >> 8) & 255, a2 = (c >> 16) & 255;
    if (ndim == 3) return a0 | (a1 << 8) | (a2 << 16);
    if (ndim == 2) return a0 | (centre << 8) | (a1 << 16);
    return centre | (centre << 8) | (a0 << 16);
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

// STEP 3 entry streams with G consecutive targets per group: per target its
// own expansion (unit phase, o = 13), then its list, the final entry
// flagged; each group padded with zero-slot entries to a multiple of
// kM2lBatch.  toff: G target element offsets per group; gent: group ranges
// of soff / code.  Returns the group count.
int m2l_streams(const fgt::Plan& P, int G, std::vector<long long>& toff, std::vector<int>& gent,
                std::vector<long long>& soff, std::vector<int>& code) {
    const long long se = 2LL * P.nd * P.nexp;
    const int nt = (int)P.m2l_tgt.size();
    const int ng = (nt + G - 1) / G;
    toff.assign((size_t)ng * G, 0);
    gent.assign(ng + 1, 0);
    soff.clear();
    code.clear();
    for (int gi = 0; gi < ng; ++gi) {
        gent[gi] = (int)soff.size();
        for (int tk = gi * G; tk < std::min(nt, (gi + 1) * G); ++tk) {
            toff[tk] = se * P.m2l_tgt[tk];
            soff.push_back(toff[tk]);
            code.push_back(13);
            for (int j = P.m2l_off[tk]; j < P.m2l_off[tk + 1]; ++j) {
                soff.push_back(se * P.m2l_src[j]);
                code.push_back(map_m2l(P.ndim, P.m2l_idx[j]));
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
    return ng;
}

// STEP 3 entry streams over sibling bricks (3-D): the targets that share a
// parent form one brick (the root, when it is the cutoff level, one alone); each source of
// their lists (and each target's own expansion) becomes one entry at its
// position in the 4x4x4 neighbourhood, with the mask of the siblings whose
// list holds it.  Per brick: toff[8] loc element offsets (-1 = no target),
// entries ordered by position, padded with zero-slot entries.  Returns the
// brick count, or -1 when the lists do not fit that shape (a position
// holding two different sources), in which case the grouped streams serve.
// Code bits: 0-5 position, 6-13 target mask, 14 the source box is a leaf,
// 15 the target at this position is a leaf.  *geometric: every entry's mask
// is exactly the siblings present whose 3x3x3 neighbourhood holds its
// position, less the leaf targets for a leaf source other than themselves
// (the list rule of the cutoff level: leaf-leaf pairs are near field), so
// the separable form applies.
int m2l_bricks(const fgt::Plan& P, std::vector<long long>& toff, std::vector<int>& gent,
               std::vector<long long>& soff, std::vector<int>& code, bool* geometric = nullptr) {
    if (geometric) *geometric = false;
    if (P.ndim != 3 || !P.do_pw) return -1;
    const fgt::TreeView& T = P.tree;
    const long long se = 2LL * P.nd * P.nexp;
    const int nt = (int)P.m2l_tgt.size();
    std::vector<int> box_of_slot(P.nslots, -1);
    for (int b = 0; b < T.nboxes; ++b)
        if (P.slot_of_box[b] >= 0) box_of_slot[P.slot_of_box[b]] = b;
    std::unordered_map<int, int> brick_of_parent;
    std::vector<std::array<int, 8>> tg;
    for (int tk = 0; tk < nt; ++tk) {
        const int b = box_of_slot[P.m2l_tgt[tk]];
        if (b < 0) return -1;
        // the root (cutoff level 0) has no parent: a brick of its own, at
        // brick position 0
        const int par = T.parent[b];
        int tp = 0;
        if (par >= 0)
            for (int a = 0; a < 3; ++a)
                if (T.centers[3 * b + a] > T.centers[3 * par + a]) tp |= 1 << a;
        auto it = brick_of_parent.find(par >= 0 ? par : -1 - b);
        if (it == brick_of_parent.end()) {
            it = brick_of_parent.emplace(par >= 0 ? par : -1 - b, (int)tg.size()).first;
            std::array<int, 8> none;
            none.fill(-1);
            tg.push_back(none);
        }
        if (tg[it->second][tp] >= 0) return -1;
        tg[it->second][tp] = tk;
    }
    const int nb = (int)tg.size();
    bool geo = true;
    toff.assign((size_t)nb * 8, -1);
    gent.assign(nb + 1, 0);
    soff.clear();
    code.clear();
    for (int k = 0; k < nb; ++k) {
        gent[k] = (int)soff.size();
        int src_at[64], mask_at[64];
        for (int i = 0; i < 64; ++i) src_at[i] = -1, mask_at[i] = 0;
        for (int tp = 0; tp < 8; ++tp) {
            const int tk = tg[k][tp];
            if (tk < 0) continue;
            toff[(size_t)8 * k + tp] = se * P.m2l_tgt[tk];
            // a source at offset j (target minus source, per axis, in boxes)
            // sits at brick position tp - j, stored + 1
            auto add = [&](int slot, int jx, int jy, int jz) {
                const int px = (tp & 1) - jx + 1, py = ((tp >> 1) & 1) - jy + 1, pz = ((tp >> 2) & 1) - jz + 1;
                const int pos = px + 4 * py + 16 * pz;
                if ((src_at[pos] >= 0 && src_at[pos] != slot) || ((mask_at[pos] >> tp) & 1)) return false;
                src_at[pos] = slot;
                mask_at[pos] |= 1 << tp;
                return true;
            };
            if (!add(P.m2l_tgt[tk], 0, 0, 0)) return -1;
            for (int j = P.m2l_off[tk]; j < P.m2l_off[tk + 1]; ++j) {
                const int idx = P.m2l_idx[j];
                if (!add(P.m2l_src[j], idx % 3 - 1, (idx / 3) % 3 - 1, idx / 9 - 1)) return -1;
            }
        }
        auto leaf_slot = [&](int slot) { return T.nchild[box_of_slot[slot]] == 0; };
        for (int pos = 0; pos < 64; ++pos)
            if (mask_at[pos]) {
                const bool leaf_src = leaf_slot(src_at[pos]);
                // the target at this position (its own entry), if any
                int own = -1;
                if ((pos & 3) >= 1 && (pos & 3) <= 2 && ((pos >> 2) & 3) >= 1 && ((pos >> 2) & 3) <= 2 &&
                    (pos >> 4) >= 1 && (pos >> 4) <= 2)
                    own = ((pos & 3) - 1) + 2 * (((pos >> 2) & 3) - 1) + 4 * ((pos >> 4) - 1);
                const bool leaf_tgt = own >= 0 && tg[k][own] >= 0 && leaf_slot(P.m2l_tgt[tg[k][own]]);
                soff.push_back(se * src_at[pos]);
                code.push_back(pos | mask_at[pos] << 6 | (leaf_src ? 1 << 14 : 0) | (leaf_tgt ? 1 << 15 : 0));
                int want = 0;
                for (int tp = 0; tp < 8; ++tp) {
                    if (tg[k][tp] < 0) continue;
                    bool in = true;
                    for (int a = 0; a < 3; ++a) {
                        const int d = ((pos >> (2 * a)) & 3) - (((tp >> a) & 1) + 1);
                        if (d < -1 || d > 1) in = false;
                    }
                    if (in && leaf_src && tp != own && leaf_slot(P.m2l_tgt[tg[k][tp]])) in = false;
                    if (in) want |= 1 << tp;
                }
                if (want != mask_at[pos]) geo = false;
                // an own entry must hold the target's own expansion
                if (own >= 0 && tg[k][own] >= 0 && src_at[pos] != P.m2l_tgt[tg[k][own]]) geo = false;
            }
        while ((soff.size() - gent[k]) % fgt::kM2lBatch) {
            soff.push_back(se * P.nslots);   // the zero slot, applied to no target
            code.push_back(0);
        }
    }
    gent[nb] = (int)soff.size();
    if (geometric) *geometric = geo;
    return nb;
}

// On the 3-D fast path a leaf below the first plane-wave level never holds
// an expansion of its own: STEP 1 writes its parent's (fused with STEP 2)
// and STEP 5 evaluates its parent's through phased tables.  Only non-leaf
// boxes and the leaves of the first plane-wave level keep slots; every slot
// list of the plan is renumbered (a dropped leaf's slot becomes -1, used by
// the other path only).
void compact_slots(fgt::Plan& P) {
    const fgt::TreeView& T = P.tree;
    std::vector<int> remap(P.nslots, -1);
    int ns = 0;
    for (int l = P.nlevstart; l <= P.nlevels; ++l)
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
            const int sl = P.slot_of_box[b];
            if (sl >= 0 && (T.nchild[b] > 0 || l == P.nlevstart)) remap[sl] = ns++;
        }
    auto rm = [&](std::vector<int>& v) {
        for (int& x : v) x = x >= 0 ? remap[x] : -1;
    };
    rm(P.slot_of_box);
    rm(P.leaf_slot);
    rm(P.par_slot);
    rm(P.pc_parent);
    rm(P.pc_child);
    rm(P.m2l_tgt);
    rm(P.m2l_src);
    for (int x : P.m2l_tgt)
        if (x < 0) throw std::runtime_error("STEP 3 target lost its slot");
    for (int x : P.m2l_src)
        if (x < 0) throw std::runtime_error("STEP 3 source lost its slot");
    P.nslots = ns;
}

// own_io = false: the caller brings device fvals / pot (no-allocation
// entry point), so the solver keeps none of its own.
// Precision a mode asks for on one problem (fgt_batch_set_precision modes):
// 0 FP64; 1 FP32 where eps >= 1e-6; 2 BF16 hi/lo where eps >= 1e-3; 3 the
// cheapest one eps allows.  ok: the problem is on the 3-D fast path and
// forms plane-wave expansions.
// STEP 3 elements per thread: FGT_M2L_EPT when set (2 only where allowed),
// else 1 (two per thread measured slower for FP64 at norder 16: 1.05 vs
// 0.91 ms); batch classes time both in the autotuner
int m2l_ept_default(bool ept2_ok) {
    return fgt::m2l_ept_env() == 2 && ept2_ok ? 2 : 1;
}

int mode_precision(int mode, double eps, bool ok) {
    if (!ok || mode == 0) return 0;
    if (mode == 1) return eps >= kEpsFp32 ? 1 : 0;
    if (mode == 2) return eps >= kEpsBf16 ? 2 : 0;
    return eps >= kEpsBf16 ? 2 : (eps >= kEpsFp32 ? 1 : 0);
}

void build_solver(fgt_solver& S, const fgt::PlanInput& in, bool own_io = true, fgt::Plan* planned = nullptr,
                  int prec_mode = 0) {
    if (planned) {
        S.plan = std::move(*planned);
    } else {
        std::string err;
        const int rc = fgt::make_plan(in, S.plan, &err);
        if (rc != 0) throw std::invalid_argument(err);
    }
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
    if (S.fast3d) compact_slots(S.plan);
    t.prec = mode_precision(prec_mode, P.eps, S.fast3d);
    t.expn_f32 = t.prec > 0 ? 1 : 0;
    if (P.do_pw) alloc_expansions(S, t.expn_f32 ? sizeof(float2) : sizeof(double2));
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
        S.h.s5 = fam;
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
        S.h.pcs_parent = pp;
        S.d_pcs_child.upload(pcc);
        S.h.pcs_child = pcc;
        S.d_pcs_sign.upload(psg);
        S.h.pcs_sign = psg;

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
        S.h.fz_lev = fzl;
        S.d_fz_slot.upload(fzs);
        S.h.fz_slot = fzs;
        S.d_fz_kids.upload(fzk);
        S.h.fz_kids = fzk;
        S.d_fz_ksg.upload(fzg);
        S.h.fz_ksg = fzg;
        S.d_s1_box.upload(s1b);
        S.h.s1_box = s1b;
        S.d_s1_lev.upload(s1l);
        S.h.s1_lev = s1l;
        S.d_s1_slot.upload(s1s);
        S.h.s1_slot = s1s;
        S.d_mp_slot.upload(mps);
        S.h.mp_slot = mps;
        S.d_mp_first.upload(mpf);
        S.h.mp_first = mpf;
        S.d_mp_count.upload(mpc);
        S.h.mp_count = mpc;
        S.d_mp_child.upload(mch);
        S.h.mp_child = mch;
        S.d_mp_sign.upload(msg);
        S.h.mp_sign = msg;
        S.d_mp_acc.upload(macc);
        S.h.mp_acc = macc;
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
        std::vector<long long> toff, soff;
        std::vector<int> gent, code;
        S.n_m2l_groups = m2l_streams(P, fgt::m2l_group(), toff, gent, soff, code);
        S.d_m2l_toff.upload(toff);
        S.h.m2l_toff = toff;
        S.d_m2l_poff.upload(gent);
        S.h.m2l_gent = gent;
        S.d_m2l_soff.upload(soff);
        S.h.m2l_soff = soff;
        S.d_m2l_so.upload(code);
        S.h.m2l_code = code;
        if (fgt::m2l_brick_env() != 0) {
            std::vector<long long> bt, bs;
            std::vector<int> bg, bc;
            bool geo = false;
            S.n_m2l_bricks = m2l_bricks(P, bt, bg, bs, bc, &geo);
            S.m2l_sep_ok = S.n_m2l_bricks > 0 && geo;
            if (S.n_m2l_bricks > 0) {
                S.d_mb_toff.upload(bt);
                S.d_mb_gent.upload(bg);
                S.d_mb_soff.upload(bs);
                S.d_mb_code.upload(bc);
            }
        }
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
            return (cx >> 8) < (cy >> 8);
        });
        for (int i = a; i < b; ++i) {
            ds[i] = P.dir_src[idx[i - a]];
            dc[i] = map_code(ndim, P.dir_code[idx[i - a]]);
        }
        for (int i = a; i < b; ++i) {
            const bool last = (i + 1 == b);
            if (last || (dc[i] >> 8) != (dc[i + 1] >> 8)) df[i] |= 1;
            if (last || (dc[i] >> 16) != (dc[i + 1] >> 16)) df[i] |= 2;
        }
    }
    S.d_dir_tgt.upload(dt);
    S.h.dir_tgt = dt;
    S.d_dir_lev.upload(dl);
    S.h.dir_lev = dl;
    S.d_dir_off.upload(doff);
    S.h.dir_off = doff;
    S.d_dir_src.upload(ds);
    S.h.dir_src = ds;
    S.d_dir_code.upload(dc);
    S.h.dir_code = dc;
    S.d_dir_flags.upload(df);
    S.h.dir_flags = df;
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

// the components of outputs pot (nd, npbox, *), grad (nd, ndim, npbox, *)
// and hess (nd, nhess, npbox, *) that ifpgh asks for
std::vector<Comp> components_on(int ndim, int nd, int ifpgh, double* pot, double* grad, double* hess) {
    const int nhess = ndim * (ndim + 1) / 2;
    auto axes = [&](int r0, int r1, int r2) {
        if (ndim == 3) return fgt::Kinds{r0, r1, r2};
        if (ndim == 2) return fgt::Kinds{r0, 0, r1};
        return fgt::Kinds{0, 0, r0};
    };
    std::vector<Comp> c;
    c.push_back({axes(0, 0, 0), {pot, nd, 0}});
    if (ifpgh >= 2)
        for (int a = 0; a < ndim; ++a)
            c.push_back({axes(a == 0, a == 1, a == 2), {grad, nd * ndim, a * nd}});
    if (ifpgh >= 3) {
        static const int h1[1][3] = {{2, 0, 0}};
        static const int h2[3][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}};
        static const int h3[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
        const int(*h)[3] = ndim == 1 ? h1 : (ndim == 2 ? h2 : h3);
        for (int k = 0; k < nhess; ++k)
            c.push_back({axes(h[k][0], h[k][1], h[k][2]), {hess, nd * nhess, k * nd}});
    }
    return c;
}

// Output components grouped by their x and y table kinds (gkd[k]), at most
// three z kinds per group (groups[k]).
void group_comps(const std::vector<Comp>& comps, std::vector<fgt::ZOut>& groups, std::vector<fgt::Kinds>& gkd) {
    for (const Comp& c : comps) {
        size_t k = 0;
        while (k < groups.size() && !(gkd[k].x == c.kd.x && gkd[k].y == c.kd.y && groups[k].nz < 3)) ++k;
        if (k == groups.size()) {
            groups.push_back(fgt::ZOut{});
            gkd.push_back(c.kd);
        }
        fgt::ZOut& z = groups[k];
        z.kz[z.nz] = c.kd.z;
        z.od[z.nz] = c.od;
        ++z.nz;
    }
}

// STEP 6 for a list of output components: on the 3-D register-blocked
// path the components that share their x and y table kinds go in one pass
// (their x and y sums are formed once; gnd_tens_prod_to_pghloc,
// bfgt_local.f:756-920, shares the partial sums the same way), elsewhere
// one pass per component.
void launch_direct_comps(const fgt::Geom& g, const fgt::DevTables& t, int ntask, const int* tgt, const int* lev,
                         const int* off, const int* src, const int* code, const int* flags, const double* fvals,
                         const std::vector<Comp>& comps, hipStream_t st) {
    std::vector<fgt::ZOut> groups;
    std::vector<fgt::Kinds> gkd;
    group_comps(comps, groups, gkd);
    for (size_t k = 0; k < groups.size(); ++k) {
        if (fgt::launch_direct_3d_z(g, t, ntask, tgt, lev, off, src, code, flags, fvals, gkd[k], groups[k], st)) {
            hcheck(hipGetLastError(), "k_direct3 launch");
            continue;
        }
        for (int q = 0; q < groups[k].nz; ++q)
            fgt::launch_direct(g, t, ntask, tgt, lev, off, src, code, flags, fvals,
                               fgt::Kinds{gkd[k].x, gkd[k].y, groups[k].kz[q]}, groups[k].od[q], st);
    }
}

// STEP 5 (3-D task kernel) for a list of output components: the one-leaf
// tasks of components that share their x and y table kinds run once per
// group (g3d_pw2pgh, bfgt_pwrouts.f:906-1000, likewise forms each pair of x
// and y partial sums once for all the z kinds), the sibling tasks once per
// component.
void launch_pw2pot_comps(const fgt::Geom& g, const fgt::DevTables& t, int ntask, int nsib, const fgt::PotTask* tasks,
                         const double2* expn, const std::vector<Comp>& comps, hipStream_t st) {
    std::vector<fgt::ZOut> groups;
    std::vector<fgt::Kinds> gkd;
    group_comps(comps, groups, gkd);
    for (size_t k = 0; k < groups.size(); ++k)
        fgt::launch_pw2pot_tasks_z(g, t, ntask, nsib, tasks, expn, gkd[k], groups[k], st);
}

std::vector<Comp> components(fgt_solver& S, double* pot_override = nullptr) {
    const fgt::Plan& P = S.plan;
    if (pot_override) return components_on(P.ndim, P.nd, 1, pot_override, nullptr, nullptr);
    return components_on(P.ndim, P.nd, S.ifpgh, S.d_pot.as<double>(), S.d_grad.as<double>(),
                         S.d_hess.as<double>());
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

// kernel_zero: clear the outputs with a fill kernel instead of
// hipMemsetAsync (work that may be captured into a graph records no memset
// nodes).
void execute(fgt_solver& S, hipStream_t st, bool timed, bool fork_near = false,
             const ExecOverride* ov = nullptr, bool kernel_zero = false) {
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
    if (!ov && kernel_zero) {
        fgt::launch_zero(S.d_pot.p, S.d_pot.bytes, st);
        if (S.ifpgh >= 2) fgt::launch_zero(S.d_grad.p, S.d_grad.bytes, st);
        if (S.ifpgh >= 3) fgt::launch_zero(S.d_hess.p, S.d_hess.bytes, st);
    } else if (!ov) {
        hcheck(hipMemsetAsync(S.d_pot.p, 0, S.d_pot.bytes, st), "hipMemsetAsync pot");
        if (S.ifpgh >= 2) hcheck(hipMemsetAsync(S.d_grad.p, 0, S.d_grad.bytes, st), "hipMemsetAsync grad");
        if (S.ifpgh >= 3) hcheck(hipMemsetAsync(S.d_hess.p, 0, S.d_hess.bytes, st), "hipMemsetAsync hess");
    }
    const std::vector<Comp> comps = components(S, ov ? ov->pot : nullptr);
    auto near_field = [&](hipStream_t sn) {
        launch_direct_comps(g, t, (int)P.dir_tgt.size(), S.d_dir_tgt.as<int>(), S.d_dir_lev.as<int>(),
                            S.d_dir_off.as<int>(), S.d_dir_src.as<int>(), S.d_dir_code.as<int>(),
                            S.d_dir_flags.as<int>(), fv, comps, sn);
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
        if (S.n_m2l_bricks > 0 && S.m2l_sep_ok && fgt::m2l_brick_env() == 2)
            fgt::launch_m2l_sep(g, t, S.n_m2l_bricks, 2LL * P.nd * P.nexp * P.nslots, S.d_mb_toff.as<long long>(),
                                S.d_mb_gent.as<int>(), S.d_mb_soff.as<long long>(), S.d_mb_code.as<int>(), ex, st);
        else if (S.n_m2l_bricks > 0)
            fgt::launch_m2l_brick(g, t, S.n_m2l_bricks, 2LL * P.nd * P.nexp * P.nslots, S.d_mb_toff.as<long long>(),
                                  S.d_mb_gent.as<int>(), S.d_mb_soff.as<long long>(), S.d_mb_code.as<int>(), ex, st);
        else
            fgt::launch_m2l27(g, t, S.n_m2l_groups, fgt::m2l_group(), 2LL * P.nd * P.nexp * P.nslots,
                              S.d_m2l_toff.as<long long>(), S.d_m2l_poff.as<int>(), S.d_m2l_soff.as<long long>(),
                              S.d_m2l_so.as<int>(), ex, st, fgt::Multi{},
                              m2l_ept_default((((long long)P.nd * P.nexp) & 1) == 0));
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
        // STEP 5 over the output components (pot, grad, hess entries)
        if (fork) {
            hcheck(hipStreamWaitEvent(st, S.ev_join, 0), "hipStreamWaitEvent");
            mark(7);   // STEP 5 starts here (after the wait for STEP 6)
        }
        if (S.fast3d)
            launch_pw2pot_comps(g, t, S.n_s5, S.n_s5_sib, S.d_s5_tasks.as<fgt::PotTask>(), ex, comps, st);
        else
            for (const Comp& c : comps)
                fgt::launch_pw2pot(g, t, (int)P.leaf_box.size(), S.d_leaf_box.as<int>(), S.d_leaf_lev.as<int>(),
                                   S.d_leaf_slot.as<int>(), ex, c.kd, c.od, st);
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

// ------------------------------------------------------------------ batch
// Many independent problems (uneven sizes, orders, delta) solved together.
// Problems on the 3-D register-blocked path are grouped into shape classes
// (norder, densities, precision, plane-wave count rounded up to 16); every
// stage of a class is one launch over the tasks of all its problems, the
// kernels reading each task's tables, values and expansions through a
// per-problem descriptor.  STEP 2 and STEP 4 keep each problem's level
// order: phase p of STEP 2 holds the parents p levels above each problem's
// finest parent level, phase p of STEP 4 the parents p levels below each
// problem's first plane-wave level.  Other problems (2-D, other orders)
// run their own executes on side streams.  The whole sequence is captured
// once into a HIP graph and replayed.

// One shape class of a batch.
struct BClass {
    int n = 0, nd = 0, prec = 0, npw_cap = 0;
    fgt::Geom g{};          // geometry with the class's largest plane-wave count
    fgt::DevTables t{};     // class-wide fields; tables per problem come from pp
    std::vector<int> members;
    DBuf pp;
    int group = fgt::kM2lGroup;
    int m2l_ept = 1;          // STEP 3 elements per thread
    int m2l_brick = 0;        // STEP 3 over sibling bricks (every member allows them)
    bool m2l_ept2_ok = false; // every member's nd * nexp is even
    DBuf fz_lev, fz_slot, fz_kids, fz_ksg, fz_prob;
    int n_fz = 0;
    DBuf s1_box, s1_lev, s1_slot, s1_prob;
    int n_s1 = 0;
    DBuf mp_slot, mp_first, mp_count, mp_child, mp_sign, mp_acc, mp_prob, mp_klev;
    std::vector<int> mp_phase;    // phase p: entries [mp_phase[p], mp_phase[p+1])
    DBuf m2l_toff, m2l_gent, m2l_soff, m2l_code, m2l_prob;
    int n_m2l = 0;
    DBuf pcs_parent, pcs_child, pcs_sign, pcs_prob, pcs_klev;
    std::vector<int> pcs_phase;
    DBuf s5, s5_prob;
    int n_s5 = 0, n_s5_sib = 0;
    DBuf dir_tgt, dir_lev, dir_off, dir_src, dir_code, dir_flags, dir_prob;
    int n_dir = 0;
    double stage_ms[6] = {0, 0, 0, 0, 0, 0};   // STEPs 1-6 of the last profiled execute
};

// Tuned per-class choices, kept for the process: key (norder, densities,
// precision, plane-wave bucket) -> (STEP 3 targets per block, STEP 1 and
// STEP 5 x-stage variants).
struct Tuned {
    int group, xs1, xs5, m2l_ept, m2l_brick;
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, Tuned> g_tuned;

}  // namespace

struct fgt_batch {
    std::vector<std::unique_ptr<fgt_solver>> solvers;
    std::vector<std::unique_ptr<BClass>> classes;
    std::vector<int> loose;            // problems executed one by one
    std::vector<hipStream_t> streams;  // [0] stages, [1] near field, [2..] loose problems
    hipEvent_t ev_fork = nullptr, ev_near = nullptr;
    std::vector<hipEvent_t> ev_loose;
    // further class chains beside streams[0] (classes are independent: each
    // chain runs its classes' STEPs 1-5 in order), their join events, and
    // one near-field event per class (STEP 5 of a class waits for its STEP 6)
    std::vector<hipStream_t> chains;
    std::vector<hipEvent_t> ev_chain, ev_cnear;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool use_graph = true;
    double last_ms = 0.0;
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~fgt_batch() {
        drop_graph();
        for (auto s : streams) (void)hipStreamDestroy(s);
        for (auto s : chains) (void)hipStreamDestroy(s);
        for (auto e : ev_loose) (void)hipEventDestroy(e);
        for (auto e : ev_chain) (void)hipEventDestroy(e);
        for (auto e : ev_cnear) (void)hipEventDestroy(e);
        for (auto e : {ev_fork, ev_near, t0, t1})
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {

bool class_member(const fgt_solver& S) {
    const fgt::Plan& P = S.plan;
    return P.ndim == 3 && (P.n == 8 || P.n == 12 || P.n == 16) && (S.fast3d || !P.do_pw);
}

// STEP 3 lists of a class at targets-per-block G (rebuilt when G changes)
void class_m2l(fgt_batch& B, BClass& C) {
    std::vector<long long> toff, soff;
    std::vector<int> gent(1, 0), code, prob;
    for (int q = 0; q < (int)C.members.size() && C.m2l_brick; ++q) {
        const fgt_solver& S = *B.solvers[C.members[q]];
        if (S.plan.do_pw && S.n_m2l_bricks <= 0) C.m2l_brick = 0;   // a member without bricks
        if (S.plan.do_pw && C.m2l_brick == 2 && !S.m2l_sep_ok) C.m2l_brick = 1;   // not separable
    }
    for (int q = 0; q < (int)C.members.size(); ++q) {
        const fgt_solver& S = *B.solvers[C.members[q]];
        if (!S.plan.do_pw) continue;
        std::vector<long long> to, so;
        std::vector<int> ge, co;
        const int ng = C.m2l_brick ? m2l_bricks(S.plan, to, ge, so, co)
                                   : m2l_streams(S.plan, C.group, to, ge, so, co);
        const int base = (int)soff.size();
        toff.insert(toff.end(), to.begin(), to.end());
        for (int gi = 0; gi < ng; ++gi) {
            gent.back() = base + ge[gi];
            gent.push_back(0);
            prob.push_back(q);
        }
        soff.insert(soff.end(), so.begin(), so.end());
        code.insert(code.end(), co.begin(), co.end());
        gent.back() = (int)soff.size();
    }
    C.n_m2l = (int)prob.size();
    C.m2l_toff.upload(toff);
    C.m2l_gent.upload(gent);
    C.m2l_soff.upload(soff);
    C.m2l_code.upload(code);
    C.m2l_prob.upload(prob);
}

void build_classes(fgt_batch& B) {
    B.drop_graph();
    B.classes.clear();
    B.loose.clear();
    std::map<std::tuple<int, int, int, int>, int> idx;
    for (int i = 0; i < (int)B.solvers.size(); ++i) {
        const fgt_solver& S = *B.solvers[i];
        if (!class_member(S)) {
            B.loose.push_back(i);
            continue;
        }
        const int cap = S.plan.do_pw ? (S.plan.npw + 15) / 16 * 16 : 0;
        const auto key = std::make_tuple(S.plan.n, S.plan.nd, S.tabs.prec, cap);
        auto it = idx.find(key);
        if (it == idx.end()) {
            it = idx.emplace(key, (int)B.classes.size()).first;
            B.classes.push_back(std::make_unique<BClass>());
            BClass& C = *B.classes.back();
            C.n = S.plan.n;
            C.nd = S.plan.nd;
            C.prec = S.tabs.prec;
            C.npw_cap = cap;
        }
        B.classes[it->second]->members.push_back(i);
    }
    for (auto& cp : B.classes) {
        BClass& C = *cp;
        // geometry and class-wide table fields from the member with the most plane waves
        int big = C.members[0];
        for (int i : C.members)
            if (B.solvers[i]->plan.npw > B.solvers[big]->plan.npw) big = i;
        C.g = B.solvers[big]->geom;
        C.t = B.solvers[big]->tabs;
        C.t.p2pw = C.t.p2pwS = C.t.pw2pot = C.t.pw2potT = C.t.pw2potS = C.t.wms = C.t.wm2l = nullptr;
        C.t.loctab = nullptr;
        C.t.pw2potS_ks = C.t.loctab_ks = 0;   // the value tables (kind 0) only
        C.m2l_ept2_ok = true;
        for (int i : C.members)
            if (((long long)B.solvers[i]->plan.nd * B.solvers[i]->plan.nexp) & 1) C.m2l_ept2_ok = false;
        C.m2l_ept = m2l_ept_default(C.m2l_ept2_ok);
        // brick entry streams (separable bricks with FGT_M2L_BRICK=2);
        // class_m2l steps down where a member cannot
        C.m2l_brick = fgt::m2l_brick_env() == 0 ? 0 : (fgt::m2l_brick_env() == 2 ? 2 : 1);
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(std::make_tuple(C.n, C.nd, C.prec, C.npw_cap));
            if (it != g_tuned.end()) {
                C.group = it->second.group;
                C.t.xs1 = it->second.xs1;
                C.t.xs5 = it->second.xs5;
                if (it->second.m2l_ept == 1 || C.m2l_ept2_ok) C.m2l_ept = it->second.m2l_ept;
                if (fgt::m2l_brick_env() < 0) C.m2l_brick = it->second.m2l_brick;
            }
        }
        std::vector<fgt::ProbPtrs> pp;
        std::vector<int> fzl, fzs, fzk, fzg, fzp, s1b, s1l, s1s, s1p;
        std::vector<int> mch, msg, pcsp, pcsc, pcsg;
        std::vector<fgt::PotTask> sib, own;
        std::vector<int> sibp, ownp;
        int maxph2 = 0, maxph4 = 0;
        for (int q = 0; q < (int)C.members.size(); ++q) {
            fgt_solver& S = *B.solvers[C.members[q]];
            const fgt::Plan& P = S.plan;
            fgt::ProbPtrs r{};
            r.g = S.geom;
            r.p2pw = S.d_p2pw.as<double2>();
            r.p2pwS = S.d_p2pwS.as<double2>();
            r.pw2potS = S.d_pw2potS.as<double2>();
            r.wms = S.d_wms.as<double2>();
            r.wm2l = S.d_wm2l.as<double2>();
            r.loctab = S.d_loctab.as<double>();
            r.fvals = S.d_fvals.as<double>();
            r.pot = S.d_pot.as<double>();
            r.expn = S.d_expn.p;
            r.zero_off = 2LL * P.nd * P.nexp * P.nslots;
            pp.push_back(r);
            if (!P.do_pw) continue;
            const auto& h = S.h;
            fzl.insert(fzl.end(), h.fz_lev.begin(), h.fz_lev.end());
            fzs.insert(fzs.end(), h.fz_slot.begin(), h.fz_slot.end());
            fzk.insert(fzk.end(), h.fz_kids.begin(), h.fz_kids.end());
            fzg.insert(fzg.end(), h.fz_ksg.begin(), h.fz_ksg.end());
            fzp.insert(fzp.end(), h.fz_lev.size(), q);
            s1b.insert(s1b.end(), h.s1_box.begin(), h.s1_box.end());
            s1l.insert(s1l.end(), h.s1_lev.begin(), h.s1_lev.end());
            s1s.insert(s1s.end(), h.s1_slot.begin(), h.s1_slot.end());
            s1p.insert(s1p.end(), h.s1_box.size(), q);
            maxph2 = std::max(maxph2, P.nlevels - P.nlevstart);
            maxph4 = std::max(maxph4, P.nlevels - P.nlevstart);
            sib.insert(sib.end(), h.s5.begin(), h.s5.begin() + S.n_s5_sib);
            sibp.insert(sibp.end(), S.n_s5_sib, q);
            own.insert(own.end(), h.s5.begin() + S.n_s5_sib, h.s5.end());
            ownp.insert(ownp.end(), h.s5.size() - S.n_s5_sib, q);
        }
        C.pp.upload(pp);
        C.n_fz = (int)fzp.size();
        C.fz_lev.upload(fzl);
        C.fz_slot.upload(fzs);
        C.fz_kids.upload(fzk);
        C.fz_ksg.upload(fzg);
        C.fz_prob.upload(fzp);
        C.n_s1 = (int)s1p.size();
        C.s1_box.upload(s1b);
        C.s1_lev.upload(s1l);
        C.s1_slot.upload(s1s);
        C.s1_prob.upload(s1p);
        // STEP 2 phases (children arrays appended once per problem)
        std::vector<int> ms, mf, mc, ma, mp, mk;
        std::vector<int> chbase(C.members.size(), 0);
        for (int q = 0; q < (int)C.members.size(); ++q) {
            const fgt_solver& S = *B.solvers[C.members[q]];
            chbase[q] = (int)mch.size();
            mch.insert(mch.end(), S.h.mp_child.begin(), S.h.mp_child.end());
            msg.insert(msg.end(), S.h.mp_sign.begin(), S.h.mp_sign.end());
        }
        C.mp_phase.assign(1, 0);
        for (int ph = 0; ph < maxph2; ++ph) {
            for (int q = 0; q < (int)C.members.size(); ++q) {
                const fgt_solver& S = *B.solvers[C.members[q]];
                const fgt::Plan& P = S.plan;
                if (!P.do_pw) continue;
                const int l = P.nlevels - 1 - ph;
                if (l < P.nlevstart) continue;
                for (int e = S.mp_level_off[l]; e < S.mp_level_off[l + 1]; ++e) {
                    ms.push_back(S.h.mp_slot[e]);
                    mf.push_back(chbase[q] + S.h.mp_first[e]);
                    mc.push_back(S.h.mp_count[e]);
                    ma.push_back(S.h.mp_acc[e]);
                    mp.push_back(q);
                    mk.push_back(P.nlevels - l);
                }
            }
            C.mp_phase.push_back((int)ms.size());
        }
        C.mp_slot.upload(ms);
        C.mp_first.upload(mf);
        C.mp_count.upload(mc);
        C.mp_acc.upload(ma);
        C.mp_prob.upload(mp);
        C.mp_klev.upload(mk);
        C.mp_child.upload(mch);
        C.mp_sign.upload(msg);
        // STEP 4 phases
        std::vector<int> pp4, pk4;
        C.pcs_phase.assign(1, 0);
        for (int ph = 0; ph < maxph4; ++ph) {
            for (int q = 0; q < (int)C.members.size(); ++q) {
                const fgt_solver& S = *B.solvers[C.members[q]];
                const fgt::Plan& P = S.plan;
                if (!P.do_pw) continue;
                const int l = P.nlevstart + ph;
                if (l > P.nlevels - 1) continue;
                for (int e = S.pcs_level_off[l]; e < S.pcs_level_off[l + 1]; ++e) {
                    pcsp.push_back(S.h.pcs_parent[e]);
                    pcsc.push_back(S.h.pcs_child[e]);
                    pcsg.push_back(S.h.pcs_sign[e]);
                    pp4.push_back(q);
                    pk4.push_back(P.nlevels - l);
                }
            }
            C.pcs_phase.push_back((int)pcsp.size());
        }
        C.pcs_parent.upload(pcsp);
        C.pcs_child.upload(pcsc);
        C.pcs_sign.upload(pcsg);
        C.pcs_prob.upload(pp4);
        C.pcs_klev.upload(pk4);
        // STEP 3
        class_m2l(B, C);
        // STEP 5: every sibling task first, then the one-leaf tasks
        C.n_s5_sib = (int)sib.size();
        sib.insert(sib.end(), own.begin(), own.end());
        sibp.insert(sibp.end(), ownp.begin(), ownp.end());
        C.n_s5 = (int)sib.size();
        C.s5.upload(sib);
        C.s5_prob.upload(sibp);
        // STEP 6: every target of every problem, longest list first
        struct DT {
            int q, k, len;
        };
        std::vector<DT> order;
        for (int q = 0; q < (int)C.members.size(); ++q) {
            const fgt_solver& S = *B.solvers[C.members[q]];
            for (int k = 0; k < (int)S.h.dir_tgt.size(); ++k)
                order.push_back({q, k, S.h.dir_off[k + 1] - S.h.dir_off[k]});
        }
        std::stable_sort(order.begin(), order.end(), [](const DT& a, const DT& b) { return a.len > b.len; });
        std::vector<int> dt, dl, doff(1, 0), ds, dc, df, dp;
        for (const DT& o : order) {
            const auto& h = B.solvers[C.members[o.q]]->h;
            dt.push_back(h.dir_tgt[o.k]);
            dl.push_back(h.dir_lev[o.k]);
            dp.push_back(o.q);
            for (int j = h.dir_off[o.k]; j < h.dir_off[o.k + 1]; ++j) {
                ds.push_back(h.dir_src[j]);
                dc.push_back(h.dir_code[j]);
                df.push_back(h.dir_flags[j]);
            }
            doff.push_back((int)ds.size());
        }
        C.n_dir = (int)dt.size();
        C.dir_tgt.upload(dt);
        C.dir_lev.upload(dl);
        C.dir_off.upload(doff);
        C.dir_src.upload(ds);
        C.dir_code.upload(dc);
        C.dir_flags.upload(df);
        C.dir_prob.upload(dp);
    }
}

fgt::Multi multi(const BClass& C, const DBuf& prob, int off = 0, const DBuf* klev = nullptr) {
    fgt::Multi m;
    m.pp = C.pp.as<fgt::ProbPtrs>();
    m.tprob = prob.as<int>() + off;
    m.tklev = klev ? klev->as<int>() + off : nullptr;
    return m;
}

// Stages of one class (timed: per-stage events on st, for the tuner)
void class_step1(const BClass& C, hipStream_t st) {
    if (C.n_fz > 0 && !fgt::p2pw_fused_3d(C.g, C.t, C.n_fz, C.fz_lev.as<int>(), C.fz_slot.as<int>(),
                                          C.fz_kids.as<int>(), C.fz_ksg.as<int>(), nullptr, nullptr, st,
                                          multi(C, C.fz_prob)))
        throw std::runtime_error("fused STEP 1+2 kernel not available for a batch class");
    if (C.n_s1 > 0)
        fgt::launch_p2pw(C.g, C.t, C.n_s1, C.s1_box.as<int>(), C.s1_lev.as<int>(), C.s1_slot.as<int>(), nullptr,
                         nullptr, st, multi(C, C.s1_prob));
}
void class_step2(const BClass& C, hipStream_t st) {
    for (size_t ph = 0; ph + 1 < C.mp_phase.size(); ++ph) {
        const int a = C.mp_phase[ph], b = C.mp_phase[ph + 1];
        if (b > a)
            fgt::launch_merge(C.g, C.t, b - a, C.mp_slot.as<int>() + a, C.mp_first.as<int>() + a,
                              C.mp_count.as<int>() + a, C.mp_child.as<int>(), C.mp_sign.as<int>(),
                              C.mp_acc.as<int>() + a, 1, nullptr, st, multi(C, C.mp_prob, a, &C.mp_klev));
    }
}
void class_step3(const BClass& C, hipStream_t st) {
    if (C.m2l_brick == 2)
        fgt::launch_m2l_sep(C.g, C.t, C.n_m2l, 0, C.m2l_toff.as<long long>(), C.m2l_gent.as<int>(),
                            C.m2l_soff.as<long long>(), C.m2l_code.as<int>(), nullptr, st, multi(C, C.m2l_prob));
    else if (C.m2l_brick)
        fgt::launch_m2l_brick(C.g, C.t, C.n_m2l, 0, C.m2l_toff.as<long long>(), C.m2l_gent.as<int>(),
                              C.m2l_soff.as<long long>(), C.m2l_code.as<int>(), nullptr, st, multi(C, C.m2l_prob));
    else
        fgt::launch_m2l27(C.g, C.t, C.n_m2l, C.group, 0, C.m2l_toff.as<long long>(), C.m2l_gent.as<int>(),
                          C.m2l_soff.as<long long>(), C.m2l_code.as<int>(), nullptr, st, multi(C, C.m2l_prob),
                          C.m2l_ept);
}
void class_step4(const BClass& C, hipStream_t st) {
    for (size_t ph = 0; ph + 1 < C.pcs_phase.size(); ++ph) {
        const int a = C.pcs_phase[ph], b = C.pcs_phase[ph + 1];
        if (b > a)
            fgt::launch_split(C.g, C.t, b - a, C.pcs_parent.as<int>() + a, C.pcs_child.as<int>() + a,
                              C.pcs_sign.as<int>() + a, 1, nullptr, st, multi(C, C.pcs_prob, a, &C.pcs_klev));
    }
}
void class_step5(const BClass& C, hipStream_t st) {
    if (C.n_s5 > 0)
        fgt::launch_pw2pot_tasks(C.g, C.t, C.n_s5, C.n_s5_sib, C.s5.as<fgt::PotTask>(), nullptr, fgt::Kinds{0, 0, 0},
                                 fgt::OutDesc{nullptr, C.nd, 0}, st, multi(C, C.s5_prob));
}
void class_step6(const BClass& C, hipStream_t st) {
    if (C.n_dir > 0)
        fgt::launch_direct(C.g, C.t, C.n_dir, C.dir_tgt.as<int>(), C.dir_lev.as<int>(), C.dir_off.as<int>(),
                           C.dir_src.as<int>(), C.dir_code.as<int>(), C.dir_flags.as<int>(), nullptr,
                           fgt::Kinds{0, 0, 0}, fgt::OutDesc{nullptr, C.nd, 0}, st, multi(C, C.dir_prob));
}

// Stage timing of a profiled execute: events around each class stage.
struct StageEvents {
    std::vector<hipEvent_t> ev;   // 2 per (class, stage)
    ~StageEvents() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// Enqueue one whole batch transform (graph capture records this sequence);
// prof != nullptr: record events around every stage of every class.
// Class chains of a batch: the classes are dealt to the chains largest
// first, each to the chain with the least work so far (work ~ output nodes
// of the members), so the small classes run beside the large ones instead
// of after them.  Headline batch (133 problems, 9 classes): one chain
// 39.6 ms, two 36.9, three 35.6, four 37.1.
constexpr int kClassChains = 3;

std::vector<int> class_chains(const fgt_batch& B) {
    const size_t nc = B.classes.size();
    std::vector<double> cost(nc, 0.0);
    for (size_t c = 0; c < nc; ++c)
        for (int i : B.classes[c]->members) {
            const fgt::Plan& P = B.solvers[i]->plan;
            cost[c] += (double)P.nleaves * P.npbox * P.nd;
        }
    std::vector<size_t> order(nc);
    for (size_t c = 0; c < nc; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<int> chain(nc, 0);
    std::vector<double> load(kClassChains, 0.0);
    for (size_t c : order) {
        int k = 0;
        for (int j = 1; j < kClassChains; ++j)
            if (load[j] < load[k]) k = j;
        chain[c] = k;
        load[k] += cost[c];
    }
    return chain;
}

// The chain streams and events (created before any graph capture: the
// capture records work, not object creation)
void batch_chain_objects(fgt_batch& B) {
    if (B.chains.empty()) {
        for (int k = 1; k < kClassChains; ++k) {
            hipStream_t s2;
            hcheck(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking), "hipStreamCreate");
            B.chains.push_back(s2);
            hipEvent_t e;
            hcheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            B.ev_chain.push_back(e);
        }
    }
    while (B.ev_cnear.size() < B.classes.size()) {
        hipEvent_t e;
        hcheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        B.ev_cnear.push_back(e);
    }
}

void batch_enqueue(fgt_batch& B, StageEvents* prof = nullptr) {
    // a profiled execute runs everything on one stream (near field and
    // loose problems included), so each per-stage event pair times its
    // launches alone
    hipStream_t st = B.streams[0], sn = prof ? B.streams[0] : B.streams[1];
    const size_t nc = B.classes.size();
    if (prof) {
        prof->ev.assign(12 * nc, nullptr);
        for (auto& e : prof->ev) hcheck(hipEventCreate(&e), "hipEventCreate");
    }
    if ((int)B.chains.size() != kClassChains - 1 || B.ev_cnear.size() < nc)
        throw std::logic_error("batch chain streams not set up");
    std::vector<hipStream_t> ch{st};
    ch.insert(ch.end(), B.chains.begin(), B.chains.end());
    const std::vector<int> chain = prof ? std::vector<int>(nc, 0) : class_chains(B);
    auto mark = [&](size_t c, int step, int end, hipStream_t s) {
        if (prof) hcheck(hipEventRecord(prof->ev[12 * c + 2 * step + end], s), "hipEventRecord");
    };
    for (auto& S : B.solvers) fgt::launch_zero(S->d_pot.p, S->d_pot.bytes, st);
    hcheck(hipEventRecord(B.ev_fork, st), "hipEventRecord");
    for (size_t k = 1; k < B.streams.size(); ++k)
        hcheck(hipStreamWaitEvent(B.streams[k], B.ev_fork, 0), "hipStreamWaitEvent");
    for (hipStream_t s2 : B.chains) hcheck(hipStreamWaitEvent(s2, B.ev_fork, 0), "hipStreamWaitEvent");
    // near field of every class beside STEPs 1-4
    for (size_t c = 0; c < nc; ++c) {
        mark(c, 5, 0, sn);
        class_step6(*B.classes[c], sn);
        mark(c, 5, 1, sn);
        hcheck(hipEventRecord(B.ev_cnear[c], sn), "hipEventRecord");
    }
    hcheck(hipEventRecord(B.ev_near, sn), "hipEventRecord");
    // problems off the class path, spread over the side streams (largest first)
    const size_t nside = B.streams.size() - 2;
    for (size_t i = 0; i < B.loose.size(); ++i) {
        fgt_solver& S = *B.solvers[B.loose[i]];
        execute(S, prof ? st : B.streams[2 + i % nside], false, false, nullptr, true);
    }
    for (size_t c = 0; c < nc; ++c) {
        const BClass& C = *B.classes[c];
        hipStream_t s = ch[chain[c]];
        mark(c, 0, 0, s);
        class_step1(C, s);
        mark(c, 0, 1, s);
        mark(c, 1, 0, s);
        class_step2(C, s);
        mark(c, 1, 1, s);
        mark(c, 2, 0, s);
        class_step3(C, s);
        mark(c, 2, 1, s);
        mark(c, 3, 0, s);
        class_step4(C, s);
        mark(c, 3, 1, s);
    }
    // STEP 5 of a class after its near field (both add into its potentials)
    for (size_t c = 0; c < nc; ++c) {
        hipStream_t s = ch[chain[c]];
        hcheck(hipStreamWaitEvent(s, B.ev_cnear[c], 0), "hipStreamWaitEvent");
        mark(c, 4, 0, s);
        class_step5(*B.classes[c], s);
        mark(c, 4, 1, s);
    }
    for (size_t k = 0; k < B.chains.size(); ++k) {
        hcheck(hipEventRecord(B.ev_chain[k], B.chains[k]), "hipEventRecord");
        hcheck(hipStreamWaitEvent(st, B.ev_chain[k], 0), "hipStreamWaitEvent");
    }
    hcheck(hipStreamWaitEvent(st, B.ev_near, 0), "hipStreamWaitEvent");
    for (size_t k = 2; k < B.streams.size(); ++k) {
        hcheck(hipEventRecord(B.ev_loose[k - 2], B.streams[k]), "hipEventRecord");
        hcheck(hipStreamWaitEvent(st, B.ev_loose[k - 2], 0), "hipStreamWaitEvent");
    }
}

// Device time of f(st): the fastest of reps runs after a warm-up (a run
// slowed by other work on the device does not decide a variant)
template <class F>
float time_ms(hipStream_t st, int reps, F&& f) {
    std::vector<hipEvent_t> ev(reps + 1, nullptr);
    for (auto& e : ev) hcheck(hipEventCreate(&e), "hipEventCreate");
    f(st);   // warm-up
    hcheck(hipEventRecord(ev[0], st), "hipEventRecord");
    for (int r = 0; r < reps; ++r) {
        f(st);
        hcheck(hipEventRecord(ev[r + 1], st), "hipEventRecord");
    }
    hcheck(hipEventSynchronize(ev[reps]), "hipEventSynchronize");
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        float ms = 0;
        hcheck(hipEventElapsedTime(&ms, ev[r], ev[r + 1]), "hipEventElapsedTime");
        best = std::min(best, ms);
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    return best;
}

// Picks, per class, the STEP 3 targets per block and the STEP 1 / STEP 5 x
// stage variant with the shortest device time on the batch's own data.
void batch_autotune(fgt_batch& B, int reps) {
    hipStream_t st = B.streams[0];
    hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    B.drop_graph();
    batch_chain_objects(B);
    batch_enqueue(B);   // expansions and potentials valid for the stage timings
    hcheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (auto& cp : B.classes) {
        BClass& C = *cp;
        if (C.n_fz + C.n_s1 + C.n_m2l + C.n_s5 == 0) continue;
        if (C.n_m2l > 0) {
            int best = C.group, best_e = C.m2l_ept, best_b = 0;
            float bt = 1e30f;
            // sibling bricks (entry streams; the separable form only when
            // FGT_M2L_BRICK=2 asks for it), where every member allows them,
            // then the groups
            const int benv = fgt::m2l_brick_env();
            for (int form : {2, 1}) {
                if (benv == 0 || (benv < 0 && form == 2) || (benv > 0 && benv != form)) continue;
                C.m2l_brick = form;
                class_m2l(B, C);
                if (C.m2l_brick == form) {
                    const float ms = time_ms(st, reps, [&](hipStream_t s) { class_step3(C, s); });
                    if (ms < bt) bt = ms, best_b = form;
                }
            }
            C.m2l_brick = 0;
            for (int E : {1, 2}) {
                if (benv > 0 && best_b) break;   // bricks forced
                if (E == 2 && !C.m2l_ept2_ok) continue;
                if (fgt::m2l_ept_env() && E != fgt::m2l_ept_env()) continue;
                C.m2l_ept = E;
                for (int G : {2, 3, 4, 6, 8}) {
                    C.group = G;
                    class_m2l(B, C);
                    const float ms = time_ms(st, reps, [&](hipStream_t s) { class_step3(C, s); });
                    if (ms < bt) bt = ms, best = G, best_e = E, best_b = 0;
                }
            }
            C.group = best;
            C.m2l_ept = best_e;
            C.m2l_brick = best_b;
            class_m2l(B, C);
        }
        {
            std::vector<int> cand{0};
            if (C.n >= 12 && C.t.npw <= 64) cand.push_back(1);
            if (C.prec >= 1 && C.t.npw <= 80) cand.push_back(2);
            int best = cand[0];
            float bt = 1e30f;
            for (int x : cand) {
                C.t.xs5 = x;
                const float ms = time_ms(st, reps, [&](hipStream_t s) { class_step5(C, s); });
                if (ms < bt) bt = ms, best = x;
            }
            C.t.xs5 = best;
        }
        if (C.prec >= 1) {
            int best = 0;
            float bt = 1e30f;
            for (int x : {0, 2}) {
                C.t.xs1 = x;
                const float ms = time_ms(st, reps, [&](hipStream_t s) { class_step1(C, s); });
                if (ms < bt) bt = ms, best = x;
            }
            C.t.xs1 = best;
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[std::make_tuple(C.n, C.nd, C.prec, C.npw_cap)] = Tuned{C.group, C.t.xs1, C.t.xs5, C.m2l_ept, C.m2l_brick};
    }
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
    o->m2l_bricks = s->n_m2l_bricks > 0 ? s->n_m2l_bricks : -1;
    o->m2l_form = s->n_m2l_bricks <= 0 ? 0 : (s->m2l_sep_ok && fgt::m2l_brick_env() == 2 ? 2 : 1);
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

// -------------------------------------------------------- drop-in cache
// Repeated boxfgt_ calls on one tree keep their plan and device buffers:
// the last few solvers stay resident, keyed by the call's parameters and
// the tree arrays (a 64-bit hash finds the entry, a stored copy of the
// arrays confirms it, so a collision can never return another tree's
// plan).  Only fvals move and STEPs 1-6 run on a hit.  Entries beyond the
// count or byte limit are dropped least recently used first.
// ------------------------------------------------------- tree refinement
// boxfgt with ifnewtree = 1 evaluates the potential on every box its
// refinement creates (boxfgt.f:1130-1227): the plane-wave part from the
// parent's local expansion shifted to the new box (gnd_shiftpw_loc +
// gnd_pw2pgh), the near field from the original leaf's modified list 1 with
// the mrefinelev = 3 tables (gnd_find_loctab_ind + gnd_tens_prod_to_pghloc).
// Here each generation of new boxes is one batch on the device: local
// expansions of the refined boxes are gathered into a workspace (copied
// from the solve's slots, or split from the parent's slot for leaves below
// the cutoff level), split into the new boxes with the STEP 4 kernel and
// evaluated with the STEP 5 kernel; the near field runs through the STEP 6
// kernels with the 193-offset tables.
struct Refiner {
    fgt_solver& S;
    DBuf d_loctab, d_locband;
    std::vector<int> loccols;   // [l][slot][2]
    fgt::DevTables tr;          // the solve's tables with the refinement near-field tables
    int nlt = 0, nslot = 0;
    DBuf prev;                  // the previous generation's workspace
    std::unordered_map<int, int> prev_slot;
    std::vector<std::vector<int>> list1;   // per original leaf, filled on demand
    std::vector<char> have_list1;

    explicit Refiner(fgt_solver& s) : S(s) {
        const fgt::Plan& P = S.plan;
        const int n = P.n, nl = P.nlevels;
        nlt = fgt::loc_nloctab(3);
        nslot = 2 * nlt + 1;
        const size_t nn = (size_t)n * n, per = nn * nslot;
        const size_t ks = per * (nl + 1);
        std::vector<double> tab(3 * ks, 0.0);
        std::vector<int> band((size_t)2 * n * nslot * (nl + 1), 0);
        loccols.assign((size_t)2 * nslot * (nl + 1), 0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int l = 0; l <= nl; ++l)
            fgt::make_loctab_level(P.eps, P.ipoly, n, P.delta, P.tree.boxsize[l], &tab[per * l],
                                   &band[(size_t)2 * n * nslot * l], &loccols[(size_t)2 * nslot * l],
                                   &tab[ks + per * l], &tab[2 * ks + per * l], 3);
        d_loctab.upload(tab);
        d_locband.upload(band);
        tr = S.tabs;
        tr.loctab = d_loctab.as<double>();
        tr.locband = d_locband.as<int>();
        tr.loctab_ks = ks;
        tr.loc_slots = nslot;
        list1.resize(P.nboxes);
        have_list1.assign(P.nboxes, 0);
    }

    const std::vector<int>& list1_of(int b) {
        if (!have_list1[b]) {
            fgt::modified_list1(S.plan, b, list1[b]);
            have_list1[b] = 1;
        }
        return list1[b];
    }

    // split sign of child c relative to parent p (kernel axes), STEP 4's convention
    int split_sign(const std::vector<double>& cen, int ndim, int c, int p) const {
        int k = 0;
        for (int a = 0; a < ndim; ++a)
            if (cen[(size_t)ndim * p + a] - cen[(size_t)ndim * c + a] < 0) k |= 1 << a;
        return map_sign(ndim, k) ^ real_mask(ndim);
    }

    void eval(const fgt::NewTree& T, const std::vector<fgt::NewBox>& kids, std::vector<double>& np,
              std::vector<double>& ng, std::vector<double>& nh) {
        const fgt::Plan& P = S.plan;
        const fgt::Geom& g = S.geom;
        const int ndim = P.ndim, nd = P.nd, mc = 1 << ndim;
        const int nk = (int)kids.size();
        const int nb0 = P.nboxes;
        const hipStream_t st = nullptr;
        // ---- plane waves: local expansions of the refined boxes, then of the new boxes
        const size_t slot_el = 2 * (size_t)nd * P.nexp;   // double2 per slot
        std::vector<int> parents;
        std::unordered_map<int, int> aslot;   // refined box -> workspace slot
        for (const auto& k : kids)
            if (P.do_pw && P.ifpwexp[k.origin] && !aslot.count(k.parent)) {
                aslot[k.parent] = -1;
                parents.push_back(k.parent);
            }
        std::unordered_map<int, int> tmp;     // solve slot of a grandparent -> workspace slot
        int ns = 0;
        for (int p : parents)
            if (p < nb0 && T.level[p] > P.nlevstart && !tmp.count(T.parent[p])) tmp[T.parent[p]] = ns++;
        for (int p : parents) aslot[p] = ns++;
        std::vector<int> bslot(nk, -1);
        for (int i = 0; i < nk; ++i)
            if (aslot.count(kids[i].parent)) bslot[i] = ns++;
        DBuf ws;
        if (ns > 0) ws.alloc(sizeof(double2) * slot_el * ns);
        double2* W = ws.as<double2>();
        const double2* X = S.d_expn.as<double2>();
        const size_t sb = sizeof(double2) * slot_el;
        for (const auto& kv : tmp) {
            const int sl = P.slot_of_box[kv.first];
            if (sl < 0) throw std::logic_error("refinement: parent without a plane-wave slot");
            hcheck(hipMemcpyAsync(W + slot_el * kv.second, X + slot_el * sl, sb, hipMemcpyDeviceToDevice, st), "copy loc");
        }
        // entries per split level code: (source slot, target slot, sign)
        std::map<int, std::array<std::vector<int>, 3>> splits;
        for (int p : parents) {
            const int a = aslot[p];
            if (p < nb0) {
                if (T.level[p] < P.nlevstart) throw std::logic_error("refinement: plane-wave leaf above the cutoff level");
                if (T.level[p] == P.nlevstart) {
                    const int sl = P.slot_of_box[p];
                    if (sl < 0) throw std::logic_error("refinement: leaf without a plane-wave slot");
                    hcheck(hipMemcpyAsync(W + slot_el * a, X + slot_el * sl, sb, hipMemcpyDeviceToDevice, st),
                           "copy loc");
                } else {
                    const int q = T.parent[p];
                    auto& e = splits[P.nlevels - T.level[q]];
                    e[0].push_back(tmp[q]);
                    e[1].push_back(a);
                    e[2].push_back(split_sign(T.centers, ndim, p, q));
                }
            } else {
                auto it = prev_slot.find(p);
                if (it == prev_slot.end()) throw std::logic_error("refinement: refined box without its expansion");
                hcheck(hipMemcpyAsync(W + slot_el * a, prev.as<double2>() + slot_el * it->second, sb,
                                      hipMemcpyDeviceToDevice, st), "copy loc");
            }
        }
        DBuf d_a, d_b, d_c;
        auto run_splits = [&]() {
            for (auto& kv : splits) {
                const int ne = (int)kv.second[0].size();
                d_a.upload(kv.second[0]);
                d_b.upload(kv.second[1]);
                d_c.upload(kv.second[2]);
                fgt::launch_split(g, S.tabs, ne, d_a.as<int>(), d_b.as<int>(), d_c.as<int>(), kv.first, W, st);
                hcheck(hipStreamSynchronize(st), "refinement split");
            }
            splits.clear();
        };
        run_splits();
        std::vector<int> pw_box, pw_lev, pw_slot;
        for (int i = 0; i < nk; ++i) {
            if (bslot[i] < 0) continue;
            const int p = kids[i].parent;
            auto& e = splits[P.nlevels - T.level[p]];
            e[0].push_back(aslot[p]);
            e[1].push_back(bslot[i]);
            e[2].push_back(split_sign(T.centers, ndim, kids[i].id, p));
            pw_box.push_back(i);
            pw_lev.push_back(kids[i].level);
            pw_slot.push_back(bslot[i]);
        }
        run_splits();
        // ---- outputs of the new boxes, local index i
        const int ifpgh = T.ifpgh;
        DBuf d_p, d_g, d_h;
        d_p.alloc(sizeof(double) * T.wp() * nk);
        hcheck(hipMemsetAsync(d_p.p, 0, d_p.bytes, st), "zero");
        if (ifpgh >= 2) {
            d_g.alloc(sizeof(double) * T.wg() * nk);
            hcheck(hipMemsetAsync(d_g.p, 0, d_g.bytes, st), "zero");
        }
        if (ifpgh >= 3) {
            d_h.alloc(sizeof(double) * T.wh() * nk);
            hcheck(hipMemsetAsync(d_h.p, 0, d_h.bytes, st), "zero");
        }
        const std::vector<Comp> comps = components_on(ndim, nd, ifpgh, d_p.as<double>(), d_g.as<double>(),
                                                      d_h.as<double>());
        if (!pw_box.empty()) {
            DBuf db, dl, dsl;
            db.upload(pw_box);
            dl.upload(pw_lev);
            dsl.upload(pw_slot);
            for (const Comp& c : comps)
                fgt::launch_pw2pot(g, S.tabs, (int)pw_box.size(), db.as<int>(), dl.as<int>(), dsl.as<int>(), W, c.kd,
                                   c.od, st);
        }
        // ---- near field: the original leaf's list, tables at the new box's level
        std::vector<int> dt, dlv, doff{0}, dsrc, dcode, dfl;
        const double bs0 = P.bs0;
        for (int i = 0; i < nk; ++i) {
            const int c = kids[i].id, lc = kids[i].level;
            std::vector<std::pair<int, int>> ent;   // (mapped code, source)
            for (int src : list1_of(kids[i].origin)) {
                // gnd_find_loctab_ind (bfgt_local.f:1624-1660), mrefinelev = 3
                const double bsq = P.tree.boxsize[P.tree.level[src]] / 32.0;
                int code = 0;
                bool empty = false;
                for (int a = 0; a < ndim; ++a) {
                    double dx = (T.centers[(size_t)ndim * c + a] - P.tree.centers[(size_t)ndim * src + a]) / bsq;
                    if (P.iper_eff == 1) {
                        const double dp = dx - bs0 / bsq, dm = dx + bs0 / bsq;
                        if (std::fabs(dx) > std::fabs(dp)) dx = dp;
                        if (std::fabs(dx) > std::fabs(dm)) dx = dm;
                    }
                    const int ix = static_cast<int>(dx);
                    if (ix < -nlt || ix > nlt) throw std::logic_error("refinement: near-field offset beyond the tables");
                    const int slot = ix + nlt;
                    const int* cr = &loccols[(size_t)2 * (nslot * lc + slot)];
                    if (cr[0] > cr[1]) empty = true;
                    code |= slot << (8 * a);
                }
                if (empty) continue;   // the reference returns early for an empty band
                ent.push_back({map_code(ndim, code, nlt), src});
            }
            if (ent.empty()) continue;
            std::stable_sort(ent.begin(), ent.end(),
                             [](const std::pair<int, int>& x, const std::pair<int, int>& y) { return (x.first >> 8) < (y.first >> 8); });
            for (size_t j = 0; j < ent.size(); ++j) {
                dsrc.push_back(ent[j].second);
                dcode.push_back(ent[j].first);
                const bool last = j + 1 == ent.size();
                int f = 0;
                if (last || (ent[j].first >> 8) != (ent[j + 1].first >> 8)) f |= 1;
                if (last || (ent[j].first >> 16) != (ent[j + 1].first >> 16)) f |= 2;
                dfl.push_back(f);
            }
            dt.push_back(i);
            dlv.push_back(lc);
            doff.push_back((int)dsrc.size());
        }
        if (!dt.empty()) {
            DBuf a1, a2, a3, a4, a5, a6;
            a1.upload(dt);
            a2.upload(dlv);
            a3.upload(doff);
            a4.upload(dsrc);
            a5.upload(dcode);
            a6.upload(dfl);
            launch_direct_comps(g, tr, (int)dt.size(), a1.as<int>(), a2.as<int>(), a3.as<int>(), a4.as<int>(),
                                a5.as<int>(), a6.as<int>(), S.d_fvals.as<double>(), comps, st);
        }
        hcheck(hipStreamSynchronize(st), "refinement evaluation");
        hcheck(hipMemcpy(np.data(), d_p.p, d_p.bytes, hipMemcpyDeviceToHost), "download");
        if (ifpgh >= 2) hcheck(hipMemcpy(ng.data(), d_g.p, d_g.bytes, hipMemcpyDeviceToHost), "download");
        if (ifpgh >= 3) hcheck(hipMemcpy(nh.data(), d_h.p, d_h.bytes, hipMemcpyDeviceToHost), "download");
        // the new boxes' expansions seed the next generation
        prev_slot.clear();
        for (int i = 0; i < nk; ++i)
            if (bslot[i] >= 0) prev_slot[kids[i].id] = bslot[i];
        prev = std::move(ws);
        (void)mc;
    }
};

// boxfgt with ifnewtree = 1 after the solve: outputs already accumulated
// into the caller's arrays.  Outputs formed only for the targets (ifpgh <
// ifpghtarg) live in zeroed host arrays of the caller's box capacity, as the
// reference's grad2 / hess2 (boxfgt.f:207-272).  timeinfo(9..12): refine,
// coarsen + rebalance, 0, targets (seconds).
void rewrite_tree(fgt_solver& S, double eps, int cap, int ifpgh_caller, int ifpgh_dev, int* itree, const int* iptr,
                  double* centers, const double* boxsize, int* nlevels, double* pot, double* grad, double* hess,
                  int nt, const double* targs, int ifpghtarg, double* pote, double* grade, double* hesse,
                  double* timeinfo) {
    const fgt::Plan& P = S.plan;
    const double ta = now_s();
    fgt::NewTree T;
    T.ndim = P.ndim;
    T.mc = 1 << P.ndim;
    T.mnb = P.ndim == 1 ? 3 : (P.ndim == 2 ? 9 : 27);
    T.nd = P.nd;
    T.n = P.n;
    T.npbox = P.npbox;
    T.ipoly = P.ipoly;
    T.iperiod = P.iper_eff;
    T.ifpgh = ifpgh_dev;
    T.cap = cap;
    T.nlevels = P.nlevels;
    std::vector<double> g2, h2;
    T.pot = pot;
    if (ifpgh_dev >= 2) {
        if (ifpgh_caller >= 2) {
            T.grad = grad;
        } else {
            g2.assign((size_t)cap * P.nd * P.ndim * P.npbox, 0.0);
            download(S.d_grad, g2.data(), 1, "download grad");
            T.grad = g2.data();
        }
    }
    if (ifpgh_dev >= 3) {
        if (ifpgh_caller >= 3) {
            T.hess = hess;
        } else {
            h2.assign((size_t)cap * P.nd * (P.ndim * (P.ndim + 1) / 2) * P.npbox, 0.0);
            download(S.d_hess, h2.data(), 1, "download hess");
            T.hess = h2.data();
        }
    }
    fgt::nt_read(T, itree, iptr, centers, boxsize);
    Refiner R(S);
    fgt::NewTreeStats st;
    fgt::nt_rewrite(
        T, eps, P.ifpwexp,
        [&](const fgt::NewTree& t, const std::vector<fgt::NewBox>& kids, std::vector<double>& np,
            std::vector<double>& ng, std::vector<double>& nh) { R.eval(t, kids, np, ng, nh); },
        &st);
    fgt::nt_write(T, itree, iptr, centers, nlevels);
    const double tb = now_s();
    if (timeinfo) {
        timeinfo[8] = tb - ta;
    }
    if (nt <= 0) return;
    // STEP 11 on the rewritten tree: a solver shell holding only the tree
    // and the outputs
    fgt_solver Q;
    fgt::Plan& Pq = Q.plan;
    Pq.nd = P.nd;
    Pq.ndim = P.ndim;
    Pq.n = P.n;
    Pq.ipoly = P.ipoly;
    Pq.npbox = P.npbox;
    Pq.nboxes = T.count();
    Pq.nlevels = T.nlevels;
    fgt::TreeView& V = Pq.tree;
    V.ndim = T.ndim;
    V.nboxes = T.count();
    V.nlevels = T.nlevels;
    V.mc = T.mc;
    V.mnbors = T.mnb;
    V.lfirst = T.lfirst;
    V.llast = T.llast;
    V.level = T.level;
    V.parent = T.parent;
    V.nchild = T.nchild;
    V.child = T.child;
    V.centers = T.centers;
    V.boxsize = T.boxsize;
    Q.ifpgh = ifpghtarg;
    auto up = [&](DBuf& d, const double* h, size_t per) {
        d.alloc(sizeof(double) * per * T.count());
        hcheck(hipMemcpy(d.p, h, d.bytes, hipMemcpyHostToDevice), "upload outputs");
    };
    up(Q.d_pot, T.pot, T.wp());
    if (ifpghtarg >= 2) up(Q.d_grad, T.grad, T.wg());
    if (ifpghtarg >= 3) up(Q.d_hess, T.hess, T.wh());
    eval_targets(Q, nt, targs, ifpghtarg, pote, grade, hesse, nullptr);
    if (timeinfo) timeinfo[11] = now_s() - tb;
}

constexpr int kDropinCacheEntries = 4;
constexpr size_t kDropinCacheBytes = size_t(32) << 30;

struct DropinKey {
    int nd, ndim, ipoly, iperiod, norder, nboxes, nlevels, ltree;
    double delta, eps;
    std::vector<int> itree, iptr;
    std::vector<double> centers, boxsize;
    unsigned long long hash = 0;
};

unsigned long long mix_hash(unsigned long long h, const void* p, size_t n) {
    // FNV-1a over 8-byte words (tail bytes one at a time)
    const unsigned char* c = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        unsigned long long w;
        std::memcpy(&w, c + i, 8);
        h = (h ^ w) * 1099511628211ULL;
    }
    for (; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    return h;
}

DropinKey dropin_key(const fgt::PlanInput& in) {
    DropinKey k{in.nd, in.ndim, in.ipoly, in.iperiod, in.norder, in.nboxes, in.nlevels, in.ltree, in.delta, in.eps,
                {}, {}, {}, {}, 0};
    k.itree.assign(in.itree, in.itree + in.ltree);
    k.iptr.assign(in.iptr, in.iptr + 8);
    k.centers.assign(in.centers, in.centers + (size_t)in.ndim * in.nboxes);
    k.boxsize.assign(in.boxsize, in.boxsize + in.nlevels + 1);
    unsigned long long h = 1469598103934665603ULL;
    const int sc[8] = {k.nd, k.ndim, k.ipoly, k.iperiod, k.norder, k.nboxes, k.nlevels, k.ltree};
    h = mix_hash(h, sc, sizeof sc);
    h = mix_hash(h, &k.delta, sizeof k.delta);
    h = mix_hash(h, &k.eps, sizeof k.eps);
    h = mix_hash(h, k.itree.data(), sizeof(int) * k.itree.size());
    h = mix_hash(h, k.iptr.data(), sizeof(int) * k.iptr.size());
    h = mix_hash(h, k.centers.data(), sizeof(double) * k.centers.size());
    h = mix_hash(h, k.boxsize.data(), sizeof(double) * k.boxsize.size());
    k.hash = h;
    return k;
}

bool same_key(const DropinKey& a, const DropinKey& b) {
    return a.hash == b.hash && a.nd == b.nd && a.ndim == b.ndim && a.ipoly == b.ipoly && a.iperiod == b.iperiod &&
           a.norder == b.norder && a.nboxes == b.nboxes && a.nlevels == b.nlevels && a.ltree == b.ltree &&
           a.delta == b.delta && a.eps == b.eps && a.itree == b.itree && a.iptr == b.iptr &&
           a.centers == b.centers && a.boxsize == b.boxsize;
}

struct DropinEntry {
    DropinKey key;
    std::unique_ptr<fgt_solver> S;
    size_t bytes = 0;
    unsigned long long last = 0;
};

struct DropinCache {
    std::mutex m;
    std::vector<DropinEntry> e;
    unsigned long long tick = 0;
    long long hits = 0, misses = 0;
};
DropinCache& dropin_cache() {
    static DropinCache c;
    return c;
}

// ----------------------------------------------------------------- server
// Request serving on one GPU.  A request names a problem (tree arrays,
// parameters, source values) by a caller key; resident problems (plan,
// tables, lists, values, expansion workspace on the device) are cached by
// key under a device-memory budget.  Admission: a request for a resident,
// idle problem starts as soon as one of the server's streams is free; a
// request for a problem that is not resident plans it on the host and, when
// its estimated footprint fits the budget after evicting idle problems
// least recently used first, builds it on the device -- otherwise it waits
// (strictly in arrival order among requests that need memory) until running
// requests finish.  Each finished request's potential is copied on the
// device into a result buffer of its own, so the problem is free for the
// next request at once; results stay on the device until fetched or
// released.  All streams belong to the library.

struct SrvProblem {
    fgt_problem spec{};                 // the caller's arrays (kept alive by the caller)
    std::unique_ptr<fgt_solver> S;      // null when not resident
    std::unique_ptr<fgt::Plan> plan;    // host plan waiting for admission
    size_t est = 0, bytes = 0;          // estimated / measured device bytes when resident
    bool busy = false;
    double last_use = 0;
    const double* fv_ptr = nullptr;     // values resident on the device
    long long fv_ver = -1;
};

struct SrvRequest {
    long long key = 0;
    const double* fvals = nullptr;
    long long fver = 0;
    double t_submit = 0, t_start = -1, t_done = -1;
    int state = 0;            // 0 waiting, 1 running, 2 done, 3 failed
    int stream = -1;
    int flags = 0;            // bit0 plan-cache hit, bit1 planned on this request, bit2 waited for memory,
                              // bit3 values uploaded
    hipEvent_t ev = nullptr;
    void* result = nullptr;   // device potential (nd, npbox, nboxes), from the result pool
    size_t result_bytes = 0;
    std::string err;
};

}  // namespace

struct fgt_server {
    double budget = 0;
    size_t used = 0;
    int mode = 0;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_req;        // request running on each stream (-1 free)
    std::unordered_map<long long, SrvProblem> probs;
    std::vector<std::unique_ptr<SrvRequest>> reqs;
    std::deque<int> waiting;
    std::vector<int> finished;          // completed since the last poll
    // released result buffers kept for reuse by size (hipFree would
    // synchronise the device between requests); counted in `used`
    std::multimap<size_t, void*> pool;
    long long plans = 0, evictions = 0, mem_waits = 0, completed = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    ~fgt_server() {
        (void)hipDeviceSynchronize();
        for (auto& r : reqs) {
            if (r && r->ev) (void)hipEventDestroy(r->ev);
            if (r && r->result) (void)hipFree(r->result);
        }
        for (auto& kv : pool) (void)hipFree(kv.second);
        probs.clear();
        for (auto st : streams) (void)hipStreamDestroy(st);
    }
};

namespace {

fgt::PlanInput plan_input(const fgt_problem& q) {
    int npbox = 1;
    for (int a = 0; a < q.ndim && a < 3; ++a) npbox *= q.norder;
    return fgt::PlanInput{q.nd, q.ndim, q.delta, q.eps, q.ipoly, q.iperiod, q.norder, npbox, q.nboxes,
                          q.nlevels, q.ltree, q.itree, q.iptr, q.centers, q.boxsize};
}

// Device bytes a problem will hold (values, potential, expansion slots as
// the 3-D fast path compacts them, tables and lists), from its host plan.
size_t estimate_bytes(const fgt::Plan& P, int mode) {
    const fgt::TreeView& T = P.tree;
    const size_t io = 2 * sizeof(double) * (size_t)P.nd * P.npbox * P.nboxes;
    size_t slots = (size_t)P.nslots;
    const bool fast = P.do_pw && P.ndim == 3 && (P.n == 8 || P.n == 12 || P.n == 16);
    if (fast) {
        slots = 0;
        for (int l = P.nlevstart; l <= P.nlevels; ++l)
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b)
                if (P.slot_of_box[b] >= 0 && (T.nchild[b] > 0 || l == P.nlevstart)) ++slots;
    }
    const size_t elem = mode_precision(mode, P.eps, fast) > 0 ? sizeof(float2) : sizeof(double2);
    const size_t expn = P.do_pw ? elem * 2 * (size_t)P.nd * P.nexp * (slots + 1) : 0;
    const size_t tabs = sizeof(double) * P.loctab.size() + 10 * sizeof(double2) * (P.p2pw.size() + P.pw2pot.size()) +
                        sizeof(int) * (P.dir_src.size() * 3 + P.locband.size() + 16 * (size_t)P.nboxes);
    return io + expn + tabs + (size_t(1) << 20);
}

void srv_evict(fgt_server& V, long long key) {
    SrvProblem& e = V.probs.at(key);
    V.used -= std::min(V.used, e.bytes);
    e.S.reset();
    e.bytes = 0;
    e.fv_ptr = nullptr;
    e.fv_ver = -1;
    ++V.evictions;
}

// A result buffer of `bytes` from the pool, or a new one.
void* srv_take_result(fgt_server& V, size_t bytes) {
    auto it = V.pool.find(bytes);
    if (it != V.pool.end()) {
        void* p = it->second;
        V.pool.erase(it);
        return p;
    }
    void* p = nullptr;
    hcheck(hipMalloc(&p, bytes), "hipMalloc result");
    V.used += bytes;
    return p;
}

void srv_give_result(fgt_server& V, SrvRequest& r) {
    if (!r.result) return;
    V.pool.emplace(r.result_bytes, r.result);
    r.result = nullptr;
}

// Make room for `need` bytes: first the pooled result buffers, then idle
// resident problems least recently used first (never `keep`); false if
// that cannot free enough.
bool srv_make_room(fgt_server& V, size_t need, long long keep) {
    while ((double)(V.used + need) > V.budget && !V.pool.empty()) {
        auto it = V.pool.begin();
        (void)hipFree(it->second);
        V.used -= std::min(V.used, it->first);
        V.pool.erase(it);
    }
    while ((double)(V.used + need) > V.budget) {
        long long victim = 0;
        double oldest = 1e300;
        bool found = false;
        for (auto& kv : V.probs) {
            const SrvProblem& e = kv.second;
            if (kv.first == keep || !e.S || e.busy) continue;
            if (e.last_use < oldest) oldest = e.last_use, victim = kv.first, found = true;
        }
        if (!found) return false;
        srv_evict(V, victim);
    }
    return true;
}

void srv_collect(fgt_server& V) {
    for (size_t k = 0; k < V.streams.size(); ++k) {
        const int id = V.stream_req[k];
        if (id < 0) continue;
        SrvRequest& r = *V.reqs[id];
        const hipError_t q = hipEventQuery(r.ev);
        if (q == hipErrorNotReady) continue;
        hcheck(q, "request completion");
        r.t_done = V.now();
        r.state = 2;
        V.stream_req[k] = -1;
        SrvProblem& e = V.probs.at(r.key);
        e.busy = false;
        e.last_use = r.t_done;
        V.finished.push_back(id);
        ++V.completed;
    }
}

// Starts one request on stream k (its problem is resident and idle).
void srv_launch(fgt_server& V, int id, int k) {
    SrvRequest& r = *V.reqs[id];
    SrvProblem& e = V.probs.at(r.key);
    fgt_solver& S = *e.S;
    hipStream_t st = V.streams[k];
    if (e.fv_ptr != r.fvals || e.fv_ver != r.fver) {
        hcheck(hipMemcpyAsync(S.d_fvals.p, r.fvals, S.d_fvals.bytes, hipMemcpyHostToDevice, st), "upload fvals");
        e.fv_ptr = r.fvals;
        e.fv_ver = r.fver;
        r.flags |= 8;
    }
    execute(S, st, false, false);
    r.result = srv_take_result(V, S.d_pot.bytes);
    r.result_bytes = S.d_pot.bytes;
    hcheck(hipMemcpyAsync(r.result, S.d_pot.p, S.d_pot.bytes, hipMemcpyDeviceToDevice, st), "result copy");
    hcheck(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming), "hipEventCreate");
    hcheck(hipEventRecord(r.ev, st), "hipEventRecord");
    r.state = 1;
    r.stream = k;
    r.t_start = V.now();
    e.busy = true;
    e.last_use = r.t_start;
    V.stream_req[k] = id;
}

void srv_admit(fgt_server& V) {
    bool mem_blocked = false;
    for (auto it = V.waiting.begin(); it != V.waiting.end();) {
        int k = 0;
        while (k < (int)V.streams.size() && V.stream_req[k] >= 0) ++k;
        if (k == (int)V.streams.size()) break;   // every stream busy
        const int id = *it;
        SrvRequest& r = *V.reqs[id];
        SrvProblem& e = V.probs.at(r.key);
        if (e.busy || mem_blocked) {
            ++it;   // its problem runs another request, or an earlier request waits for memory
            continue;
        }
        auto fail = [&](const std::string& m) {
            r.state = 3;
            r.err = m;
            V.finished.push_back(id);
            it = V.waiting.erase(it);
        };
        if (!e.S && !e.plan) {
            auto pl = std::make_unique<fgt::Plan>();
            std::string err;
            if (fgt::make_plan(plan_input(e.spec), *pl, &err) != 0) {
                fail(err);
                continue;
            }
            if (e.est == 0) e.est = estimate_bytes(*pl, V.mode);
            e.plan = std::move(pl);
            r.flags |= 2;
            ++V.plans;
        }
        // device memory this request needs: its problem (if not resident)
        // and a result buffer (unless one of that size is pooled)
        const size_t rb = sizeof(double) * (size_t)e.spec.nd * plan_input(e.spec).npbox * e.spec.nboxes;
        const size_t need = (e.S ? 0 : e.est) + (V.pool.count(rb) ? 0 : rb);
        if ((double)need > V.budget) {
            fail("problem needs more device memory than the server's budget");
            continue;
        }
        if (need > 0 && !srv_make_room(V, need, r.key)) {
            bool running = false;
            for (int q : V.stream_req) running = running || q >= 0;
            if (!running) {
                // nothing will free memory: the budget is held by results
                // the caller has not fetched or released
                fail("not enough device memory within the budget (held by unreleased results)");
                continue;
            }
            // wait for running requests to finish; later requests wait
            // behind this one
            if (!(r.flags & 4)) ++V.mem_waits;
            r.flags |= 4;
            mem_blocked = true;
            ++it;
            continue;
        }
        if (!e.S) {
            auto S = std::make_unique<fgt_solver>();
            g_tally = 0;
            build_solver(*S, plan_input(e.spec), true, e.plan.get(), V.mode);
            e.plan.reset();
            e.bytes = g_tally;
            e.est = e.bytes;   // measured: exact for its next admission
            V.used += e.bytes;
            e.S = std::move(S);
        } else {
            r.flags |= 1;
        }
        srv_launch(V, id, k);
        it = V.waiting.erase(it);
    }
}

void srv_poll(fgt_server& V) {
    srv_collect(V);
    srv_admit(V);
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
        for (int i = 0; i < nprob; ++i) {
            const fgt_problem& q = probs[i];
            if (!q.itree || !q.iptr || !q.centers || !q.boxsize) throw std::invalid_argument("null tree array");
            int npbox = 1;
            for (int a = 0; a < q.ndim && a < 3; ++a) npbox *= q.norder;
            fgt::PlanInput in{q.nd, q.ndim, q.delta, q.eps, q.ipoly, q.iperiod, q.norder, npbox, q.nboxes,
                              q.nlevels, q.ltree, q.itree, q.iptr, q.centers, q.boxsize};
            auto S = std::make_unique<fgt_solver>();
            build_solver(*S, in);
            B->solvers.push_back(std::move(S));
        }
        // stream 0: class stages; 1: near field; 2..: problems off the class path
        for (int k = 0; k < 2 + nstreams; ++k) {
            hipStream_t st;
            hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
            B->streams.push_back(st);
        }
        hcheck(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming), "hipEventCreate");
        hcheck(hipEventCreateWithFlags(&B->ev_near, hipEventDisableTiming), "hipEventCreate");
        B->ev_loose.assign(nstreams, nullptr);
        for (auto& e : B->ev_loose) hcheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        hcheck(hipEventCreate(&B->t0), "hipEventCreate");
        hcheck(hipEventCreate(&B->t1), "hipEventCreate");
        // loose problems largest first (they go round-robin on the side streams)
        build_classes(*B);
        std::stable_sort(B->loose.begin(), B->loose.end(), [&](int x, int y) {
            return B->solvers[x]->plan.nleaves > B->solvers[y]->plan.nleaves;
        });
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
        // expansion buffers and class keys changed: regroup
        const std::vector<int> loose_order = b->loose;
        build_classes(*b);
        std::stable_sort(b->loose.begin(), b->loose.end(), [&](int x, int y) {
            return b->solvers[x]->plan.nleaves > b->solvers[y]->plan.nleaves;
        });
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

int fgt_batch_set_graph(fgt_batch* b, int on) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        b->use_graph = on != 0;
        b->drop_graph();
    });
}

int fgt_batch_autotune(fgt_batch* b, int reps) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (reps < 1 || reps > 100) throw std::invalid_argument("reps must be 1..100");
        batch_autotune(*b, reps);
    });
}

int fgt_batch_class_info(const fgt_batch* b, int k, int* info) {
    return guarded([&] {
        if (!b || !info) throw std::invalid_argument("null argument");
        if (k < 0 || k >= (int)b->classes.size()) throw std::invalid_argument("no such class");
        const BClass& C = *b->classes[k];
        const int v[12] = {C.n, C.nd, C.prec, C.npw_cap, (int)C.members.size(), C.group, C.t.xs1, C.t.xs5,
                           C.n_s5, C.n_dir, C.m2l_ept, C.m2l_brick};
        std::memcpy(info, v, sizeof v);
    });
}

int fgt_batch_profile(fgt_batch* b) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        batch_chain_objects(*b);
        StageEvents prof;
        batch_enqueue(*b, &prof);
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        for (size_t c = 0; c < b->classes.size(); ++c)
            for (int k = 0; k < 6; ++k) {
                float ms = 0;
                hcheck(hipEventElapsedTime(&ms, prof.ev[12 * c + 2 * k], prof.ev[12 * c + 2 * k + 1]),
                       "hipEventElapsedTime");
                b->classes[c]->stage_ms[k] = ms;
            }
    });
}

int fgt_batch_class_stage_ms(const fgt_batch* b, int k, double* ms) {
    return guarded([&] {
        if (!b || !ms) throw std::invalid_argument("null argument");
        if (k < 0 || k >= (int)b->classes.size()) throw std::invalid_argument("no such class");
        for (int i = 0; i < 6; ++i) ms[i] = b->classes[k]->stage_ms[i];
    });
}

int fgt_batch_class_members(const fgt_batch* b, int k, int* members) {
    return guarded([&] {
        if (!b || !members) throw std::invalid_argument("null argument");
        if (k < 0 || k >= (int)b->classes.size()) throw std::invalid_argument("no such class");
        const BClass& C = *b->classes[k];
        std::memcpy(members, C.members.data(), sizeof(int) * C.members.size());
    });
}

int fgt_batch_execute(fgt_batch* b) {
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        hipStream_t st = b->streams[0];
        // earlier work on any stream (uploads, other executes) comes first
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        batch_chain_objects(*b);
        if (b->use_graph && !b->gexec) {
            hcheck(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                batch_enqueue(*b);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hcheck(hipStreamEndCapture(st, &b->graph), "hipStreamEndCapture");
            hcheck(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        }
        hcheck(hipEventRecord(b->t0, st), "hipEventRecord");
        if (b->use_graph) hcheck(hipGraphLaunch(b->gexec, st), "hipGraphLaunch");
        else batch_enqueue(*b);
        hcheck(hipEventRecord(b->t1, st), "hipEventRecord");
        hcheck(hipEventSynchronize(b->t1), "hipEventSynchronize");
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

int fgt_batch_nclasses(const fgt_batch* b) { return b ? (int)b->classes.size() : -1; }

void fgt_batch_destroy(fgt_batch* b) { delete b; }

void* fgt_solver_fvals_device(fgt_solver* s) { return s ? s->d_fvals.p : nullptr; }
void* fgt_solver_pot_device(fgt_solver* s) { return s ? s->d_pot.p : nullptr; }

void fgt_solver_destroy(fgt_solver* s) { delete s; }

int fgt_server_create(double budget_bytes, int nstreams, int precision_mode, fgt_server** out) {
    if (!out) return 2;
    *out = nullptr;
    return guarded([&] {
        if (!(budget_bytes > 0)) throw std::invalid_argument("budget must be > 0");
        if (nstreams < 1 || nstreams > 16) throw std::invalid_argument("nstreams must be 1..16");
        if (precision_mode < 0 || precision_mode > 3) throw std::invalid_argument("precision mode must be 0..3");
        auto V = std::make_unique<fgt_server>();
        V->budget = budget_bytes;
        V->mode = precision_mode;
        for (int k = 0; k < nstreams; ++k) {
            hipStream_t st;
            hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
            V->streams.push_back(st);
        }
        V->stream_req.assign(nstreams, -1);
        *out = V.release();
    });
}

int fgt_server_submit(fgt_server* v, const fgt_problem* p, long long key, const double* fvals, long long fvals_version,
                      int* req_id) {
    return guarded([&] {
        if (!v || !p || !fvals || !req_id) throw std::invalid_argument("null argument");
        if (!p->itree || !p->iptr || !p->centers || !p->boxsize) throw std::invalid_argument("null tree array");
        auto it = v->probs.find(key);
        if (it == v->probs.end()) {
            it = v->probs.emplace(key, SrvProblem{}).first;
            it->second.spec = *p;
        } else {
            const fgt_problem& q = it->second.spec;
            if (q.nd != p->nd || q.ndim != p->ndim || q.norder != p->norder || q.nboxes != p->nboxes ||
                q.nlevels != p->nlevels || q.delta != p->delta || q.eps != p->eps || q.iperiod != p->iperiod ||
                q.ipoly != p->ipoly)
                throw std::invalid_argument("key already names a different problem");
            it->second.spec = *p;   // the latest arrays (same tree)
        }
        auto r = std::make_unique<SrvRequest>();
        r->key = key;
        r->fvals = fvals;
        r->fver = fvals_version;
        r->t_submit = v->now();
        const int id = (int)v->reqs.size();
        v->reqs.push_back(std::move(r));
        v->waiting.push_back(id);
        *req_id = id;
        srv_poll(*v);
    });
}

int fgt_server_poll(fgt_server* v, int* done, int max_done, int* ndone) {
    return guarded([&] {
        if (!v || !ndone || (max_done > 0 && !done)) throw std::invalid_argument("null argument");
        srv_poll(*v);
        const int n = std::min<int>(max_done, (int)v->finished.size());
        for (int i = 0; i < n; ++i) done[i] = v->finished[i];
        v->finished.erase(v->finished.begin(), v->finished.begin() + n);
        *ndone = n;
    });
}

int fgt_server_drain(fgt_server* v, double timeout_s) {
    return guarded([&] {
        if (!v) throw std::invalid_argument("null server");
        const double t_end = v->now() + timeout_s;
        for (;;) {
            srv_poll(*v);
            bool busy = !v->waiting.empty();
            for (int r : v->stream_req) busy = busy || r >= 0;
            if (!busy) break;
            if (v->now() > t_end) throw std::runtime_error("server did not drain in time");
        }
    });
}

int fgt_server_request_info(const fgt_server* v, int id, double* times, int* flags) {
    if (!v || id < 0 || id >= (int)v->reqs.size()) return 2;
    const SrvRequest& r = *v->reqs[id];
    if (times) {
        times[0] = r.t_submit;
        times[1] = r.t_start;
        times[2] = r.t_done;
    }
    if (flags) *flags = r.flags | (r.state << 8);
    return 0;
}

int fgt_server_result(fgt_server* v, int id, double* pot) {
    return guarded([&] {
        if (!v || !pot || id < 0 || id >= (int)v->reqs.size()) throw std::invalid_argument("bad argument");
        SrvRequest& r = *v->reqs[id];
        if (r.state == 3) throw std::invalid_argument("request failed: " + r.err);
        if (r.state != 2) throw std::invalid_argument("request not finished");
        if (!r.result) throw std::invalid_argument("result already released");
        DBuf view;   // borrowed: the pool keeps the buffer
        view.p = r.result;
        view.bytes = r.result_bytes;
        view.from_arena = true;
        download(view, pot, 0, "download result");
        view.p = nullptr;
        srv_give_result(*v, r);
    });
}

int fgt_server_release(fgt_server* v, int id) {
    return guarded([&] {
        if (!v || id < 0 || id >= (int)v->reqs.size()) throw std::invalid_argument("bad argument");
        SrvRequest& r = *v->reqs[id];
        if (r.state == 1) throw std::invalid_argument("request still running");
        srv_give_result(*v, r);
        if (r.ev) {
            (void)hipEventDestroy(r.ev);
            r.ev = nullptr;
        }
    });
}

int fgt_server_stats(const fgt_server* v, fgt_server_info* o) {
    if (!v || !o) return 2;
    std::memset(o, 0, sizeof *o);
    o->budget_bytes = v->budget;
    o->used_bytes = (double)v->used;
    for (const auto& kv : v->probs) o->resident += kv.second.S ? 1 : 0;
    o->problems = (long long)v->probs.size();
    o->requests = (long long)v->reqs.size();
    o->completed = v->completed;
    o->waiting = (long long)v->waiting.size();
    for (int r : v->stream_req) o->running += r >= 0;
    o->plans = v->plans;
    o->evictions = v->evictions;
    o->mem_waits = v->mem_waits;
    return 0;
}

void fgt_server_destroy(fgt_server* v) { delete v; }

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
        if (*ifnewtree != 0 && *ifnewtree != 1) throw std::invalid_argument("ifnewtree must be 0 or 1");
        const double t0 = now_s();
        fgt::PlanInput in{*nd, *ndim, *delta, *eps, *ipoly, *iperiod, *norder, *npbox, *nboxes, *nlevels,
                          *ltree, itree, iptr, centers, boxsize};
        // the plan and device buffers of an earlier call on the same tree
        // and parameters are reused (see DropinCache)
        DropinCache& cache = dropin_cache();
        std::unique_lock<std::mutex> lk(cache.m);
        DropinKey key = dropin_key(in);
        DropinEntry* hit = nullptr;
        for (auto& en : cache.e)
            if (same_key(en.key, key)) hit = &en;
        if (hit) {
            ++cache.hits;
        } else {
            ++cache.misses;
            auto Sn = std::make_unique<fgt_solver>();
            g_tally = 0;
            build_solver(*Sn, in);
            DropinEntry en;
            en.key = std::move(key);
            en.bytes = g_tally;
            en.S = std::move(Sn);
            // keep at most kDropinCacheEntries (and kDropinCacheBytes): drop
            // the least recently used first
            size_t total = en.bytes;
            for (auto& o : cache.e) total += o.bytes;
            while (!cache.e.empty() && ((int)cache.e.size() >= kDropinCacheEntries || total > kDropinCacheBytes)) {
                auto old = std::min_element(cache.e.begin(), cache.e.end(),
                                            [](const DropinEntry& a, const DropinEntry& b) { return a.last < b.last; });
                total -= old->bytes;
                cache.e.erase(old);
            }
            cache.e.push_back(std::move(en));
            hit = &cache.e.back();
        }
        hit->last = ++cache.tick;
        fgt_solver& S = *hit->S;
        // gradients / Hessians are formed when the caller or the targets need
        // them (boxfgt.f:202-272 routes ifpgh < ifpghtarg the same way)
        const int ifpgh_dev = std::max(*ifpgh, nt > 0 ? *ifpghtarg : 0);
        {
            // a cached solver keeps the previous call's output set
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
        double t3 = now_s();
        if (*ifnewtree == 1) {
            // STEPs 8-10: refine, coarsen and rebalance the tree around the
            // computed outputs, then STEP 11 on the rewritten tree
            rewrite_tree(S, *eps, *nboxes, *ifpgh, ifpgh_dev, itree, iptr, centers, boxsize, nlevels, pot, grad,
                         hess, nt, targs, *ifpghtarg, pote, grade, hesse, timeinfo);
            if (timeinfo) {
                timeinfo[0] = t1 - t0;
                for (int i = 0; i < 6; ++i) timeinfo[1 + i] = S.stage_ms[i] * 1e-3;
                timeinfo[7] = t3 - t2;
            }
            return;
        }
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

int fgt_dropin_cache_stats(long long* hits, long long* misses, int* entries) {
    DropinCache& c = dropin_cache();
    std::lock_guard<std::mutex> lk(c.m);
    if (hits) *hits = c.hits;
    if (misses) *misses = c.misses;
    if (entries) *entries = (int)c.e.size();
    return 0;
}

int fgt_dropin_cache_clear(void) {
    return guarded([&] {
        DropinCache& c = dropin_cache();
        std::lock_guard<std::mutex> lk(c.m);
        hcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        c.e.clear();
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
