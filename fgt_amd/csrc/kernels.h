// kernels.h -- launch interface of the box-FGT HIP kernels (gfx950).
//
// Geometry is expressed with three axes (x, y, z); the plane-wave half
// spectrum always lives on z.  A 3-D problem uses (n,n,n) nodes and
// (npw,npw,nh) frequencies; a 2-D problem maps (x,y) -> (x,z) with a unit
// y axis; a 1-D problem maps x -> z.  Unit axes carry the 1x1 identity
// table and unit phases, so one set of kernels serves every dimension.
//
// Expansion layout in HBM (one "slot" = mp then loc, nd densities each):
//   E[ind][k3][k1][k2], k2 fastest  (k3 < Pz = nh, k1 < Px, k2 < Py)
#pragma once

#include <hip/hip_runtime.h>

namespace fgt {

struct Geom {
    int nd;
    int nx, ny, nz;      // nodes per axis
    int px, py, pz;      // frequencies per axis (pz = half spectrum)
    int npw;             // full 1-D frequency count (odd => zero-frequency halving)
    int npbox;           // nx*ny*nz
    long long nexp;      // px*py*pz
    int ndim;
};

#ifdef __HIPCC__
// Expansion storage: FP64 (double2) or, in the mixed-precision form, FP32
// (float2) complex numbers; the arithmetic is FP64 either way.
__device__ __forceinline__ double2 ldE(const double2& v) { return v; }
__device__ __forceinline__ double2 ldE(const float2& v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void stE(double2* p, double2 v) { *p = v; }
__device__ __forceinline__ void stE(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

// Read-only 1-D tables viewed through the constant address space: loads
// there are invariant for the kernel, so wave-uniform reads become scalar
// loads into SGPRs whether the table pointer is a kernel argument or comes
// from a per-problem descriptor in memory (where the compiler could not
// otherwise rule out a clobbering store).  CTab2: complex (double2) entries,
// CTab1: real.
typedef const double __attribute__((address_space(4))) cdouble_t;
typedef double dvec2_t __attribute__((ext_vector_type(2)));
typedef const dvec2_t __attribute__((address_space(4))) cdvec2_t;
struct CTab2 {
    cdvec2_t* p;
    __device__ __forceinline__ CTab2 operator+(long long o) const { return CTab2{p + o}; }
    __device__ __forceinline__ double2 operator[](long long i) const {
        const dvec2_t v = p[i];   // one 16-byte load
        return make_double2(v.x, v.y);
    }
};
struct CTab1 {
    cdouble_t* p;
    __device__ __forceinline__ CTab1 operator+(long long o) const { return CTab1{p + o}; }
    __device__ __forceinline__ double operator[](long long i) const { return p[i]; }
};
__device__ __forceinline__ CTab2 ctab(const double2* q) { return CTab2{(cdvec2_t*)q}; }
__device__ __forceinline__ CTab1 ctab(const double* q) { return CTab1{(cdouble_t*)q}; }

// Workgroups are dealt to the 8 XCDs round-robin; give each XCD a contiguous
// range of tasks so that tasks sharing data (children of one parent,
// neighbouring targets, one element range of many targets) meet in the same L2.
__device__ __forceinline__ int xcd_task(int b, int nb) {
    const int x = b & 7, i = b >> 3, per = nb >> 3, rem = nb & 7;
    return x * per + min(x, rem) + i;
}
// Runs of C consecutive tasks per XCD, dealt round-robin: the XCDs work on
// the same region of the task list at any time (balanced when task costs
// drift along the list) while each run (e.g. the children of one parent)
// shares one L2.  Blocks past the last whole round keep their own index.
template <int C>
__device__ __forceinline__ int xcd_run_task(int b, int nb) {
    const int whole = nb / (8 * C) * (8 * C);
    if (b >= whole) return b;
    const int x = b & 7, i = b >> 3;
    return (i / C) * (8 * C) + x * C + (i % C);
}
#endif

// One output component of STEP 5 / STEP 6 (potential, one gradient or one
// Hessian entry): value of node `node` of box b, density ind, goes to
// p[(b*npbox + node)*ns + coff + ind] (ns = nd * components of the array).
struct OutDesc {
    double* p;
    int ns;
    int coff;
};

// Per kernel axis (x, y, z): which 1-D table kind a component uses
// (0 value, 1 first derivative, 2 second derivative).
struct Kinds {
    int x, y, z;
};

// Output components that differ only in their z table kind (up to three:
// value, first, second derivative) share one STEP 5 or STEP 6 pass: the x
// and y sums are formed once and each z kind zo.kz[q] contracts them into
// its own output zo.od[q].
struct ZOut {
    int nz;
    int kz[3];
    OutDesc od[3];
};

// One task of the 3-D STEP 5 kernel: the local expansion in `slot`,
// evaluated at up to four leaves of level `lev` that share the x table
// variant vx (0 plain table; 1 / 2 table times the + / conjugate
// parent->child split phase).  Leaf c: box[c] (-1 = unused), y variant
// vy[code[c] & 1], z variant code[c] >> 1.  A leaf evaluating its own
// expansion is a task with one leaf and all variants 0; the leaves of one
// parent with the same x split sign form one task, so the x contraction of
// the parent's expansion is done once for all of them.
struct PotTask {
    int slot, lev, vx, nvy;
    int vy[2];
    int box[4];
    int code[4];
};

// Several problems of one shape class (3-D, same norder, density count and
// precision; plane-wave counts may differ) in one launch.  Task i of a launch belongs
// to problem tprob[i]; the kernel then takes that problem's tables, values,
// outputs and expansion workspace from pp[tprob[i]] instead of its
// single-problem arguments.  tklev[i]: the merge/split level code of task i
// (STEPs 2 and 4, whose phase rows depend on the problem's depth).  All
// pointers device pointers; pp == nullptr means a single-problem launch.
struct ProbPtrs {
    Geom g;                   // the problem's geometry (plane-wave count, nexp)
    const double2* p2pw;      // value->PW tables [l][k][m]
    const double2* p2pwS;     // the same times the merge phases [l][v][k][m]
    const double2* pw2potS;   // PW->value tables [l][v][k][i] (value kind)
    const double2* wms;       // merge/split 1-D phases [klev-1][k]
    const double2* wm2l;      // M2L 1-D phases [o][k]
    const double* loctab;     // near-field tables [l][slot][k][j] (value kind)
    const double* fvals;      // (nd, npbox, nboxes)
    double* pot;              // (nd, npbox, nboxes), accumulated into
    void* expn;               // expansion slots (FP64 or FP32 complex)
    long long zero_off;       // element offset of the all-zero slot
};
struct Multi {
    const ProbPtrs* pp = nullptr;
    const int* tprob = nullptr;
    const int* tklev = nullptr;
};

// Device-resident tables (all pointers are device pointers).
struct DevTables {
    // per level l: value->PW  T[l][k][m] (k < npw, m < n) complex (double2)
    const double2* p2pw;
    // per kind and level l: PW->value  U[kind][l][i][k] (i < n, k < npw),
    // and transposed UT[kind][l][k][i]
    const double2* pw2pot;
    const double2* pw2potT;
    // pw2potT and its products with the parent->child split phase of each
    // level: pw2potS[kind][l][v][k][i], v = 0 plain, 1 = + phase, 2 = conjugate
    const double2* pw2potS;
    // p2pw and its products with the child->parent merge phase of each level:
    // p2pwS[l][v][k][m], v = 0 plain, 1 = + phase, 2 = conjugate phase
    const double2* p2pwS;
    // merge/split 1-D phases wms[klev-1][k], M2L 1-D phases wm2l[o][k]
    const double2* wms;
    const double2* wm2l;
    // identity table for unit axes (one complex 1.0)
    const double2* one;
    // near-field masked tables loc[kind][l][slot][k][j], bands [l][slot][k][2];
    // loc_slots table slots per level (13 for the solve, 193 for the
    // refinement of a new tree); a pair's table slots are packed 8 bits per
    // kernel axis: x | y << 8 | z << 16
    const double* loctab;
    const int* locband;
    int loc_slots = 13;
    // kind strides (elements): pw2pot / pw2potT, pw2potS, loctab
    size_t pw2pot_ks, pw2potS_ks, loctab_ks;
    int n;     // norder
    int npw;
    // 1: expansions are stored as FP32 complex (mixed precision; 3-D
    // register-blocked path only), 0: FP64
    int expn_f32;
    // contraction precision of the mixed-precision forms: 0 FP64, 1 FP32
    // matrix cores, 2 BF16 hi/lo matrix cores (prec >= 1 implies expn_f32)
    int prec;
    // x-stage variant of STEP 1 / STEP 5 (tuned per shape class; -1 = the
    // default policy): 0 VALU, 1 FP64 matrix cores (STEP 5, norder >= 12,
    // npw <= 64), 2 the reduced-precision matrix cores (prec >= 1)
    int xs1 = -1, xs5 = -1;
    size_t lds_limit;  // device LDS bytes per block
};

// Register-blocked 3-D STEP 1 / STEP 5 kernels (norder 8, 12, 16); return
// false (nothing launched) when the shape is not covered.
bool launch_p2pw_3d(const Geom& g, const DevTables& t, int ntask, const int* box, const int* lev,
                    const int* slot, const double* fvals, double2* expn, hipStream_t s, size_t lds_limit,
                    const Multi& mu = Multi{});
// tasks [0, nsib): sibling tasks (up to four leaves); [nsib, ntask): one leaf
bool launch_pw2pot_3d(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                      const double2* expn, Kinds kd, OutDesc od, hipStream_t s, size_t lds_limit,
                      const Multi& mu = Multi{});
// the same for the components of zo (x and y kinds kd.x, kd.y): the
// one-leaf tasks share their x and y stages across the z kinds
bool launch_pw2pot_3d_z(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                        const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s, size_t lds_limit,
                        const Multi& mu = Multi{});
bool pw2pot_3d_supported(const Geom& g, const DevTables& t);
// true when FGT_PW2POT_VALU=1: STEP 5 keeps its x stage on the VALU
bool pw2pot_valu_only();
// Fused STEP 1 + 2 for parents with leaf children (3-D, N in {8,12,16}):
// plev/pslot per parent, kids/ksg 8 per parent (leaf child box or -1, merge
// sign bits).  Writes (overwrites) the parent's mp; returns false if not covered.
bool p2pw_fused_3d(const Geom& g, const DevTables& t, int ntask, const int* plev, const int* pslot,
                   const int* kids, const int* ksg, const double* fvals, double2* expn, hipStream_t s,
                   const Multi& mu = Multi{});

// Stage launchers (stream-ordered; all arguments device pointers).
void launch_p2pw(const Geom& g, const DevTables& t, int ntask, const int* box,
                 const int* lev, const int* slot, const double* fvals,
                 double2* expn, hipStream_t s, const Multi& mu = Multi{});
// par_acc (nullable): per parent, nonzero = add to the parent's current mp
void launch_merge(const Geom& g, const DevTables& t, int ntask, const int* par_slot,
                  const int* par_first, const int* par_count, const int* pc_child,
                  const int* pc_sign, const int* par_acc, int klev, double2* expn, hipStream_t s,
                  const Multi& mu = Multi{});
// STEP 3, kM2lGroup targets per block.  toff[i]: element offset of target
// i's slot (slot * 2 * nd * nexp); gent: per group, the range of its entry
// stream (each target's own expansion, then its list; padded to a multiple
// of kM2lBatch); soff[j]: element offset of the entry's source slot (padding
// points at an all-zero slot, whose element offset is zero_off); scode[j] =
// o | last << 5 with o = ox + 3 oy + 9 oz in 0..26 and last set on each
// target's final entry.
constexpr int kM2lGroup = 4;   // best at 4-5 with six blocks per CU (8 was best at four)
constexpr int kM2lGroupMax = 12;
// targets per STEP 3 block: kM2lGroup, or FGT_M2L_GROUP (1..kM2lGroupMax,
// tuning runs)
int m2l_group();
constexpr int kM2lBatch = 4;
// longest entry stream of a group: kM2lGroupMax targets x (own + 26 list
// entries), padded
constexpr int kM2lMaxEntries = kM2lGroupMax * 27 + kM2lBatch;
// ept: elements per thread, 1 or 2 (2 needs an even nd * nexp in every
// problem of the launch)
void launch_m2l27(const Geom& g, const DevTables& t, int ngroup, int group, long long zero_off,
                  const long long* toff, const int* gent, const long long* soff, const int* scode, double2* expn,
                  hipStream_t s, const Multi& mu = Multi{}, int ept = 1);
// FGT_M2L_EPT (1 or 2) forces the elements per thread of STEP 3; 0 = policy
int m2l_ept_env();
// STEP 3 over sibling bricks (3-D): per brick 8 loc element offsets (-1 =
// none), an entry stream of source mp offsets and codes (position 6 bits |
// target mask << 6), padded to a multiple of kM2lBatch
constexpr int kM2lBrickMax = 64 + kM2lBatch;
void launch_m2l_brick(const Geom& g, const DevTables& t, int nbrick, long long zero_off, const long long* toff,
                      const int* gent, const long long* soff, const int* scode, double2* expn, hipStream_t s,
                      const Multi& mu = Multi{});
// The same bricks as a separable stencil (the brick's lists must be
// geometric: every source holds every sibling whose 3x3x3 neighbourhood
// contains it): the phase of offset (ox, oy, oz) is w_x(ox) w_y(oy)
// w_z(oz) with w(0) = conj(w(2)), w(1) = 1, so the sum over a sibling's 27
// neighbours runs as three 3-tap passes along x, y and z over the brick's
// 4x4x4 source grid, each source loaded once.  Same arrays as
// launch_m2l_brick.
void launch_m2l_sep(const Geom& g, const DevTables& t, int nbrick, long long zero_off, const long long* toff,
                    const int* gent, const long long* soff, const int* scode, double2* expn, hipStream_t s,
                    const Multi& mu = Multi{});
// FGT_M2L_BRICK: 0 keeps STEP 3 on the target groups, 1 on sibling bricks
// (entry streams), 2 on sibling bricks as a separable stencil, where a
// problem allows them; unset (-1) = policy (entry-stream bricks: the
// separable form does a sixth of the arithmetic but was measured slower,
// 1.17 vs 0.74 ms on a norder-16 periodic problem, bound by the cache
// traffic of the source loads that both forms share)
int m2l_brick_env();
void launch_split(const Geom& g, const DevTables& t, int nentry, const int* pc_parent,
                  const int* pc_child, const int* pc_sign, int klev, double2* expn,
                  hipStream_t s, const Multi& mu = Multi{});
void launch_pw2pot(const Geom& g, const DevTables& t, int ntask, const int* box,
                   const int* lev, const int* slot, const double2* expn,
                   Kinds kd, OutDesc od, hipStream_t s);
// 3-D STEP 5 over PotTask lists (throws when the shape is not covered)
void launch_pw2pot_tasks(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                         const double2* expn, Kinds kd, OutDesc od, hipStream_t s, const Multi& mu = Multi{});
void launch_pw2pot_tasks_z(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                           const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s);
// flags (nullable): per entry bit0 = last of its (y,z) table group, bit1 =
// last of its z group; entries must then be grouped in that order.
void launch_direct(const Geom& g, const DevTables& t, int ntask, const int* tgt,
                   const int* lev, const int* off, const int* src, const int* code,
                   const int* flags, const double* fvals, Kinds kd, OutDesc od, hipStream_t s,
                   const Multi& mu = Multi{});
bool launch_direct_3d(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev,
                      const int* off, const int* src, const int* code, const int* flags,
                      const double* fvals, Kinds kd, OutDesc od, hipStream_t s, const Multi& mu = Multi{});
// 3-D register-blocked path only (norder 8, 12, 16, with flags); false
// when not covered (nothing launched)
bool launch_direct_3d_z(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev,
                        const int* off, const int* src, const int* code, const int* flags,
                        const double* fvals, Kinds kd, const ZOut& zo, hipStream_t s);

// Device workspace of the no-allocation entry point: a file-scope device
// array of kScratchBytes per GPU.  scratch_region() returns its address on
// the current device (looked up once per device); launch_zero() clears a
// byte range (8-byte aligned) with a fill kernel on stream s.
constexpr size_t kScratchBytes = size_t(128) << 20;
void* scratch_region(size_t* bytes);
void launch_zero(void* p, size_t bytes, hipStream_t s);

// Extra targets (STEP 11): values (ncd, npbox, nboxes) on the listed leaves
// -> coefficients coefs[b][q][i] (u = values->coefficients, u(k,j) at
// u[k + j*n]); then per task (leaf tbox, sorted targets tfirst..+tcount,
// tcount <= 256) the expansion at each target, out[perm[t]*ncd + q].
void launch_vals2coefs(int ndim, int n, int ncd, const double* u, int nleaf, const int* leaves, const double* vals,
                       double* coefs, hipStream_t s);
void launch_evalt(int ndim, int n, int ipoly, int ncd, int ntask, const int* tbox, const int* tfirst,
                  const int* tcount, const double* centers, const double* bsize_of_box, const double* targs_sorted,
                  const int* perm, const double* coefs, double* out, hipStream_t s);

}  // namespace fgt
