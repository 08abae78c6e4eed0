// kernels.hip -- box-FGT stage kernels for gfx950 (FP64).
//
// Stage map (reference loops they replace):
//   k_p2pw    STEP 1, values -> PW per PW leaf    (boxfgt.f:710-726,
//             g3d_tens_prod_to_pw bfgt_pwrouts.f:153-222)
//   k_p2pw2   STEP 1 for 2-D                      (g2d_tens_prod_to_pw bfgt_pwrouts.f:98-146)
//   k_merge   STEP 2, child mp -> parent mp       (boxfgt.f:746-765)
//   k_m2l     STEP 3, mp -> loc at cutoff level   (boxfgt.f:787-804)
//   k_split   STEP 4, parent loc -> child loc     (boxfgt.f:827-844)
//   k_pw2pot  STEP 5, loc -> node values          (boxfgt.f:865-881,
//             g3d_pw2pot bfgt_pwrouts.f:719-790)
//   k_pw2pot2 STEP 5 for 2-D                      (g2d_pw2pot bfgt_pwrouts.f:473-526)
//   k_direct  STEP 6, banded near field           (boxfgt.f:904-928,
//             g3d_tens_prod_to_potloc_fast bfgt_local.f:1322-1618)
//   k_direct2 STEP 6 for 2-D                      (g2d_tens_prod_to_potloc_fast bfgt_local.f:1230-1315)
#include "kernels.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>

namespace fgt {

namespace {

constexpr int kBlock = 256;
// rows of grid y per launch (HIP's y dimension is limited to 65535)
constexpr int kMaxGridY = 65535;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
    // c + a*b
    return make_double2(c.x + a.x * b.x - a.y * b.y, c.y + a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

// Separable phase of element (k1,k2,k3) for sign mask s (bit0 x, bit1 y,
// bit2 z; bit set = conjugate), formed as (x*y)*z like the reference tables.
__device__ __forceinline__ double2 ms_phase(const double2* w1, int s, int k1, int k2, int k3,
                                            const Geom& g) {
    double2 wx = g.px > 1 ? w1[k1] : make_double2(1.0, 0.0);
    double2 wy = g.py > 1 ? w1[k2] : make_double2(1.0, 0.0);
    double2 wz = w1[k3];
    if (s & 1) wx = conjd(wx);
    if (s & 2) wy = conjd(wy);
    if (s & 4) wz = conjd(wz);
    return cmul(cmul(wx, wy), wz);
}

__device__ __forceinline__ double2 m2l_phase(const double2* w, int npw, int idx, int k1, int k2,
                                             int k3, const Geom& g) {
    const int ox = idx % 3, oy = (idx / 3) % 3, oz = idx / 9;
    double2 wx = g.px > 1 ? w[ox * npw + k1] : make_double2(1.0, 0.0);
    double2 wy = g.py > 1 ? w[oy * npw + k2] : make_double2(1.0, 0.0);
    double2 wz = w[oz * npw + k3];
    return cmul(cmul(wx, wy), wz);
}

// ---------------------------------------------------------------- STEP 1
__global__ void __launch_bounds__(kBlock) k_p2pw(Geom g, DevTables t, const int* __restrict__ box,
                                                 const int* __restrict__ lev,
                                                 const int* __restrict__ slot,
                                                 const double* __restrict__ fvals,
                                                 double2* __restrict__ expn) {
    extern __shared__ double2 sm[];
    const int task = blockIdx.x;
    const int b = box[task], l = lev[task], s = slot[task];
    const int n = t.n, npw = t.npw;
    double2* sT = sm;                    // [npw][n]
    double2* ff = sT + npw * n;          // [m2][m1]  nx*ny
    double2* ff2 = ff + g.nx * g.ny;     // [k2][m1]  py*nx
    const double2* T = t.p2pw + (size_t)l * npw * n;
    for (int i = threadIdx.x; i < npw * n; i += blockDim.x) sT[i] = T[i];
    const int nxy = g.nx * g.ny;
    const int m1 = threadIdx.x % g.nx, m2 = threadIdx.x / g.nx;
    double freg[16];
    const double one = 1.0;
    for (int ind = 0; ind < g.nd; ++ind) {
        __syncthreads();
        if ((int)threadIdx.x < nxy)
            for (int m3 = 0; m3 < g.nz; ++m3)
                freg[m3] = fvals[((size_t)b * g.npbox + m1 + g.nx * (m2 + g.ny * m3)) * g.nd + ind];
        double2* E = expn + ((size_t)s * 2 * g.nd + ind) * g.nexp;
        for (int k3 = 0; k3 < g.pz; ++k3) {
            if ((int)threadIdx.x < nxy) {
                double2 acc = make_double2(0.0, 0.0);
                for (int m3 = 0; m3 < g.nz; ++m3) {
                    const double2 tz = sT[k3 * n + m3];
                    acc.x += tz.x * freg[m3];
                    acc.y += tz.y * freg[m3];
                }
                ff[m2 * g.nx + m1] = acc;
            }
            __syncthreads();
            for (int o = threadIdx.x; o < g.py * g.nx; o += blockDim.x) {
                const int a1 = o % g.nx, k2 = o / g.nx;
                double2 acc = make_double2(0.0, 0.0);
                if (g.ny == 1) {
                    acc = ff[a1];
                } else {
                    for (int b2 = 0; b2 < g.ny; ++b2) acc = cfma(sT[k2 * n + b2], ff[b2 * g.nx + a1], acc);
                }
                ff2[k2 * g.nx + a1] = acc;
            }
            __syncthreads();
            double2* Ek = E + (size_t)k3 * g.px * g.py;
            for (int o = threadIdx.x; o < g.px * g.py; o += blockDim.x) {
                const int k2 = o % g.py, k1 = o / g.py;
                double2 acc = make_double2(0.0, 0.0);
                if (g.nx == 1) {
                    acc = ff2[k2];
                } else {
                    for (int a1 = 0; a1 < g.nx; ++a1) acc = cfma(sT[k1 * n + a1], ff2[k2 * g.nx + a1], acc);
                }
                Ek[o] = acc;
            }
            __syncthreads();
        }
    }
    (void)one;
}

// ---------------------------------------------------------------- STEP 1, 2-D
// g2d_tens_prod_to_pw (bfgt_pwrouts.f:98-146): E[k3][k1] = sum_{m1, m2}
// T[k1][m1] T[k3][m2] f[m1][m2] over the half axis k3.  The block stages
// the level's table and the leaf's n x n values in LDS and spreads each
// contraction over all its threads: the half axis with one thread per
// (m1, k3), then x with one thread per output (k1 fastest, so the
// expansion row is written coalesced).  (The generic kernel ran one
// barrier-separated pass per half-axis frequency with n active threads.)
// LDS: sT [k][m] (npw*n complex), ff [m1][k3] (n*nh complex), sf [m2][m1]
// (n*n real).
__global__ void __launch_bounds__(kBlock) k_p2pw2(Geom g, DevTables t, const int* __restrict__ box,
                                                  const int* __restrict__ lev, const int* __restrict__ slot,
                                                  const double* __restrict__ fvals, double2* __restrict__ expn) {
    extern __shared__ double2 sm[];
    const int task = blockIdx.x;
    const int b = box[task], l = lev[task], sl = slot[task];
    const int n = t.n, npw = t.npw, nh = g.pz;
    double2* sT = sm;
    double2* ff = sT + (size_t)npw * n;
    double* sf = reinterpret_cast<double*>(ff + (size_t)n * nh);
    const double2* __restrict__ T = t.p2pw + (size_t)l * npw * n;
    const int tid = threadIdx.x;
    for (int i = tid; i < npw * n; i += blockDim.x) sT[i] = T[i];
    for (int ind = 0; ind < g.nd; ++ind) {
        for (int i = tid; i < n * n; i += blockDim.x) sf[i] = fvals[((size_t)b * g.npbox + i) * g.nd + ind];
        __syncthreads();
        for (int o = tid; o < n * nh; o += blockDim.x) {
            const int m1 = o % n, k3 = o / n;
            const double2* __restrict__ tz = sT + (size_t)k3 * n;
            double2 acc = make_double2(0.0, 0.0);
            for (int m2 = 0; m2 < n; ++m2) {
                const double f = sf[m1 + n * m2];
                acc.x += tz[m2].x * f;
                acc.y += tz[m2].y * f;
            }
            ff[m1 * nh + k3] = acc;
        }
        __syncthreads();
        double2* __restrict__ E = expn + ((size_t)sl * 2 * g.nd + ind) * g.nexp;
        for (int o = tid; o < npw * nh; o += blockDim.x) {
            const int k1 = o % npw, k3 = o / npw;
            const double2* __restrict__ tx = sT + (size_t)k1 * n;
            double2 acc = make_double2(0.0, 0.0);
            for (int m1 = 0; m1 < n; ++m1) acc = cfma(tx[m1], ff[m1 * nh + k3], acc);
            E[o] = acc;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- STEP 2
template <typename ET>
__global__ void __launch_bounds__(kBlock) k_merge(Geom g, DevTables t, const int* __restrict__ par_slot,
                                                  const int* __restrict__ par_first,
                                                  const int* __restrict__ par_count,
                                                  const int* __restrict__ pc_child,
                                                  const int* __restrict__ pc_sign,
                                                  const int* __restrict__ par_acc, int klev,
                                                  ET* __restrict__ expn_a, Multi mu) {
    const int task = blockIdx.y;
    ET* __restrict__ expn = expn_a;
    const double2* w1 = t.wms + (size_t)(klev - 1) * t.npw;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[task]];
        g = q.g;
        expn = static_cast<ET*>(q.expn);
        w1 = q.wms + (size_t)(mu.tklev[task] - 1) * q.g.npw;
    }
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long tot = (long long)g.nd * g.nexp;
    if (e >= tot) return;
    const int r = (int)(e % g.nexp);
    const int k2 = r % g.py, k1 = (r / g.py) % g.px, k3 = r / (g.py * g.px);
    const int first = par_first[task], cnt = par_count[task];
    // par_acc[task] != 0: the parent already holds its leaf children's part
    double2 acc = (par_acc && par_acc[task]) ? ldE(expn[(size_t)par_slot[task] * 2 * tot + e])
                                             : make_double2(0.0, 0.0);
    for (int i = 0; i < cnt; ++i) {
        const int cs = pc_child[first + i];
        const double2 v = ldE(expn[(size_t)cs * 2 * tot + e]);
        acc = cfma(v, ms_phase(w1, pc_sign[first + i], k1, k2, k3, g), acc);
    }
    stE(expn + (size_t)par_slot[task] * 2 * tot + e, acc);
}

// ---------------------------------------------------------------- STEP 3
// STEP 3, grouped: a block owns one element range (one element per thread)
// and kM2lGroup consecutive targets (Morton neighbours, so their sources
// repeat and hit in L2).  Each thread forms the w_x w_y products and the
// three w_z of its element once and reuses them for every list entry of the
// group.  The 1-D phases are conj(z), 1, z for offsets 0, 1, 2, so the nine
// products w_x w_y are a b, a conj(b), a, b, 1 (a = z[k1], b = z[k2]) or
// their conjugates: five LDS rows, and a conjugate is an (exact) sign flip
// of the imaginary part -- 20 KB of LDS per block instead of 36 KB, six
// blocks per CU instead of four.  The group's entries (each target's own expansion first, then its
// list) form one stream; source loads go kM2lBatch at a time into three
// register batches, the batch two ahead issued before the current one is
// applied, so a wave keeps loads in flight across its whole stream.  The
// entry's offset o = ox + 3 oy + 9 oz is wave-uniform: it picks the w_x w_y
// row in LDS and its conjugation, and w_z by a uniform select (no branches).

// (ox, oy) -> row of s_fxy, 3 bits each at oxy = ox + 3 oy:
//   (0,0) ab*  (1,0) b*  (2,0) (a* b)*  (0,1) a*  (1,1) 1  (2,1) a  (0,2) a* b  (1,2) b  (2,2) ab
constexpr unsigned kM2lRow = (0u << 0) | (3u << 3) | (1u << 6) | (2u << 9) | (4u << 12) | (2u << 15) |
                             (1u << 18) | (3u << 21) | (0u << 24);
// conjugated at oxy = 0, 1, 2 (oy = 0) and 3 ((ox, oy) = (0, 1))
constexpr unsigned kM2lConj = 0x0fu;

// EPT elements per thread (1 or 2, consecutive): with 2 a source load is
// one 16-byte (FP32) or two 16-byte (FP64) accesses per lane, so a wave
// keeps twice the bytes in flight per entry; the problem's nd * nexp must
// then be even (every slot offset and element pair stays aligned and inside
// its slot).
template <typename ET, int EPT> struct M2lVec;
template <> struct M2lVec<float2, 1> {
    __device__ static void load(const float2* p, double2* v) { v[0] = ldE(*p); }
};
template <> struct M2lVec<double2, 1> {
    __device__ static void load(const double2* p, double2* v) { v[0] = *p; }
};
template <> struct M2lVec<float2, 2> {
    __device__ static void load(const float2* p, double2* v) {
        const float4 w = *reinterpret_cast<const float4*>(p);
        v[0] = make_double2(w.x, w.y);
        v[1] = make_double2(w.z, w.w);
    }
};
template <> struct M2lVec<double2, 2> {
    __device__ static void load(const double2* p, double2* v) {
        v[0] = p[0];
        v[1] = p[1];
    }
};

template <typename ET, int EPT>
__global__ void __launch_bounds__(kBlock) k_m2l27(Geom g, DevTables t, int ngroup, int group, long long zero_off,
                                                  const long long* __restrict__ toff,
                                                  const int* __restrict__ gent,
                                                  const long long* __restrict__ soff,
                                                  const int* __restrict__ scode,
                                                  ET* __restrict__ expn_a, Multi mu) {
    const int lin = xcd_task(blockIdx.x, gridDim.x);
    const int grp = lin % ngroup, chunk = lin / ngroup;
    ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[grp]];
        g = q.g;
        t.npw = q.g.npw;
        expn = static_cast<ET*>(q.expn);
        t.wm2l = q.wm2l;
        zero_off = q.zero_off;
    }
    const long long tot = (long long)g.nd * g.nexp;
    // the grid covers the class's largest problem: a block past this
    // problem's elements has nothing to do (uniform exit, before any barrier)
    if ((long long)chunk * blockDim.x * EPT >= tot) return;
    const int npw = t.npw;
    long long e = ((long long)chunk * blockDim.x + threadIdx.x) * EPT;
    const bool ok = e < tot;
    if (!ok) e = tot - EPT;   // valid duplicate, never stored
    // w_x w_y of each element in LDS (one column per thread and element,
    // read at a wave-uniform row), w_z in registers picked by a uniform
    // select.  Rows: 0 a b, 1 conj(a) b, 2 a, 3 b; row 4 (one) is a
    // constant.  Per (ox, oy) the row is kM2lRow and the conjugation
    // kM2lConj (a = w_x(2), b = w_y(2)).
    __shared__ double2 s_fxy[4][EPT * kBlock];
    const int tid = threadIdx.x;
    double2 fz0[EPT], fz1[EPT], fz2[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int r = (int)((e + q) % g.nexp);
        const int k2 = r % g.py, k1 = (r / g.py) % g.px, k3 = r / (g.py * g.px);
        const double2 a = g.px > 1 ? t.wm2l[2 * npw + k1] : make_double2(1.0, 0.0);
        const double2 b = g.py > 1 ? t.wm2l[2 * npw + k2] : make_double2(1.0, 0.0);
        s_fxy[0][q * kBlock + tid] = cmul(a, b);
        s_fxy[1][q * kBlock + tid] = cmul(conjd(a), b);
        s_fxy[2][q * kBlock + tid] = a;
        s_fxy[3][q * kBlock + tid] = b;
        fz0[q] = t.wm2l[k3];
        fz1[q] = t.wm2l[npw + k3];
        fz2[q] = t.wm2l[2 * npw + k3];
    }
    const ET* __restrict__ eb = expn + e;
    // the group's entry stream, staged in LDS: read back at a wave-uniform
    // address (broadcast), it keeps the vector-memory counter for the source
    // loads alone
    __shared__ long long s_soff[kM2lMaxEntries + 2 * kM2lBatch];
    __shared__ int s_code[kM2lMaxEntries];
    const int g0 = gent[grp], n = gent[grp + 1] - g0;   // a positive multiple of kM2lBatch
    for (int i = threadIdx.x; i < n + 2 * kM2lBatch; i += blockDim.x) {
        s_soff[i] = i < n ? soff[g0 + i] : zero_off;   // two batches of read-ahead past the end
        if (i < n) s_code[i] = scode[g0 + i];
    }
    __syncthreads();
    const int j0 = 0, j1 = n;
    const long long* __restrict__ to = toff + (size_t)grp * group;
    int ti = 0;
    double2 acc[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) acc[q] = make_double2(0.0, 0.0);
    double2 va[kM2lBatch][EPT], vb[kM2lBatch][EPT], vc[kM2lBatch][EPT];
#define M2L_LOAD(v, j)                                                                   \
    _Pragma("unroll") for (int u = 0; u < kM2lBatch; ++u) M2lVec<ET, EPT>::load(eb + s_soff[(j) + u], v[u]);
    // code = o | last-entry-of-its-target << 5
#define M2L_APPLY(v, j)                                                                  \
    _Pragma("unroll") for (int u = 0; u < kM2lBatch; ++u) {                              \
        const int code = __builtin_amdgcn_readfirstlane(s_code[(j) + u]);                \
        const int o = code & 31, oz = o / 9, oxy = o - 9 * oz;                           \
        const int row = (kM2lRow >> (3 * oxy)) & 7;                                      \
        const long long cj = (long long)((kM2lConj >> oxy) & 1) << 63;                   \
        _Pragma("unroll") for (int q = 0; q < EPT; ++q) {                                \
            const double2 zf = oz == 0 ? fz0[q] : (oz == 1 ? fz1[q] : fz2[q]);           \
            double2 fxy = s_fxy[row & 3][q * kBlock + tid];                              \
            if (row == 4) fxy = make_double2(1.0, 0.0);                                  \
            fxy.y = __longlong_as_double(__double_as_longlong(fxy.y) ^ cj);              \
            acc[q] = cfma(cmul(v[u][q], fxy), zf, acc[q]);                               \
        }                                                                                \
        if (code >> 5) {                                                                 \
            if (ok) {                                                                    \
                _Pragma("unroll") for (int q = 0; q < EPT; ++q)                          \
                    stE(expn + to[ti] + tot + e + q, acc[q]);                            \
            }                                                                            \
            ++ti;                                                                        \
            _Pragma("unroll") for (int q = 0; q < EPT; ++q) acc[q] = make_double2(0.0, 0.0); \
        }                                                                                \
    }
    // The batch two ahead is always loaded before the current one is
    // applied (past the end of the stream it reads the zero slot), so the
    // wait for a batch leaves exactly two batches in flight on every path.
    constexpr int B = kM2lBatch;
    M2L_LOAD(va, j0);
    M2L_LOAD(vb, j0 + B);
    for (int j = j0; j < j1; j += 3 * B) {
        M2L_LOAD(vc, j + 2 * B);
        M2L_APPLY(va, j);
        if (j + B >= j1) break;
        M2L_LOAD(va, j + 3 * B);
        M2L_APPLY(vb, j + B);
        if (j + 2 * B >= j1) break;
        M2L_LOAD(vb, j + 4 * B);
        M2L_APPLY(vc, j + 2 * B);
    }
#undef M2L_APPLY
#undef M2L_LOAD
}

// STEP 3 over sibling bricks (3-D): a block owns one element range and the
// (up to) eight cutoff-level children of one parent, whose M2L sources lie
// in the 4x4x4 boxes around them.  Each source is loaded once per brick and
// applied to every sibling whose list holds it: an entry's code carries the
// source position p (2 bits per axis, box offset + 1 from the brick's
// first child) and an 8-bit target mask; for target t at brick position
// tp the offset index is o = sum_a (tp_a + 2 - p_a) 3^a, the same phase
// selection as k_m2l27 (w_z from three more LDS rows).  Every target's own expansion is the entry at its
// own position (o = 13, unit phase).  toff: 8 loc offsets per brick (-1 =
// no target there); entries padded with zero-slot entries of mask 0.
template <typename ET>
__global__ void __launch_bounds__(kBlock) k_m2l_brick(Geom g, DevTables t, int nbrick, long long zero_off,
                                                      const long long* __restrict__ toff,
                                                      const int* __restrict__ gent,
                                                      const long long* __restrict__ soff,
                                                      const int* __restrict__ scode,
                                                      ET* __restrict__ expn_a, Multi mu) {
    const int lin = xcd_task(blockIdx.x, gridDim.x);
    const int brk = lin % nbrick, chunk = lin / nbrick;
    ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[brk]];
        g = q.g;
        t.npw = q.g.npw;
        expn = static_cast<ET*>(q.expn);
        t.wm2l = q.wm2l;
        zero_off = q.zero_off;
    }
    const long long tot = (long long)g.nd * g.nexp;
    if ((long long)chunk * blockDim.x >= tot) return;   // past this problem's elements
    const int npw = t.npw;
    long long e = (long long)chunk * blockDim.x + threadIdx.x;
    const bool ok = e < tot;
    if (!ok) e = tot - 1;   // valid duplicate, never stored
    const int r = (int)(e % g.nexp);
    const int k2 = r % g.py, k1 = (r / g.py) % g.px, k3 = r / (g.py * g.px);
    // rows 0-4: the w_x w_y products (row 4 = 1); row 5: c = w_z(2).  w_z
    // for oz = 0, 1, 2 is conj(c), 1, c (the table's rows are exactly that),
    // read at a uniform row (5 with a sign flip, 4, 5) instead of selected
    // from registers per pair (STEP 3 at a norder-16 periodic problem 0.745
    // -> 0.696 ms); six rows rather than eight leave room for six blocks per
    // CU in LDS instead of four
    __shared__ double2 s_fxy[6][kBlock];
    const int tid = threadIdx.x;
    {
        const double2 a = t.wm2l[2 * npw + k1];
        const double2 b = t.wm2l[2 * npw + k2];
        s_fxy[0][tid] = cmul(a, b);
        s_fxy[1][tid] = cmul(conjd(a), b);
        s_fxy[2][tid] = a;
        s_fxy[3][tid] = b;
        s_fxy[4][tid] = make_double2(1.0, 0.0);
        s_fxy[5][tid] = t.wm2l[2 * npw + k3];
    }
    const ET* __restrict__ eb = expn + e;
    __shared__ long long s_soff[kM2lBrickMax + 2 * kM2lBatch];
    __shared__ int s_code[kM2lBrickMax];
    const int g0 = gent[brk], n = gent[brk + 1] - g0;   // a positive multiple of kM2lBatch
    for (int i = threadIdx.x; i < n + 2 * kM2lBatch; i += blockDim.x) {
        s_soff[i] = i < n ? soff[g0 + i] : zero_off;
        if (i < n) s_code[i] = scode[g0 + i];
    }
    __syncthreads();
    double2 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = make_double2(0.0, 0.0);
    double2 va[kM2lBatch], vb[kM2lBatch], vc[kM2lBatch];
#define M2L_LOAD(v, j)                                                                   \
    _Pragma("unroll") for (int u = 0; u < kM2lBatch; ++u) v[u] = ldE(eb[s_soff[(j) + u]]);
#define M2L_APPLY(v, j)                                                                  \
    _Pragma("unroll") for (int u = 0; u < kM2lBatch; ++u) {                              \
        const int code = __builtin_amdgcn_readfirstlane(s_code[(j) + u]);                \
        const int px = code & 3, py = (code >> 2) & 3, pz = (code >> 4) & 3;            \
        const int mask = (code >> 6) & 255;                                              \
        _Pragma("unroll") for (int q = 0; q < 8; ++q) {                                  \
            if ((mask >> q) & 1) {                                                       \
                const int ox = (q & 1) + 2 - px, oy = ((q >> 1) & 1) + 2 - py;           \
                const int oz = ((q >> 2) & 1) + 2 - pz, oxy = ox + 3 * oy;               \
                double2 zf = s_fxy[oz == 1 ? 4 : 5][tid];                                \
                zf.y = __longlong_as_double(__double_as_longlong(zf.y) ^                 \
                                            ((long long)(oz == 0) << 63));               \
                double2 fxy = s_fxy[(kM2lRow >> (3 * oxy)) & 7][tid];                    \
                fxy.y = __longlong_as_double(__double_as_longlong(fxy.y) ^               \
                                             ((long long)((kM2lConj >> oxy) & 1) << 63)); \
                acc[q] = cfma(cmul(v[u], fxy), zf, acc[q]);                              \
            }                                                                            \
        }                                                                                \
    }
    constexpr int B = kM2lBatch;
    M2L_LOAD(va, 0);
    M2L_LOAD(vb, B);
    for (int j = 0; j < n; j += 3 * B) {
        M2L_LOAD(vc, j + 2 * B);
        M2L_APPLY(va, j);
        if (j + B >= n) break;
        M2L_LOAD(va, j + 3 * B);
        M2L_APPLY(vb, j + B);
        if (j + 2 * B >= n) break;
        M2L_LOAD(vb, j + 4 * B);
        M2L_APPLY(vc, j + 2 * B);
    }
#undef M2L_APPLY
#undef M2L_LOAD
    if (ok) {
        const long long* __restrict__ to = toff + (size_t)brk * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long long o = to[q];
            if (o >= 0) stE(expn + o + tot + e, acc[q]);
        }
    }
}

// STEP 3 over a sibling brick as a separable stencil (see launch_m2l_sep).
// Source grid positions p = (px, py, pz) in 0..3, siblings q in 0..1 per
// axis at p = q + 1; sibling q takes the source at p with the offset
// o_a = q_a + 2 - p_a in 0..2.  Rows (py, pz) of four sources go x first:
//   X_qx = w(2) v(qx) + v(qx + 1) + conj(w(2)) v(qx + 2)
// then into the plane sums Y(qx, qy) += w_y(qy + 2 - py) X_qx, and after a
// plane's four rows into acc(q) += w_z(qz + 2 - pz) Y(qx, qy).  Every index
// is a compile-time constant of the unrolled loops, so each phase is a
// register (no selects).  Rows without a source are skipped (uniform
// branches); loads run two rows ahead of the arithmetic.
//
// A leaf target's list leaves out the leaf sources other than itself (their
// pairs are near field).  The first sweep applies every source to every
// sibling; a second sweep over the leaf sources alone, with the values
// negated, goes to the leaf targets only, which then take their own
// expansion back.  Entry code bits: 0-5 position, 6-13 target mask, 14 the
// source is a leaf, 15 the target at this position is a leaf.
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
// w u + conj(w) v
__device__ __forceinline__ double2 csym(double2 w, double2 u, double2 v) {
    const double sx = u.x + v.x, sy = u.y + v.y, dx = u.x - v.x, dy = u.y - v.y;
    return make_double2(w.x * sx - w.y * dy, w.x * sy + w.y * dx);
}
// acc += w^(O) x with w^(2) = w, w^(1) = 1, w^(0) = conj(w)
template <int O>
__device__ __forceinline__ void tap(double2& acc, const double2 w, const double2 x) {
    if constexpr (O == 1) {
        acc.x += x.x;
        acc.y += x.y;
    } else if constexpr (O == 2) {
        acc = cfma(x, w, acc);
    } else {
        acc = cfma(x, conjd(w), acc);
    }
}

template <typename ET>
__global__ void __launch_bounds__(kBlock) k_m2l_sep(Geom g, DevTables t, int nbrick, long long zero_off,
                                                    const long long* __restrict__ toff,
                                                    const int* __restrict__ gent,
                                                    const long long* __restrict__ soff,
                                                    const int* __restrict__ scode,
                                                    ET* __restrict__ expn_a, Multi mu) {
    const int lin = xcd_task(blockIdx.x, gridDim.x);
    const int brk = lin % nbrick, chunk = lin / nbrick;
    ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[brk]];
        g = q.g;
        t.npw = q.g.npw;
        expn = static_cast<ET*>(q.expn);
        t.wm2l = q.wm2l;
        zero_off = q.zero_off;
    }
    const long long tot = (long long)g.nd * g.nexp;
    if ((long long)chunk * blockDim.x >= tot) return;   // past this problem's elements
    const int npw = t.npw;
    long long e = (long long)chunk * blockDim.x + threadIdx.x;
    const bool ok = e < tot;
    if (!ok) e = tot - 1;   // valid duplicate, never stored
    const int r = (int)(e % g.nexp);
    const int k2 = r % g.py, k1 = (r / g.py) % g.px, k3 = r / (g.py * g.px);
    const double2 wx = t.wm2l[2 * npw + k1], wy = t.wm2l[2 * npw + k2], wz = t.wm2l[2 * npw + k3];
    // the brick's entries as a dense 4x4x4 grid of source offsets, with
    // 4-bit presence masks per row (py, pz) for all and for leaf sources,
    // and the mask of leaf targets
    __shared__ long long s_poff[64];
    __shared__ int s_rows[2][16];
    __shared__ int s_leaft;
    const int tid = threadIdx.x;
    if (tid < 64) s_poff[tid] = zero_off;
    if (tid < 32) s_rows[tid >> 4][tid & 15] = 0;
    if (tid == 0) s_leaft = 0;
    __syncthreads();
    const int g0 = gent[brk], n = gent[brk + 1] - g0;
    for (int i = tid; i < n; i += blockDim.x) {
        const int c = scode[g0 + i];
        if ((c >> 6) & 255) {   // padding entries carry no target
            const int pos = c & 63;
            s_poff[pos] = soff[g0 + i];
            atomicOr(&s_rows[0][pos >> 2], 1 << (pos & 3));
            if ((c >> 14) & 1) atomicOr(&s_rows[1][pos >> 2], 1 << (pos & 3));
            if ((c >> 15) & 1) {
                const int qx = (pos & 3) - 1, qy = ((pos >> 2) & 3) - 1, qz = (pos >> 4) - 1;
                atomicOr(&s_leaft, 1 << (qx + 2 * qy + 4 * qz));
            }
        }
    }
    __syncthreads();
    const int leaft = __builtin_amdgcn_readfirstlane(s_leaft);
    const ET* __restrict__ eb = expn + e;
    double2 acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = make_double2(0.0, 0.0);
    // one sweep: LEAF = 0 every source to every sibling; LEAF = 1 the leaf
    // sources, negated, to the leaf targets
    auto sweep = [&](auto leaf_tag) {
        constexpr int LEAF = decltype(leaf_tag)::value;
        double2 Y[4];
        ET v[3][4];
        int rb[3];
        auto load_row = [&](int row, int buf) {
            const int m = __builtin_amdgcn_readfirstlane(s_rows[LEAF][row]);
            rb[buf] = m;
            if (m) {
#pragma unroll
                for (int px = 0; px < 4; ++px)
                    if (!LEAF || ((m >> px) & 1)) v[buf][px] = eb[s_poff[4 * row + px]];
            }
        };
        load_row(0, 0);
        load_row(1, 1);
#pragma unroll
        for (int row = 0; row < 16; ++row) {
            const int py = row & 3, pz = row >> 2;
            if (row + 2 < 16) load_row(row + 2, (row + 2) % 3);
            if (py == 0) {
#pragma unroll
                for (int k = 0; k < 4; ++k) Y[k] = make_double2(0.0, 0.0);
            }
            const int b = row % 3;
            const int m = rb[b];
            if (m) {
                double2 u[4];
#pragma unroll
                for (int px = 0; px < 4; ++px) {
                    u[px] = ldE(v[b][px]);
                    if (LEAF) {
                        if ((m >> px) & 1) u[px] = make_double2(-u[px].x, -u[px].y);
                        else u[px] = make_double2(0.0, 0.0);
                    }
                }
                double2 X[2];
                X[0] = cadd(u[1], csym(wx, u[0], u[2]));
                X[1] = cadd(u[2], csym(wx, u[1], u[3]));
                // Y(qx, qy) at k = qx + 2 qy, offset oy = qy + 2 - py
#pragma unroll
                for (int qx = 0; qx < 2; ++qx) {
                    switch (py) {
                        case 0: tap<2>(Y[qx], wy, X[qx]); break;
                        case 1: tap<1>(Y[qx], wy, X[qx]); tap<2>(Y[qx + 2], wy, X[qx]); break;
                        case 2: tap<0>(Y[qx], wy, X[qx]); tap<1>(Y[qx + 2], wy, X[qx]); break;
                        default: tap<0>(Y[qx + 2], wy, X[qx]); break;
                    }
                }
            }
            if (py == 3) {
                // acc(q) at q = k + 4 qz, offset oz = qz + 2 - pz
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const bool lo = !LEAF || ((leaft >> k) & 1), hi = !LEAF || ((leaft >> (k + 4)) & 1);
                    switch (pz) {
                        case 0: if (lo) tap<2>(acc[k], wz, Y[k]); break;
                        case 1:
                            if (lo) tap<1>(acc[k], wz, Y[k]);
                            if (hi) tap<2>(acc[k + 4], wz, Y[k]);
                            break;
                        case 2:
                            if (lo) tap<0>(acc[k], wz, Y[k]);
                            if (hi) tap<1>(acc[k + 4], wz, Y[k]);
                            break;
                        default: if (hi) tap<0>(acc[k + 4], wz, Y[k]); break;
                    }
                }
            }
        }
    };
    sweep(std::integral_constant<int, 0>{});
    if (leaft) {
        sweep(std::integral_constant<int, 1>{});
        // the leaf targets' own expansions, taken out with the leaf sources
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if ((leaft >> q) & 1) {
                const double2 own = ldE(eb[s_poff[((q & 1) + 1) + 4 * (((q >> 1) & 1) + 1) + 16 * ((q >> 2) + 1)]]);
                acc[q] = cadd(acc[q], own);
            }
    }
    if (ok) {
        const long long* __restrict__ to = toff + (size_t)brk * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const long long o = to[q];
            if (o >= 0) stE(expn + o + tot + e, acc[q]);
        }
    }
}

// ---------------------------------------------------------------- STEP 4
// Each thread takes kSplitEPT elements (all their loads issued before the
// stores): a quarter of the blocks of one element per thread, where most
// blocks of non-first entries only exit.
constexpr int kSplitEPT = 4;

template <typename ET>
__global__ void __launch_bounds__(kBlock) k_split(Geom g, DevTables t, int nentry, const int* __restrict__ pc_parent,
                                                  const int* __restrict__ pc_child,
                                                  const int* __restrict__ pc_sign, int klev,
                                                  ET* __restrict__ expn_a, Multi mu) {
    // the entries of one parent are consecutive: the block of the first
    // reads the parent's elements once and writes every child of the run;
    // the blocks of the others have nothing to do
    const int entry = blockIdx.y;
    auto same = [&](int a, int b) {
        return pc_parent[a] == pc_parent[b] && (!mu.pp || mu.tprob[a] == mu.tprob[b]);
    };
    if (entry > 0 && same(entry - 1, entry)) return;
    ET* __restrict__ expn = expn_a;
    const double2* w1 = t.wms + (size_t)(klev - 1) * t.npw;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[entry]];
        g = q.g;
        expn = static_cast<ET*>(q.expn);
        w1 = q.wms + (size_t)(mu.tklev[entry] - 1) * q.g.npw;
    }
    const long long tot = (long long)g.nd * g.nexp;
    const long long e0 = (long long)blockIdx.x * kSplitEPT * blockDim.x + threadIdx.x;
    if (e0 >= tot) return;
    const ET* __restrict__ src = expn + ((size_t)pc_parent[entry] * 2 + 1) * tot;
    double2 v[kSplitEPT], pa[kSplitEPT], pb[kSplitEPT], wz[kSplitEPT];
#pragma unroll
    for (int k = 0; k < kSplitEPT; ++k) {
        const long long e = e0 + (long long)k * blockDim.x;
        const long long ec = e < tot ? e : tot - 1;   // valid duplicate, never stored
        v[k] = ldE(src[ec]);
        const int r = (int)(ec % g.nexp);
        const int k2 = r % g.py, k1 = (r / g.py) % g.px, k3 = r / (g.py * g.px);
        // ms_phase per child from the two x-y products (a conjugate is exact)
        const double2 wx = g.px > 1 ? w1[k1] : make_double2(1.0, 0.0);
        const double2 wy = g.py > 1 ? w1[k2] : make_double2(1.0, 0.0);
        wz[k] = w1[k3];
        pa[k] = cmul(wx, wy);
        pb[k] = cmul(conjd(wx), wy);
    }
    for (int j = entry; j < nentry && (j == entry || same(entry, j)); ++j) {
        const int sg = pc_sign[j];
        ET* __restrict__ dst = expn + ((size_t)pc_child[j] * 2 + 1) * tot;
#pragma unroll
        for (int k = 0; k < kSplitEPT; ++k) {
            const long long e = e0 + (long long)k * blockDim.x;
            const double2 xy = (sg & 1) ? ((sg & 2) ? conjd(pa[k]) : pb[k]) : ((sg & 2) ? conjd(pb[k]) : pa[k]);
            const double2 ph = cmul(xy, (sg & 4) ? conjd(wz[k]) : wz[k]);
            if (e < tot) stE(dst + e, cmul(v[k], ph));
        }
    }
}

// ---------------------------------------------------------------- STEP 5, 2-D
// g2d_pw2pot (bfgt_pwrouts.f:473-526): pot[kk1][kk2] += 2 Re sum_{k1, k3}
// Ux[kk1][k1] Uz[kk2][k3] loc[k3][k1] over the half axis k3 (the last one
// halved when npw is odd).  A 2-D leaf's expansion is small (npw x nh
// complex), so the block stages it in LDS with coalesced loads and spreads
// both contractions over all its threads: x with one thread per (kk1, k3),
// then the half axis with one thread per output node.  (The generic kernel
// gave the half axis one barrier-separated iteration each with n active
// threads; 2-D problems ran ~40x slower per node than 3-D ones.)
// LDS: sE [k3][k1] (nh*npw complex), sff [kk1][k3] (n*nh complex).
__global__ void __launch_bounds__(kBlock) k_pw2pot2(Geom g, DevTables t, const int* __restrict__ box,
                                                    const int* __restrict__ lev, const int* __restrict__ slot,
                                                    const double2* __restrict__ expn, Kinds kd, OutDesc od) {
    extern __shared__ double2 sm[];
    const int task = blockIdx.x;
    const int b = box[task], l = lev[task], sl = slot[task];
    const int n = t.n, npw = t.npw, nh = g.pz;
    double2* sE = sm;
    double2* sff = sE + (size_t)nh * npw;
    const size_t lo = (size_t)l * n * npw;
    const double2* __restrict__ Ux = t.pw2pot + kd.x * t.pw2pot_ks + lo;   // [i][k]
    const double2* __restrict__ Uz = t.pw2pot + kd.z * t.pw2pot_ks + lo;
    const bool halve_last = (npw % 2) == 1;
    const int tid = threadIdx.x;
    for (int ind = 0; ind < g.nd; ++ind) {
        const double2* __restrict__ E = expn + ((size_t)sl * 2 * g.nd + g.nd + ind) * g.nexp;
        for (int i = tid; i < nh * npw; i += blockDim.x) sE[i] = E[i];
        __syncthreads();
        for (int o = tid; o < n * nh; o += blockDim.x) {
            const int kk1 = o % n, k3 = o / n;
            const double2* __restrict__ u = Ux + (size_t)kk1 * npw;
            const double2* __restrict__ e = sE + (size_t)k3 * npw;
            double2 v0 = make_double2(0.0, 0.0), v1 = make_double2(0.0, 0.0);
            int k1 = 0;
            for (; k1 + 1 < npw; k1 += 2) {
                v0 = cfma(u[k1], e[k1], v0);
                v1 = cfma(u[k1 + 1], e[k1 + 1], v1);
            }
            if (k1 < npw) v0 = cfma(u[k1], e[k1], v0);
            sff[kk1 * nh + k3] = make_double2(v0.x + v1.x, v0.y + v1.y);
        }
        __syncthreads();
        for (int o = tid; o < n * n; o += blockDim.x) {
            const int kk1 = o % n, kk2 = o / n;
            const double2* __restrict__ u = Uz + (size_t)kk2 * npw;
            const double2* __restrict__ f = sff + (size_t)kk1 * nh;
            double acc = 0.0;
            for (int k3 = 0; k3 < nh; ++k3) {
                const double w = (halve_last && k3 == nh - 1) ? 0.5 : 1.0;
                acc += (u[k3].x * f[k3].x - u[k3].y * f[k3].y) * w;
            }
            od.p[((size_t)b * g.npbox + kk1 + n * kk2) * od.ns + od.coff + ind] += 2.0 * acc;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- STEP 5
// Tables per axis by kind (kd); the z table is staged in LDS, the x/y
// tables come from LDS when they are the same kind, else from global memory.
__global__ void __launch_bounds__(kBlock) k_pw2pot(Geom g, DevTables t, const int* __restrict__ box,
                                                   const int* __restrict__ lev,
                                                   const int* __restrict__ slot,
                                                   const double2* __restrict__ expn,
                                                   Kinds kd, OutDesc od) {
    extern __shared__ double2 sm[];
    const int task = blockIdx.x;
    const int b = box[task], l = lev[task], s = slot[task];
    const int n = t.n, npw = t.npw;
    double2* sU = sm;                 // [n][npw]
    double2* ff = sU + n * npw;       // [kk1][k2]  nx*py
    const size_t lo = (size_t)l * n * npw;
    const double2* U = t.pw2pot + kd.z * t.pw2pot_ks + lo;
    for (int i = threadIdx.x; i < n * npw; i += blockDim.x) sU[i] = U[i];
    const double2* Ux = kd.x == kd.z ? sU : t.pw2pot + kd.x * t.pw2pot_ks + lo;
    const double2* Uy = kd.y == kd.z ? sU : t.pw2pot + kd.y * t.pw2pot_ks + lo;
    const int nxy = g.nx * g.ny;
    const int kk1 = threadIdx.x % g.nx, kk2 = threadIdx.x / g.nx;
    const bool halve_last = (npw % 2) == 1;
    for (int ind = 0; ind < g.nd; ++ind) {
        double acc[16];
        for (int i = 0; i < 16; ++i) acc[i] = 0.0;
        const double2* E = expn + ((size_t)s * 2 * g.nd + g.nd + ind) * g.nexp;
        for (int k3 = 0; k3 < g.pz; ++k3) {
            __syncthreads();
            const double2* Ek = E + (size_t)k3 * g.px * g.py;
            for (int o = threadIdx.x; o < g.nx * g.py; o += blockDim.x) {
                const int k2 = o % g.py, a1 = o / g.py;
                double2 v = make_double2(0.0, 0.0);
                if (g.nx == 1) {
                    v = Ek[k2];
                } else {
                    for (int k1 = 0; k1 < g.px; ++k1) v = cfma(Ux[a1 * npw + k1], Ek[(size_t)k1 * g.py + k2], v);
                }
                ff[a1 * g.py + k2] = v;
            }
            __syncthreads();
            if ((int)threadIdx.x < nxy) {
                double2 v = make_double2(0.0, 0.0);
                if (g.ny == 1) {
                    v = ff[kk1 * g.py];
                } else {
                    for (int k2 = 0; k2 < g.py; ++k2) v = cfma(Uy[kk2 * npw + k2], ff[kk1 * g.py + k2], v);
                }
                const double wgt = (halve_last && k3 == g.pz - 1) ? 0.5 : 1.0;
                for (int kk3 = 0; kk3 < g.nz; ++kk3) {
                    const double2 u = sU[kk3 * npw + k3];
                    acc[kk3] += (u.x * v.x - u.y * v.y) * wgt;
                }
            }
        }
        if ((int)threadIdx.x < nxy)
            for (int kk3 = 0; kk3 < g.nz; ++kk3)
                od.p[((size_t)b * g.npbox + kk1 + g.nx * (kk2 + g.ny * kk3)) * od.ns + od.coff + ind] +=
                    2.0 * acc[kk3];
    }
}

// ---------------------------------------------------------------- STEP 6
// One block per target leaf; sources of its list are processed in order and
// accumulated in registers; per source the contraction runs x, y (per z
// slab, through LDS) and z (in registers).
__global__ void __launch_bounds__(kBlock) k_direct(Geom g, DevTables t, const int* __restrict__ tgt,
                                                   const int* __restrict__ lev,
                                                   const int* __restrict__ off,
                                                   const int* __restrict__ src,
                                                   const int* __restrict__ code,
                                                   const double* __restrict__ fvals,
                                                   Kinds kd, OutDesc od) {
    extern __shared__ double dsm[];
    const int task = blockIdx.x;
    const int tb = tgt[task], l = lev[task];
    const int n = t.n;
    const int nxy = g.nx * g.ny;
    double* sf = dsm;                  // npbox
    double* h = sf + g.npbox;          // [j2][k1]  nx*ny
    double* sA = h + nxy;              // 3 tables [k][j]  n*n each
    int* sB = (int*)(sA + 3 * n * n);  // 3 bands [k][2]
    const size_t nn = (size_t)n * n;
    const int kax[3] = {kd.x, kd.y, kd.z};
    const int* bandl = t.locband + (size_t)l * t.loc_slots * 2 * n;
    const int k1 = threadIdx.x % g.nx, k2 = threadIdx.x / g.nx;
    for (int ind = 0; ind < g.nd; ++ind) {
        double acc[16];
        for (int i = 0; i < 16; ++i) acc[i] = 0.0;
        for (int j = off[task]; j < off[task + 1]; ++j) {
            const int sb = src[j], c = code[j];
            __syncthreads();
            for (int i = threadIdx.x; i < g.npbox; i += blockDim.x)
                sf[i] = fvals[((size_t)sb * g.npbox + i) * g.nd + ind];
            for (int a = 0; a < 3; ++a) {
                const int slot = (c >> (8 * a)) & 255;
                const double* A = t.loctab + kax[a] * t.loctab_ks + ((size_t)l * t.loc_slots + slot) * nn;
                const int* B = bandl + (size_t)slot * 2 * n;
                for (int i = threadIdx.x; i < n * n; i += blockDim.x) sA[a * nn + i] = A[i];
                for (int i = threadIdx.x; i < 2 * n; i += blockDim.x) sB[a * 2 * n + i] = B[i];
            }
            __syncthreads();
            const double* Ax = sA;
            const double* Ay = sA + nn;
            const double* Az = sA + 2 * nn;
            for (int j3 = 0; j3 < g.nz; ++j3) {
                for (int o = threadIdx.x; o < nxy; o += blockDim.x) {
                    const int a1 = o % g.nx, j2 = o / g.nx;
                    double v = 0.0;
                    const double* fs = sf + g.nx * (j2 + g.ny * j3);
                    if (g.nx == 1) {
                        v = fs[0];
                    } else {
                        const int lo = sB[2 * a1], hi = sB[2 * a1 + 1];
                        for (int j1 = lo; j1 <= hi; ++j1) v += Ax[a1 * n + j1] * fs[j1];
                    }
                    h[j2 * g.nx + a1] = v;
                }
                __syncthreads();
                if ((int)threadIdx.x < nxy) {
                    double v = 0.0;
                    if (g.ny == 1) {
                        v = h[k1];
                    } else {
                        const int lo = sB[2 * n + 2 * k2], hi = sB[2 * n + 2 * k2 + 1];
                        for (int j2 = lo; j2 <= hi; ++j2) v += Ay[k2 * n + j2] * h[j2 * g.nx + k1];
                    }
                    for (int k3 = 0; k3 < g.nz; ++k3) acc[k3] += Az[k3 * n + j3] * v;
                }
                __syncthreads();
            }
        }
        if ((int)threadIdx.x < nxy)
            for (int k3 = 0; k3 < g.nz; ++k3)
                od.p[((size_t)tb * g.npbox + k1 + g.nx * (k2 + g.ny * k3)) * od.ns + od.coff + ind] += acc[k3];
    }
}

// STEP 6 for 2-D (g2d_tens_prod_to_potloc_fast, bfgt_local.f:1230-1315): per source,
// h[k1][j3] = sum_j1 Ax[k1][j1] f[j1][j3] over the x band, then
// pot[k1][k3] += sum_j3 Az[k3][j3] h[k1][j3], each with one thread per
// (k1, j3) / (k1, k3) pair; the n x n outputs stay in registers across the
// target's sources.  Two barriers per source instead of two per source row.
// LDS: sf [j3][j1] (n*n), h [j3][k1] (n*n), Ax, Az (n*n each), x band (2n).
__global__ void __launch_bounds__(kBlock) k_direct2(Geom g, DevTables t, const int* __restrict__ tgt,
                                                    const int* __restrict__ lev, const int* __restrict__ off,
                                                    const int* __restrict__ src, const int* __restrict__ code,
                                                    const double* __restrict__ fvals, Kinds kd, OutDesc od) {
    extern __shared__ double dsm[];
    const int task = blockIdx.x;
    const int tb = tgt[task], l = lev[task];
    const int n = t.n;
    const size_t nn = (size_t)n * n;
    double* sf = dsm;
    double* h = sf + nn;
    double* Ax = h + nn;
    double* Az = Ax + nn;
    int* band = (int*)(Az + nn);
    const int* bandl = t.locband + (size_t)l * t.loc_slots * 2 * n;
    const int tid = threadIdx.x;
    const bool act = tid < (int)nn;
    const int r = tid % n, c = tid / n;   // (k1, j3) in the first stage, (k1, k3) in the second
    for (int ind = 0; ind < g.nd; ++ind) {
        double acc = 0.0;
        for (int j = off[task]; j < off[task + 1]; ++j) {
            const int sb = src[j], cd = code[j];
            const int slx = cd & 255, slz = (cd >> 16) & 255;
            __syncthreads();
            const double* __restrict__ A = t.loctab + kd.x * t.loctab_ks + ((size_t)l * t.loc_slots + slx) * nn;
            const double* __restrict__ Z = t.loctab + kd.z * t.loctab_ks + ((size_t)l * t.loc_slots + slz) * nn;
            if (act) {
                sf[tid] = fvals[((size_t)sb * g.npbox + tid) * g.nd + ind];
                Ax[tid] = A[tid];
                Az[tid] = Z[tid];
            }
            if (tid < 2 * n) band[tid] = bandl[(size_t)slx * 2 * n + tid];
            __syncthreads();
            if (act) {
                const int lo = band[2 * r], hi = band[2 * r + 1];
                const double* __restrict__ fs = sf + (size_t)n * c;
                double v = 0.0;
                for (int j1 = lo; j1 <= hi; ++j1) v += Ax[r * n + j1] * fs[j1];
                h[c * n + r] = v;
            }
            __syncthreads();
            if (act) {
                const double* __restrict__ az = Az + (size_t)c * n;
                for (int j3 = 0; j3 < n; ++j3) acc += az[j3] * h[j3 * n + r];
            }
        }
        if (act) od.p[((size_t)tb * g.npbox + tid) * od.ns + od.coff + ind] += acc;
    }
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
        throw std::runtime_error(buf);
    }
}

unsigned blocks_for(long long tot) { return (unsigned)((tot + kBlock - 1) / kBlock); }

}  // namespace

void launch_p2pw(const Geom& g, const DevTables& t, int ntask, const int* box, const int* lev,
                 const int* slot, const double* fvals, double2* expn, hipStream_t s, const Multi& mu) {
    if (ntask <= 0) return;
    if (launch_p2pw_3d(g, t, ntask, box, lev, slot, fvals, expn, s, t.lds_limit, mu)) {
        check(hipGetLastError(), "k_p2pw3 launch");
        return;
    }
    if (t.expn_f32) throw std::runtime_error("FP32 expansions need the 3-D register-blocked kernels");
    if (g.ndim == 2) {
        const size_t lds2 = sizeof(double2) * ((size_t)t.npw * t.n + (size_t)t.n * g.pz) + sizeof(double) * t.n * t.n;
        if (lds2 <= t.lds_limit) {
            hipLaunchKernelGGL(k_p2pw2, dim3(ntask), dim3(kBlock), lds2, s, g, t, box, lev, slot, fvals, expn);
            check(hipGetLastError(), "k_p2pw2 launch");
            return;
        }
    }
    const size_t lds = sizeof(double2) * ((size_t)t.npw * t.n + (size_t)g.nx * g.ny + (size_t)g.py * g.nx);
    hipLaunchKernelGGL(k_p2pw, dim3(ntask), dim3(kBlock), lds, s, g, t, box, lev, slot, fvals, expn);
    check(hipGetLastError(), "k_p2pw launch");
}

void launch_merge(const Geom& g, const DevTables& t, int ntask, const int* par_slot,
                  const int* par_first, const int* par_count, const int* pc_child, const int* pc_sign,
                  const int* par_acc, int klev, double2* expn, hipStream_t s, const Multi& mu) {
    // tasks go on grid y, which holds at most kMaxGridY rows: longer lists
    // are launched in chunks
    for (int a = 0; a < ntask; a += kMaxGridY) {
        const int cnt = std::min(kMaxGridY, ntask - a);
        const dim3 grid(blocks_for((long long)g.nd * g.nexp), cnt);
        const int* acc = par_acc ? par_acc + a : nullptr;
        Multi ma = mu;
        if (ma.pp) ma.tprob += a, ma.tklev += a;
        if (t.expn_f32)
            hipLaunchKernelGGL(k_merge<float2>, grid, dim3(kBlock), 0, s, g, t, par_slot + a, par_first + a,
                               par_count + a, pc_child, pc_sign, acc, klev, reinterpret_cast<float2*>(expn), ma);
        else
            hipLaunchKernelGGL(k_merge<double2>, grid, dim3(kBlock), 0, s, g, t, par_slot + a, par_first + a,
                               par_count + a, pc_child, pc_sign, acc, klev, expn, ma);
        check(hipGetLastError(), "k_merge launch");
    }
}

int m2l_group() {
    static const int v = [] {
        const char* e = std::getenv("FGT_M2L_GROUP");
        const int g = e ? std::atoi(e) : 0;
        return g >= 1 && g <= kM2lGroupMax ? g : kM2lGroup;
    }();
    return v;
}

void launch_m2l_brick(const Geom& g, const DevTables& t, int nbrick, long long zero_off, const long long* toff,
                      const int* gent, const long long* soff, const int* scode, double2* expn, hipStream_t s,
                      const Multi& mu) {
    if (nbrick <= 0) return;
    if (g.ndim != 3) throw std::runtime_error("k_m2l_brick: 3-D only");
    const long long nblk = (long long)blocks_for((long long)g.nd * g.nexp) * nbrick;
    if (nblk > 0x7fffffffLL) throw std::runtime_error("k_m2l_brick: grid too large");
    if (t.expn_f32)
        hipLaunchKernelGGL(k_m2l_brick<float2>, dim3((unsigned)nblk), dim3(kBlock), 0, s, g, t, nbrick, zero_off,
                           toff, gent, soff, scode, reinterpret_cast<float2*>(expn), mu);
    else
        hipLaunchKernelGGL(k_m2l_brick<double2>, dim3((unsigned)nblk), dim3(kBlock), 0, s, g, t, nbrick, zero_off,
                           toff, gent, soff, scode, expn, mu);
    check(hipGetLastError(), "k_m2l_brick launch");
}

void launch_m2l_sep(const Geom& g, const DevTables& t, int nbrick, long long zero_off, const long long* toff,
                    const int* gent, const long long* soff, const int* scode, double2* expn, hipStream_t s,
                    const Multi& mu) {
    if (nbrick <= 0) return;
    if (g.ndim != 3) throw std::runtime_error("k_m2l_sep: 3-D only");
    const long long nblk = (long long)blocks_for((long long)g.nd * g.nexp) * nbrick;
    if (nblk > 0x7fffffffLL) throw std::runtime_error("k_m2l_sep: grid too large");
    if (t.expn_f32)
        hipLaunchKernelGGL(k_m2l_sep<float2>, dim3((unsigned)nblk), dim3(kBlock), 0, s, g, t, nbrick, zero_off,
                           toff, gent, soff, scode, reinterpret_cast<float2*>(expn), mu);
    else
        hipLaunchKernelGGL(k_m2l_sep<double2>, dim3((unsigned)nblk), dim3(kBlock), 0, s, g, t, nbrick, zero_off,
                           toff, gent, soff, scode, expn, mu);
    check(hipGetLastError(), "k_m2l_sep launch");
}

int m2l_brick_env() {
    static const int v = [] {
        const char* e = std::getenv("FGT_M2L_BRICK");
        return e ? std::atoi(e) : -1;
    }();
    return v;
}

int m2l_ept_env() {
    static const int v = [] {
        const char* e = std::getenv("FGT_M2L_EPT");
        const int x = e ? std::atoi(e) : 0;
        return x == 1 || x == 2 ? x : 0;
    }();
    return v;
}

void launch_m2l27(const Geom& g, const DevTables& t, int ngroup, int group, long long zero_off,
                  const long long* toff, const int* gent, const long long* soff, const int* scode, double2* expn,
                  hipStream_t s, const Multi& mu, int ept) {
    if (ngroup <= 0) return;
    if (ept != 1 && ept != 2) throw std::runtime_error("k_m2l27: elements per thread must be 1 or 2");
    const long long tot = (long long)g.nd * g.nexp;
    if (ept == 2 && !mu.pp && (tot & 1)) throw std::runtime_error("k_m2l27: two elements per thread need even nd*nexp");
    const long long nblk = (long long)blocks_for((tot + ept - 1) / ept) * ngroup;
    if (nblk > 0x7fffffffLL) throw std::runtime_error("k_m2l27: grid too large");
    const dim3 grid((unsigned)nblk);
    if (t.expn_f32) {
        float2* x = reinterpret_cast<float2*>(expn);
        if (ept == 2)
            hipLaunchKernelGGL((k_m2l27<float2, 2>), grid, dim3(kBlock), 0, s, g, t, ngroup, group, zero_off, toff,
                               gent, soff, scode, x, mu);
        else
            hipLaunchKernelGGL((k_m2l27<float2, 1>), grid, dim3(kBlock), 0, s, g, t, ngroup, group, zero_off, toff,
                               gent, soff, scode, x, mu);
    } else {
        if (ept == 2)
            hipLaunchKernelGGL((k_m2l27<double2, 2>), grid, dim3(kBlock), 0, s, g, t, ngroup, group, zero_off, toff,
                               gent, soff, scode, expn, mu);
        else
            hipLaunchKernelGGL((k_m2l27<double2, 1>), grid, dim3(kBlock), 0, s, g, t, ngroup, group, zero_off, toff,
                               gent, soff, scode, expn, mu);
    }
    check(hipGetLastError(), "k_m2l27 launch");
}

void launch_split(const Geom& g, const DevTables& t, int nentry, const int* pc_parent,
                  const int* pc_child, const int* pc_sign, int klev, double2* expn, hipStream_t s,
                  const Multi& mu) {
    for (int a = 0; a < nentry; a += kMaxGridY) {
        const int cnt = std::min(kMaxGridY, nentry - a);
        const dim3 grid(blocks_for(((long long)g.nd * g.nexp + kSplitEPT - 1) / kSplitEPT), cnt);
        Multi ma = mu;
        if (ma.pp) ma.tprob += a, ma.tklev += a;
        if (t.expn_f32)
            hipLaunchKernelGGL(k_split<float2>, grid, dim3(kBlock), 0, s, g, t, cnt, pc_parent + a, pc_child + a,
                               pc_sign + a, klev, reinterpret_cast<float2*>(expn), ma);
        else
            hipLaunchKernelGGL(k_split<double2>, grid, dim3(kBlock), 0, s, g, t, cnt, pc_parent + a, pc_child + a,
                               pc_sign + a, klev, expn, ma);
        check(hipGetLastError(), "k_split launch");
    }
}

void launch_pw2pot_tasks(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                         const double2* expn, Kinds kd, OutDesc od, hipStream_t s, const Multi& mu) {
    if (ntask <= 0) return;
    if (!launch_pw2pot_3d(g, t, ntask, nsib, tasks, expn, kd, od, s, t.lds_limit, mu))
        throw std::runtime_error("3-D STEP 5 kernel not available for this shape");
    check(hipGetLastError(), "k_pw2pot3 launch");
}

void launch_pw2pot_tasks_z(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                           const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s) {
    if (ntask <= 0) return;
    if (!launch_pw2pot_3d_z(g, t, ntask, nsib, tasks, expn, kd, zo, s, t.lds_limit))
        throw std::runtime_error("3-D STEP 5 kernel not available for this shape");
    check(hipGetLastError(), "k_pw2pot3 launch");
}

void launch_pw2pot(const Geom& g, const DevTables& t, int ntask, const int* box, const int* lev,
                   const int* slot, const double2* expn, Kinds kd, OutDesc od, hipStream_t s) {
    if (ntask <= 0) return;
    if (t.expn_f32) throw std::runtime_error("FP32 expansions need the 3-D register-blocked kernels");
    if (g.ndim == 2) {
        const size_t lds2 = sizeof(double2) * ((size_t)g.pz * t.npw + (size_t)t.n * g.pz);
        if (lds2 <= t.lds_limit) {
            hipLaunchKernelGGL(k_pw2pot2, dim3(ntask), dim3(kBlock), lds2, s, g, t, box, lev, slot, expn, kd, od);
            check(hipGetLastError(), "k_pw2pot2 launch");
            return;
        }
    }
    const size_t lds = sizeof(double2) * ((size_t)t.n * t.npw + (size_t)g.nx * g.py);
    hipLaunchKernelGGL(k_pw2pot, dim3(ntask), dim3(kBlock), lds, s, g, t, box, lev, slot, expn, kd, od);
    check(hipGetLastError(), "k_pw2pot launch");
}

void launch_direct(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev,
                   const int* off, const int* src, const int* code, const int* flags, const double* fvals,
                   Kinds kd, OutDesc od, hipStream_t s, const Multi& mu) {
    if (ntask <= 0) return;
    if (launch_direct_3d(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, od, s, mu)) {
        check(hipGetLastError(), "k_direct3 launch");
        return;
    }
    if (g.ndim == 2) {
        const size_t lds2 = sizeof(double) * 4 * (size_t)t.n * t.n + sizeof(int) * 2 * (size_t)t.n;
        hipLaunchKernelGGL(k_direct2, dim3(ntask), dim3(kBlock), lds2, s, g, t, tgt, lev, off, src, code, fvals,
                           kd, od);
        check(hipGetLastError(), "k_direct2 launch");
        return;
    }
    const size_t lds = sizeof(double) * ((size_t)g.npbox + (size_t)g.nx * g.ny + 3 * (size_t)t.n * t.n) +
                       sizeof(int) * 6 * (size_t)t.n;
    hipLaunchKernelGGL(k_direct, dim3(ntask), dim3(kBlock), lds, s, g, t, tgt, lev, off, src, code, fvals,
                       kd, od);
    check(hipGetLastError(), "k_direct launch");
}

}  // namespace fgt
