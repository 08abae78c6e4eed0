// kernels_pw3.hip -- register-blocked 3-D plane-wave kernels for gfx950 (FP64).
//
// STEP 1 (values -> PW, g3d_tens_prod_to_pw, bfgt_pwrouts.f:153-222) and
// STEP 5 (PW -> values, g3d_pw2pot, bfgt_pwrouts.f:719-790) for norder
// N in {8, 12, 16}, one 256-thread block (4 waves) per task (a leaf, a
// parent with leaf children, or up to four sibling leaves).
//
// Design: on the VALU, every contraction is arranged so that the 1-D table
// operand is wave-uniform (its address depends only on loop counters and the
// wave index), which lets it come from the scalar cache into SGPRs; the other
// operand lives in VGPRs, so each v_fma_f64 has one scalar and one vector
// source and LDS only carries the transposes between stages.  STEP 5's x
// contraction for N = 12 and 16 runs on the FP64 matrix cores instead.  The half
// (z) axis is processed in rounds of four slabs, one per wave.
#include "kernels.h"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace fgt {

int lowp_stages(int prec);

namespace {

__device__ __forceinline__ double2 zfma(double2 t, double2 f, double2 c) {
    return make_double2(c.x + t.x * f.x - t.y * f.y, c.y + t.x * f.y + t.y * f.x);
}

// Complex 16x16x4 tile product on the FP64 matrix cores: (re, im) += A B with
// A, B complex, as four v_mfma_f64_16x16x4f64 (Ar Br - Ai Bi, Ar Bi + Ai Br).
// Operand layout of that instruction: A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16]; result lane l, element v = D[l/16 + 4v][l%16].
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void cmfma(double2 a, double2 b, dbl4& re, dbl4& im) {
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im, 0, 0, 0);
}

// Reduced-precision matrix-core contractions (mixed-precision forms).  A
// complex product C[row][kk] = sum_k A[row][k] U[kk][k] becomes a real GEMM
// with columns (kk, part) -- part 0 the real, 1 the imaginary part of C:
//   C = Ar * B1 + Ai * B2,  B1[k][(kk,0|1)] = (Ur | Ui),  B2 = (-Ui | Ur),
// so a column tile of 16 holds 8 complex outputs (norder 8 fills one tile).
// FP32: v_mfma_f32_16x16x4_f32, A lane l = A[l%16][l/16], B lane l =
// B[l/16][l%16], result lane l element v = D[4(l/16)+v][l%16].
// BF16 hi/lo: v_mfma_f32_16x16x32_bf16 with K = (k, part) interleaved, lane l
// holding K entries 8(l/16)..8(l/16)+7 of row (A) or column (B) l%16; each
// FP32 operand x is split as hi = bf16(x), lo = bf16(x - hi) and the product
// is Ah Bh + Ah Bl + Al Bh (relative error ~1e-5), accumulated in FP32.
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
// largest plane-wave count of the reduced-precision stages (row tiles of 16)
constexpr int kLowpMaxPw = 80;

__device__ __forceinline__ flt4 mfma32(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma_bf(bf8 a, bf8 b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// hi/lo split of eight FP32 values
__device__ __forceinline__ void split_bf(const float (&x)[8], bf8& hi, bf8& lo) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const __bf16 h = (__bf16)x[i];
        hi[i] = h;
        lo[i] = (__bf16)(x[i] - (float)h);
    }
}
// B entry of column c = (kk, part) for K entry (k, kpart) of the expanded
// real GEMM, from the complex table value u = U[kk][k]
__device__ __forceinline__ float bcol(double2 u, int part, int kpart) {
    return kpart == 0 ? (part == 0 ? (float)u.x : (float)u.y) : (part == 0 ? -(float)u.y : (float)u.x);
}

// x stage of STEP 1 in reduced precision on the matrix cores (one wave):
//   E[k1][k2] = sum_{kk < KN} B(kk, k1) A(k2, kk)
// rows k2 (tiles of 16), columns (k1, part) interleaved (a tile holds 8
// frequencies k1), K = kk (KN complex terms: N values of m1, or 2N for the
// fused form's two x signs).  fa(k2, kk) / fb(kk, k1) return the FP64
// operands (stage buffer in LDS / 1-D table); fw(k2, k1, v) stores one
// complex result.  The wave does column tiles ct0, ct0 + ctstep, ... in
// pairs (FP32: a last tile without a partner alone).  Each lane trades its part
// with the neighbouring lane (the other part of the same output) so every
// result is written once as a complex pair: part-0 lanes write rows v = 0, 1
// of their four, part-1 lanes 2, 3.
constexpr int kLowpColGroup = 2;

// column tiles c0, c0 + ctstep, ... (CG of them, all below CT)
template <int KN, int XS, int CG, class FA, class FB, class FW>
__device__ __forceinline__ void x1_lowp_group(int npw, int lane, int c0, int ctstep, FA fa, FB fb, FW fw) {
    const int MT = (npw + 15) >> 4;
    const int CT = (2 * npw + 15) >> 4;
    const int li = lane & 15, lq = lane >> 4;
    {
        int k1c[CG], part[CG];
        bool cval[CG];
#pragma unroll
        for (int gi = 0; gi < CG; ++gi) {
            const int c = 16 * (c0 + gi * ctstep) + li;
            part[gi] = c & 1;
            cval[gi] = (c >> 1) < npw && c0 + gi * ctstep < CT;
            k1c[gi] = cval[gi] ? (c >> 1) : 0;
        }
        // FP32: the column tiles' table operands of every K step, read once
        // for all row tiles (the BF16 form re-reads them per row tile)
        constexpr int KS = (KN + 3) / 4;
        float tb1[CG][XS == 2 ? KS : 1], tb2[CG][XS == 2 ? KS : 1];
        if constexpr (XS == 2) {
#pragma unroll
            for (int st = 0; st < KS; ++st) {
                const int kk = 4 * st + lq;
                const bool kv = kk < KN;
                const int kkc = kv ? kk : KN - 1;
#pragma unroll
                for (int gi = 0; gi < CG; ++gi) {
                    const double2 u = fb(kkc, k1c[gi]);
                    const bool ok = kv && cval[gi];
                    tb1[gi][st] = ok ? bcol(u, part[gi], 0) : 0.f;
                    tb2[gi][st] = ok ? bcol(u, part[gi], 1) : 0.f;
                }
            }
        }
        // one row tile at a time (few live accumulators, high occupancy)
        for (int mt = 0; mt < MT; ++mt) {
            // two accumulator chains per tile (real / imaginary A parts, or
            // hi*hi / the correction products) so dependent MFMAs interleave
            flt4 acc[CG][2];
#pragma unroll
            for (int gi = 0; gi < CG; ++gi) acc[gi][0] = acc[gi][1] = flt4{0.f, 0.f, 0.f, 0.f};
            const int k2r = min(16 * mt + li, npw - 1);
            if constexpr (XS == 2) {
#pragma unroll
                for (int st = 0; st < KS; ++st) {
                    const int kk = 4 * st + lq;
                    const double2 a = fa(k2r, kk < KN ? kk : KN - 1);
#pragma unroll
                    for (int gi = 0; gi < CG; ++gi) {
                        acc[gi][0] = mfma32((float)a.x, tb1[gi][st], acc[gi][0]);
                        acc[gi][1] = mfma32((float)a.y, tb2[gi][st], acc[gi][1]);
                    }
                }
            } else {
#pragma unroll
                for (int st = 0; st < (KN + 15) / 16; ++st) {
                    float ax[8];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int kk = 16 * st + 4 * lq + j;
                        const double2 a = fa(k2r, kk < KN ? kk : KN - 1);
                        ax[2 * j] = (float)a.x;
                        ax[2 * j + 1] = (float)a.y;
                    }
                    bf8 ah, al;
                    split_bf(ax, ah, al);
#pragma unroll
                    for (int gi = 0; gi < CG; ++gi) {
                        float bx[8];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int kk = 16 * st + 4 * lq + j;
                            const bool ok = kk < KN && cval[gi];
                            const double2 u = fb(kk < KN ? kk : KN - 1, k1c[gi]);
                            bx[2 * j] = ok ? bcol(u, part[gi], 0) : 0.f;
                            bx[2 * j + 1] = ok ? bcol(u, part[gi], 1) : 0.f;
                        }
                        bf8 bh, bl;
                        split_bf(bx, bh, bl);
                        acc[gi][0] = mfma_bf(ah, bh, acc[gi][0]);
                        acc[gi][1] = mfma_bf(ah, bl, acc[gi][1]);
                        acc[gi][1] = mfma_bf(al, bh, acc[gi][1]);
                    }
                }
            }
            // lane holds rows 16 mt + 4 lq + v of column (k1, part)
#pragma unroll
            for (int gi = 0; gi < CG; ++gi) {
                float o[4], q[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    o[v] = acc[gi][0][v] + acc[gi][1][v];
                    q[v] = __shfl_xor(o[v], 1);
                }
                if (cval[gi])
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int v = 2 * part[gi] + h;
                        const int k2 = 16 * mt + 4 * lq + v;
                        if (k2 < npw)
                            fw(k2, k1c[gi], part[gi] ? make_float2(q[v], o[v]) : make_float2(o[v], q[v]));
                    }
            }
        }
    }
}

template <int KN, int XS, class FA, class FB, class FW>
__device__ __forceinline__ void x1_lowp(int npw, int lane, int ct0, int ctstep, FA fa, FB fb, FW fw) {
    constexpr int CG = kLowpColGroup;
    const int CT = (2 * npw + 15) >> 4;
    for (int c0 = ct0; c0 < CT; c0 += CG * ctstep) {
        // (BF16: the one-tile form's registers would cost a block per CU)
        if (XS != 2 || c0 + (CG - 1) * ctstep < CT)
            x1_lowp_group<KN, XS, CG>(npw, lane, c0, ctstep, fa, fb, fw);
        else
            x1_lowp_group<KN, XS, 1>(npw, lane, c0, ctstep, fa, fb, fw);
    }
}

// ---------------------------------------------------------------- STEP 1
// LDS: sff [m2][s][m1] (N*4*N complex) and sff2 [s][m1][k2] (4*N*npw
// complex) share one buffer: the y stage reads its sff column into registers
// and passes a barrier before writing sff2 (more blocks fit on a CU).
template <int N, int XS, typename ET>
__global__ void __launch_bounds__(256) k_p2pw3(Geom g, const double2* __restrict__ tab_a, int npw,
                                              const int* __restrict__ box, const int* __restrict__ lev,
                                              const int* __restrict__ slot,
                                              const double* __restrict__ fvals_a,
                                              ET* __restrict__ expn_a, Multi mu) {
    extern __shared__ double2 sm[];
    double2* sff = sm;
    double2* sff2 = sm;
    const int task = blockIdx.x;
    const double2* __restrict__ tab = tab_a;
    const double* __restrict__ fvals = fvals_a;
    ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[task]];
        g = q.g;
        npw = q.g.npw;
        tab = q.p2pw;
        fvals = q.fvals;
        expn = static_cast<ET*>(q.expn);
    }
    const int b = box[task], l = lev[task], sl = slot[task];
    const CTab2 T = ctab(tab) + (long long)l * npw * N;   // T[k][m]
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int nh = g.pz, nd = g.nd;
    for (int ind = 0; ind < nd; ++ind) {
        double fr[N];
        if (tid < N * N) {
#pragma unroll
            for (int m3 = 0; m3 < N; ++m3)
                fr[m3] = fvals[((size_t)b * g.npbox + tid + N * N * m3) * nd + ind];
        }
        ET* __restrict__ E = expn + ((size_t)sl * 2 * nd + ind) * g.nexp;
        for (int r = 0; r < nh; r += 4) {
            // z: ff[s][m1][m2] = sum_m3 T[k3_s][m3] f[m1,m2,m3], thread = (m1,m2)
            if (tid < N * N) {
                const int m1 = tid % N, m2 = tid / N;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int k3 = r + s;
                    if (k3 < nh) {
                        double ar = 0.0, ai = 0.0;
#pragma unroll
                        for (int m3 = 0; m3 < N; ++m3) {
                            const double2 t = T[k3 * N + m3];
                            ar += t.x * fr[m3];
                            ai += t.y * fr[m3];
                        }
                        sff[(m2 * 4 + s) * N + m1] = make_double2(ar, ai);
                    }
                }
            }
            __syncthreads();
            // y: ff2[s][m1][k2] = sum_m2 T[k2][m2] ff[s][m1][m2]; lane = (m1, s), wave strides k2
            {
                const int m1 = lane & 15, s = lane >> 4;
                const bool act = (m1 < N) && (r + s < nh);
                double2 ff[N];
#pragma unroll
                for (int m2 = 0; m2 < N; ++m2)
                    ff[m2] = act ? sff[(m2 * 4 + s) * N + m1] : make_double2(0.0, 0.0);
                __syncthreads();   // sff consumed; sff2 (same memory) may be written
                for (int k2 = wave; k2 < npw; k2 += 4) {
                    // two partial sums: four independent FMA chains
                    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll
                    for (int m2 = 0; m2 < N; m2 += 2) {
                        a0 = zfma(T[k2 * N + m2], ff[m2], a0);
                        a1 = zfma(T[k2 * N + m2 + 1], ff[m2 + 1], a1);
                    }
                    if (act) sff2[(s * N + m1) * npw + k2] = make_double2(a0.x + a1.x, a0.y + a1.y);
                }
            }
            __syncthreads();
            // x: E[k3][k1][k2] = sum_m1 T[k1][m1] ff2[s][m1][k2]; wave = slab, lane = k2
            if constexpr (XS >= 2) {
                const int k3 = r + wave;
                if (k3 < nh) {
                    ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    x1_lowp<N, XS>(
                        npw, lane, 0, 1,
                        [&](int k2, int m1) { return sff2[(wave * N + m1) * npw + k2]; },
                        [&](int m1, int k1) { return T[k1 * N + m1]; },
                        [&](int k2, int k1, float2 v) { Ek[(size_t)k1 * npw + k2] = v; });
                }
            } else {
                const int k3 = r + wave;
                if (k3 < nh) {
                    ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    for (int c = 0; c < npw; c += 64) {
                        const int k2 = c + lane;
                        const bool act = k2 < npw;
                        double2 f2[N];
#pragma unroll
                        for (int m1 = 0; m1 < N; ++m1)
                            f2[m1] = act ? sff2[(wave * N + m1) * npw + k2] : make_double2(0.0, 0.0);
                        for (int k1 = 0; k1 < npw; ++k1) {
                            double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll
                            for (int m1 = 0; m1 < N; m1 += 2) {
                                a0 = zfma(T[k1 * N + m1], f2[m1], a0);
                                a1 = zfma(T[k1 * N + m1 + 1], f2[m1 + 1], a1);
                            }
                            if (act) stE(Ek + (size_t)k1 * npw + k2, make_double2(a0.x + a1.x, a0.y + a1.y));
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------- STEP 1 + STEP 2
// Parent expansion formed directly from its leaf children (STEP 1 of each
// leaf child and its STEP 2 shift into the parent, fused):
//   mp_P = sum_c (Tx o w_x^{c}) (Ty o w_y^{c}) (Tz o w_z^{c}) f_c
// with w^{c} the child->parent merge phase of each axis.  Children are grouped
// by their x and y signs: the z stage runs per child and is summed into four
// (sx,sy) groups, the y stage runs per x sign over both y signs, the x stage
// runs once over both x signs (K = 2N) -- four times fewer x-stage flops than
// eight separate leaf transforms, and no leaf expansion reaches HBM.
// tabS[l][v][k][m]: child-level table, v = 0 plain, 1 = + phase, 2 = conjugate.
// Task lists: kids (up to 8 leaf children), ksg (merge sign bits, x = bit0).
// Rounds of SR slabs.  SR = 2: wave w takes slab w & 1 and half (w >> 1) of
// the x stage's k1 range; the y stage's lanes are (m1 < 16, slab, sx).  SR = 4
// (norder 8, matrix-core x stage, npw <= 48): wave w takes slab w, and the y
// stage's 64 lanes are (m1 < 8, slab, sx) with none idle.
// Each parent is split over kFusedParts blocks, each taking a contiguous run
// of rounds (the parts of one parent are adjacent in launch order and land on
// one XCD, so the children's values they all read meet in its L2).
// LDS: sH [4 groups][SR slabs][m2][m1] and sF [2 sx][SR slabs][m1][npw]
// share one buffer: sH is fully read into registers before sF is written.
constexpr int kFusedParts = 4;

template <int N, int XS, typename ET, int SR = 2>
__global__ void __launch_bounds__(256) k_p2pwF3(Geom g, const double2* __restrict__ tabS_a, int npw,
                                               const int* __restrict__ plev, const int* __restrict__ pslot,
                                               const int* __restrict__ kids, const int* __restrict__ ksg,
                                               const double* __restrict__ fvals_a,
                                               ET* __restrict__ expn_a, Multi mu) {
    static_assert(SR == 2 || (SR == 4 && N == 8 && XS >= 2), "four-slab rounds: norder 8, matrix-core x stage");
    extern __shared__ double2 sm[];
    double2* sH = sm;                     // 4*SR*N*N
    double2* sF = sm;                     // 2*SR*N*npw (aliases sH)
    const int lin = xcd_task(blockIdx.x, gridDim.x);
    const int task = lin / kFusedParts, part = lin % kFusedParts;
    const double2* __restrict__ tabS = tabS_a;
    const double* __restrict__ fvals = fvals_a;
    ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[task]];
        g = q.g;
        npw = q.g.npw;
        tabS = q.p2pwS;
        fvals = q.fvals;
        expn = static_cast<ET*>(q.expn);
    }
    const int l = plev[task] + 1, sl = pslot[task];
    const int per = npw * N;
    const int base0 = __builtin_amdgcn_readfirstlane(3 * l * per);
    const CTab2 T = ctab(tabS);
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int nh = g.pz, nd = g.nd;
    int nk = 0;
    while (nk < 8 && kids[8 * task + nk] >= 0) ++nk;   // leaf children are packed first
    const int nround = (nh + SR - 1) / SR;
    const int rb = SR * (part * nround / kFusedParts), re = SR * ((part + 1) * nround / kFusedParts);
    for (int ind = 0; ind < nd; ++ind) {
        ET* __restrict__ E = expn + ((size_t)sl * 2 * nd + ind) * g.nexp;
        // norder 8: one wave per group g = (sx, sy) of children; the values
        // of its (up to two) children -- the z sign tells them apart -- stay
        // in registers for every round
        double frg[2][N];
        bool has0 = false, has1 = false;
        if constexpr (N * N == 64) {
            for (int c = 0; c < nk; ++c) {
                const int sg = ksg[8 * task + c];
                if ((sg & 3) != wave) continue;
                const int b = kids[8 * task + c];
                const double* __restrict__ fb = fvals + ((size_t)b * g.npbox + lane) * nd + ind;
                if ((sg >> 2) & 1) {
                    has1 = true;
#pragma unroll
                    for (int m3 = 0; m3 < N; ++m3) frg[1][m3] = fb[(size_t)N * N * m3 * nd];
                } else {
                    has0 = true;
#pragma unroll
                    for (int m3 = 0; m3 < N; ++m3) frg[0][m3] = fb[(size_t)N * N * m3 * nd];
                }
            }
        }
        for (int r = rb; r < re; r += SR) {
            // z: H[g][s][m2][m1] = sum_{c in g} sum_m3 Tz_c[k3_s][m3] f_c[m1,m2,m3]
            if constexpr (N * N == 64) {
                const int q = wave;
                double2 H[SR];
#pragma unroll
                for (int s = 0; s < SR; ++s) H[s] = make_double2(0.0, 0.0);
#pragma unroll
                for (int z = 0; z < 2; ++z) {
                    if (!(z ? has1 : has0)) continue;
                    const CTab2 Tz = T + base0 + (1 + z) * per;
#pragma unroll
                    for (int s = 0; s < SR; ++s) {
                        const int k3 = min(r + s, nh - 1);
                        double ar = 0.0, ai = 0.0;
#pragma unroll
                        for (int m3 = 0; m3 < N; ++m3) {
                            const double2 t = Tz[k3 * N + m3];
                            ar += t.x * frg[z][m3];
                            ai += t.y * frg[z][m3];
                        }
                        H[s].x += ar;
                        H[s].y += ai;
                    }
                }
#pragma unroll
                for (int s = 0; s < SR; ++s) sH[((q * SR + s) * N) * N + lane] = H[s];
            } else if (tid < N * N) {
                double2 H[4][2];
#pragma unroll
                for (int q = 0; q < 4; ++q) H[q][0] = H[q][1] = make_double2(0.0, 0.0);
                for (int c = 0; c < nk; ++c) {
                    const int b = kids[8 * task + c], sg = ksg[8 * task + c];
                    const int vz = 1 + ((sg >> 2) & 1);
                    const CTab2 Tz = T + base0 + vz * per;
                    double fr[N];
#pragma unroll
                    for (int m3 = 0; m3 < N; ++m3)
                        fr[m3] = fvals[((size_t)b * g.npbox + tid + N * N * m3) * nd + ind];
                    const int q = sg & 3;
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int k3 = min(r + s, nh - 1);
                        double ar = 0.0, ai = 0.0;
#pragma unroll
                        for (int m3 = 0; m3 < N; ++m3) {
                            const double2 t = Tz[k3 * N + m3];
                            ar += t.x * fr[m3];
                            ai += t.y * fr[m3];
                        }
                        // q is the same for all lanes (uniform per child)
#pragma unroll
                        for (int qq = 0; qq < 4; ++qq)
                            if (qq == q) { H[qq][s].x += ar; H[qq][s].y += ai; }
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int s = 0; s < 2; ++s) sH[((q * 2 + s) * N) * N + tid] = H[q][s];
            }
            __syncthreads();
            // y: F[sx][s][m1][k2] = sum_sy sum_m2 Ty^{sy}[k2][m2] H[(sx,sy)][s][m2][m1]
            // lane = (m1 < MS, s < SR, sx in 0..1), wave strides k2
            {
                constexpr int MS = 32 / SR;
                const int m1 = lane % MS, s = (lane / MS) % SR, sx = lane / (MS * SR);
                const bool act = (m1 < N) && (r + s < nh);
                double2 hp[N], hm[N];
#pragma unroll
                for (int m2 = 0; m2 < N; ++m2) {
                    hp[m2] = act ? sH[(((sx + 0) * SR + s) * N + m2) * N + m1] : make_double2(0.0, 0.0);
                    hm[m2] = act ? sH[(((sx + 2) * SR + s) * N + m2) * N + m1] : make_double2(0.0, 0.0);
                }
                const CTab2 Typ = T + base0 + 1 * per;
                const CTab2 Tym = T + base0 + 2 * per;
                __syncthreads();   // sH consumed; sF (same memory) may be written
                for (int k2 = wave; k2 < npw; k2 += 4) {
                    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll
                    for (int m2 = 0; m2 < N; ++m2) {
                        a0 = zfma(Typ[k2 * N + m2], hp[m2], a0);
                        a1 = zfma(Tym[k2 * N + m2], hm[m2], a1);
                    }
                    if (act) sF[((sx * SR + s) * N + m1) * npw + k2] = make_double2(a0.x + a1.x, a0.y + a1.y);
                }
            }
            __syncthreads();
            // x: E[k3][k1][k2] = sum_sx sum_m1 Tx^{sx}[k1][m1] F[sx][s][m1][k2]
            if constexpr (XS >= 2) {
                // K = (sx, m1); SR = 2: the two waves of a slab take alternate
                // column tiles, SR = 4: one wave per slab takes them all
                const int s = SR == 4 ? wave : (wave & 1), hlf = SR == 4 ? 0 : wave >> 1;
                const int k3 = r + s;
                if (k3 < nh) {
                    ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    const CTab2 Txp = T + base0 + 1 * per;
                    const CTab2 Txm = T + base0 + 2 * per;
                    x1_lowp<2 * N, XS>(
                        npw, lane, hlf, SR == 4 ? 1 : 2,
                        [&](int k2, int kk) {
                            const int sx = kk >= N ? 1 : 0;
                            return sF[((sx * SR + s) * N + (kk - N * sx)) * npw + k2];
                        },
                        [&](int kk, int k1) { return kk >= N ? Txm[k1 * N + kk - N] : Txp[k1 * N + kk]; },
                        [&](int k2, int k1, float2 v) { Ek[(size_t)k1 * npw + k2] = v; });
                }
            } else {
                const int s = wave & 1, hlf = wave >> 1;
                const int k3 = r + s;
                if (k3 < nh) {
                    ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    const int k1a = hlf ? (npw + 1) / 2 : 0, k1b = hlf ? npw : (npw + 1) / 2;
                    const CTab2 Txp = T + base0 + 1 * per;
                    const CTab2 Txm = T + base0 + 2 * per;
                    for (int c = 0; c < npw; c += 64) {
                        const int k2 = c + lane;
                        const bool act = k2 < npw;
                        double2 fp[N], fm[N];
#pragma unroll
                        for (int m1 = 0; m1 < N; ++m1) {
                            fp[m1] = act ? sF[((0 * 2 + s) * N + m1) * npw + k2] : make_double2(0.0, 0.0);
                            fm[m1] = act ? sF[((1 * 2 + s) * N + m1) * npw + k2] : make_double2(0.0, 0.0);
                        }
                        for (int k1 = k1a; k1 < k1b; ++k1) {
                            double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll
                            for (int m1 = 0; m1 < N; ++m1) {
                                a0 = zfma(Txp[k1 * N + m1], fp[m1], a0);
                                a1 = zfma(Txm[k1 * N + m1], fm[m1], a1);
                            }
                            if (act) stE(Ek + (size_t)k1 * npw + k2, make_double2(a0.x + a1.x, a0.y + a1.y));
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// XS 4 (FP32, one column tile: norder 8): the x table operands of every K
// step, converted once per task (the table depends only on the task's
// level, x variant and kind) for all its rounds and row tiles (registers:
// the largest plane-wave count's steps, uniform guards past npw; 147 VGPRs,
// three waves per SIMD, against four for XS 2, which converts them per row
// tile: STEP 5 of 21 norder-8, npw <= 32 problems 1.80 -> 1.47 ms)
template <int N, int XS>
struct XHoist {
    static constexpr bool on = XS == 4 && (2 * N + 15) / 16 == 1;
    static constexpr int KSH = on ? kLowpMaxPw / 4 : 1;
    float b0[KSH], b1[KSH];
    __device__ __forceinline__ void load(const CTab2 UX, int npw, int lane) {
        if constexpr (on) {
            const int li = lane & 15, lq = lane >> 4;
            const bool cval = li < 2 * N;
            const int cpart = li >= N ? 1 : 0;
            const int ckk = cval ? li - N * cpart : 0;
            const int ks = (npw + 3) >> 2;
#pragma unroll
            for (int st = 0; st < KSH; ++st) {
                if (st < ks) {
                    const int k1 = 4 * st + lq;
                    const bool kv = k1 < npw;
                    const double2 u = UX[(kv ? k1 : npw - 1) * N + ckk];
                    const bool ok = kv && cval;
                    b0[st] = ok ? bcol(u, cpart, 0) : 0.f;
                    b1[st] = ok ? bcol(u, cpart, 1) : 0.f;
                }
            }
        }
    }
};

// x stage of STEP 5 in reduced precision on the matrix cores (wave w takes
// the round's slabs w, w + 4, ... of SR):
//   ff[k2][kk1] = sum_k1 Ux[kk1][k1] E[k3][k1][k2],
// rows k2 (row tiles of 16, up to 80 frequencies), columns (kk1, part), K =
// k1.  XS 2: FP32 operands; XS 3: BF16 hi/lo operands.  The result goes to
// sff[(k2*SR + slab)*KW + (kk1 ^ (k2 & (KW-1)))], KW = 64 / SR (FP32
// complex: the two accumulator chains' FP32 sum, stored exactly; for SR = 4
// the layout of the FP64 x stages), where the y stage reads it.  Lanes past npw read a valid duplicate row (discarded); K
// entries past npw get zero table values.
template <int N, int XS, int SR>
__device__ __forceinline__ void pw2pot_x_lowp(const float2* __restrict__ E, const CTab2 UX,
                                              int npw, int nh, int r, int wave, int lane, float2* sff,
                                              const XHoist<N, XS>& hx) {
    constexpr int KW = 64 / SR;
    static_assert(KW >= N, "a (k2, slab) row of sff holds the N values kk1");
    if (r + wave >= nh) return;
    constexpr int NT = (2 * N + 15) / 16;
    const int MT = (npw + 15) >> 4;
    const int li = lane & 15, lq = lane >> 4;
    int ckk[NT], cpart[NT];
    bool cval[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int c = 16 * t + li;
        cval[t] = c < 2 * N;
        cpart[t] = c >= N ? 1 : 0;
        ckk[t] = cval[t] ? c - N * cpart[t] : 0;
    }
    float* __restrict__ sffd = reinterpret_cast<float*>(sff);
    constexpr bool HOIST = XHoist<N, XS>::on;
    constexpr int KSH = XHoist<N, XS>::KSH;
    for (int sl = wave; sl < SR; sl += 4) {
        const int k3 = r + sl;
        if (k3 >= nh) break;
        const float2* __restrict__ Ek = E + (size_t)k3 * npw * npw;
        // one row tile (16 frequencies k2) at a time, two accumulator chains per
        // column tile
        for (int mt = 0; mt < MT; ++mt) {
            const int k2r = min(16 * mt + li, npw - 1);
            flt4 acc[NT][2];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t][0] = acc[t][1] = flt4{0.f, 0.f, 0.f, 0.f};
            if constexpr (HOIST) {
                // E rows in chunks of four K steps, the next chunk loaded before
                // the current one's products (chunks unrolled to the largest
                // plane-wave count, uniform guards past npw)
                const int ks = (npw + 3) >> 2;
                constexpr int NCH = KSH / 4;
                float2 e[2][4];
#pragma unroll
                for (int j = 0; j < 4; ++j) e[0][j] = Ek[(size_t)min(4 * j + lq, npw - 1) * npw + k2r];
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    if (4 * c < ks) {
                        if (c + 1 < NCH && 4 * (c + 1) < ks) {
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                e[(c + 1) & 1][j] = Ek[(size_t)min(4 * (4 * (c + 1) + j) + lq, npw - 1) * npw + k2r];
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            if (4 * c + j < ks) {
                                acc[0][0] = mfma32(e[c & 1][j].x, hx.b0[4 * c + j], acc[0][0]);
                                acc[0][1] = mfma32(e[c & 1][j].y, hx.b1[4 * c + j], acc[0][1]);
                            }
                        }
                    }
                }
            } else if constexpr (XS == 2) {
                // FP32: K steps in chunks of KC, the next chunk's operand row and
                // table entries loaded (from valid clamped addresses, zeroed past
                // npw) before the current chunk's products, into two register sets
                // (two steps per chunk where two column tiles hold the registers)
                constexpr int KC = NT == 1 ? 4 : 2;
                const int ks = (npw + 3) >> 2;
                const int nchunk = (ks + KC - 1) / KC;
                float2 ea[KC], eb[KC];
                float ba[NT][KC][2], bb[NT][KC][2];
                auto load = [&](float2 (&e)[KC], float (&bv)[NT][KC][2], int c) {
#pragma unroll
                    for (int j = 0; j < KC; ++j) {
                        const int k1 = 4 * (KC * c + j) + lq;
                        const bool kv = k1 < npw;
                        const int k1c = kv ? k1 : npw - 1;
                        e[j] = Ek[(size_t)k1c * npw + k2r];
#pragma unroll
                        for (int t = 0; t < NT; ++t) {
                            const double2 u = UX[k1c * N + ckk[t]];
                            const bool ok = kv && cval[t];
                            bv[t][j][0] = ok ? bcol(u, cpart[t], 0) : 0.f;
                            bv[t][j][1] = ok ? bcol(u, cpart[t], 1) : 0.f;
                        }
                    }
                };
                auto apply = [&](const float2 (&e)[KC], const float (&bv)[NT][KC][2]) {
#pragma unroll
                    for (int j = 0; j < KC; ++j)
#pragma unroll
                        for (int t = 0; t < NT; ++t) {
                            acc[t][0] = mfma32(e[j].x, bv[t][j][0], acc[t][0]);
                            acc[t][1] = mfma32(e[j].y, bv[t][j][1], acc[t][1]);
                        }
                };
                load(ea, ba, 0);
                for (int c = 0; c < nchunk; c += 2) {
                    load(eb, bb, min(c + 1, nchunk - 1));
                    __builtin_amdgcn_sched_barrier(0);
                    apply(ea, ba);
                    if (c + 1 >= nchunk) break;
                    load(ea, ba, min(c + 2, nchunk - 1));
                    __builtin_amdgcn_sched_barrier(0);
                    apply(eb, bb);
                }
            } else {
                // K = (k1, part): 16 frequencies per step; lane K entries 8 lq .. 8 lq + 7
                const int ks = (npw + 15) >> 4;
                for (int st = 0; st < ks; ++st) {
                    float ax[8];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int k1c = min(16 * st + 4 * lq + j, npw - 1);
                        const float2 e = Ek[(size_t)k1c * npw + k2r];
                        ax[2 * j] = e.x;
                        ax[2 * j + 1] = e.y;
                    }
                    bf8 ah, al;
                    split_bf(ax, ah, al);
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        float bx[8];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int k1 = 16 * st + 4 * lq + j;
                            const bool ok = k1 < npw && cval[t];
                            const double2 u = UX[min(k1, npw - 1) * N + ckk[t]];
                            bx[2 * j] = ok ? bcol(u, cpart[t], 0) : 0.f;
                            bx[2 * j + 1] = ok ? bcol(u, cpart[t], 1) : 0.f;
                        }
                        bf8 bh, bl;
                        split_bf(bx, bh, bl);
                        acc[t][0] = mfma_bf(ah, bh, acc[t][0]);
                        acc[t][1] = mfma_bf(ah, bl, acc[t][1]);
                        acc[t][1] = mfma_bf(al, bh, acc[t][1]);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < NT; ++t)
                if (cval[t])
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k2 = 16 * mt + 4 * lq + v;
                        if (k2 < npw)
                            sffd[2 * ((k2 * SR + sl) * KW + (ckk[t] ^ (k2 & (KW - 1)))) + cpart[t]] =
                                acc[t][0][v] + acc[t][1][v];
                    }
        }
    }
}

// ---------------------------------------------------------------- STEP 5
// Tables: tabS[kind][l][v][k][i] (frequency k, node i), v = 0 the plain
// PW->value table, 1 / 2 the table times the + / conjugate parent->child
// split phase.  A task (PotTask) evaluates one expansion at up to four
// leaves: for a leaf evaluating its parent's expansion the phase is folded
// into the tables (STEP 4 folded into STEP 5), and the x variant depends only
// on the x split sign, so four siblings share one x contraction.
// Per round of SR k3 slabs (4; 8 for norder 8 with a matrix-core x stage):
//   x: ff[k2][s][kk1] = sum_k1 Ux[kk1][k1] E[k3][k1][k2]   (wave = slab,
//      or slabs w and w + 4)
//   y: yres[s][kk1][kk2] = sum_k2 Uy[kk2][k2] ff[k2][s][kk1], per y variant
//   z: G_c[kk3] += Re(Uz_c[kk3][k3] yres[s][zk1][zk2]), per leaf c, with one
//      thread per output column (zk1, zk2) holding all kk3 in registers.
// LDS: sff [k2][s][KW] (npw*64 complex, KW = 64 / SR; FP32 after a
// reduced-precision x stage, whose FP32 results it holds exactly, else
// FP64), yres [s][kk1][16]
// (1024 complex, entry kk2 at kk2 ^ kk1), sUr [v][s][kk3] (3*SR*N complex:
// the round's z table rows of the three variants, zero past nh).
// MF: x stage on the FP64 matrix cores (N >= 12, npw <= 64): per slab the
// complex product ff[kk1][k2] = sum_k1 U[kk1][k1] E[k1][k2] runs as
// 16x16x4 tiles (rows kk1, columns 16 k2, K = 4 k1; for N = 12 the rows
// kk1 >= N carry zero table entries).  For N = 8 half of each tile would be
// padding and the VALU form is faster (STEP 5 over 21 norder-8 problems:
// 4.34 vs 5.10 ms; norder 12: 5.17 VALU vs 4.53 ms here).  The table is read per
// lane from global memory (cached); E rows stream straight into the B
// operand.  (For STEP 1 the matrix-core form is slower: its 50 x 50 output
// pads to 64 x 64 tiles, where the VALU form wastes only the lane padding.)

// C: most leaves per task (4 for sibling tasks, 1 for a leaf's own
// expansion; with one leaf the freed registers carry the x stage's next
// K-step loads while the current one runs).
// NZ: output components sharing the x and y table kinds and differing in the
// z kind zo.kz[q] (one-leaf tasks only: all their variants are 0, so sUr
// holds the round's rows of NZ <= 3 z kinds in place of the three
// variants).  The x and y stages run once for all of them, the z stage once
// per kind, into zo.od[q].
template <int N, int XS, int C, typename ET, int NZ = 1>
__global__ void __launch_bounds__(256, 2) k_pw2pot3(Geom g, const double2* __restrict__ tabS_a, int npw,
                                                const PotTask* __restrict__ tasks, const ET* __restrict__ expn_a,
                                                size_t ks, Kinds kd, ZOut zo, Multi mu) {
    static_assert(NZ >= 1 && NZ <= 3 && (NZ == 1 || C == 1), "z kinds are grouped for one-leaf tasks only");
    extern __shared__ double2 sm[];
    using SF = typename std::conditional<(XS >= 2), float2, double2>::type;
    SF* sff = reinterpret_cast<SF*>(sm);
    double2* yres = reinterpret_cast<double2*>(sff + (size_t)npw * 64);
    double2* sUr = yres + 4 * 16 * 16;
    constexpr int NQ = N / 4;
    constexpr int PF = (N == 12) ? 6 : 8;
    // slabs per round: 8 for norder 8 with a matrix-core x stage (each wave
    // computes two slabs, and the y stage's 64 lanes are (kk1, slab) with no
    // idle half), else 4; KW: kk1 slots per (k2, slab) row of sff, per
    // (slab, kk1) group of yres
    constexpr int SR = (N == 8 && XS >= 2) ? 8 : 4;
    constexpr int KW = 64 / SR;
    const int tix = xcd_task(blockIdx.x, gridDim.x);
    const PotTask* __restrict__ T = tasks + tix;
    const double2* __restrict__ tabS = tabS_a;
    const ET* __restrict__ expn = expn_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[tix]];
        g = q.g;
        npw = q.g.npw;
        tabS = q.pw2potS;
        expn = static_cast<const ET*>(q.expn);
        zo.od[0].p = zo.od[1].p = zo.od[2].p = q.pot;
    }
    const int sl = T->slot, l = T->lev, nvy = T->nvy;
    int cbox[C], ccode[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        cbox[c] = T->box[c];
        ccode[c] = T->code[c];
    }
    // table kind per axis (value / first / second derivative) at stride ks;
    // offsets are wave-uniform so the table reads stay on the scalar path
    const int per = npw * N;
    const int ox = __builtin_amdgcn_readfirstlane((int)(kd.x * ks) + (3 * l + T->vx) * per);
    const int oz = __builtin_amdgcn_readfirstlane((int)(zo.kz[0] * ks) + 3 * l * per);
    const CTab2 UX = ctab(tabS) + ox;
    const CTab2 UZ = ctab(tabS) + oz;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int nh = g.pz, nd = g.nd;
    const int kk1 = lane & (KW - 1), s = lane / KW;
    const bool halve = (npw & 1) != 0;
    // z stage: one thread per output column (zk1, zk2), all kk3 in
    // registers; where the block holds exactly one column per thread and leaf
    // (N = 8, four leaves) each wave takes one leaf (ZW), where it holds one
    // column per thread and z kind (N = 8, NZ kinds) wave q takes kind q
    // (ZK), else every thread below N^2 takes all the leaves or kinds
    constexpr bool ZW = C > 1 && N * N * C == 256;
    constexpr bool ZK = NZ > 1 && N * N == 64;
    constexpr int CG = (ZW || ZK) ? 1 : C * NZ;
    const bool zact = ZW || (ZK ? wave < NZ : tid < N * N);
    const int zt = (ZW || ZK) ? lane : tid;
    const int zk1 = zt % N, zk2 = zt / N;
    const int wbox = ZW ? T->box[wave] : -1, wcode = ZW ? T->code[wave] : 0;
    // x-stage table operands of the task (XS 4 only)
    XHoist<N, XS> hx;
    hx.load(UX, npw, lane);
    for (int ind = 0; ind < nd; ++ind) {
        double G[CG][N];
#pragma unroll
        for (int c = 0; c < CG; ++c)
#pragma unroll
            for (int k = 0; k < N; ++k) G[c][k] = 0.0;
        const ET* __restrict__ E = expn + ((size_t)sl * 2 * nd + nd + ind) * g.nexp;
        for (int r = 0; r < nh; r += SR) {
            __syncthreads();
            // the round's z table rows, all three variants or (NZ > 1) the
            // plain rows of each z kind (zero past nh)
            if constexpr (NZ == 1) {
                if (tid < 3 * SR * N) {
                    const int v = tid / (SR * N), s2 = (tid / N) % SR, k = tid % N;
                    sUr[tid] = (r + s2 < nh) ? UZ[v * per + (r + s2) * N + k] : make_double2(0.0, 0.0);
                }
            } else {
                if (tid < SR * N * NZ) {
                    const int z = tid / (SR * N), s2 = (tid / N) % SR, k = tid % N;
                    const int kz = z == 0 ? zo.kz[0] : (z == 1 ? zo.kz[1] : zo.kz[2]);
                    const CTab2 U = ctab(tabS) + ((long long)kz * ks + 3 * l * per);
                    sUr[tid] = (r + s2 < nh) ? U[(r + s2) * N + k] : make_double2(0.0, 0.0);
                }
            }
            // x: wave = slab, lane = k2.  Rows of E are fetched PF at a time
            // (all loads issued before the FMAs that use them) so enough reads
            // are in flight to cover HBM latency; lanes past npw read a valid
            // duplicate and discard it.
            if constexpr (XS >= 2) {
                pw2pot_x_lowp<N, XS, SR>(E, UX, npw, nh, r, wave, lane, sff, hx);
            } else if constexpr (XS == 1) {
                const int k3 = r + wave;
                if (k3 < nh) {
                    const ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    const int li = lane & 15, lk = lane >> 4;
                    // rows kk1 >= N (N < 16) read a valid table entry and are zeroed
                    const int lir = li < N ? li : N - 1;
                    dbl4 are[4], aim[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) are[t] = aim[t] = dbl4{0.0, 0.0, 0.0, 0.0};
                    const int ksteps = (npw + 3) >> 2;
                    if constexpr (C == 1 && NZ == 1) {
                        // the next K-step's operands are loaded (always, from a
                        // valid clamped address) before this step's products,
                        // into two alternating register sets (with NZ > 1 the
                        // z sums take those registers)
                        double2 ua, ub, ea[4], eb[4];
#define PW2POT_XLOAD(u_, e_, st)                                                             \
    {                                                                                        \
        const int k1c_ = min(4 * (st) + lk, npw - 1);                                        \
        u_ = UX[k1c_ * N + lir];                                                             \
        _Pragma("unroll") for (int t = 0; t < 4; ++t)                                        \
            e_[t] = ldE(Ek[(size_t)k1c_ * npw + min(16 * t + li, npw - 1)]);                 \
    }
#define PW2POT_XSTEP(u_, e_, st)                                                             \
    {                                                                                        \
        const double2 um_ = (4 * (st) + lk < npw && li < N) ? u_ : make_double2(0.0, 0.0);   \
        _Pragma("unroll") for (int t = 0; t < 4; ++t) cmfma(um_, e_[t], are[t], aim[t]);     \
    }
                        PW2POT_XLOAD(ua, ea, 0);
                        for (int st = 0; st < ksteps; st += 2) {
                            PW2POT_XLOAD(ub, eb, min(st + 1, ksteps - 1));
                            // keep the loads above this step's products (the
                            // scheduler would otherwise sink them to their use)
                            __builtin_amdgcn_sched_barrier(0);
                            PW2POT_XSTEP(ua, ea, st);
                            if (st + 1 >= ksteps) break;
                            PW2POT_XLOAD(ua, ea, min(st + 2, ksteps - 1));
                            __builtin_amdgcn_sched_barrier(0);
                            PW2POT_XSTEP(ub, eb, st + 1);
                        }
#undef PW2POT_XSTEP
#undef PW2POT_XLOAD
                    } else {
                        for (int st = 0; st < ksteps; ++st) {
                            const int k1 = 4 * st + lk;
                            const bool kv = k1 < npw;
                            const int k1c = kv ? k1 : npw - 1;
                            double2 u = UX[k1c * N + lir];
                            if (!kv || li >= N) u = make_double2(0.0, 0.0);
                            double2 e[4];
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                const int k2 = 16 * t + li;
                                e[t] = ldE(Ek[(size_t)k1c * npw + (k2 < npw ? k2 : npw - 1)]);
                            }
#pragma unroll
                            for (int t = 0; t < 4; ++t) cmfma(u, e[t], are[t], aim[t]);
                        }
                    }
                    // lane holds kk1 = lk + 4v for column k2 = 16t + li
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int k2 = 16 * t + li;
                        if (k2 < npw)
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                sff[(k2 * 4 + wave) * 16 + ((lk + 4 * v) ^ (k2 & 15))] =
                                    make_double2(are[t][v], aim[t][v]);
                    }
                }
            } else {
                const int k3 = r + wave;
                if (k3 < nh) {
                    const ET* __restrict__ Ek = E + (size_t)k3 * npw * npw;
                    for (int c = 0; c < npw; c += 64) {
                        const int k2 = c + lane;
                        const bool act = k2 < npw;
                        const int k2c = act ? k2 : npw - 1;
                        double2 acc[N];
#pragma unroll
                        for (int k = 0; k < N; ++k) acc[k] = make_double2(0.0, 0.0);
                        int k1 = 0;
                        for (; k1 + PF <= npw; k1 += PF) {
                            double2 e[PF];
#pragma unroll
                            for (int u = 0; u < PF; ++u) e[u] = ldE(Ek[(size_t)(k1 + u) * npw + k2c]);
#pragma unroll
                            for (int u = 0; u < PF; ++u)
#pragma unroll
                                for (int k = 0; k < N; ++k) acc[k] = zfma(UX[(k1 + u) * N + k], e[u], acc[k]);
                        }
                        for (; k1 < npw; ++k1) {
                            const double2 e = ldE(Ek[(size_t)k1 * npw + k2c]);
#pragma unroll
                            for (int k = 0; k < N; ++k) acc[k] = zfma(UX[k1 * N + k], e, acc[k]);
                        }
                        // row (k2, slab) holds 16 entries; entry k is stored at k ^ (k2 & 15)
                        // so the 64 lanes of a write hit different banks
                        if (act)
#pragma unroll
                            for (int k = 0; k < N; ++k) sff[(k2 * 4 + wave) * 16 + (k ^ (k2 & 15))] = acc[k];
                    }
                }
            }
            __syncthreads();
            for (int j = 0; j < nvy; ++j) {
                if (j) __syncthreads();   // the z stage of the previous y variant is done with yres
                // y: lane = (kk1, s), wave owns kk2 in [wave*NQ, wave*NQ+NQ)
                const int oy = __builtin_amdgcn_readfirstlane((int)(kd.y * ks) + (3 * l + T->vy[j]) * per);
                const CTab2 UY = ctab(tabS) + oy;
                const int k3 = r + s;
                if (kk1 < N && k3 < nh) {
                    double2 y[NQ];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) y[q] = make_double2(0.0, 0.0);
                    for (int k2 = 0; k2 < npw; ++k2) {
                        const SF fs = sff[(k2 * SR + s) * KW + (kk1 ^ (k2 & (KW - 1)))];
                        const double2 f = make_double2(fs.x, fs.y);
#pragma unroll
                        for (int q = 0; q < NQ; ++q) y[q] = zfma(UY[k2 * N + wave * NQ + q], f, y[q]);
                    }
#pragma unroll
                    for (int q = 0; q < NQ; ++q) yres[(s * KW + kk1) * 16 + ((wave * NQ + q) ^ kk1)] = y[q];
                }
                __syncthreads();
                // z, for the leaves that use this y variant
                if (zact) {
#pragma unroll 1
                    for (int s2 = 0; s2 < SR; ++s2) {
                        const int k3z = r + s2;
                        if (k3z >= nh) break;
                        const double2 yv = yres[(s2 * KW + zk1) * 16 + (zk2 ^ zk1)];
                        const double w = (halve && k3z == nh - 1) ? 0.5 : 1.0;
                        const double yr = yv.x * w, yi = yv.y * w;
                        if constexpr (ZW) {
                            if (wbox >= 0 && (wcode & 1) == j) {
                                const double2* __restrict__ uz = sUr + ((wcode >> 1) * SR + s2) * N;
#pragma unroll
                                for (int kk3 = 0; kk3 < N; ++kk3) {
                                    const double2 u = uz[kk3];
                                    G[0][kk3] += u.x * yr - u.y * yi;
                                }
                            }
                        } else if constexpr (ZK) {
                            if (cbox[0] >= 0) {
                                const double2* __restrict__ uz = sUr + (wave * SR + s2) * N;
#pragma unroll
                                for (int kk3 = 0; kk3 < N; ++kk3) {
                                    const double2 u = uz[kk3];
                                    G[0][kk3] += u.x * yr - u.y * yi;
                                }
                            }
                        } else if constexpr (NZ > 1) {
                            if (cbox[0] >= 0) {
#pragma unroll
                                for (int z = 0; z < NZ; ++z) {
                                    const double2* __restrict__ uz = sUr + (z * SR + s2) * N;
#pragma unroll
                                    for (int kk3 = 0; kk3 < N; ++kk3) {
                                        const double2 u = uz[kk3];
                                        G[z][kk3] += u.x * yr - u.y * yi;
                                    }
                                }
                            }
                        } else {
#pragma unroll
                            for (int c = 0; c < C; ++c) {
                                if (cbox[c] >= 0 && (ccode[c] & 1) == j) {
                                    const double2* __restrict__ uz = sUr + ((ccode[c] >> 1) * SR + s2) * N;
#pragma unroll
                                    for (int kk3 = 0; kk3 < N; ++kk3) {
                                        const double2 u = uz[kk3];
                                        G[c][kk3] += u.x * yr - u.y * yi;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
        if (zact) {
            // the empty asm makes col opaque here, so the 64 output addresses
            // are formed now and not hoisted out of the loops (where they
            // would be held, or spilled, through the whole task)
            int col = zk1 + N * zk2;
            asm volatile("" : "+v"(col));
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                const int b = ZW ? wbox : cbox[C > 1 ? c : 0];
                // output of kind c (NZ > 1), of this wave's kind (ZK), else
                // the one output (selected field by field: a whole-struct
                // select by a run-time index would put zo in scratch memory)
                const int q = ZK ? wave : (NZ > 1 ? c : 0);
                double* const op = q == 0 ? zo.od[0].p : (q == 1 ? zo.od[1].p : zo.od[2].p);
                const int ons = q == 0 ? zo.od[0].ns : (q == 1 ? zo.od[1].ns : zo.od[2].ns);
                const int ocoff = q == 0 ? zo.od[0].coff : (q == 1 ? zo.od[1].coff : zo.od[2].coff);
                if (b >= 0) {
                    double* __restrict__ o = op + (size_t)b * g.npbox * ons + ocoff + ind;
#pragma unroll
                    for (int kk3 = 0; kk3 < N; ++kk3) o[(size_t)(col + N * N * kk3) * ons] += 2.0 * G[c][kk3];
                }
            }
        }
    }
}

// STEP 1 x stage on the reduced-precision matrix cores?
bool use_lowp1(const DevTables& t) {
    if (t.prec < 1 || t.npw > kLowpMaxPw) return false;
    return t.xs1 >= 0 ? t.xs1 == 2 : (lowp_stages(t.prec) & 1) != 0;
}

// STEP 5 x-stage variant (0 VALU, 1 FP64 matrix cores, 2 reduced precision)
int x5_variant(int n, const DevTables& t) {
    const bool lowp_ok = t.prec >= 1 && t.npw <= kLowpMaxPw;
    const bool f64mf_ok = n >= 12 && t.npw <= 64;
    if (t.xs5 >= 0) {
        if (t.xs5 == 2 && lowp_ok) return 2;
        if (t.xs5 == 1 && f64mf_ok) return 1;
        return 0;
    }
    if (lowp_ok && (lowp_stages(t.prec) & 2)) return 2;
    if (f64mf_ok && !pw2pot_valu_only()) return 1;
    return 0;
}

template <int N>
bool try_p2pw3(const Geom& g, const DevTables& t, int ntask, const int* box, const int* lev, const int* slot,
               const double* fvals, double2* expn, hipStream_t s, size_t lds_limit, const Multi& mu) {
    const size_t lds = sizeof(double2) * std::max(4 * (size_t)N * N, 4 * (size_t)N * t.npw);
    if (lds > lds_limit) return false;
    if (use_lowp1(t)) {
        float2* e32 = reinterpret_cast<float2*>(expn);
        if (t.prec == 2)
            hipLaunchKernelGGL((k_p2pw3<N, 3, float2>), dim3(ntask), dim3(256), lds, s, g, t.p2pw, t.npw, box, lev,
                               slot, fvals, e32, mu);
        else
            hipLaunchKernelGGL((k_p2pw3<N, 2, float2>), dim3(ntask), dim3(256), lds, s, g, t.p2pw, t.npw, box, lev,
                               slot, fvals, e32, mu);
    } else if (t.expn_f32) {
        hipLaunchKernelGGL((k_p2pw3<N, 0, float2>), dim3(ntask), dim3(256), lds, s, g, t.p2pw, t.npw, box, lev, slot,
                           fvals, reinterpret_cast<float2*>(expn), mu);
    } else {
        hipLaunchKernelGGL((k_p2pw3<N, 0, double2>), dim3(ntask), dim3(256), lds, s, g, t.p2pw, t.npw, box, lev,
                           slot, fvals, expn, mu);
    }
    return true;
}

// xs: the x-stage variant (>= 2: reduced precision, FP32 sff)
size_t pw2pot3_lds(int n, int npw, int xs = 0) {
    // sff, yres, sUr (three variants of up to eight slabs at norder 8, four above)
    return (xs >= 2 ? sizeof(float2) : sizeof(double2)) * (size_t)npw * 64 +
           sizeof(double2) * (4 * 16 * 16 + (n == 8 ? 24 : 12) * (size_t)n);
}

template <int N, int XS, int C, int NZ = 1>
void launch_pw2pot3_c(const Geom& g, const DevTables& t, int ntask, const PotTask* tasks, const double2* expn,
                      Kinds kd, const ZOut& zo, hipStream_t s, size_t lds, const Multi& mu) {
    if (ntask <= 0) return;
    if constexpr (XS >= 2) {
        hipLaunchKernelGGL((k_pw2pot3<N, XS, C, float2, NZ>), dim3(ntask), dim3(256), lds, s, g, t.pw2potS, t.npw,
                           tasks, reinterpret_cast<const float2*>(expn), t.pw2potS_ks, kd, zo, mu);
    } else {
        if (t.expn_f32)
            hipLaunchKernelGGL((k_pw2pot3<N, XS, C, float2, NZ>), dim3(ntask), dim3(256), lds, s, g, t.pw2potS,
                               t.npw, tasks, reinterpret_cast<const float2*>(expn), t.pw2potS_ks, kd, zo, mu);
        else
            hipLaunchKernelGGL((k_pw2pot3<N, XS, C, double2, NZ>), dim3(ntask), dim3(256), lds, s, g, t.pw2potS,
                               t.npw, tasks, expn, t.pw2potS_ks, kd, zo, mu);
    }
}

// components zo.kz[a..a+nz) of zo as a group of their own
ZOut zsub(const ZOut& zo, int a, int nz) {
    ZOut r{};
    r.nz = nz;
    for (int q = 0; q < nz; ++q) {
        r.kz[q] = zo.kz[a + q];
        r.od[q] = zo.od[a + q];
    }
    return r;
}

// tasks [0, nsib) hold up to four leaves and run once per z kind; [nsib,
// ntask) hold one leaf each and run once for all the kinds
template <int N, int XS>
void launch_pw2pot3_typed(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                          const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s, size_t lds,
                          const Multi& mu) {
    for (int q = 0; q < zo.nz; ++q) launch_pw2pot3_c<N, XS, 4>(g, t, nsib, tasks, expn, kd, zsub(zo, q, 1), s, lds, mu);
    Multi m1 = mu;
    if (m1.tprob) m1.tprob += nsib;
    const int n1 = ntask - nsib;
    const PotTask* t1 = tasks + nsib;
    if (zo.nz == 3) {
        launch_pw2pot3_c<N, XS, 1, 3>(g, t, n1, t1, expn, kd, zo, s, lds, m1);
    } else if (zo.nz == 2) {
        launch_pw2pot3_c<N, XS, 1, 2>(g, t, n1, t1, expn, kd, zo, s, lds, m1);
    } else {
        launch_pw2pot3_c<N, XS, 1, 1>(g, t, n1, t1, expn, kd, zo, s, lds, m1);
    }
}

template <int N>
bool try_pw2pot3(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                 const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s, size_t lds_limit, const Multi& mu) {
    const size_t lds = pw2pot3_lds(N, t.npw);
    if (lds > lds_limit) return false;
    if (3 * t.pw2potS_ks > 0x7fffffffULL) return false;
    const int xs = x5_variant(N, t);
    if (xs == 2) {
        // reduced-precision x stage (FP32 or BF16 hi/lo matrix cores; FP32 sff)
        const size_t lds32 = pw2pot3_lds(N, t.npw, 2);
        if (t.prec == 2) {
            launch_pw2pot3_typed<N, 3>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds32, mu);
            return true;
        }
        if constexpr (N == 8) {
            launch_pw2pot3_typed<N, 4>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds32, mu);
            return true;
        }
        launch_pw2pot3_typed<N, 2>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds32, mu);
        return true;
    }
    if constexpr (N >= 12) {
        if (xs == 1) {
            launch_pw2pot3_typed<N, 1>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds, mu);
            return true;
        }
    }
    launch_pw2pot3_typed<N, 0>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds, mu);
    return true;
}

// slabs per round of the fused STEP 1 kernel: four for norder 8 with the
// matrix-core x stage while sF (8 n npw complex) leaves room for three blocks
// per CU, else two
int p2pwF3_slabs(int n, const DevTables& t) {
    return (n == 8 && use_lowp1(t) && t.npw <= 48) ? 4 : 2;
}

template <int N>
void launch_p2pwF3_typed(const Geom& g, const DevTables& t, unsigned nblk, size_t lds, const int* plev,
                         const int* pslot, const int* kids, const int* ksg, const double* fvals, double2* expn,
                         hipStream_t s, const Multi& mu) {
    if (use_lowp1(t)) {
        float2* e32 = reinterpret_cast<float2*>(expn);
        if constexpr (N == 8) {
            if (p2pwF3_slabs(N, t) == 4) {
                if (t.prec == 2)
                    hipLaunchKernelGGL((k_p2pwF3<N, 3, float2, 4>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS,
                                       t.npw, plev, pslot, kids, ksg, fvals, e32, mu);
                else
                    hipLaunchKernelGGL((k_p2pwF3<N, 2, float2, 4>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS,
                                       t.npw, plev, pslot, kids, ksg, fvals, e32, mu);
                return;
            }
        }
        if (t.prec == 2)
            hipLaunchKernelGGL((k_p2pwF3<N, 3, float2>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS, t.npw, plev,
                               pslot, kids, ksg, fvals, e32, mu);
        else
            hipLaunchKernelGGL((k_p2pwF3<N, 2, float2>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS, t.npw, plev,
                               pslot, kids, ksg, fvals, e32, mu);
    } else if (t.expn_f32) {
        hipLaunchKernelGGL((k_p2pwF3<N, 0, float2>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS, t.npw, plev, pslot,
                           kids, ksg, fvals, reinterpret_cast<float2*>(expn), mu);
    } else {
        hipLaunchKernelGGL((k_p2pwF3<N, 0, double2>), dim3(nblk), dim3(256), lds, s, g, t.p2pwS, t.npw, plev,
                           pslot, kids, ksg, fvals, expn, mu);
    }
}

}  // namespace

// FGT_LOWP_STAGES (bit 0 STEP 1, bit 1 STEP 5): which stages run their x
// contraction on the reduced-precision matrix cores in the FP32 / BF16
// modes; the others keep the FP64 VALU contraction on FP32 expansions
// (tuning and comparison runs).
int lowp_stages(int prec) {
    static const int v = [] {
        const char* e = std::getenv("FGT_LOWP_STAGES");
        return e ? (std::atoi(e) & 3) : -1;
    }();
    // defaults: FP32 keeps STEP 1 on the VALU (its matrix-core form measured
    // slower: 50 x 50-class outputs pad to 64-row / 16-column tiles with K = 8
    // or 16); BF16 hi/lo runs both stages on the matrix cores
    return v >= 0 ? v : (prec == 2 ? 3 : 2);
}

// FGT_PW2POT_VALU=1 selects the FP64 VALU x stage of STEP 5 instead of the
// matrix-core one (for comparison runs).
bool pw2pot_valu_only() {
    static const bool v = [] {
        const char* e = std::getenv("FGT_PW2POT_VALU");
        return e && e[0] == '1';
    }();
    return v;
}

bool launch_p2pw_3d(const Geom& g, const DevTables& t, int ntask, const int* box, const int* lev,
                    const int* slot, const double* fvals, double2* expn, hipStream_t s, size_t lds_limit,
                    const Multi& mu) {
    if (g.ndim != 3 || ntask <= 0) return false;
    switch (t.n) {
        case 8: return try_p2pw3<8>(g, t, ntask, box, lev, slot, fvals, expn, s, lds_limit, mu);
        case 12: return try_p2pw3<12>(g, t, ntask, box, lev, slot, fvals, expn, s, lds_limit, mu);
        case 16: return try_p2pw3<16>(g, t, ntask, box, lev, slot, fvals, expn, s, lds_limit, mu);
        default: return false;
    }
}

bool p2pw_fused_3d(const Geom& g, const DevTables& t, int ntask, const int* plev, const int* pslot,
                   const int* kids, const int* ksg, const double* fvals, double2* expn, hipStream_t s,
                   const Multi& mu) {
    if (g.ndim != 3 || ntask <= 0) return ntask <= 0;
    const int N = t.n;
    const size_t sr = (size_t)p2pwF3_slabs(N, t);
    const size_t lds = sizeof(double2) * std::max(4 * sr * N * N, 2 * sr * N * t.npw);
    if (lds > t.lds_limit) return false;
    if ((long long)ntask * kFusedParts > 0x7fffffffLL) return false;
    const unsigned nblk = (unsigned)(ntask * kFusedParts);
    switch (N) {
        case 8: launch_p2pwF3_typed<8>(g, t, nblk, lds, plev, pslot, kids, ksg, fvals, expn, s, mu); return true;
        case 12: launch_p2pwF3_typed<12>(g, t, nblk, lds, plev, pslot, kids, ksg, fvals, expn, s, mu); return true;
        case 16: launch_p2pwF3_typed<16>(g, t, nblk, lds, plev, pslot, kids, ksg, fvals, expn, s, mu); return true;
        default: return false;
    }
}

bool pw2pot_3d_supported(const Geom& g, const DevTables& t) {
    if (g.ndim != 3 || (t.n != 8 && t.n != 12 && t.n != 16)) return false;
    const size_t lds = pw2pot3_lds(t.n, t.npw);
    const size_t lds1 = sizeof(double2) * std::max(4 * (size_t)t.n * t.n, 4 * (size_t)t.n * t.npw);
    const size_t lds2 = sizeof(double2) * std::max(8 * (size_t)t.n * t.n, 4 * (size_t)t.n * t.npw);
    return lds <= t.lds_limit && lds1 <= t.lds_limit && lds2 <= t.lds_limit;
}

bool launch_pw2pot_3d_z(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                        const double2* expn, Kinds kd, const ZOut& zo, hipStream_t s, size_t lds_limit,
                        const Multi& mu) {
    if (g.ndim != 3 || ntask <= 0 || zo.nz < 1 || zo.nz > 3) return false;
    switch (t.n) {
        case 8: return try_pw2pot3<8>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds_limit, mu);
        case 12: return try_pw2pot3<12>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds_limit, mu);
        case 16: return try_pw2pot3<16>(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds_limit, mu);
        default: return false;
    }
}

bool launch_pw2pot_3d(const Geom& g, const DevTables& t, int ntask, int nsib, const PotTask* tasks,
                      const double2* expn, Kinds kd, OutDesc od, hipStream_t s, size_t lds_limit, const Multi& mu) {
    ZOut zo{};
    zo.nz = 1;
    zo.kz[0] = kd.z;
    zo.od[0] = od;
    return launch_pw2pot_3d_z(g, t, ntask, nsib, tasks, expn, kd, zo, s, lds_limit, mu);
}

}  // namespace fgt
