// kernels_direct3.hip -- 3-D near-field (STEP 6) kernel for gfx950 (FP64).
//
// Replaces the per-pair banded contraction of g3d_tens_prod_to_potloc_fast
// (src/bfgt/bfgt_local.f:1322-1618) summed over a target's list-1 sources
// (boxfgt.f:904-928).  Sources of one target are pre-sorted on the host by
// their (z, y, x) table slots; since the near-field operator is the
// separable product A_z (x) A_y (x) A_x, sources that share the y and z
// tables are summed after their x contraction, and y-results that share the
// z table are summed before it:
//     pot += sum_z A_z ( sum_{y | z} A_y ( sum_{x | y,z} A_x f_s ) )
// which needs one x contraction per source, one y contraction per (y,z)
// group and one z contraction per z group.
// One block of N^2 threads (rounded up to whole waves: 64 / 192 / 256 for
// n = 8 / 12 / 16) per target leaf; the 1-D tables (masked to their
// significant band, zero elsewhere) are wave-uniform and read into SGPRs;
// the partial sums live in registers between LDS transposes.
#include <cstdlib>

#include "kernels.h"

namespace fgt {

namespace {

// Task order: runs of eight consecutive targets per XCD (neighbours of one
// problem, whose list-1 sources overlap, read them through one L2) or the
// launch order.  Batches take the runs (their many small near-field-only
// problems: 1.1-1.8 vs 4.4 ms for a norder-16 class), single problems the
// launch order (0.71 vs 0.88 ms on a norder-16 periodic problem);
// FGT_DIRECT_RUNS=0/1 forces one.
int direct_runs(const Multi& mu) {
    static const int v = [] {
        const char* e = std::getenv("FGT_DIRECT_RUNS");
        return e ? std::atoi(e) : -1;
    }();
    return v >= 0 ? v : (mu.pp ? 1 : 0);
}

// flags per source entry: bit0 = last of its (y,z) group, bit1 = last of its z group;
// nls table slots per level, slot codes 8 bits per axis
template <int N>
constexpr int direct3_block() { return (N * N + 63) / 64 * 64; }

// NZ output components that differ only in their z kind (ZOut) share the x
// and y sums; NZ = 1 is the one-component kernel (od = zo.od[0]).
template <int N, int NZ>
__global__ void __launch_bounds__(direct3_block<N>(), N == 16 ? (NZ == 1 ? 3 : 2) : 2)
    k_direct3(Geom g, const double* __restrict__ loctab_a, int nls, const int* __restrict__ tgt,
              const int* __restrict__ lev, const int* __restrict__ off, const int* __restrict__ src,
              const int* __restrict__ code, const int* __restrict__ flags, const double* __restrict__ fvals_a,
              size_t ks, Kinds kd, ZOut zo, Multi mu, int runs) {
    extern __shared__ double dsm[];   // N^3 doubles (transpose buffer)
    constexpr int NN = N * N;
    const int task = runs ? xcd_run_task<8>(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const double* __restrict__ loctab = loctab_a;
    const double* __restrict__ fvals = fvals_a;
    if (mu.pp) {
        const ProbPtrs& q = mu.pp[mu.tprob[task]];
        loctab = q.loctab;
        fvals = q.fvals;
        zo.od[0].p = q.pot;   // batches evaluate the potential only
    }
    const int tb = tgt[task], l = lev[task];
    // table kind per axis (value / first / second derivative): stride ks
    const CTab1 tabx = ctab(loctab) + (long long)(kd.x * ks + (size_t)l * nls * NN);
    const CTab1 taby = ctab(loctab) + (long long)(kd.y * ks + (size_t)l * nls * NN);
    CTab1 tabz[NZ];
#pragma unroll
    for (int q = 0; q < NZ; ++q) tabz[q] = ctab(loctab) + (long long)(zo.kz[q] * ks + (size_t)l * nls * NN);
    const int tid = threadIdx.x;
    const bool act = tid < NN;
    // x stage: thread = (j2, j3); y stage: thread = (k1, j3); z stage: thread = (k1, k2)
    const int a = tid % N, bq = tid / N;
    const int nd = g.nd;
    const int j0 = off[task], j1e = off[task + 1];
    for (int ind = 0; ind < nd; ++ind) {
        double X[N], Y[N], acc[NZ][N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            X[i] = Y[i] = 0.0;
#pragma unroll
            for (int q = 0; q < NZ; ++q) acc[q][i] = 0.0;
        }
        for (int j = j0; j < j1e; ++j) {
            const int sb = src[j], c = code[j], fl = flags[j];
            const CTab1 Ax = tabx + (long long)(c & 255) * NN;
            // ---- x: X[k1] += sum_j1 Ax[k1][j1] f[j1, j2=a, j3=bq]
            if (act) {
                double f[N];
                const double* fs = fvals + ((size_t)sb * g.npbox + (size_t)N * tid) * nd + ind;
                if (nd == 1) {
                    // one density: the thread's N values are contiguous and
                    // 16-byte aligned; read them as pairs
                    const double2* fs2 = reinterpret_cast<const double2*>(fs);
#pragma unroll
                    for (int i = 0; i < N / 2; ++i) {
                        const double2 v = fs2[i];
                        f[2 * i] = v.x;
                        f[2 * i + 1] = v.y;
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < N; ++i) f[i] = fs[(size_t)i * nd];
                }
#pragma unroll
                for (int k1 = 0; k1 < N; ++k1) {
                    double s = X[k1];
#pragma unroll
                    for (int i = 0; i < N; ++i) s += Ax[k1 * N + i] * f[i];
                    X[k1] = s;
                }
            }
            if (fl & 1) {
                // ---- y: Y[k2] += sum_j2 Ay[k2][j2] X[k1, j2, j3]
                const CTab1 Ay = taby + (long long)((c >> 8) & 255) * NN;
                __syncthreads();
                if (act) {
#pragma unroll
                    for (int k1 = 0; k1 < N; ++k1) dsm[(bq * N + a) * N + k1] = X[k1];   // [j3][j2][k1]
                }
                __syncthreads();
                if (act) {
                    double xv[N];   // X[k1=a, j2, j3=bq]
#pragma unroll
                    for (int j2 = 0; j2 < N; ++j2) xv[j2] = dsm[(bq * N + j2) * N + a];
#pragma unroll
                    for (int k2 = 0; k2 < N; ++k2) {
                        double s = Y[k2];
#pragma unroll
                        for (int j2 = 0; j2 < N; ++j2) s += Ay[k2 * N + j2] * xv[j2];
                        Y[k2] = s;
                    }
                }
#pragma unroll
                for (int i = 0; i < N; ++i) X[i] = 0.0;
                if (fl & 2) {
                    // ---- z: acc[q][k3] += sum_j3 Az_q[k3][j3] Y[k1, k2, j3]
                    __syncthreads();
                    if (act) {
#pragma unroll
                        for (int k2 = 0; k2 < N; ++k2) dsm[(bq * N + k2) * N + a] = Y[k2];   // [j3][k2][k1]
                    }
                    __syncthreads();
                    if (act) {
                        double yv[N];   // Y[k1=a, k2=bq, j3]
#pragma unroll
                        for (int j3 = 0; j3 < N; ++j3) yv[j3] = dsm[(j3 * N + bq) * N + a];
#pragma unroll
                        for (int q = 0; q < NZ; ++q) {
                            const CTab1 Az = tabz[q] + (long long)((c >> 16) & 255) * NN;
#pragma unroll
                            for (int k3 = 0; k3 < N; ++k3) {
                                double s = acc[q][k3];
#pragma unroll
                                for (int j3 = 0; j3 < N; ++j3) s += Az[k3 * N + j3] * yv[j3];
                                acc[q][k3] = s;
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < N; ++i) Y[i] = 0.0;
                }
            }
        }
        if (act) {
#pragma unroll
            for (int q = 0; q < NZ; ++q) {
                const OutDesc& od = zo.od[q];
#pragma unroll
                for (int k3 = 0; k3 < N; ++k3)
                    od.p[((size_t)tb * g.npbox + a + N * (bq + N * k3)) * od.ns + od.coff + ind] += acc[q][k3];
            }
        }
        __syncthreads();
    }
}

template <int N>
bool try_direct3(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev, const int* off,
                 const int* src, const int* code, const int* flags, const double* fvals, Kinds kd, const ZOut& zo,
                 hipStream_t s, const Multi& mu) {
    const size_t lds = sizeof(double) * (size_t)N * N * N;
    if (lds > t.lds_limit) return false;
    const dim3 grid(ntask), block(direct3_block<N>());
    const int runs = direct_runs(mu);
    switch (zo.nz) {
        case 1:
            hipLaunchKernelGGL((k_direct3<N, 1>), grid, block, lds, s, g, t.loctab, t.loc_slots, tgt, lev, off, src,
                               code, flags, fvals, t.loctab_ks, kd, zo, mu, runs);
            return true;
        case 2:
            hipLaunchKernelGGL((k_direct3<N, 2>), grid, block, lds, s, g, t.loctab, t.loc_slots, tgt, lev, off, src,
                               code, flags, fvals, t.loctab_ks, kd, zo, mu, runs);
            return true;
        case 3:
            if constexpr (N == 16) {
                // three z kinds exceed the registers at norder 16 (spills):
                // two passes, the first sharing two of them
                ZOut a = zo, b{};
                a.nz = 2;
                b.nz = 1;
                b.kz[0] = zo.kz[2];
                b.od[0] = zo.od[2];
                hipLaunchKernelGGL((k_direct3<N, 2>), grid, block, lds, s, g, t.loctab, t.loc_slots, tgt, lev, off,
                                   src, code, flags, fvals, t.loctab_ks, kd, a, mu, runs);
                hipLaunchKernelGGL((k_direct3<N, 1>), grid, block, lds, s, g, t.loctab, t.loc_slots, tgt, lev, off,
                                   src, code, flags, fvals, t.loctab_ks, kd, b, mu, runs);
            } else {
                hipLaunchKernelGGL((k_direct3<N, 3>), grid, block, lds, s, g, t.loctab, t.loc_slots, tgt, lev, off,
                                   src, code, flags, fvals, t.loctab_ks, kd, zo, mu, runs);
            }
            return true;
        default:
            return false;
    }
}

bool direct3_covered(const Geom& g, const DevTables& t, int ntask, const int* flags) {
    return g.ndim == 3 && ntask > 0 && flags && (t.n == 8 || t.n == 12 || t.n == 16);
}

bool direct3_dispatch(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev, const int* off,
                      const int* src, const int* code, const int* flags, const double* fvals, Kinds kd,
                      const ZOut& zo, hipStream_t s, const Multi& mu) {
    switch (t.n) {
        case 8: return try_direct3<8>(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, zo, s, mu);
        case 12: return try_direct3<12>(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, zo, s, mu);
        case 16: return try_direct3<16>(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, zo, s, mu);
        default: return false;
    }
}

}  // namespace

bool launch_direct_3d(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev,
                      const int* off, const int* src, const int* code, const int* flags,
                      const double* fvals, Kinds kd, OutDesc od, hipStream_t s, const Multi& mu) {
    if (!direct3_covered(g, t, ntask, flags)) return false;
    ZOut zo{};
    zo.nz = 1;
    zo.kz[0] = kd.z;
    zo.od[0] = od;
    return direct3_dispatch(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, zo, s, mu);
}

bool launch_direct_3d_z(const Geom& g, const DevTables& t, int ntask, const int* tgt, const int* lev,
                        const int* off, const int* src, const int* code, const int* flags,
                        const double* fvals, Kinds kd, const ZOut& zo, hipStream_t s) {
    if (!direct3_covered(g, t, ntask, flags) || zo.nz < 1 || zo.nz > 3) return false;
    return direct3_dispatch(g, t, ntask, tgt, lev, off, src, code, flags, fvals, kd, zo, s, Multi{});
}

}  // namespace fgt
