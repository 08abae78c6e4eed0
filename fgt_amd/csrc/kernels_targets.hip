// kernels_targets.hip -- extra-target evaluation (STEP 11, ifpghtarg >= 1).
//
// Values on the leaves -> orthogonal-polynomial coefficients
// (treedata_trans_nd, src/common/tree_data_routs_nd.f:16-95) -> values at
// arbitrary targets (treedata_evalt_nd, tree_data_routs_nd.f:582-686).
// The host finds each target's leaf (the reference's root-to-leaf descent)
// and groups targets by leaf; one block evaluates up to 256 targets of one
// leaf with that leaf's coefficients staged in LDS, where every thread reads
// the same entries (broadcast).
#include "kernels.h"

namespace fgt {

namespace {

// coefs[b][q][i] = sum U(i1,j1) U(i2,j2) U(i3,j3) vals[b][j][q]  (x, then y, then z)
// vals(ncd, npbox, nboxes) as the caller's arrays; U(k, j) = u[k + j*n].
__global__ void __launch_bounds__(256) k_vals2coefs(int ndim, int n, int ncd, const double* __restrict__ u,
                                                    const int* __restrict__ leaves,
                                                    const double* __restrict__ vals,
                                                    double* __restrict__ coefs) {
    extern __shared__ double sm[];
    int npbox = 1;
    for (int a = 0; a < ndim; ++a) npbox *= n;
    double* sU = sm;               // n*n
    double* c0 = sU + n * n;       // npbox
    double* c1 = c0 + npbox;       // npbox
    const int b = leaves[blockIdx.x];
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) sU[i] = u[i];
    for (int q = 0; q < ncd; ++q) {
        __syncthreads();
        for (int i = threadIdx.x; i < npbox; i += blockDim.x) c0[i] = vals[((size_t)b * npbox + i) * ncd + q];
        int stride = 1;
        double* src = c0;
        double* dst = c1;
        for (int a = 0; a < ndim; ++a) {
            __syncthreads();
            for (int i = threadIdx.x; i < npbox; i += blockDim.x) {
                const int ka = (i / stride) % n, base = i - ka * stride;
                double dd = 0.0;
                for (int k1 = 0; k1 < n; ++k1) dd += sU[ka + k1 * n] * src[base + k1 * stride];
                dst[i] = dd;
            }
            double* t = src;
            src = dst;
            dst = t;
            stride *= n;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < npbox; i += blockDim.x) coefs[((size_t)b * ncd + q) * npbox + i] = src[i];
    }
}

__device__ __forceinline__ void polys(int ipoly, double x, int n, double* p) {
    p[0] = 1.0;
    if (n == 1) return;
    p[1] = x;
    if (ipoly == 0) {
        double pkm1, pk = 1.0, pkp1 = x;
        for (int k = 1; k <= n - 2; ++k) {
            pkm1 = pk;
            pk = pkp1;
            pkp1 = ((2 * k + 1) * x * pk - k * pkm1) / (k + 1);
            p[k + 1] = pkp1;
        }
    } else {
        for (int k = 2; k < n; ++k) p[k] = 2 * x * p[k - 1] - p[k - 2];
    }
}

// task = (leaf box, first sorted target, count <= blockDim.x)
__global__ void __launch_bounds__(256) k_evalt(int ndim, int n, int ipoly, int ncd, const int* __restrict__ tbox,
                                               const int* __restrict__ tfirst, const int* __restrict__ tcount,
                                               const double* __restrict__ centers,
                                               const double* __restrict__ bsize_of_box,
                                               const double* __restrict__ targs_sorted,
                                               const int* __restrict__ perm, const double* __restrict__ coefs,
                                               double* __restrict__ out) {
    extern __shared__ double sc[];   // npbox coefficients of one (leaf, q)
    int npbox = 1;
    for (int a = 0; a < ndim; ++a) npbox *= n;
    const int b = tbox[blockIdx.x], first = tfirst[blockIdx.x], cnt = tcount[blockIdx.x];
    const bool act = (int)threadIdx.x < cnt;
    const int it = first + (act ? (int)threadIdx.x : 0);
    double p[3][16];
    const double s = 2.0 / bsize_of_box[b];
    for (int a = 0; a < 3; ++a) {
        const double x = a < ndim ? (targs_sorted[(size_t)it * ndim + a] - centers[(size_t)b * ndim + a]) * s : 0.0;
        polys(ipoly, x, n, p[a]);
    }
    for (int q = 0; q < ncd; ++q) {
        __syncthreads();
        for (int i = threadIdx.x; i < npbox; i += blockDim.x) sc[i] = coefs[((size_t)b * ncd + q) * npbox + i];
        __syncthreads();
        double v;
        if (ndim == 1) {
            v = 0.0;
            for (int k = 0; k < n; ++k) v += sc[k] * p[0][k];
        } else if (ndim == 2) {
            v = 0.0;
            for (int j = 0; j < n; ++j) {
                double t = 0.0;
                for (int k = 0; k < n; ++k) t += sc[k + n * j] * p[0][k];
                v += t * p[1][j];
            }
        } else {
            v = 0.0;
            for (int i = 0; i < n; ++i) {
                double t2 = 0.0;
                for (int j = 0; j < n; ++j) {
                    double t = 0.0;
                    for (int k = 0; k < n; ++k) t += sc[k + n * (j + n * i)] * p[0][k];
                    t2 += t * p[1][j];
                }
                v += t2 * p[2][i];
            }
        }
        if (act) out[(size_t)perm[it] * ncd + q] = v;
    }
}

}  // namespace

void launch_vals2coefs(int ndim, int n, int ncd, const double* u, int nleaf, const int* leaves, const double* vals,
                       double* coefs, hipStream_t s) {
    if (nleaf <= 0) return;
    int npbox = 1;
    for (int a = 0; a < ndim; ++a) npbox *= n;
    const size_t lds = sizeof(double) * ((size_t)n * n + 2 * (size_t)npbox);
    hipLaunchKernelGGL(k_vals2coefs, dim3(nleaf), dim3(256), lds, s, ndim, n, ncd, u, leaves, vals, coefs);
}

void launch_evalt(int ndim, int n, int ipoly, int ncd, int ntask, const int* tbox, const int* tfirst,
                  const int* tcount, const double* centers, const double* bsize_of_box, const double* targs_sorted,
                  const int* perm, const double* coefs, double* out, hipStream_t s) {
    if (ntask <= 0) return;
    int npbox = 1;
    for (int a = 0; a < ndim; ++a) npbox *= n;
    hipLaunchKernelGGL(k_evalt, dim3(ntask), dim3(256), sizeof(double) * npbox, s, ndim, n, ipoly, ncd, tbox,
                       tfirst, tcount, centers, bsize_of_box, targs_sorted, perm, coefs, out);
}

}  // namespace fgt
