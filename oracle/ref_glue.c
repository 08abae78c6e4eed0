/*
 * Test-infrastructure glue for the reference build in oracle/_ref/.
 *
 * This file is ours; it is compiled together with the unmodified Fortran
 * sources of flatironinstitute/fgt (by oracle/Makefile, into
 * oracle/_ref/libfgtref.so) so that tests can
 *   - build adaptive trees with the reference's vol_tree_mem/vol_tree_build
 *     (src/common/tree_vol_coeffs_nd.f:83,444) for a source made of
 *     anisotropic Gaussians, and
 *   - call the reference boxfgt (src/bfgt/boxfgt.f:5) on them.
 * It is the checker, never the thing measured or shipped.
 *
 * Source model (one density, nd = 1):
 *   f(y) = sum_g a_g * exp(-sum_k (y_k - c_gk)^2 / s_gk)
 * with y in [-1/2,1/2]^ndim.  With iperiod = 1 the sum also runs over the
 * 3^ndim nearest lattice images, so the source is smooth and periodic.
 * Parameter block dpars: per Gaussian, c[ndim], s[ndim], a  (2*ndim+1).
 * ipars(1) = ndim, ipars(2) = number of Gaussians, ipars(3) = iperiod.
 */
#include <math.h>
#include <complex.h>
#include <stddef.h>

/* Fortran entry points of the reference (default gfortran/flang mangling). */
extern void vol_tree_mem_(int *ndim, int *ipoly, int *iperiod, double *eps,
                          double _Complex *zk, double *boxlen, int *norder,
                          int *iptype, double *eta, void (*fun)(), int *nd,
                          double *dpars, double _Complex *zpars, int *ipars,
                          int *ifnewtree, int *nboxes, int *nlevels,
                          int *ltree, double *rintl);
extern void vol_tree_build_(int *ndim, int *ipoly, int *iperiod, double *eps,
                            double _Complex *zk, double *boxlen, int *norder,
                            int *iptype, double *eta, void (*fun)(), int *nd,
                            double *dpars, double _Complex *zpars, int *ipars,
                            double *rintl, int *nboxes, int *nlevels,
                            int *ltree, int *itree, int *iptr,
                            double *centers, double *boxsize, double *fvals);
extern void prini_(int *ip, int *iq);

void refglue_gauss_rhs_(int *nd, double *xyz, double *dpars,
                        double _Complex *zpars, int *ipars, double *f)
{
    (void)zpars;
    const int ndim = ipars[0], ng = ipars[1], iper = ipars[2];
    const int stride = 2 * ndim + 1;
    double acc = 0.0;
    int nimg = 1;
    for (int k = 0; k < ndim; ++k) nimg *= iper ? 3 : 1;
    for (int g = 0; g < ng; ++g) {
        const double *p = dpars + (size_t)g * stride;
        for (int im = 0; im < nimg; ++im) {
            int code = im;
            double rr = 0.0;
            for (int k = 0; k < ndim; ++k) {
                double shift = 0.0;
                if (iper) { shift = (double)(code % 3 - 1); code /= 3; }
                const double d = xyz[k] - p[k] - shift;
                rr += d * d / p[ndim + k];
            }
            acc += p[2 * ndim] * exp(-rr);
        }
    }
    for (int i = 0; i < *nd; ++i) f[i] = acc;
}

static double g_rintl[201];

void refglue_silence(void)
{
    int z = 0;
    prini_(&z, &z);
}

/* Sizes of the tree that vol_tree_build will produce (vol_tree_mem); with
 * ifnewtree = 1 the box capacity is three times the tree's box count
 * (tree_vol_coeffs_nd.f:432), room for boxfgt's refinement. */
void refglue_tree_mem_nt(int ndim, int ipoly, int iperiod, double eps,
                         int norder, int ng, double *dpars, int ifnewtree,
                         int *nboxes, int *nlevels, int *ltree)
{
    int ipars[3] = {ndim, ng, iperiod};
    double _Complex zk = 0.0, zpars[1] = {0.0};
    double boxlen = 1.0, eta = 0.0;
    int iptype = 2, nd = 1;
    refglue_silence();
    vol_tree_mem_(&ndim, &ipoly, &iperiod, &eps, &zk, &boxlen, &norder,
                  &iptype, &eta, (void (*)())refglue_gauss_rhs_, &nd, dpars,
                  zpars, ipars, &ifnewtree, nboxes, nlevels, ltree, g_rintl);
}

void refglue_tree_mem(int ndim, int ipoly, int iperiod, double eps,
                      int norder, int ng, double *dpars, int *nboxes,
                      int *nlevels, int *ltree)
{
    refglue_tree_mem_nt(ndim, ipoly, iperiod, eps, norder, ng, dpars, 0,
                        nboxes, nlevels, ltree);
}

/* Build the tree whose sizes the previous refglue_tree_mem call returned. */
void refglue_tree_build(int ndim, int ipoly, int iperiod, double eps,
                        int norder, int ng, double *dpars, int nboxes,
                        int nlevels, int ltree, int *itree, int *iptr,
                        double *centers, double *boxsize, double *fvals)
{
    int ipars[3] = {ndim, ng, iperiod};
    double _Complex zk = 0.0, zpars[1] = {0.0};
    double boxlen = 1.0, eta = 0.0;
    int iptype = 2, nd = 1;
    vol_tree_build_(&ndim, &ipoly, &iperiod, &eps, &zk, &boxlen, &norder,
                    &iptype, &eta, (void (*)())refglue_gauss_rhs_, &nd, dpars,
                    zpars, ipars, g_rintl, &nboxes, &nlevels, &ltree, itree,
                    iptr, centers, boxsize, fvals);
}

/* vol_tree_mem / vol_tree_build for a caller's source (the reference's
 * `fun` convention), refinement norm iptype, eta, nd densities. */
void refglue_tree_mem_fun(int ndim, int ipoly, int iperiod, double eps,
                          double zkr, int norder, int iptype, double eta, void (*fun)(),
                          int nd, double *dpars, int *ipars, int ifnewtree,
                          int *nboxes, int *nlevels, int *ltree)
{
    double _Complex zk = zkr, zpars[1] = {0.0};
    double boxlen = 1.0;
    refglue_silence();
    vol_tree_mem_(&ndim, &ipoly, &iperiod, &eps, &zk, &boxlen, &norder,
                  &iptype, &eta, fun, &nd, dpars, zpars, ipars, &ifnewtree,
                  nboxes, nlevels, ltree, g_rintl);
}

void refglue_tree_build_fun(int ndim, int ipoly, int iperiod, double eps,
                            double zkr, int norder, int iptype, double eta, void (*fun)(),
                            int nd, double *dpars, int *ipars, int nboxes,
                            int nlevels, int ltree, int *itree, int *iptr,
                            double *centers, double *boxsize, double *fvals)
{
    double _Complex zk = zkr, zpars[1] = {0.0};
    double boxlen = 1.0;
    vol_tree_build_(&ndim, &ipoly, &iperiod, &eps, &zk, &boxlen, &norder,
                    &iptype, &eta, fun, &nd, dpars, zpars, ipars, g_rintl,
                    &nboxes, &nlevels, &ltree, itree, iptr, centers, boxsize,
                    fvals);
}
