/*
 * fgt_host.h -- host-side planner pieces of libfgt_amd.so, exported so the
 * FP64 tables and lists that feed the kernels can be checked on a machine
 * without a GPU.  Each entry point replaces one reference routine (cited);
 * layouts follow the reference's Fortran arrays unless stated.
 */
#ifndef FGT_AMD_FGT_HOST_H
#define FGT_AMD_FGT_HOST_H
#ifdef __cplusplus
extern "C" {
#endif

/* fgtpwterms (src/common/fgtterms.f:72-141). */
void fgt_host_pwterms(double bsize, double delta, double eps, int iperiod,
                      double *pmax, int *npw);

/* legeexps / chebexps with itype = 2 (src/common/legeexps.f:140-222,
 * src/common/chebexps.f:12-180): x[n], w[n], u(n,n). */
void fgt_host_polyexps(int ipoly, int n, double *x, double *w, double *u);

/* besseljs3d for real z, scale = 1 (src/common/besseljs3d.f:5-168):
 * fjs[0..nterms]. */
void fgt_host_sph_bessel(int nterms, double z, double *fjs);

/* mk_poly2pw_tables (src/bfgt/bfgt_volrouts.f:8-275): tab(n,npw) complex
 * interleaved; and mk_pw2pgh_tables' potential table
 * (src/bfgt/bfgt_volrouts.f:339-401): tab(npw,n) complex interleaved. */
void fgt_host_poly2pw(int ipoly, int n, int npw, const double *ws,
                      const double *ts, double delta, double boxdim,
                      double *tab);
void fgt_host_pw2pot(int ipoly, int n, int npw, const double *ts,
                     double delta, double boxdim, double *tab);

/* mk_loctab_all with mrefinelev = 0 (src/bfgt/bfgt_volrouts.f:407-660):
 * tab(n,n,-6:6) with entries outside the significant band zeroed,
 * ind(2,n+1,-6:6) 1-based row bounds per column and the column range in
 * column n+1, as the reference's ind_loc. */
void fgt_host_loctab(double eps, int ipoly, int n, double delta,
                     double boxdim, double *tab, int *ind);

/* mk_loctab_all with mrefinelev = mref (0 or 3; boxfgt.f:601-607 uses 3
 * when the tree is refined after the solve): tab(n,n,-m:m), ind(2,n+1,-m:m),
 * m = 2^(mref+1) (mref+3); offsets no geometric case produces have an empty
 * band (zero rows, column range 0..-1). */
void fgt_host_loctab_mref(double eps, int ipoly, int n, double delta,
                          double boxdim, int mref, double *tab, int *ind);

/* mk_pw2pgh_tables (src/bfgt/bfgt_volrouts.f:339-401): tab_pw2pot,
 * tab_pw2der, tab_pw2dxx, each (npw,n) complex interleaved. */
void fgt_host_pw2pgh(int ipoly, int n, int npw, const double *ts,
                     double delta, double boxdim, double *tab, double *tabx,
                     double *tabxx);
/* mk_loctab_all's derivative tables tabx_loc / tabxx_loc
 * (src/bfgt/bfgt_volrouts.f:407-660), layout of fgt_host_loctab, zeroed
 * outside the same band as tab. */
void fgt_host_loctab_der(double eps, int ipoly, int n, double delta,
                         double boxdim, double *tabx, double *tabxx);

/* Plan summary for a tree (boxfgt.f:132-162 cutoff rules, list sizes),
 * computed on the host only. */
typedef struct {
    int npwlevel, npw, nexp, nlevstart, nslots;
    int periodic_small, all_pw;
    long long nleaves, npw_leaves, ndirect_pairs;
    long long n_m2l_pairs, n_parent_child;
    long long n_m2l_targets;   /* cutoff-level boxes receiving a local expansion */
} fgt_plan_info;

int fgt_plan_summary(int nd, int ndim, double delta, double eps, int ipoly,
                     int iperiod, int norder, int nboxes, int nlevels,
                     int ltree, const int *itree, const int *iptr,
                     const double *centers, const double *boxsize,
                     fgt_plan_info *out);

/* Direct-interaction list of the plan, flattened: for pair p, target box
 * tgt[p] and source box src[p] (1-based), table offsets per axis ixyz[3p..]
 * in -6..6 (gnd_compute_modified_list1 + gnd_find_loctab_ind,
 * src/common/tree_routs_nd.f:864-1055, src/bfgt/bfgt_local.f:1624-1660),
 * pairs with an empty band dropped as the reference's early return does.
 * Call with cap = 0 to get the count. */
long long fgt_plan_direct_pairs(int nd, int ndim, double delta, double eps,
                                int ipoly, int iperiod, int norder,
                                int nboxes, int nlevels, int ltree,
                                const int *itree, const int *iptr,
                                const double *centers, const double *boxsize,
                                long long cap, int *tgt, int *src, int *ixyz);

#ifdef __cplusplus
}
#endif
#endif
