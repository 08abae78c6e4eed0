/*
 * fgt_oracle.h -- CPU restatement of the reference box FGT (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load liboracle.so, and only as the checker
 * (or the timed CPU baseline), never as part of the product path.
 *
 * It restates, in plain C (FP64, OpenMP over boxes of one level), the
 * reference algorithm of flatironinstitute/fgt:
 *   boxfgt / bfgtmain STEPs 0-6         src/bfgt/boxfgt.f:5-280, 285-932
 *   plane-wave kernels                  src/bfgt/bfgt_pwrouts.f
 *   1-D table builders                  src/bfgt/bfgt_volrouts.f
 *   banded near-field contractions      src/bfgt/bfgt_local.f
 *   term counts / PW nodes              src/common/fgtterms.f
 *   list builders                       src/common/tree_routs_nd.f
 *   nodes, weights, u matrices          src/common/legeexps.f, chebexps.f
 *   spherical Bessel functions          src/common/besseljs3d.f
 * Each function in fgt_oracle.c cites the file:line it follows.
 */
#ifndef FGT_ORACLE_H
#define FGT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Plan scalars reported by oracle_boxfgt (for tests). */
typedef struct {
    int npwlevel;   /* cutoff level after the reference's own rules        */
    int npw;        /* 1-D plane-wave count                                 */
    int nlevstart;  /* max(npwlevel,0), or 0 on the periodic small-level path */
    int nexp;       /* complex terms per expansion per density              */
    int periodic_small; /* 1 when the periodic npwlevel<=1 branch was taken */
    double pmax;
} oracle_info;

/*
 * Box FGT, potential only (ifpgh = 1), no extra targets, ifnewtree = 0.
 * Arguments follow boxfgt (src/bfgt/boxfgt.f:5-8) with Fortran layouts:
 *   itree/iptr: reference tree arrays (1-based ids), centers(ndim,nboxes),
 *   boxsize(0:nlevels), fvals(nd,npbox,nboxes), pot(nd,npbox,nboxes) (+=).
 * d1_mode: 0 = follow the reference exactly, including its behaviour when
 *   npwlevel <= 0 (free space) or the periodic npwlevel <= 1 branch, where
 *   only the root is flagged for PW expansions (src/common/tree_routs_nd.f:
 *   1327-1330) and leaf potentials are never formed;
 *   1 = the behaviour described at src/bfgt/boxfgt.f:485-497: every box forms
 *   and evaluates its PW expansion on that path.
 * timeinfo (nullable, >= 8 doubles): seconds for STEP 0..6 in [0..6].
 * Returns 0 on success, nonzero on invalid input / allocation failure.
 */
int oracle_boxfgt(int nd, int ndim, double delta, double eps, int ipoly,
                  int iperiod, int norder, int nboxes, int nlevels,
                  const int *itree, const int *iptr, const double *centers,
                  const double *boxsize, const double *fvals, double *pot,
                  int d1_mode, double *timeinfo, oracle_info *info);

/*
 * Same with gradients and Hessians (ifpgh = 2: pot + grad, 3: + hess),
 * grad(nd,ndim,npbox,nboxes) and hess(nd,ndim*(ndim+1)/2,npbox,nboxes)
 * accumulated like pot (Hessian order 1-D xx; 2-D xx,xy,yy; 3-D
 * xx,yy,zz,xy,xz,yz, as src/bfgt/bfgt_pwrouts.f:612-1065).
 */
int oracle_boxfgt_pgh(int nd, int ndim, double delta, double eps, int ipoly,
                      int iperiod, int norder, int nboxes, int nlevels,
                      const int *itree, const int *iptr, const double *centers,
                      const double *boxsize, const double *fvals, int ifpgh,
                      double *pot, double *grad, double *hess,
                      int d1_mode, double *timeinfo, oracle_info *info);

/*
 * STEP 11 for ifpghtarg >= 1 (src/bfgt/boxfgt.f:1343-1378): values on the
 * leaves (nd,npbox,nboxes) -> polynomial coefficients -> values at nt
 * targets targ(ndim,nt), out(nd,nt) overwritten.  For grad/hess call with
 * nd*ndim / nd*nhess densities, as the reference does.
 */
int oracle_evalt(int ndim, int nd, int ipoly, int norder, int nboxes, int nlevels,
                 const int *itree, const int *iptr, const double *centers,
                 const double *boxsize, const double *vals, int nt,
                 const double *targ, double *out);

/* ---- exposed pieces, so tests can pin them one by one ---- */

/* fgtpwterms (src/common/fgtterms.f:72-141). */
void oracle_pwterms(double bsize, double delta, double eps, int iperiod,
                    double *pmax, int *npw);
/* get_pwnodes itype=0 (fgtterms.f:151-179) / get_periodic_pwnodes (191-211). */
void oracle_pwnodes(double pmax, int npw, double *ws, double *ts);
void oracle_periodic_pwnodes(double bs0, double delta, int npw, double *ws,
                             double *ts);
/* legeexps/chebexps itype=2: nodes x[n], weights w[n], u[n*n] (Fortran u(i,j)
 * at u[i + j*n]: coefficient i from value j). */
void oracle_polyexps(int ipoly, int n, double *x, double *w, double *u);
/* spherical Bessel j_0..j_nterms at real z (besseljs3d.f:5-168, scale=1). */
void oracle_sph_bessel(int nterms, double z, double *fjs);
/* tab_poly2pw(n,npw) complex, interleaved re/im (bfgt_volrouts.f:8-275). */
void oracle_poly2pw(int ipoly, int n, int npw, const double *ws,
                    const double *ts, double delta, double boxdim,
                    double *tab);
/* tab_pw2pot(npw,n) complex interleaved (bfgt_volrouts.f:339-401). */
void oracle_pw2pot(int ipoly, int n, int npw, const double *ts, double delta,
                   double boxdim, double *tab);
/* mk_loctab_all with mrefinelev = 0 (bfgt_volrouts.f:407-660):
 * tab[n,n,13] (k = -6..6 at slot k+6, Fortran tab(i,j,k) at i + j*n + slot*n*n),
 * ind[2,n+1,13] likewise (1-based row/col bounds, as the reference). */
void oracle_loctab(double eps, int ipoly, int n, double delta, double boxdim,
                   double *tab, int *ind);
/* Same plus the first/second derivative tables (nullable), same layout. */
void oracle_loctab3(double eps, int ipoly, int n, double delta, double boxdim,
                    double *tab, double *tabx, double *tabxx, int *ind);
/* tab_pw2pot, tab_pw2der, tab_pw2dxx (npw,n) complex interleaved
 * (mk_pw2pgh_tables, bfgt_volrouts.f:339-401). */
void oracle_pw2pgh(int ipoly, int n, int npw, const double *ts, double delta,
                   double boxdim, double *tab, double *tabx, double *tabxx);

#ifdef __cplusplus
}
#endif
#endif
