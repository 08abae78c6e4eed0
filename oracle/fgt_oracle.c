/*
 * fgt_oracle.c -- CPU restatement of the reference box FGT (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see fgt_oracle.h).  Plain C11 + OpenMP, FP64.
 * Box ids are kept 1-based internally, as in the reference tree arrays, and
 * every array keeps the reference's Fortran layout so that the steps read
 * like the routines they restate.  Complex numbers are C99 double complex.
 */
#include "fgt_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* ------------------------------------------------------------------ */
/* fgtterms.f                                                          */
/* ------------------------------------------------------------------ */

/* fgtpwterms, src/common/fgtterms.f:72-141. */
void oracle_pwterms(double bsize, double delta, double eps, int iperiod,
                    double *pmax, int *npw)
{
    const double pi = 4.0 * atan(1.0);
    if (iperiod == 0) {
        double rmax = bsize / sqrt(delta);
        double d = sqrt(log(1.0 / eps));
        *pmax = 2.0 * d;
        double h = 2 * pi / (rmax + d);
        int m = (int)(*pmax / h) + 1;
        *npw = 2 * m;
    } else {
        *npw = 2 * ((int)(bsize * sqrt(log(1.0 / eps) / delta) / pi) + 1) + 1;
    }
}

/* get_pwnodes with itype = 0 (midpoint rule), fgtterms.f:151-179.
 * Slot i = j + npw/2 holds node j = -npw/2 .. npw/2-1. */
void oracle_pwnodes(double pmax, int npw, double *ws, double *ts)
{
    const double pi = 4.0 * atan(1.0);
    int npw2 = npw / 2;
    double h = pmax / npw2;
    double w = h / (2.0 * sqrt(pi));
    for (int j = -npw2; j <= npw2 - 1; ++j) {
        double t = (j + 0.5) * h;
        ts[j + npw2] = t;
        ws[j + npw2] = w * exp(-t * t / 4);
    }
}

/* get_periodic_pwnodes, fgtterms.f:191-211. */
void oracle_periodic_pwnodes(double bs0, double delta, int npw, double *ws,
                             double *ts)
{
    const double pi = 4.0 * atan(1.0);
    int npw2 = npw / 2;
    double h = 2 * pi / bs0;
    double w = sqrt(pi * delta) / bs0;
    double hw = h * sqrt(delta) / 2;
    for (int j = -npw2; j <= (npw - 1) / 2; ++j) {
        ts[j + npw2] = j * h;
        ws[j + npw2] = w * exp(-(hw * j) * (hw * j));
    }
}

/* ------------------------------------------------------------------ */
/* legeexps.f / chebexps.f                                             */
/* ------------------------------------------------------------------ */

/* Legendre P_0..P_{n-1} at x by the three-term recurrence
 * (legepols, src/common/legeexps.f:597-631). */
static void legepols(double x, int nm1, double *p)
{
    p[0] = 1.0;
    if (nm1 == 0) return;
    p[1] = x;
    double pkm1, pk = 1.0, pkp1 = x;
    for (int k = 1; k <= nm1 - 1; ++k) {
        pkm1 = pk;
        pk = pkp1;
        pkp1 = ((2 * k + 1) * x * pk - k * pkm1) / (k + 1);
        p[k + 1] = pkp1;
    }
}

/* Chebyshev T_0..T_{n-1} at x (chebpols, src/common/chebexps.f:597-632). */
static void chebpols(double x, int nm1, double *p)
{
    p[0] = 1.0;
    if (nm1 == 0) return;
    p[1] = x;
    for (int k = 2; k <= nm1; ++k) p[k] = 2 * x * p[k - 1] - p[k - 2];
}

/* P_n(x) and P_n'(x). */
static void legendre_pd(int n, double x, double *pn, double *dpn)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
    for (int k = 1; k < n; ++k) {
        double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *dpn = n * (x * p1 - p0) / (x * x - 1.0);
}

/* Gauss-Legendre nodes (increasing) and weights by Newton's method on P_n;
 * the reference computes the same rule with legerts (legeexps.f:2399-2530). */
static void gauss_legendre(int n, double *x, double *w)
{
    const double pi = 4.0 * atan(1.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        /* i-th root from the top */
        double z = cos(pi * (i + 0.75) / (n + 0.5));
        double pn, dpn;
        for (int it = 0; it < 100; ++it) {
            legendre_pd(n, z, &pn, &dpn);
            double dz = pn / dpn;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        legendre_pd(n, z, &pn, &dpn);
        double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

/* legeexps / chebexps with itype = 2 (legeexps.f:140-222, chebexps.f:12-180):
 * nodes, weights and u(i,j) = coefficient i of the interpolant from value j. */
void oracle_polyexps(int ipoly, int n, double *x, double *w, double *u)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    if (ipoly == 0) {
        gauss_legendre(n, x, w);
        for (int j = 0; j < n; ++j) {
            legepols(x[j], n - 1, p);
            for (int i = 0; i < n; ++i)
                u[i + (size_t)j * n] = p[i] * w[j] * (2 * i + 1) / 2.0;
        }
    } else {
        const double pi = 4.0 * atan(1.0);
        double h = pi / (2 * n);
        for (int i = 1; i <= n; ++i) x[n - i] = cos((2 * i - 1) * h);
        for (int j = 0; j < n; ++j) {
            chebpols(x[j], n - 1, p);
            /* Fejer weights, as chebexps.f:1350-1402 accumulates them */
            double wt = 2.0;
            for (int k = 2; k <= n - 1; k += 2)
                wt -= 4.0 / ((double)k * k - 1.0) * p[k];
            w[j] = wt / n;
            for (int i = 0; i < n; ++i)
                u[i + (size_t)j * n] = (i == 0 ? 1.0 : 2.0) / n * p[i];
        }
    }
    free(p);
}

/* ------------------------------------------------------------------ */
/* besseljs3d.f                                                        */
/* ------------------------------------------------------------------ */

/* Spherical Bessel j_0..j_nterms at real z >= 0 by Miller's downward
 * recurrence normalised with j_0 or j_1 (besseljs3d.f:5-168, scale = 1). */
void oracle_sph_bessel(int nterms, double z, double *fjs)
{
    const double upbound2 = 1.0e40, upbound2inv = 1.0e-40;
    if (fabs(z) < 1.0e-200) {
        fjs[0] = 1.0;
        for (int i = 1; i <= nterms; ++i) fjs[i] = 0.0;
        return;
    }
    double zinv = 1.0 / z;
    double fjm1 = 0.0, fj0 = 1.0, fj1;
    int lwfjs = nterms + 100000, ntop = lwfjs;
    for (int i = nterms; i <= lwfjs; ++i) {
        fj1 = (2 * i + 1.0) * zinv * fj0 - fjm1;
        if (fj1 * fj1 > upbound2) { ntop = i + 1; break; }
        fjm1 = fj0;
        fj0 = fj1;
    }
    double *t = (double *)calloc((size_t)ntop + 1, sizeof(double));
    char *isc = (char *)calloc((size_t)ntop + 1, 1);
    t[ntop] = 0.0;
    t[ntop - 1] = 1.0;
    for (int i = ntop - 1; i >= 1; --i) {
        t[i - 1] = (2 * i + 1.0) * zinv * t[i] - t[i + 1];
        if (t[i - 1] * t[i - 1] > upbound2) {
            t[i] *= upbound2inv;
            t[i - 1] *= upbound2inv;
            isc[i] = 1;
        }
    }
    double sctot = 1.0;
    for (int i = 1; i <= ntop; ++i) {
        if (isc[i - 1]) sctot *= upbound2inv;
        t[i] *= sctot;
    }
    double j0 = sin(z) * zinv;
    double j1 = j0 * zinv - cos(z) * zinv;
    double zs = (fabs(j1) > fabs(j0)) ? j1 / t[1] : j0 / t[0];
    for (int i = 0; i <= nterms; ++i) fjs[i] = t[i] * zs;
    free(t);
    free(isc);
}

/* ------------------------------------------------------------------ */
/* bfgt_volrouts.f: 1-D tables                                         */
/* ------------------------------------------------------------------ */

/* mk_leg2pw_tables / mk_cheb2pw_tables (bfgt_volrouts.f:58-275):
 * tab(m,j) = sum_k tabtemp(k,j) u(k,m), tabtemp(k,j) = ws_j D
 * int_{-1}^1 p_k(x) e^{-i ts_j D x/(2 sqrt(delta))} dx / 2 ... via
 * int_{-1}^1 P_k(x) e^{-iax} dx = 2 j_k(a)/i^k. */
void oracle_poly2pw(int ipoly, int n, int npw, const double *ws,
                    const double *ts, double delta, double boxdim,
                    double *tab_out)
{
    cplx *tabtemp = (cplx *)calloc((size_t)n * npw, sizeof(cplx));
    double *fjs = (double *)malloc(sizeof(double) * (n + 1));
    double *x = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    double *ul = (double *)malloc(sizeof(double) * n * n);
    double *uc = (double *)malloc(sizeof(double) * n * n);
    double *v = (double *)malloc(sizeof(double) * n * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double fac = boxdim / (2 * sqrt(delta));
    oracle_polyexps(0, n, x, w, ul);
    if (ipoly == 1) {
        /* v(j,i): Legendre coefficients of T_{i-1} (bfgt_volrouts.f:216-231) */
        double *chebv = (double *)malloc(sizeof(double) * n * n);
        for (int k = 0; k < n; ++k) {
            chebpols(x[k], n - 1, p);
            for (int i = 0; i < n; ++i) chebv[i + k * n] = p[i];
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double dd = 0;
                for (int k = 0; k < n; ++k) dd += ul[j + k * n] * chebv[i + k * n];
                v[j + i * n] = dd;
            }
        free(chebv);
    }
    for (int j = npw / 2; j < npw; ++j) {
        double dd = ws[j] * boxdim;
        if (ipoly == 0) {
            if (fabs(ts[j]) < 1e-12) {
                tabtemp[0 + (size_t)j * n] = dd;
                for (int m = 1; m < n; ++m) tabtemp[m + (size_t)j * n] = 0;
            } else {
                oracle_sph_bessel(n, ts[j] * fac, fjs);
                cplx eyem = 1.0;
                for (int m = 0; m < n; ++m) {
                    cplx val = dd * fjs[m] / eyem;
                    tabtemp[m + (size_t)j * n] = val;
                    tabtemp[m + (size_t)(npw - 1 - j) * n] = conj(val);
                    eyem *= I;
                }
            }
        } else {
            if (fabs(ts[j]) < 1e-12) {
                for (int m = 0; m < n; m += 2)
                    tabtemp[m + (size_t)j * n] = dd / (1.0 - (double)m * m);
                for (int m = 1; m < n; m += 2) tabtemp[m + (size_t)j * n] = 0;
            } else {
                cplx *legv = (cplx *)malloc(sizeof(cplx) * n);
                oracle_sph_bessel(n, ts[j] * fac, fjs);
                cplx eyem = 1.0;
                for (int m = 0; m < n; ++m) {
                    legv[m] = dd * fjs[m] / eyem;
                    eyem *= I;
                }
                for (int i = 0; i < n; ++i) {
                    cplx cd = 0;
                    for (int k = 0; k < n; ++k) cd += v[k + i * n] * legv[k];
                    tabtemp[i + (size_t)j * n] = cd;
                }
                free(legv);
            }
            if (j != npw - 1 - j)
                for (int i = 0; i < n; ++i)
                    tabtemp[i + (size_t)(npw - 1 - j) * n] = conj(tabtemp[i + (size_t)j * n]);
        }
    }
    const double *u = ul;
    if (ipoly == 1) {
        oracle_polyexps(1, n, x, w, uc);
        u = uc;
    }
    for (int i = 0; i < npw; ++i)
        for (int j = 0; j < n; ++j) {
            cplx cd = 0;
            for (int k = 0; k < n; ++k) cd += tabtemp[k + (size_t)i * n] * u[k + j * n];
            tab_out[2 * (j + (size_t)i * n)] = creal(cd);
            tab_out[2 * (j + (size_t)i * n) + 1] = cimag(cd);
        }
    free(tabtemp); free(fjs); free(x); free(w); free(ul); free(uc); free(v); free(p);
}

/* mk_pw2pgh_tables, potential part (bfgt_volrouts.f:339-401):
 * tab(j,i) = exp(i ts_j xs_i D/(2 sqrt(delta))), conjugate for the mirror. */
void oracle_pw2pot(int ipoly, int n, int npw, const double *ts, double delta,
                   double boxdim, double *tab_out)
{
    double *x = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    double *u = (double *)malloc(sizeof(double) * n * n);
    oracle_polyexps(ipoly, n, x, w, u);
    double dsq = boxdim / 2 / sqrt(delta);
    int npw2 = npw / 2;
    for (int i = 0; i < n; ++i) {
        double xx = x[i] * dsq;
        for (int j1 = npw2; j1 < npw; ++j1) {
            cplx e = cexp(I * ts[j1] * xx);
            size_t a = j1 + (size_t)i * npw, b = (npw - 1 - j1) + (size_t)i * npw;
            tab_out[2 * a] = creal(e);
            tab_out[2 * a + 1] = cimag(e);
            tab_out[2 * b] = creal(e);
            tab_out[2 * b + 1] = -cimag(e);
        }
    }
    free(x); free(w); free(u);
}

/* mk_pw2pgh_tables (bfgt_volrouts.f:339-401): the potential table and its
 * first/second derivative versions, tab_x(j,i) = t_j tab(j,i) dfac and
 * tab_xx(j,i) = t_j tab_x(j,i) dfac with dfac = i (D/(2 sqrt(delta))) (2/D);
 * the upper half j >= npw/2 is formed, the lower half is its conjugate. */
void oracle_pw2pgh(int ipoly, int n, int npw, const double *ts, double delta,
                   double boxdim, double *tab_out, double *tabx_out, double *tabxx_out)
{
    double *x = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    double *u = (double *)malloc(sizeof(double) * n * n);
    oracle_polyexps(ipoly, n, x, w, u);
    double dsq = boxdim / 2 / sqrt(delta);
    cplx dfac = I * dsq * 2 / boxdim;
    int npw2 = npw / 2;
    for (int i = 0; i < n; ++i) {
        double xx = x[i] * dsq;
        for (int j1 = npw2; j1 < npw; ++j1) {
            cplx e = cexp(I * ts[j1] * xx);
            cplx ex = ts[j1] * e * dfac;
            cplx exx = ts[j1] * ex * dfac;
            size_t a = j1 + (size_t)i * npw, b = (npw - 1 - j1) + (size_t)i * npw;
            tab_out[2 * a] = creal(e);   tab_out[2 * a + 1] = cimag(e);
            tabx_out[2 * a] = creal(ex); tabx_out[2 * a + 1] = cimag(ex);
            tabxx_out[2 * a] = creal(exx); tabxx_out[2 * a + 1] = cimag(exx);
            tab_out[2 * b] = creal(e);   tab_out[2 * b + 1] = -cimag(e);
            tabx_out[2 * b] = creal(ex); tabx_out[2 * b + 1] = -cimag(ex);
            tabxx_out[2 * b] = creal(exx); tabxx_out[2 * b + 1] = -cimag(exx);
        }
    }
    free(x); free(w); free(u);
}

/* mk_legetab_recur (bfgt_volrouts.f:1095-1188): five-term recurrence for
 * I_n(t)/lambda = (1/lambda) int_{-1}^1 P_n(x) e^{-(t-x)^2/lambda^2} dx and the
 * matching first/second-derivative integrals, lambda <= 1/8. */
static void legetab_recur(int m, double targ, double lambda, double boxdim,
                          double *fint, double *fx, double *fxx)
{
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[100], fpp[100];
    double lambda2 = lambda * lambda;
    double a = (-1 - targ) / lambda, b = (1 - targ) / lambda;
    double erfa = erf(a), erfb = erf(b);
    double expa = exp(-a * a), expb = exp(-b * b);
    double d1 = lambda * (expb - expa), d2 = lambda * (expb + expa);
    fint[0] = sqrtpih * (erfb - erfa);
    if (m > 1) fint[1] = -d1 / 2 + targ * fint[0];
    if (m > 2) {
        fint[2] = -d2 - targ * d1 + (lambda2 - 2.0 / 3 + 2 * targ * targ) * fint[0];
        fint[2] = 0.75 * fint[2];
    }
    if (m > 3) {
        fint[3] = -d1 + 2 * targ * (4.0 * fint[2] / 3 - fint[0] / 3) - (0.4 - 4 * lambda2) * fint[1];
        fint[3] = 5.0 * fint[3] / 8;
    }
    for (int k = 5; k <= m; ++k) {
        int n = k - 3;
        double v = targ * (fint[k - 2] - fint[k - 4])
                   + (2 * n + 1) * (lambda2 / 2 + 1.0 / ((2 * n + 3) * (2 * n - 1))) * fint[k - 3]
                   + (n - 1.0) * fint[k - 5] / (2 * n - 1);
        fint[k - 1] = (2 * n + 3.0) * v / (n + 2);
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fint[0] * lambda;
    for (int k = 3; k <= m; ++k) fp[k - 1] = fp[k - 3] + (2 * k - 3) * fint[k - 2] * lambda;
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    for (int k = 3; k <= m; ++k) fpp[k - 1] = fpp[k - 3] + (2 * k - 3) * fp[k - 2];
    int isign = 1;
    for (int k = 1; k <= m; ++k) {
        double px1 = (k - 1) * k / 2.0;
        fxx[k - 1] = fpp[k - 1] - px1 * (expb + isign * expa) - 2 * (b * expb - isign * a * expa) / lambda;
        isign = -isign;
    }
    double ww = boxdim * lambda / 2;
    for (int k = 0; k < m; ++k) {
        fint[k] *= ww;
        fxx[k] = fxx[k] * 2 / boxdim;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lambda;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lambda;
}

/* mk_chebtab_recur (bfgt_volrouts.f:1191-1294). */
static void chebtab_recur(int m, double targ, double lambda, double boxdim,
                          double *fint, double *fx, double *fxx)
{
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[100], fpp[100];
    double lambda2 = lambda * lambda;
    double a = (-1 - targ) / lambda, b = (1 - targ) / lambda;
    double erfa = erf(a), erfb = erf(b);
    double expa = exp(-a * a), expb = exp(-b * b);
    double d1 = lambda * (expb - expa), d2 = lambda * (expb + expa);
    fint[0] = sqrtpih * (erfb - erfa);
    if (m > 1) fint[1] = -d1 / 2 + targ * fint[0];
    if (m > 2) fint[2] = -d2 - targ * d1 + (lambda2 - 1 + 2 * targ * targ) * fint[0];
    if (m > 3) fint[3] = -d1 + 2 * targ * fint[2] - (1 - 4 * lambda2) * fint[1];
    int isign = 1;
    for (int k = 5; k <= m; ++k) {
        double dc = (isign == 1) ? d2 : d1;
        int n = k - 1;
        double dd = (n - 1.0) / (n - 3);
        fint[k - 1] = 2 * targ * fint[k - 2] + 2 * ((n - 1) * lambda2 + 1.0 / (n - 3)) * fint[k - 3]
                      + dd * (fint[k - 5] - 2 * targ * fint[k - 4]) - (1 - dd) * dc;
        isign = -isign;
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fint[0] * lambda;
    if (m > 2) fp[2] = 4 * fint[1] * lambda;
    for (int k = 4; k <= m; ++k) fp[k - 1] = (fp[k - 3] / (k - 3) + 2 * fint[k - 2] * lambda) * (k - 1);
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    if (m > 2) fpp[2] = 4 * fint[0] * lambda;
    for (int k = 4; k <= m; ++k) fpp[k - 1] = (fpp[k - 3] / (k - 3) + 2 * fp[k - 2]) * (k - 1);
    isign = 1;
    for (int k = 1; k <= m; ++k) {
        double tp1 = (double)(k - 1) * (k - 1);
        fxx[k - 1] = fpp[k - 1] - tp1 * (expb + isign * expa) - 2 * (b * expb - isign * a * expa) / lambda;
        isign = -isign;
    }
    double ww = boxdim * lambda / 2, sc = 2 / boxdim;
    for (int k = 0; k < m; ++k) {
        fint[k] *= ww;
        fxx[k] *= sc;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lambda;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lambda;
}

/* mk_loctab (bfgt_volrouts.f:920-1060): tab(m,j) = int_{source} p_m(x)
 * exp(-(targ_j - x)^2/delta) dx for source box at [-1,1] (scaled). */
static void mk_loctab(int ipoly, int n, double delta, const double *tnodes,
                      double tboxdim, double tboxcen, double scale,
                      double *tab, double *tabx, double *tabxx, int nquad,
                      const double *xq, const double *wq, const double *polyv)
{
    double targs[64], fint[100], fx[100], fxx[100];
    for (int i = 0; i < n; ++i) targs[i] = tnodes[i] / scale + tboxcen;
    double bs = tboxdim * scale;
    double sigma = 4 * delta / (bs * bs);
    double lambda = sqrt(sigma);
    double reps = 1e-16;
    double dmax = sqrt(log(1 / reps) * sigma);
    if (lambda <= 0.125) {
        for (int j = 0; j < n; ++j) {
            double targ = targs[j], dis = 0;
            if (targ < -1) dis = -1 - targ;
            if (targ > 1) dis = targ - 1;
            if (dis > dmax) {
                for (int m = 0; m < n; ++m) tab[m + j * n] = tabx[m + j * n] = tabxx[m + j * n] = 0;
            } else {
                if (ipoly == 0) legetab_recur(n, targ, lambda, bs, fint, fx, fxx);
                else chebtab_recur(n, targ, lambda, bs, fint, fx, fxx);
                for (int m = 0; m < n; ++m) {
                    tab[m + j * n] = fint[m];
                    tabx[m + j * n] = fx[m];
                    tabxx[m + j * n] = fxx[m];
                }
            }
        }
    } else {
        double *wexp = (double *)malloc(sizeof(double) * nquad * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < nquad; ++i) {
                double dx = targs[j] - xq[i];
                wexp[i + j * nquad] = (fabs(dx) < dmax) ? exp(-dx * dx / sigma) * wq[i] : 0.0;
            }
        double bsh = bs / 2;
        for (int j = 0; j < n; ++j) {
            double targ = targs[j], dis = 0;
            if (targ < -1) dis = -1 - targ;
            if (targ > 1) dis = targ - 1;
            if (dis > dmax) {
                for (int m = 0; m < n; ++m) tab[m + j * n] = tabx[m + j * n] = tabxx[m + j * n] = 0;
                continue;
            }
            for (int m = 0; m < n; ++m) {
                double rs = 0, rsx = 0, rsxx = 0;
                for (int i = 0; i < nquad; ++i) {
                    double dx = targ - xq[i];
                    double dd = -2 * dx / sigma * 2 / bs;
                    double pw = polyv[i + m * nquad] * wexp[i + j * nquad];
                    rs += pw;
                    rsx += pw * dd;
                    rsxx += pw * (dd * dd - 2 / delta);
                }
                tab[m + j * n] = rs * bsh;
                tabx[m + j * n] = rsx * bsh;
                tabxx[m + j * n] = rsxx * bsh;
            }
        }
        free(wexp);
    }
}

/* compute_sparse_pattern (bfgt_volrouts.f:1300-1378); indc is (2,n+1),
 * 1-based row bounds per column, with (0,-1) meaning "no entries". */
static void sparse_pattern(double eps, int n, const double *a, int *indc)
{
    for (int i = 0; i <= n; ++i) { indc[2 * i] = 0; indc[2 * i + 1] = -1; }
    double dmax = 0;
    for (int i = 0; i < n * n; ++i) if (dmax < fabs(a[i])) dmax = fabs(a[i]);
    if (dmax <= 2e-16) return;
    double thr = eps * dmax / n;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j)
            if (fabs(a[j + i * n]) > thr) { indc[2 * i] = j + 1; break; }
        for (int j = n - 1; j >= 0; --j)
            if (fabs(a[j + i * n]) > thr) { indc[2 * i + 1] = j + 1; break; }
    }
    for (int i = 0; i < n; ++i) if (indc[2 * i] > 0) { indc[2 * n] = i + 1; break; }
    for (int i = n - 1; i >= 0; --i) if (indc[2 * i + 1] > 0) { indc[2 * n + 1] = i + 1; break; }
}

/* Union of the three patterns as mk_loctab_all forms it (bfgt_volrouts.f:611-647). */
static void merge_patterns(int n, const int *ind, const int *indx,
                           const int *indxx, int *out)
{
    for (int j = 0; j <= n; ++j) {
        int i2 = ind[2 * j + 1];
        if (i2 < indx[2 * j + 1]) i2 = indx[2 * j + 1];
        if (i2 < indxx[2 * j + 1]) i2 = indxx[2 * j + 1];
        out[2 * j + 1] = i2;
        int i1 = ind[2 * j];
        if (indx[2 * j] != 0 && i1 > indx[2 * j]) i1 = indx[2 * j];
        if (indxx[2 * j] != 0 && i1 > indxx[2 * j]) i1 = indxx[2 * j];
        if (i2 > 0 && i1 == 0) i1 = 1;
        out[2 * j] = i1;
    }
}

/* mk_loctab_all with mrefinelev = 0, nloctab = 6 (bfgt_volrouts.f:407-660). */
void oracle_loctab3(double eps, int ipoly, int n, double delta, double boxdim,
                    double *tab_loc, double *tabx_loc, double *tabxx_loc, int *ind_loc);

void oracle_loctab(double eps, int ipoly, int n, double delta, double boxdim,
                   double *tab_loc, int *ind_loc)
{
    oracle_loctab3(eps, ipoly, n, delta, boxdim, tab_loc, NULL, NULL, ind_loc);
}

/* mk_loctab_all (bfgt_volrouts.f:407-660) with the first/second derivative
 * tables tabx_loc, tabxx_loc (nullable) in the layout of tab_loc. */
void oracle_loctab3(double eps, int ipoly, int n, double delta, double boxdim,
                    double *tab_loc, double *tabx_loc, double *tabxx_loc, int *ind_loc)
{
    const int nq = 50;
    double tn[64], tw[64], tu[64 * 64];
    oracle_polyexps(ipoly, n, tn, tw, tu);
    double xq[50], wq[50];
    gauss_legendre(nq, xq, wq);
    double *polyv = (double *)malloc(sizeof(double) * nq * n);
    double p[64];
    for (int i = 0; i < nq; ++i) {
        if (ipoly == 0) legepols(xq[i], n - 1, p); else chebpols(xq[i], n - 1, p);
        for (int j = 0; j < n; ++j) polyv[i + j * nq] = p[j];
    }
    size_t nn = (size_t)n * n;
    double *tab = (double *)malloc(sizeof(double) * nn * 3);
    double *tabx = tab + nn, *tabxx = tab + 2 * nn;
    double *lx = (double *)malloc(sizeof(double) * nn * 13 * 2);
    double *lxx = lx + nn * 13;
    memset(tab_loc, 0, sizeof(double) * nn * 13);
    memset(lx, 0, sizeof(double) * nn * 26);
    for (int i = 0; i < 2 * (n + 1) * 13; ++i) ind_loc[i] = 0;
    const int m0 = 1;
    for (int i0 = m0; i0 >= -1; --i0) {
        for (int j0 = 1; j0 <= (1 << m0); ++j0) {
            int k;
            if (i0 > 0) k = -(1 << (m0 - i0)) - (1 << (m0 - i0 + 1)) * (j0 - 1);
            else if (i0 == 0) k = -(1 << (m0 - i0 + 1)) * (j0 - 1);
            else k = -(1 << m0) - (1 << (m0 + 1)) * (j0 - 1);
            if (k < -6) break;
            double scale = pow(2.0, i0), cen = k / (double)(1 << m0);
            mk_loctab(ipoly, n, delta, tn, boxdim, cen, scale, tab, tabx, tabxx,
                      nq, xq, wq, polyv);
            double *tk = tab_loc + (k + 6) * nn, *txk = lx + (k + 6) * nn, *txxk = lxx + (k + 6) * nn;
            /* u^T * tab: acts on values (dgemm 't','n') */
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    double s0 = 0, s1 = 0, s2 = 0;
                    for (int m = 0; m < n; ++m) {
                        double um = tu[m + i * n];
                        s0 += um * tab[m + j * n];
                        s1 += um * tabx[m + j * n];
                        s2 += um * tabxx[m + j * n];
                    }
                    tk[i + j * n] = s0;
                    txk[i + j * n] = s1;
                    txxk[i + j * n] = s2;
                }
            /* mirror symmetry for targets on the other side (in place for k=0) */
            double *tm = tab_loc + (-k + 6) * nn, *txm = lx + (-k + 6) * nn, *txxm = lxx + (-k + 6) * nn;
            for (int j = 0; j < n; ++j)
                for (int m = 0; m < n; ++m) {
                    tm[m + j * n] = tk[(n - 1 - m) + (n - 1 - j) * n];
                    txm[m + j * n] = -txk[(n - 1 - m) + (n - 1 - j) * n];
                    txxm[m + j * n] = txxk[(n - 1 - m) + (n - 1 - j) * n];
                }
            int a[2 * 65], b[2 * 65], c[2 * 65];
            sparse_pattern(eps, n, tm, a);
            sparse_pattern(eps, n, txm, b);
            sparse_pattern(eps, n, txxm, c);
            merge_patterns(n, a, b, c, ind_loc + (-k + 6) * 2 * (n + 1));
            sparse_pattern(eps, n, tk, a);
            sparse_pattern(eps, n, txk, b);
            sparse_pattern(eps, n, txxk, c);
            merge_patterns(n, a, b, c, ind_loc + (k + 6) * 2 * (n + 1));
        }
    }
    if (tabx_loc) memcpy(tabx_loc, lx, sizeof(double) * nn * 13);
    if (tabxx_loc) memcpy(tabxx_loc, lxx, sizeof(double) * nn * 13);
    free(polyv); free(tab); free(lx);
}

/* ------------------------------------------------------------------ */
/* tree_routs_nd.f: lists                                              */
/* ------------------------------------------------------------------ */

typedef struct {
    int ndim, nboxes, nlevels, mc, mnbors;
    const int *laddr;   /* laddr(2,0:nlevels) */
    const int *ilevel, *iparent, *nchild, *ichild;
    int *ncoll, *coll;  /* may be a private copy */
    const double *centers, *boxsize;
} tree_t;

#define LADDR(t, i, l) ((t)->laddr[2 * (l) + (i) - 1])
#define NCHILD(t, b) ((t)->nchild[(b) - 1])
#define ICHILD(t, j, b) ((t)->ichild[(size_t)(t)->mc * ((b) - 1) + (j) - 1])
#define NCOLL(t, b) ((t)->ncoll[(b) - 1])
#define COLL(t, j, b) ((t)->coll[(size_t)(t)->mnbors * ((b) - 1) + (j) - 1])
#define CEN(t, k, b) ((t)->centers[(size_t)(t)->ndim * ((b) - 1) + (k) - 1])

/* computecoll (tree_routs_nd.f:106-242). */
static void computecoll(tree_t *t, int iperiod)
{
    double bs0 = t->boxsize[0];
    for (int i = 1; i <= t->nboxes; ++i) {
        NCOLL(t, i) = 0;
        for (int j = 1; j <= t->mnbors; ++j) COLL(t, j, i) = -1;
    }
    NCOLL(t, 1) = 1;
    COLL(t, 1, 1) = 1;
    for (int ilev = 1; ilev <= t->nlevels; ++ilev) {
        for (int ibox = LADDR(t, 1, ilev); ibox <= LADDR(t, 2, ilev); ++ibox) {
            int dad = t->iparent[ibox - 1];
            for (int i = 1; i <= NCOLL(t, dad); ++i) {
                int jbox = COLL(t, i, dad);
                for (int j = 1; j <= t->mc; ++j) {
                    int kbox = ICHILD(t, j, jbox);
                    if (kbox <= 0) continue;
                    int ok = 1;
                    for (int k = 1; k <= t->ndim; ++k) {
                        double dis = fabs(CEN(t, k, kbox) - CEN(t, k, ibox));
                        if (iperiod == 1) { double dp1 = bs0 - dis; if (dp1 < dis) dis = dp1; }
                        if (dis > 1.05 * t->boxsize[ilev]) { ok = 0; break; }
                    }
                    if (ok) {
                        NCOLL(t, ibox) += 1;
                        COLL(t, NCOLL(t, ibox), ibox) = kbox;
                    }
                }
            }
        }
    }
}

/* gnd_find_pwexp_boxes (tree_routs_nd.f:1266-1366). */
static void find_pwexp_boxes(const tree_t *t, int npwlevel, int *ifpwexp)
{
    for (int i = 0; i < t->nboxes; ++i) ifpwexp[i] = 0;
    if (npwlevel <= 0) { ifpwexp[0] = 1; return; }
    if (npwlevel > t->nlevels) return;
    int ilev = npwlevel;
    for (int ibox = LADDR(t, 1, ilev); ibox <= LADDR(t, 2, ilev); ++ibox) {
        if (NCHILD(t, ibox) == 0) {
            for (int j = 1; j <= NCOLL(t, ibox); ++j) {
                int jbox = COLL(t, j, ibox);
                if (NCHILD(t, jbox) > 0) { ifpwexp[ibox - 1] = 1; break; }
            }
        } else {
            ifpwexp[ibox - 1] = 1;
        }
    }
    for (int l = npwlevel + 1; l <= t->nlevels; ++l)
        for (int ibox = LADDR(t, 1, l); ibox <= LADDR(t, 2, l); ++ibox) ifpwexp[ibox - 1] = 1;
}

/* gnd_compute_modified_list1 (tree_routs_nd.f:864-1055); list1 has room
 * for mnl entries per box. */
static void modified_list1(const tree_t *t, int npwlevel, const int *ifpwexp,
                           int iperiod, int mnl, int *nlist1, int *list1)
{
    double bs0 = t->boxsize[0];
    for (int i = 0; i < t->nboxes; ++i) nlist1[i] = 0;
    if (NCHILD(t, 1) == 0) { nlist1[0] = 1; list1[0] = 1; }
    int nlevend = t->nlevels;
    if (npwlevel < t->nlevels) nlevend = npwlevel;
#define ADD1(b, v) do { list1[(size_t)mnl * ((b) - 1) + nlist1[(b) - 1]] = (v); nlist1[(b) - 1]++; } while (0)
    for (int ilev = 1; ilev <= nlevend; ++ilev) {
#pragma omp parallel for schedule(dynamic)
        for (int ibox = LADDR(t, 1, ilev); ibox <= LADDR(t, 2, ilev); ++ibox) {
            if (NCHILD(t, ibox) != 0) continue;
            int dad = t->iparent[ibox - 1];
            for (int i = 1; i <= NCOLL(t, ibox); ++i) {
                int jbox = COLL(t, i, ibox);
                if (NCHILD(t, jbox) == 0) {
                    if (ifpwexp[ibox - 1] != 1 || jbox != ibox) ADD1(ibox, jbox);
                } else if (ilev < npwlevel) {
                    double distest = 1.05 * (t->boxsize[ilev] + t->boxsize[ilev + 1]) / 2.0;
                    for (int j = 1; j <= t->mc; ++j) {
                        int kbox = ICHILD(t, j, jbox);
                        if (NCHILD(t, kbox) != 0) continue;
                        int ok = 1;
                        for (int k = 1; k <= t->ndim; ++k) {
                            double dis = fabs(CEN(t, k, kbox) - CEN(t, k, ibox));
                            if (iperiod == 1) { double dp1 = bs0 - dis; if (dp1 < dis) dis = dp1; }
                            if (dis >= distest) { ok = 0; break; }
                        }
                        if (ok) ADD1(ibox, kbox);
                    }
                }
            }
            for (int i = 1; i <= NCOLL(t, dad); ++i) {
                int jbox = COLL(t, i, dad);
                if (NCHILD(t, jbox) != 0) continue;
                double distest = 1.05 * (t->boxsize[ilev] + t->boxsize[ilev - 1]) / 2.0;
                int ok = 1;
                for (int k = 1; k <= t->ndim; ++k) {
                    double dis = fabs(CEN(t, k, jbox) - CEN(t, k, ibox));
                    if (iperiod == 1) { double dp1 = bs0 - dis; if (dp1 < dis) dis = dp1; }
                    if (dis >= distest) { ok = 0; break; }
                }
                if (ok) ADD1(ibox, jbox);
            }
        }
    }
#undef ADD1
}

/* gnd_compute_listpw (tree_routs_nd.f:1562-1628). */
static void compute_listpw(const tree_t *t, int npwlevel, int mnlistpw,
                           int *nlistpw, int *listpw)
{
    for (int i = 0; i < t->nboxes; ++i) nlistpw[i] = 0;
    if (npwlevel > t->nlevels) return;
    int ilev = npwlevel > 0 ? npwlevel : 0;
    for (int ibox = LADDR(t, 1, ilev); ibox <= LADDR(t, 2, ilev); ++ibox) {
        for (int i = 1; i <= NCOLL(t, ibox); ++i) {
            int jbox = COLL(t, i, ibox);
            if (jbox != ibox && (NCHILD(t, jbox) > 0 || NCHILD(t, ibox) > 0)) {
                listpw[(size_t)mnlistpw * (ibox - 1) + nlistpw[ibox - 1]] = jbox;
                nlistpw[ibox - 1]++;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* bfgt_pwrouts.f / bfgt_local.f: per-box kernels                       */
/* ------------------------------------------------------------------ */

/* g{1,2,3}d_tens_prod_to_pw (bfgt_pwrouts.f:57-222): one density. */
static void tens_prod_to_pw(int ndim, int n, int nd, int ind, const double *f,
                            int npw, const cplx *T, cplx *pw, cplx *work)
{
    int nh = (npw + 1) / 2;
#define FV(i) f[(size_t)nd * (i) + ind]
    if (ndim == 1) {
        for (int k1 = 0; k1 < nh; ++k1) {
            cplx cd = 0;
            for (int m1 = 0; m1 < n; ++m1) cd += T[m1 + k1 * n] * FV(m1);
            pw[k1] = cd;
        }
    } else if (ndim == 2) {
        cplx *ff = work; /* ff(n,nh) */
        for (int m1 = 0; m1 < n; ++m1)
            for (int k2 = 0; k2 < nh; ++k2) {
                cplx cd = 0;
                for (int m2 = 0; m2 < n; ++m2) cd += T[m2 + k2 * n] * FV(m1 + n * m2);
                ff[m1 + n * k2] = cd;
            }
        for (int k2 = 0; k2 < nh; ++k2)
            for (int k1 = 0; k1 < npw; ++k1) {
                cplx cd = 0;
                for (int m1 = 0; m1 < n; ++m1) cd += T[m1 + k1 * n] * ff[m1 + n * k2];
                pw[k1 + npw * k2] = cd;
            }
    } else {
        cplx *ff = work;                        /* ff(n,n,nh)   */
        cplx *ff2 = work + (size_t)n * n * nh;  /* ff2(n,npw,nh) */
        for (int k3 = 0; k3 < nh; ++k3)
            for (int m2 = 0; m2 < n; ++m2)
                for (int m1 = 0; m1 < n; ++m1) {
                    cplx cd = 0;
                    for (int m3 = 0; m3 < n; ++m3) cd += T[m3 + k3 * n] * FV(m1 + n * (m2 + n * m3));
                    ff[m1 + n * (m2 + n * k3)] = cd;
                }
        for (int k3 = 0; k3 < nh; ++k3)
            for (int k2 = 0; k2 < npw; ++k2)
                for (int m1 = 0; m1 < n; ++m1) {
                    cplx cd = 0;
                    for (int m2 = 0; m2 < n; ++m2) cd += T[m2 + k2 * n] * ff[m1 + n * (m2 + n * k3)];
                    ff2[m1 + n * (k2 + (size_t)npw * k3)] = cd;
                }
        for (int k3 = 0; k3 < nh; ++k3)
            for (int k2 = 0; k2 < npw; ++k2)
                for (int k1 = 0; k1 < npw; ++k1) {
                    cplx cd = 0;
                    for (int m1 = 0; m1 < n; ++m1) cd += T[m1 + k1 * n] * ff2[m1 + n * (k2 + (size_t)npw * k3)];
                    pw[k1 + npw * (k2 + (size_t)npw * k3)] = cd;
                }
    }
#undef FV
}

/* g{1,2,3}d_pw2pot (bfgt_pwrouts.f:305-790): pot += 2 Re(...), one density,
 * zero frequency halved when npw is odd. */
/* One output component of gnd_pw2pgh (bfgt_pwrouts.f:233-299; 1-D 308-462,
 * 2-D 473-708, 3-D 719-1065): out += 2 Re sum Tx Ty Tz pw, contracting x,
 * then y, then the half axis (zero frequency halved when npw is odd).
 * Tax[a] is the table of axis a: tab_pw2pot, tab_pw2der or tab_pw2dxx, so
 * pot, each gradient and each Hessian component is one call with the
 * matching tables (the reference forms them together from shared partial
 * sums; each component's operation sequence is the same).
 * out[ostride * node + ooff] receives node node = k1 + n k2 + n^2 k3. */
static void pw2pot(int ndim, int n, int npw, const cplx *pw, const cplx *const *Tax,
                   double *out, size_t ostride, size_t ooff, cplx *work)
{
    int nh = (npw + 1) / 2, npw2 = npw / 2;
#define PV(i) out[ostride * (size_t)(i) + ooff]
    if (ndim == 1) {
        const cplx *T = Tax[0];
        for (int k1 = 0; k1 < n; ++k1) {
            cplx cd = 0;
            for (int m1 = 0; m1 < npw2; ++m1) cd += T[m1 + npw * k1] * pw[m1];
            if (nh > npw2) cd += T[nh - 1 + npw * k1] * pw[nh - 1] / 2;
            PV(k1) += creal(cd) * 2;
        }
    } else if (ndim == 2) {
        const cplx *Tx = Tax[0], *Ty = Tax[1];
        cplx *ff = work; /* ff(n,nh) */
        for (int m2 = 0; m2 < nh; ++m2)
            for (int k1 = 0; k1 < n; ++k1) {
                cplx cd = 0;
                for (int m1 = 0; m1 < npw; ++m1) cd += Tx[m1 + npw * k1] * pw[m1 + npw * m2];
                ff[k1 + n * m2] = cd;
            }
        for (int k2 = 0; k2 < n; ++k2)
            for (int k1 = 0; k1 < n; ++k1) {
                cplx cd = 0;
                for (int m2 = 0; m2 < npw2; ++m2) cd += Ty[m2 + npw * k2] * ff[k1 + n * m2];
                if (nh > npw2) cd += Ty[nh - 1 + npw * k2] * ff[k1 + n * (nh - 1)] / 2;
                PV(k1 + n * k2) += creal(cd) * 2;
            }
    } else {
        const cplx *Tx = Tax[0], *Ty = Tax[1], *Tz = Tax[2];
        cplx *ff = work;                          /* ff(n,npw,nh) */
        cplx *ff2 = work + (size_t)n * npw * nh;  /* ff2(n,n,nh)  */
        for (int m3 = 0; m3 < nh; ++m3)
            for (int m2 = 0; m2 < npw; ++m2)
                for (int k1 = 0; k1 < n; ++k1) {
                    cplx cd = 0;
                    for (int m1 = 0; m1 < npw; ++m1) cd += Tx[m1 + npw * k1] * pw[m1 + npw * (m2 + (size_t)npw * m3)];
                    ff[k1 + n * (m2 + (size_t)npw * m3)] = cd;
                }
        for (int m3 = 0; m3 < nh; ++m3)
            for (int k2 = 0; k2 < n; ++k2)
                for (int k1 = 0; k1 < n; ++k1) {
                    cplx cd = 0;
                    for (int m2 = 0; m2 < npw; ++m2) cd += Ty[m2 + npw * k2] * ff[k1 + n * (m2 + (size_t)npw * m3)];
                    ff2[k1 + n * (k2 + n * m3)] = cd;
                }
        for (int k3 = 0; k3 < n; ++k3)
            for (int k2 = 0; k2 < n; ++k2)
                for (int k1 = 0; k1 < n; ++k1) {
                    cplx cd = 0;
                    for (int m3 = 0; m3 < npw2; ++m3) cd += Tz[m3 + npw * k3] * ff2[k1 + n * (k2 + n * m3)];
                    if (nh > npw2) cd += Tz[nh - 1 + npw * k3] * ff2[k1 + n * (k2 + n * (nh - 1))] / 2;
                    PV(k1 + n * (k2 + n * k3)) += creal(cd) * 2;
                }
    }
#undef PV
}

/* Output components of ifpgh (boxfgt.f:35-38 and the pg/pgh routines):
 * per-axis table kind (0 value, 1 first derivative, 2 second derivative).
 * Component 0 is pot; then grad(ndim); then the Hessian in the reference's
 * order: 1-D xx; 2-D xx, xy, yy (bfgt_pwrouts.f:612-708); 3-D xx, yy, zz,
 * xy, xz, yz (bfgt_pwrouts.f:906-1065). */
static int pgh_components(int ndim, int ifpgh, int kinds[10][3])
{
    static const int h1[1][3] = {{2, 0, 0}};
    static const int h2[3][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}};
    static const int h3[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    int nc = 0;
    kinds[nc][0] = kinds[nc][1] = kinds[nc][2] = 0; ++nc;
    if (ifpgh >= 2)
        for (int a = 0; a < ndim; ++a, ++nc) {
            kinds[nc][0] = kinds[nc][1] = kinds[nc][2] = 0;
            kinds[nc][a] = 1;
        }
    if (ifpgh >= 3) {
        int nhess = ndim * (ndim + 1) / 2;
        const int (*h)[3] = ndim == 1 ? h1 : (ndim == 2 ? h2 : h3);
        for (int i = 0; i < nhess; ++i, ++nc)
            for (int a = 0; a < 3; ++a) kinds[nc][a] = h[i][a];
    }
    return nc;
}

/* gnd_find_loctab_ind with mrefinelev = 0 (bfgt_local.f:1624-1660). */
static void find_loctab_ind(int ndim, int iperiod, const double *tc,
                            const double *sc, double sbs, double bs0, int *ixyz)
{
    double bs = sbs / 4;
    for (int i = 0; i < ndim; ++i) {
        double dx = (tc[i] - sc[i]) / bs;
        if (iperiod == 1) {
            double dxp1 = dx - bs0 / bs, dxm1 = dx + bs0 / bs;
            if (fabs(dx) > fabs(dxp1)) dx = dxp1;
            if (fabs(dx) > fabs(dxm1)) dx = dxm1;
        }
        ixyz[i] = (int)dx;
    }
}

/* gnd_tens_prod_to_pghloc (bfgt_local.f:14-140; pot 152-1618, pg/pgh
 * 644-920), one density, one output component: out += (A_x (x) A_y (x) A_z) f
 * with each A the table of the component's kind on that axis (tab_loc,
 * tabx_loc or tabxx_loc) restricted to the band of ind_loc.  The contraction
 * order is x, then y, then z, as in the reference's pg/pgh routines; its
 * potential-only routine picks the order by band width, which changes only
 * rounding. */
static void tens_prod_to_potloc(int ndim, int n, int nd, int ind,
                                const double *f, double *out, size_t ostride, size_t ooff,
                                const double *const *tabk, const int *kind, const int *indl,
                                const int *ixyz, double *work)
{
    const double *A[3];
    const int *B[3];
    for (int a = 0; a < ndim; ++a) {
        int slot = ixyz[a] + 6;
        A[a] = tabk[kind[a]] + (size_t)slot * n * n;
        B[a] = indl + (size_t)slot * 2 * (n + 1);
        int nx = B[a][2 * n + 1] - B[a][2 * n] + 1;
        if (nx == 0) return;
    }
    int npb = 1;
    for (int a = 0; a < ndim; ++a) npb *= n;
    double *cur = work, *nxt = work + npb;
    for (int i = 0; i < npb; ++i) cur[i] = f[(size_t)nd * i + ind];
    int stride = 1;
    for (int a = 0; a < ndim; ++a) {
        const double *Aa = A[a];
        const int *Ba = B[a];
        int kmin = Ba[2 * n], kmax = Ba[2 * n + 1];
        for (int i = 0; i < npb; ++i) nxt[i] = 0;
        for (int base = 0; base < npb; ++base) {
            int ka = (base / stride) % n;
            if (ka != 0) continue;
            /* base indexes a line along axis a */
            for (int k = kmin; k <= kmax; ++k) {
                double cd = 0;
                for (int j = Ba[2 * (k - 1)]; j <= Ba[2 * (k - 1) + 1]; ++j)
                    cd += Aa[(j - 1) + (size_t)(k - 1) * n] * cur[base + (j - 1) * stride];
                nxt[base + (k - 1) * stride] = cd;
            }
        }
        double *tmp = cur; cur = nxt; nxt = tmp;
        stride *= n;
    }
    for (int i = 0; i < npb; ++i) out[ostride * (size_t)i + ooff] += cur[i];
}

/* ------------------------------------------------------------------ */
/* boxfgt / bfgtmain                                                   */
/* ------------------------------------------------------------------ */

static int ipow(int b, int e) { int r = 1; while (e-- > 0) r *= b; return r; }

/* Merge/split phase tables W(j,k,klev) (bfgt_pwrouts.f:1305-1543), j with
 * the x frequency fastest and the half (last) axis slowest; k = sign pattern
 * with the x bit fastest (bit set = conjugate). */
static void mk_merge_split(int ndim, double xmin, int npw, const double *ts,
                           int nmax, cplx *w)
{
    int nh = (npw + 1) / 2;
    int nexp = nh * ipow(npw, ndim - 1), mc = 1 << ndim;
    cplx *ww = (cplx *)malloc(sizeof(cplx) * npw * (nmax > 0 ? nmax : 1));
    for (int j = 0; j < npw; ++j) {
        cplx z = cexp(I * ts[j] * xmin);
        for (int k1 = 0; k1 < nmax; ++k1) { ww[j + npw * k1] = z; z = z * z; }
    }
    for (int k1 = 0; k1 < nmax; ++k1)
        for (int jj = 0; jj < nexp; ++jj) {
            int jx[3] = {0, 0, 0}, r = jj;
            /* digits: x (fastest, npw), y (npw), last axis (nh) */
            for (int a = 0; a < ndim; ++a) {
                int base = (a == ndim - 1) ? nh : npw;
                jx[a] = r % base;
                r /= base;
            }
            for (int k = 0; k < mc; ++k) {
                cplx v = 1.0;
                for (int a = 0; a < ndim; ++a) {
                    cplx e = ww[jx[a] + npw * k1];
                    if ((k >> a) & 1) e = conj(e);
                    v = (a == 0) ? e : v * e;
                }
                w[jj + (size_t)nexp * (k + (size_t)mc * k1)] = v;
            }
        }
    free(ww);
}

/* M2L phases at the cutoff level, nmax = 1 (bfgt_pwrouts.f:1076-1293);
 * k = sum_a (off_a + 1) 3^a with the x offset fastest. */
static void mk_m2l(int ndim, double xmin, int npw, const double *ts, cplx *w)
{
    int nh = (npw + 1) / 2;
    int nexp = nh * ipow(npw, ndim - 1), nk = ipow(3, ndim);
    cplx *ww = (cplx *)malloc(sizeof(cplx) * npw * 3);
    for (int j = 0; j < npw; ++j) {
        cplx z = cexp(I * ts[j] * xmin);
        ww[j + npw * 1] = 1.0;
        ww[j + npw * 2] = z;
        ww[j + npw * 0] = conj(z);
    }
    for (int k = 0; k < nk; ++k)
        for (int jj = 0; jj < nexp; ++jj) {
            int r = jj, kr = k;
            cplx v = 1.0;
            for (int a = 0; a < ndim; ++a) {
                int base = (a == ndim - 1) ? nh : npw;
                int ja = r % base; r /= base;
                int oa = kr % 3; kr /= 3;
                cplx e = ww[ja + npw * oa];
                v = (a == 0) ? e : v * e;
            }
            w[jj + (size_t)nexp * k] = v;
        }
    free(ww);
}

int oracle_boxfgt(int nd, int ndim, double delta, double eps, int ipoly,
                  int iperiod, int norder, int nboxes, int nlevels,
                  const int *itree, const int *iptr, const double *centers,
                  const double *boxsize, const double *fvals, double *pot,
                  int d1_mode, double *timeinfo, oracle_info *info)
{
    return oracle_boxfgt_pgh(nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels, itree, iptr,
                             centers, boxsize, fvals, 1, pot, NULL, NULL, d1_mode, timeinfo, info);
}

/* Output of component c for box b (1-based) and density ind: base pointer
 * and (stride, offset) within the box, Fortran pot(nd,npbox), grad(nd,ndim,
 * npbox), hess(nd,nhess,npbox). */
static double *comp_out(int c, int ndim, int nd, int npbox, int b, int ind, double *pot,
                        double *grad, double *hess, size_t *ostride, size_t *ooff)
{
    int nhess = ndim * (ndim + 1) / 2;
    if (c == 0) {
        *ostride = nd; *ooff = ind;
        return pot + (size_t)nd * npbox * (b - 1);
    }
    if (c <= ndim) {
        *ostride = (size_t)nd * ndim; *ooff = (size_t)(c - 1) * nd + ind;
        return grad + (size_t)nd * ndim * npbox * (b - 1);
    }
    *ostride = (size_t)nd * nhess; *ooff = (size_t)(c - 1 - ndim) * nd + ind;
    return hess + (size_t)nd * nhess * npbox * (b - 1);
}

int oracle_boxfgt_pgh(int nd, int ndim, double delta, double eps, int ipoly,
                      int iperiod, int norder, int nboxes, int nlevels,
                      const int *itree, const int *iptr, const double *centers,
                      const double *boxsize, const double *fvals, int ifpgh,
                      double *pot, double *grad, double *hess,
                      int d1_mode, double *timeinfo, oracle_info *info)
{
    if (ifpgh < 1 || ifpgh > 3 || (ifpgh >= 2 && !grad) || (ifpgh >= 3 && !hess)) return 1;
    if (ndim < 1 || ndim > 3 || norder < 1 || norder > 32 || nd < 1 ||
        (ipoly != 0 && ipoly != 1) || (iperiod != 0 && iperiod != 1) ||
        nboxes < 1 || nlevels < 0 || !(delta > 0) || !(eps > 0 && eps < 1))
        return 1;
    double tstep[8] = {0};
    double t0 = now_s();
    const int n = norder, mc = 1 << ndim, mnbors = ipow(3, ndim);
    const int npbox = ipow(n, ndim);
    const double bs0 = boxsize[0];

    tree_t T;
    T.ndim = ndim; T.nboxes = nboxes; T.nlevels = nlevels; T.mc = mc; T.mnbors = mnbors;
    T.laddr = itree + (iptr[0] - 1);
    T.ilevel = itree + (iptr[1] - 1);
    T.iparent = itree + (iptr[2] - 1);
    T.nchild = itree + (iptr[3] - 1);
    T.ichild = itree + (iptr[4] - 1);
    int *ncoll = (int *)malloc(sizeof(int) * nboxes);
    int *coll = (int *)malloc(sizeof(int) * (size_t)mnbors * nboxes);
    memcpy(ncoll, itree + (iptr[5] - 1), sizeof(int) * nboxes);
    memcpy(coll, itree + (iptr[6] - 1), sizeof(int) * (size_t)mnbors * nboxes);
    T.ncoll = ncoll; T.coll = coll;
    T.centers = centers; T.boxsize = boxsize;

    /* cutoff level and npw (boxfgt.f:132-162) */
    double dcutoff = sqrt(delta * log(1.0 / eps));
    int npwlevel = nlevels + 1;
    for (int i = nlevels; i >= 0; --i)
        if (boxsize[i] >= dcutoff) { npwlevel = i; break; }
    if (boxsize[nlevels] > dcutoff) npwlevel = nlevels + 1;
    if (boxsize[0] <= dcutoff) npwlevel = 0;
    int npw = 0;
    double pmax = 0;
    if (npwlevel <= nlevels) {
        if (npwlevel <= 1)
            oracle_pwterms(2 * boxsize[npwlevel > 0 ? npwlevel : 0], delta, eps, iperiod, &pmax, &npw);
        else
            oracle_pwterms(2 * boxsize[npwlevel], delta, eps, 0, &pmax, &npw);
    }

    /* STEP 0 (boxfgt.f:464-688) */
    double *ws = (double *)calloc(npw > 0 ? npw : 1, sizeof(double));
    double *ts = (double *)calloc(npw > 0 ? npw : 1, sizeof(double));
    double delta0 = delta;
    int iper = iperiod, periodic_small = 0;
    if (npwlevel <= 1 && iperiod == 1) {
        periodic_small = 1;
        oracle_periodic_pwnodes(bs0, delta, npw, ws, ts);
        iper = 0;
        npwlevel = 0;
        delta0 = 1.0;
        computecoll(&T, 0);
    } else if (npw > 0) {
        oracle_pwnodes(pmax, npw, ws, ts);
    }
    double xq[64], wq[64], uq[64 * 64];
    oracle_polyexps(ipoly, n, xq, wq, uq);
    int nlevend = nlevels;
    if (npwlevel < nlevels) nlevend = npwlevel;

    int *ifpwexp = (int *)malloc(sizeof(int) * nboxes);
    find_pwexp_boxes(&T, npwlevel, ifpwexp);
    int all_pw = 0;
    if (d1_mode == 1 && npwlevel <= 0 && nlevels >= 1) {
        all_pw = 1;
        for (int i = 0; i < nboxes; ++i) ifpwexp[i] = 1;
    }
    const int mnl = 1 + mnbors + mnbors * mc; /* generous bound on list1 */
    int *nlist1 = (int *)calloc(nboxes, sizeof(int));
    int *list1 = (int *)calloc((size_t)mnl * nboxes, sizeof(int));
    modified_list1(&T, npwlevel, ifpwexp, iper, mnl, nlist1, list1);

    int kinds[10][3];
    const int ncomp = pgh_components(ndim, ifpgh, kinds);
    size_t nn = (size_t)n * n;
    size_t ltsz = nn * 13 * (nlevels + 1);
    double *tab_loc = (double *)malloc(sizeof(double) * ltsz * 3);
    int *ind_loc = (int *)malloc(sizeof(int) * 2 * (n + 1) * 13 * (nlevels + 1));
    for (int l = 0; l <= nlevels; ++l)
        oracle_loctab3(eps, ipoly, n, delta, boxsize[l], tab_loc + nn * 13 * l,
                       tab_loc + ltsz + nn * 13 * l, tab_loc + 2 * ltsz + nn * 13 * l,
                       ind_loc + (size_t)2 * (n + 1) * 13 * l);

    int nlevstart = npwlevel > 0 ? npwlevel : 0;
    int nh = (npw + 1) / 2;
    int nexp = npw > 0 ? nh * ipow(npw, ndim - 1) : 0;
    cplx *tab_p2pw = NULL, *tab_pw2pot = NULL, *wms = NULL, *wm2l = NULL, *rml = NULL;
    int *nlistpw = NULL, *listpw = NULL;
    size_t *iaddr = NULL;
    int nmax = 0;
    if (info) {
        info->npwlevel = npwlevel; info->npw = npw; info->nlevstart = nlevstart;
        info->nexp = nexp; info->periodic_small = periodic_small; info->pmax = pmax;
    }
    int do_pw = npwlevel < nlevels;
    if (do_pw) {
        tab_p2pw = (cplx *)malloc(sizeof(cplx) * (size_t)n * npw * (nlevels + 1));
        /* tab_pw2pot, tab_pw2der, tab_pw2dxx per level (boxfgt.f:639-647) */
        tab_pw2pot = (cplx *)malloc(sizeof(cplx) * (size_t)n * npw * (nlevels + 1) * 3);
        for (int l = nlevstart; l <= nlevels; ++l) {
            size_t pl = (size_t)n * npw * (nlevels + 1);
            oracle_poly2pw(ipoly, n, npw, ws, ts, delta0, boxsize[l],
                           (double *)(tab_p2pw + (size_t)n * npw * l));
            oracle_pw2pgh(ipoly, n, npw, ts, delta0, boxsize[l],
                          (double *)(tab_pw2pot + (size_t)n * npw * l),
                          (double *)(tab_pw2pot + pl + (size_t)n * npw * l),
                          (double *)(tab_pw2pot + 2 * pl + (size_t)n * npw * l));
        }
        nmax = nlevels - nlevstart;
        wms = (cplx *)malloc(sizeof(cplx) * (size_t)nexp * mc * (nmax > 0 ? nmax : 1));
        mk_merge_split(ndim, boxsize[nlevels] / sqrt(delta0) / 2, npw, ts, nmax, wms);
        wm2l = (cplx *)malloc(sizeof(cplx) * (size_t)nexp * mnbors);
        mk_m2l(ndim, boxsize[nlevstart] / sqrt(delta0), npw, ts, wm2l);
        nlistpw = (int *)calloc(nboxes, sizeof(int));
        listpw = (int *)calloc((size_t)(mnbors - 1) * nboxes, sizeof(int));
        compute_listpw(&T, npwlevel, mnbors - 1, nlistpw, listpw);
        /* expansion storage for levels nlevstart..nlevels (gnd_mpalloc,
         * boxfgt.f:1442-1541): mp then loc, nd*nexp complex each */
        int first = LADDR(&T, 1, nlevstart), last = LADDR(&T, 2, nlevels);
        size_t nb = (size_t)(last - first + 1), per = (size_t)nd * nexp;
        rml = (cplx *)calloc(2 * nb * per, sizeof(cplx));
        iaddr = (size_t *)malloc(sizeof(size_t) * 2 * nboxes);
        if (!rml || !iaddr) {
            free(rml); free(iaddr);
            rml = NULL; iaddr = NULL;
        } else {
            for (int b = first; b <= last; ++b) {
                iaddr[2 * (b - 1)] = (size_t)(b - first) * per;
                iaddr[2 * (b - 1) + 1] = (nb + (size_t)(b - first)) * per;
            }
        }
    }
    tstep[0] = now_s() - t0;
    int rc = 0;
    if (do_pw && !rml) { rc = 2; goto done; }

    if (do_pw) {
        size_t wsz1 = (size_t)n * n * nh + (size_t)n * npw * nh;
        size_t wsz5 = (size_t)n * npw * nh + (size_t)n * n * nh;
        size_t wsz = wsz1 > wsz5 ? wsz1 : wsz5;
        /* STEP 1 (boxfgt.f:710-726) */
        t0 = now_s();
        for (int ilev = nlevels; ilev >= nlevstart; --ilev) {
            const cplx *Tl = tab_p2pw + (size_t)n * npw * ilev;
#pragma omp parallel
            {
                cplx *work = (cplx *)malloc(sizeof(cplx) * wsz);
#pragma omp for schedule(dynamic)
                for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                    if (ifpwexp[ibox - 1] != 1 || NCHILD(&T, ibox) != 0) continue;
                    for (int ind = 0; ind < nd; ++ind)
                        tens_prod_to_pw(ndim, n, nd, ind, fvals + (size_t)nd * npbox * (ibox - 1), npw, Tl,
                                        rml + iaddr[2 * (ibox - 1)] + (size_t)ind * nexp, work);
                }
                free(work);
            }
        }
        tstep[1] = now_s() - t0;
        /* STEP 2 (boxfgt.f:746-765) */
        t0 = now_s();
        for (int ilev = nlevels - 1; ilev >= nlevstart; --ilev) {
            int klev = nlevels - ilev;
#pragma omp parallel for schedule(dynamic)
            for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                int nch = NCHILD(&T, ibox);
                cplx *mp = rml + iaddr[2 * (ibox - 1)];
                if (nch > 0) memset(mp, 0, sizeof(cplx) * (size_t)nd * nexp);
                for (int i = 1; i <= nch; ++i) {
                    int jbox = ICHILD(&T, i, ibox);
                    int k = 0;
                    for (int a = 0; a < ndim; ++a)
                        if (CEN(&T, a + 1, ibox) - CEN(&T, a + 1, jbox) < 0) k += 1 << a;
                    const cplx *w = wms + (size_t)nexp * (k + (size_t)mc * (klev - 1));
                    const cplx *src = rml + iaddr[2 * (jbox - 1)];
                    for (size_t e = 0; e < (size_t)nd * nexp; ++e) mp[e] += src[e] * w[e % nexp];
                }
            }
        }
        tstep[2] = now_s() - t0;
        /* STEP 3 (boxfgt.f:787-804) */
        t0 = now_s();
        {
            int ilev = nlevstart;
            double xmin = boxsize[nlevstart];
#pragma omp parallel for schedule(dynamic)
            for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                cplx *loc = rml + iaddr[2 * (ibox - 1) + 1];
                memcpy(loc, rml + iaddr[2 * (ibox - 1)], sizeof(cplx) * (size_t)nd * nexp);
                for (int j = 0; j < nlistpw[ibox - 1]; ++j) {
                    int jbox = listpw[(size_t)(mnbors - 1) * (ibox - 1) + j];
                    int ind = 0, inc = 1;
                    for (int a = 0; a < ndim; ++a) {
                        double d = CEN(&T, a + 1, ibox) - CEN(&T, a + 1, jbox);
                        int jx = (int)lround(d / xmin);
                        if (iper == 1) {
                            int jxp1 = (int)lround((d - bs0) / xmin);
                            int jxm1 = (int)lround((d + bs0) / xmin);
                            if (abs(jx) > abs(jxp1)) jx = jxp1;
                            if (abs(jx) > abs(jxm1)) jx = jxm1;
                        }
                        ind += (jx + 1) * inc;
                        inc *= 3;
                    }
                    const cplx *w = wm2l + (size_t)nexp * ind;
                    const cplx *src = rml + iaddr[2 * (jbox - 1)];
                    for (size_t e = 0; e < (size_t)nd * nexp; ++e) loc[e] += src[e] * w[e % nexp];
                }
            }
        }
        tstep[3] = now_s() - t0;
        /* STEP 4 (boxfgt.f:827-844) */
        t0 = now_s();
        for (int ilev = nlevstart; ilev <= nlevels - 1; ++ilev) {
            int klev = nlevels - ilev;
#pragma omp parallel for schedule(dynamic)
            for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                int nch = NCHILD(&T, ibox);
                const cplx *locp = rml + iaddr[2 * (ibox - 1) + 1];
                for (int i = 1; i <= nch; ++i) {
                    int jbox = ICHILD(&T, i, ibox);
                    int k = 0;
                    for (int a = 0; a < ndim; ++a)
                        if (CEN(&T, a + 1, jbox) - CEN(&T, a + 1, ibox) < 0) k += 1 << a;
                    const cplx *w = wms + (size_t)nexp * (k + (size_t)mc * (klev - 1));
                    cplx *dst = rml + iaddr[2 * (jbox - 1) + 1];
                    for (size_t e = 0; e < (size_t)nd * nexp; ++e) dst[e] = locp[e] * w[e % nexp];
                }
            }
        }
        tstep[4] = now_s() - t0;
        /* STEP 5 (boxfgt.f:865-881) */
        t0 = now_s();
        for (int ilev = nlevstart; ilev <= nlevels; ++ilev) {
            size_t pl = (size_t)n * npw * (nlevels + 1);
            const cplx *Tk[3] = {tab_pw2pot + (size_t)n * npw * ilev, tab_pw2pot + pl + (size_t)n * npw * ilev,
                                 tab_pw2pot + 2 * pl + (size_t)n * npw * ilev};
#pragma omp parallel
            {
                cplx *work = (cplx *)malloc(sizeof(cplx) * wsz);
#pragma omp for schedule(dynamic)
                for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                    if (ifpwexp[ibox - 1] != 1 || NCHILD(&T, ibox) != 0) continue;
                    for (int ind = 0; ind < nd; ++ind)
                        for (int c = 0; c < ncomp; ++c) {
                            const cplx *Tax[3] = {Tk[kinds[c][0]], Tk[kinds[c][1]], Tk[kinds[c][2]]};
                            size_t os, oo;
                            double *out = comp_out(c, ndim, nd, npbox, ibox, ind, pot, grad, hess, &os, &oo);
                            pw2pot(ndim, n, npw, rml + iaddr[2 * (ibox - 1) + 1] + (size_t)ind * nexp, Tax,
                                   out, os, oo, work);
                        }
                }
                free(work);
            }
        }
        tstep[5] = now_s() - t0;
    }

    /* STEP 6 (boxfgt.f:904-928) */
    t0 = now_s();
    if (!all_pw) {
        for (int ilev = 0; ilev <= nlevend; ++ilev) {
            const double *tabk[3] = {tab_loc + nn * 13 * ilev, tab_loc + ltsz + nn * 13 * ilev,
                                     tab_loc + 2 * ltsz + nn * 13 * ilev};
            const int *indl = ind_loc + (size_t)2 * (n + 1) * 13 * ilev;
#pragma omp parallel
            {
                double *work = (double *)malloc(sizeof(double) * 2 * npbox);
#pragma omp for schedule(dynamic)
                for (int ibox = LADDR(&T, 1, ilev); ibox <= LADDR(&T, 2, ilev); ++ibox) {
                    for (int j = 0; j < nlist1[ibox - 1]; ++j) {
                        int jbox = list1[(size_t)mnl * (ibox - 1) + j];
                        int jlev = T.ilevel[jbox - 1];
                        int ixyz[3];
                        find_loctab_ind(ndim, iper, centers + (size_t)ndim * (ibox - 1),
                                        centers + (size_t)ndim * (jbox - 1), boxsize[jlev], bs0, ixyz);
                        for (int ind = 0; ind < nd; ++ind)
                            for (int c = 0; c < ncomp; ++c) {
                                size_t os, oo;
                                double *out = comp_out(c, ndim, nd, npbox, ibox, ind, pot, grad, hess, &os, &oo);
                                tens_prod_to_potloc(ndim, n, nd, ind, fvals + (size_t)nd * npbox * (jbox - 1),
                                                    out, os, oo, tabk, kinds[c], indl, ixyz, work);
                            }
                    }
                }
                free(work);
            }
        }
    }
    tstep[6] = now_s() - t0;

done:
    if (timeinfo) for (int i = 0; i < 7; ++i) timeinfo[i] = tstep[i];
    free(ncoll); free(coll); free(ws); free(ts); free(ifpwexp); free(nlist1); free(list1);
    free(tab_loc); free(ind_loc); free(tab_p2pw); free(tab_pw2pot); free(wms); free(wm2l);
    free(nlistpw); free(listpw); free(rml); free(iaddr);
    return rc;
}

/* ------------------------------------------------------------------ */
/* STEP 11: extra targets                                              */
/* ------------------------------------------------------------------ */

/* treedata_trans_nd with itype = 0 (tree_data_routs_nd.f:16-95, via
 * ortho_trans_nd, tensor_prod_routs_nd.f:21-300): values -> orthogonal
 * polynomial coefficients on one leaf, x then y then z:
 * c(k,j,i) = sum u(k,k1) u(j,j1) u(i,i1) f(k1,j1,i1); then
 * treedata_evalt_nd (tree_data_routs_nd.f:582-686): each target descends
 * from the root to its leaf (find_childbox_ind, pts_tree_nd.f:713-740: child
 * bit a set when t_a - c_a >= 0; pts_tree_sort, pts_tree_nd.f:1226-1256) and
 * the expansion is evaluated at (t - c) 2/boxsize (ortho_evalt_nd,
 * tensor_prod_routs_nd.f:1726-1907: x, then y, then z).
 * vals(nd,npbox,nboxes) -> out(nd,nt), overwritten. */
int oracle_evalt(int ndim, int nd, int ipoly, int norder, int nboxes, int nlevels,
                 const int *itree, const int *iptr, const double *centers,
                 const double *boxsize, const double *vals, int nt,
                 const double *targ, double *out)
{
    if (ndim < 1 || ndim > 3 || nd < 1 || norder < 1 || norder > 32 || nt < 0) return 1;
    const int n = norder, mc = 1 << ndim, npbox = ipow(n, ndim);
    const int *ilevel = itree + (iptr[1] - 1), *nchild = itree + (iptr[3] - 1);
    const int *ichild = itree + (iptr[4] - 1);
    double xq[64], wq[64], u[64 * 64];
    oracle_polyexps(ipoly, n, xq, wq, u);
    (void)nboxes; (void)nlevels;
#pragma omp parallel
    {
        double *c = (double *)malloc(sizeof(double) * 2 * npbox);
        double *w = c + npbox;
#pragma omp for schedule(dynamic, 64)
        for (int it = 0; it < nt; ++it) {
            const double *t = targ + (size_t)ndim * it;
            int b = 1;
            while (nchild[b - 1] > 0) {
                int k = 0;
                for (int a = 0; a < ndim; ++a)
                    if (t[a] - centers[(size_t)ndim * (b - 1) + a] >= 0) k |= 1 << a;
                b = ichild[(size_t)mc * (b - 1) + k];
            }
            const double sc = 2.0 / boxsize[ilevel[b - 1]];
            double p[3][64];
            for (int a = 0; a < ndim; ++a) {
                double x = (t[a] - centers[(size_t)ndim * (b - 1) + a]) * sc;
                if (ipoly == 0) legepols(x, n - 1, p[a]); else chebpols(x, n - 1, p[a]);
            }
            for (int ind = 0; ind < nd; ++ind) {
                /* coefficients of this leaf, density ind */
                for (int i = 0; i < npbox; ++i) c[i] = vals[((size_t)(b - 1) * npbox + i) * nd + ind];
                int stride = 1;
                for (int a = 0; a < ndim; ++a) {
                    for (int i = 0; i < npbox; ++i) {
                        int ka = (i / stride) % n, base = i - ka * stride;
                        double dd = 0;
                        for (int k1 = 0; k1 < n; ++k1) dd += u[ka + k1 * n] * c[base + k1 * stride];
                        w[i] = dd;
                    }
                    memcpy(c, w, sizeof(double) * npbox);
                    stride *= n;
                }
                /* evaluate: contract x, then y, then z */
                int len = npbox;
                for (int a = 0; a < ndim; ++a) {
                    len /= n;
                    for (int r = 0; r < len; ++r) {
                        double pp = 0;
                        for (int k = 0; k < n; ++k) pp += c[r * n + k] * p[a][k];
                        w[r] = pp;
                    }
                    memcpy(c, w, sizeof(double) * len);
                }
                out[(size_t)it * nd + ind] = c[0];
            }
        }
        free(c);
    }
    return 0;
}
