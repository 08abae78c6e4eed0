// plan.cpp -- host planner for the box FGT (FP64 tables, lists, work lists).
//
// Rules and tables follow the reference (file:line cited per function); the
// organisation (flat work lists per stage and level, 0-based ids, per-level
// masked near-field tables) is this library's own, shaped for the kernels.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fgt {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

int ipow(int b, int e) {
    int r = 1;
    while (e-- > 0) r *= b;
    return r;
}

// Legendre P_0..P_{m} at x (three-term recurrence).
void legendre_all(double x, int m, double* p) {
    p[0] = 1.0;
    if (m == 0) return;
    p[1] = x;
    for (int k = 1; k < m; ++k) p[k + 1] = ((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1);
}

// Chebyshev T_0..T_{m} at x.
void chebyshev_all(double x, int m, double* p) {
    p[0] = 1.0;
    if (m == 0) return;
    p[1] = x;
    for (int k = 2; k <= m; ++k) p[k] = 2 * x * p[k - 1] - p[k - 2];
}

// Gauss-Legendre rule on [-1,1], nodes increasing (Newton on P_n from the
// Chebyshev-like initial guesses; the reference uses legerts,
// src/common/legeexps.f:2399-2530, for the same rule).
void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        double pn = 0, dpn = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 1; k < n; ++k) {
                double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            pn = (n == 0) ? 1.0 : p1;
            dpn = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = pn / dpn;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        {
            double p0 = 1.0, p1 = z;
            for (int k = 1; k < n; ++k) {
                double p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
                p0 = p1;
                p1 = p2;
            }
            dpn = n * (z * p1 - p0) / (z * z - 1.0);
        }
        double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

}  // namespace

// fgtpwterms (src/common/fgtterms.f:72-141).
void pw_terms(double bsize, double delta, double eps, int iperiod, double* pmax, int* npw) {
    if (iperiod == 0) {
        double rmax = bsize / std::sqrt(delta);
        double d = std::sqrt(std::log(1.0 / eps));
        *pmax = 2.0 * d;
        double h = 2 * kPi / (rmax + d);
        *npw = 2 * (static_cast<int>(*pmax / h) + 1);
    } else {
        *npw = 2 * (static_cast<int>(bsize * std::sqrt(std::log(1.0 / eps) / delta) / kPi) + 1) + 1;
    }
}

// legeexps / chebexps, itype = 2 (src/common/legeexps.f:140-222,
// src/common/chebexps.f:12-180): u[i + j*n] = coefficient i from value j.
void poly_exps(int ipoly, int n, double* x, double* w, double* u) {
    std::vector<double> p(n + 1);
    if (ipoly == 0) {
        gauss_legendre(n, x, w);
        for (int j = 0; j < n; ++j) {
            legendre_all(x[j], n - 1, p.data());
            for (int i = 0; i < n; ++i) u[i + j * n] = p[i] * w[j] * (2 * i + 1) / 2.0;
        }
    } else {
        const double h = kPi / (2 * n);
        for (int i = 1; i <= n; ++i) x[n - i] = std::cos((2 * i - 1) * h);
        for (int j = 0; j < n; ++j) {
            chebyshev_all(x[j], n - 1, p.data());
            double wt = 2.0;
            for (int k = 2; k <= n - 1; k += 2) wt -= 4.0 / (double(k) * k - 1.0) * p[k];
            w[j] = wt / n;
            for (int i = 0; i < n; ++i) u[i + j * n] = (i == 0 ? 1.0 : 2.0) / n * p[i];
        }
    }
}

// Spherical Bessel j_0..j_nterms at real z >= 0 by Miller's backward
// recurrence (the method of src/common/besseljs3d.f:5-168, scale = 1).
void sph_bessel(int nterms, double z, double* fjs) {
    const double up2 = 1.0e40, up2inv = 1.0e-40;
    if (std::fabs(z) < 1.0e-200) {
        fjs[0] = 1.0;
        for (int i = 1; i <= nterms; ++i) fjs[i] = 0.0;
        return;
    }
    const double zinv = 1.0 / z;
    double fm1 = 0.0, f0 = 1.0;
    int ntop = nterms + 100000;
    for (int i = nterms; i <= nterms + 100000; ++i) {
        double f1 = (2 * i + 1.0) * zinv * f0 - fm1;
        if (f1 * f1 > up2) {
            ntop = i + 1;
            break;
        }
        fm1 = f0;
        f0 = f1;
    }
    std::vector<double> t(ntop + 1, 0.0);
    std::vector<char> rescaled(ntop + 1, 0);
    t[ntop] = 0.0;
    t[ntop - 1] = 1.0;
    for (int i = ntop - 1; i >= 1; --i) {
        t[i - 1] = (2 * i + 1.0) * zinv * t[i] - t[i + 1];
        if (t[i - 1] * t[i - 1] > up2) {
            t[i] *= up2inv;
            t[i - 1] *= up2inv;
            rescaled[i] = 1;
        }
    }
    double sc = 1.0;
    for (int i = 1; i <= ntop; ++i) {
        if (rescaled[i - 1]) sc *= up2inv;
        t[i] *= sc;
    }
    const double j0 = std::sin(z) * zinv;
    const double j1 = j0 * zinv - std::cos(z) * zinv;
    const double s = (std::fabs(j1) > std::fabs(j0)) ? j1 / t[1] : j0 / t[0];
    for (int i = 0; i <= nterms; ++i) fjs[i] = t[i] * s;
}

// Table value -> plane wave for one level (mk_leg2pw_tables /
// mk_cheb2pw_tables, src/bfgt/bfgt_volrouts.f:58-275), written as
// out[k*n + m] (frequency k, value m).
void make_p2pw(int ipoly, int n, int npw, const double* ws, const double* ts,
               double delta, double boxdim, cplx* out) {
    const cplx I(0.0, 1.0);
    std::vector<cplx> ft(static_cast<size_t>(n) * npw);  // ft[k*n + m]: transform of p_m at k
    std::vector<double> fj(n + 1), xl(n), wl(n), ul(n * n), xc(n), wc(n), uc(n * n), p(n + 1);
    poly_exps(0, n, xl.data(), wl.data(), ul.data());
    std::vector<double> c2l;  // Legendre coefficients of T_i: c2l[j + i*n]
    if (ipoly == 1) {
        c2l.assign(n * n, 0.0);
        std::vector<double> tv(n * n);
        for (int k = 0; k < n; ++k) {
            chebyshev_all(xl[k], n - 1, p.data());
            for (int i = 0; i < n; ++i) tv[i + k * n] = p[i];
        }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double s = 0;
                for (int k = 0; k < n; ++k) s += ul[j + k * n] * tv[i + k * n];
                c2l[j + i * n] = s;
            }
    }
    const double fac = boxdim / (2 * std::sqrt(delta));
    for (int k = npw / 2; k < npw; ++k) {
        const double dd = ws[k] * boxdim;
        cplx* col = &ft[static_cast<size_t>(k) * n];
        cplx* mir = &ft[static_cast<size_t>(npw - 1 - k) * n];
        if (std::fabs(ts[k]) < 1e-12) {
            // zero frequency: int p_m = 2 delta_{m0} (Legendre), 2/(1-m^2) even m (Chebyshev)
            for (int m = 0; m < n; ++m) {
                if (ipoly == 0) col[m] = (m == 0) ? cplx(dd) : cplx(0.0);
                else col[m] = (m % 2 == 0) ? cplx(dd / (1.0 - double(m) * m)) : cplx(0.0);
            }
            if (ipoly == 1 && k != npw - 1 - k)
                for (int m = 0; m < n; ++m) mir[m] = std::conj(col[m]);
            continue;
        }
        sph_bessel(n, ts[k] * fac, fj.data());
        // int_{-1}^1 P_m(x) e^{-i a x} dx = 2 j_m(a) / i^m
        std::vector<cplx> leg(n);
        cplx im = 1.0;
        for (int m = 0; m < n; ++m) {
            leg[m] = dd * fj[m] / im;
            im *= I;
        }
        if (ipoly == 0) {
            for (int m = 0; m < n; ++m) {
                col[m] = leg[m];
                mir[m] = std::conj(leg[m]);
            }
        } else {
            for (int i = 0; i < n; ++i) {
                cplx s = 0;
                for (int j = 0; j < n; ++j) s += c2l[j + i * n] * leg[j];
                col[i] = s;
            }
            if (k != npw - 1 - k)
                for (int i = 0; i < n; ++i) mir[i] = std::conj(col[i]);
        }
    }
    const double* u = ul.data();
    if (ipoly == 1) {
        poly_exps(1, n, xc.data(), wc.data(), uc.data());
        u = uc.data();
    }
    // compose with u (values -> coefficients) so the table acts on values
    for (int k = 0; k < npw; ++k)
        for (int m = 0; m < n; ++m) {
            cplx s = 0;
            for (int c = 0; c < n; ++c) s += ft[static_cast<size_t>(k) * n + c] * u[c + m * n];
            out[static_cast<size_t>(k) * n + m] = s;
        }
}

// Plane wave -> node values for one level (mk_pw2pgh_tables, potential part,
// src/bfgt/bfgt_volrouts.f:339-401), out[i*npw + k].
void make_pw2pot(int n, int npw, const double* ts, const double* xs, double delta,
                 double boxdim, cplx* out) {
    const double dsq = boxdim / 2 / std::sqrt(delta);
    for (int i = 0; i < n; ++i) {
        const double x = xs[i] * dsq;
        for (int k = npw / 2; k < npw; ++k) {
            const cplx e = std::exp(cplx(0.0, ts[k] * x));
            out[static_cast<size_t>(i) * npw + k] = e;
            out[static_cast<size_t>(i) * npw + (npw - 1 - k)] = std::conj(e);
        }
    }
}

// Plane wave -> node values and their derivatives (mk_pw2pgh_tables,
// src/bfgt/bfgt_volrouts.f:339-401): for the upper half k >= npw/2,
// e = exp(i t_k x_i D/(2 sqrt delta)), der = t_k e dfac, dxx = t_k der dfac
// with dfac = i (D/(2 sqrt delta)) (2/D); the lower half is the conjugate.
void make_pw2pgh(int n, int npw, const double* ts, const double* xs, double delta,
                 double boxdim, cplx* out0, cplx* out1, cplx* out2) {
    const double dsq = boxdim / 2 / std::sqrt(delta);
    const cplx dfac = cplx(0.0, 1.0) * dsq * 2.0 / boxdim;
    for (int i = 0; i < n; ++i) {
        const double x = xs[i] * dsq;
        for (int k = npw / 2; k < npw; ++k) {
            const cplx e = std::exp(cplx(0.0, ts[k] * x));
            const cplx ex = ts[k] * e * dfac;
            const cplx exx = ts[k] * ex * dfac;
            const size_t a = static_cast<size_t>(i) * npw + k, b = static_cast<size_t>(i) * npw + (npw - 1 - k);
            out0[a] = e;
            out0[b] = std::conj(e);
            if (out1) { out1[a] = ex; out1[b] = std::conj(ex); }
            if (out2) { out2[a] = exx; out2[b] = std::conj(exx); }
        }
    }
}

namespace {

// I_m(t)/lambda integrals of P_m e^{-(t-x)^2/lambda^2} over [-1,1] by the
// five-term recurrence, and their first/second derivative forms
// (mk_legetab_recur, src/bfgt/bfgt_volrouts.f:1095-1188), lambda <= 1/8.
void lege_recur(int m, double t, double lam, double bdim, double* fi, double* fx, double* fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[100], fpp[100];
    const double l2 = lam * lam;
    const double a = (-1 - t) / lam, b = (1 - t) / lam;
    const double ea = std::exp(-a * a), eb = std::exp(-b * b);
    const double d1 = lam * (eb - ea), d2 = lam * (eb + ea);
    fi[0] = sqrtpih * (std::erf(b) - std::erf(a));
    if (m > 1) fi[1] = -d1 / 2 + t * fi[0];
    if (m > 2) fi[2] = 0.75 * (-d2 - t * d1 + (l2 - 2.0 / 3 + 2 * t * t) * fi[0]);
    if (m > 3) fi[3] = 5.0 * (-d1 + 2 * t * (4.0 * fi[2] / 3 - fi[0] / 3) - (0.4 - 4 * l2) * fi[1]) / 8;
    for (int k = 5; k <= m; ++k) {
        const int q = k - 3;
        const double v = t * (fi[k - 2] - fi[k - 4]) +
                         (2 * q + 1) * (l2 / 2 + 1.0 / ((2 * q + 3) * (2 * q - 1))) * fi[k - 3] +
                         (q - 1.0) * fi[k - 5] / (2 * q - 1);
        fi[k - 1] = (2 * q + 3.0) * v / (q + 2);
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fi[0] * lam;
    for (int k = 3; k <= m; ++k) fp[k - 1] = fp[k - 3] + (2 * k - 3) * fi[k - 2] * lam;
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    for (int k = 3; k <= m; ++k) fpp[k - 1] = fpp[k - 3] + (2 * k - 3) * fp[k - 2];
    int sg = 1;
    for (int k = 1; k <= m; ++k) {
        fxx[k - 1] = fpp[k - 1] - (k - 1) * k / 2.0 * (eb + sg * ea) - 2 * (b * eb - sg * a * ea) / lam;
        sg = -sg;
    }
    const double ww = bdim * lam / 2;
    for (int k = 0; k < m; ++k) {
        fi[k] *= ww;
        fxx[k] = fxx[k] * 2 / bdim;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lam;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lam;
}

// Chebyshev analogue (mk_chebtab_recur, src/bfgt/bfgt_volrouts.f:1191-1294).
void cheb_recur(int m, double t, double lam, double bdim, double* fi, double* fx, double* fxx) {
    const double sqrtpih = 0.886226925452757940959713778283913;
    double fp[100], fpp[100];
    const double l2 = lam * lam;
    const double a = (-1 - t) / lam, b = (1 - t) / lam;
    const double ea = std::exp(-a * a), eb = std::exp(-b * b);
    const double d1 = lam * (eb - ea), d2 = lam * (eb + ea);
    fi[0] = sqrtpih * (std::erf(b) - std::erf(a));
    if (m > 1) fi[1] = -d1 / 2 + t * fi[0];
    if (m > 2) fi[2] = -d2 - t * d1 + (l2 - 1 + 2 * t * t) * fi[0];
    if (m > 3) fi[3] = -d1 + 2 * t * fi[2] - (1 - 4 * l2) * fi[1];
    int sg = 1;
    for (int k = 5; k <= m; ++k) {
        const double dc = (sg == 1) ? d2 : d1;
        const int q = k - 1;
        const double dd = (q - 1.0) / (q - 3);
        fi[k - 1] = 2 * t * fi[k - 2] + 2 * ((q - 1) * l2 + 1.0 / (q - 3)) * fi[k - 3] +
                    dd * (fi[k - 5] - 2 * t * fi[k - 4]) - (1 - dd) * dc;
        sg = -sg;
    }
    fp[0] = 0;
    if (m > 1) fp[1] = fi[0] * lam;
    if (m > 2) fp[2] = 4 * fi[1] * lam;
    for (int k = 4; k <= m; ++k) fp[k - 1] = (fp[k - 3] / (k - 3) + 2 * fi[k - 2] * lam) * (k - 1);
    fpp[0] = 0;
    if (m > 1) fpp[1] = 0;
    if (m > 2) fpp[2] = 4 * fi[0] * lam;
    for (int k = 4; k <= m; ++k) fpp[k - 1] = (fpp[k - 3] / (k - 3) + 2 * fp[k - 2]) * (k - 1);
    sg = 1;
    for (int k = 1; k <= m; ++k) {
        fxx[k - 1] = fpp[k - 1] - double(k - 1) * (k - 1) * (eb + sg * ea) - 2 * (b * eb - sg * a * ea) / lam;
        sg = -sg;
    }
    const double ww = bdim * lam / 2, sc = 2 / bdim;
    for (int k = 0; k < m; ++k) {
        fi[k] *= ww;
        fxx[k] *= sc;
    }
    for (int k = 2; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d2 / lam;
    for (int k = 1; k <= m; k += 2) fx[k - 1] = fp[k - 1] - d1 / lam;
}

// Row bounds (1-based, 0/-1 = none) of the significant entries of each
// column of an n x n column-major matrix, plus the column range in slot n
// (compute_sparse_pattern, src/bfgt/bfgt_volrouts.f:1300-1378).
void band_pattern(double eps, int n, const double* a, int* ind) {
    for (int i = 0; i <= n; ++i) {
        ind[2 * i] = 0;
        ind[2 * i + 1] = -1;
    }
    double amax = 0;
    for (int i = 0; i < n * n; ++i) amax = std::max(amax, std::fabs(a[i]));
    if (amax <= 2e-16) return;
    const double thr = eps * amax / n;
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r)
            if (std::fabs(a[r + c * n]) > thr) { ind[2 * c] = r + 1; break; }
        for (int r = n - 1; r >= 0; --r)
            if (std::fabs(a[r + c * n]) > thr) { ind[2 * c + 1] = r + 1; break; }
    }
    for (int c = 0; c < n; ++c)
        if (ind[2 * c] > 0) { ind[2 * n] = c + 1; break; }
    for (int c = n - 1; c >= 0; --c)
        if (ind[2 * c + 1] > 0) { ind[2 * n + 1] = c + 1; break; }
}

}  // namespace

// Near-field tables for one target level (mk_loctab_all with mrefinelev = 0
// and mk_loctab, src/bfgt/bfgt_volrouts.f:407-660,920-1060).  Output:
//   tab[slot][k][j] masked to the band, band[slot][k] = (lo, hi) 0-based,
//   cols[slot] = (first, last) 0-based target columns (first > last: empty).
int loc_nloctab(int mref) { return (1 << (mref + 1)) * (mref + 3); }

void make_loctab_level(double eps, int ipoly, int n, double delta, double boxdim,
                       double* tab, int* band, int* cols, double* tabx, double* tabxx, int mref) {
    constexpr int nq = 50;
    const int nlt = loc_nloctab(mref), nslot = 2 * nlt + 1;
    std::vector<double> tn(n), tw(n), tu(n * n), xq(nq), wq(nq), pv(nq * n), p(n + 1);
    poly_exps(ipoly, n, tn.data(), tw.data(), tu.data());
    gauss_legendre(nq, xq.data(), wq.data());
    for (int i = 0; i < nq; ++i) {
        if (ipoly == 0) legendre_all(xq[i], n - 1, p.data());
        else chebyshev_all(xq[i], n - 1, p.data());
        for (int j = 0; j < n; ++j) pv[i + j * nq] = p[j];
    }
    const size_t nn = static_cast<size_t>(n) * n;
    // full (unmasked) tables for pot/grad/hess kinds, Fortran layout (m=row, j=col)
    std::vector<double> full(nn * nslot * 3, 0.0);
    auto F = [&](int kind, int k) { return &full[(static_cast<size_t>(kind) * nslot + (k + nlt)) * nn]; };
    std::vector<double> raw(nn * 3), wexp(static_cast<size_t>(nq) * n);
    double fi[100], fx[100], fxx[100];
    // geometric cases (mk_loctab_all, bfgt_volrouts.f:565-580): source /
    // target size ratio 2^i0 for i0 = m0 .. -1 (m0 = mref + 1), target
    // centres k / 2^m0 source half-widths apart, k <= 0 up to -nloctab; +k
    // by mirror symmetry.  mref = 0 gives k = -1, -3 (ratio 2), 0, -4 (1),
    // -2, -6 (1/2).
    const int m0 = mref + 1;
    std::vector<int> ks;
    std::vector<double> scales;
    for (int i0 = m0; i0 >= -1; --i0)
        for (int j0 = 1; j0 <= (1 << m0); ++j0) {
            int k;
            if (i0 > 0) k = -(1 << (m0 - i0)) - (1 << (m0 - i0 + 1)) * (j0 - 1);
            else if (i0 == 0) k = -(1 << (m0 + 1)) * (j0 - 1);
            else k = -(1 << m0) - (1 << (m0 + 1)) * (j0 - 1);
            if (k < -nlt) break;
            ks.push_back(k);
            scales.push_back(std::ldexp(1.0, i0));
        }
    std::vector<char> made(nslot, 0);
    for (size_t c = 0; c < ks.size(); ++c) {
        const int k = ks[c];
        made[k + nlt] = made[-k + nlt] = 1;
        const double scale = scales[c], cen = std::ldexp(static_cast<double>(k), -m0);
        double tg[64];
        for (int j = 0; j < n; ++j) tg[j] = tn[j] / scale + cen;
        const double bs = boxdim * scale;
        const double sigma = 4 * delta / (bs * bs), lam = std::sqrt(sigma);
        const double dmax = std::sqrt(std::log(1.0 / 1e-16) * sigma);
        double* r0 = raw.data();
        double* r1 = r0 + nn;
        double* r2 = r1 + nn;
        if (lam <= 0.125) {
            for (int j = 0; j < n; ++j) {
                double dis = 0;
                if (tg[j] < -1) dis = -1 - tg[j];
                if (tg[j] > 1) dis = tg[j] - 1;
                if (dis > dmax) {
                    for (int m = 0; m < n; ++m) r0[m + j * n] = r1[m + j * n] = r2[m + j * n] = 0;
                    continue;
                }
                if (ipoly == 0) lege_recur(n, tg[j], lam, bs, fi, fx, fxx);
                else cheb_recur(n, tg[j], lam, bs, fi, fx, fxx);
                for (int m = 0; m < n; ++m) {
                    r0[m + j * n] = fi[m];
                    r1[m + j * n] = fx[m];
                    r2[m + j * n] = fxx[m];
                }
            }
        } else {
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < nq; ++i) {
                    const double dx = tg[j] - xq[i];
                    wexp[i + j * nq] = (std::fabs(dx) < dmax) ? std::exp(-dx * dx / sigma) * wq[i] : 0.0;
                }
            const double bsh = bs / 2;
            for (int j = 0; j < n; ++j) {
                double dis = 0;
                if (tg[j] < -1) dis = -1 - tg[j];
                if (tg[j] > 1) dis = tg[j] - 1;
                if (dis > dmax) {
                    for (int m = 0; m < n; ++m) r0[m + j * n] = r1[m + j * n] = r2[m + j * n] = 0;
                    continue;
                }
                for (int m = 0; m < n; ++m) {
                    double s0 = 0, s1 = 0, s2 = 0;
                    for (int i = 0; i < nq; ++i) {
                        const double dx = tg[j] - xq[i];
                        const double dd = -2 * dx / sigma * 2 / bs;
                        const double pw = pv[i + m * nq] * wexp[i + j * nq];
                        s0 += pw;
                        s1 += pw * dd;
                        s2 += pw * (dd * dd - 2 / delta);
                    }
                    r0[m + j * n] = s0 * bsh;
                    r1[m + j * n] = s1 * bsh;
                    r2[m + j * n] = s2 * bsh;
                }
            }
        }
        // act on values: T_k = u^T raw
        for (int kind = 0; kind < 3; ++kind) {
            double* dst = F(kind, k);
            const double* src = raw.data() + kind * nn;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    double s = 0;
                    for (int m = 0; m < n; ++m) s += tu[m + i * n] * src[m + j * n];
                    dst[i + j * n] = s;
                }
            // mirror: T_{-k}(m,j) = +-T_k(n-1-m, n-1-j) (sign - for the first
            // derivative); for k = 0 this runs in place, in the same order
            double* mir = F(kind, -k);
            const double sg = (kind == 1) ? -1.0 : 1.0;
            for (int j = 0; j < n; ++j)
                for (int m = 0; m < n; ++m) mir[m + j * n] = sg * dst[(n - 1 - m) + (n - 1 - j) * n];
        }
    }
    // bands: union of the pot/grad/hess patterns, as mk_loctab_all forms it
    // (offsets no geometric case produces stay empty)
    for (int k = -nlt; k <= nlt; ++k) {
        const int slot = k + nlt;
        int a[2 * 65], b[2 * 65], c[2 * 65], u[2 * 65];
        if (!made[slot]) {
            for (int j = 0; j <= n; ++j) { u[2 * j] = 0; u[2 * j + 1] = -1; }
        } else {
            band_pattern(eps, n, F(0, k), a);
            band_pattern(eps, n, F(1, k), b);
            band_pattern(eps, n, F(2, k), c);
            for (int j = 0; j <= n; ++j) {
                int i2 = std::max(a[2 * j + 1], std::max(b[2 * j + 1], c[2 * j + 1]));
                int i1 = a[2 * j];
                if (b[2 * j] != 0 && i1 > b[2 * j]) i1 = b[2 * j];
                if (c[2 * j] != 0 && i1 > c[2 * j]) i1 = c[2 * j];
                if (i2 > 0 && i1 == 0) i1 = 1;
                u[2 * j] = i1;
                u[2 * j + 1] = i2;
            }
        }
        // column range and masked table in [k][j] layout
        cols[2 * slot] = u[2 * n] - 1;
        cols[2 * slot + 1] = u[2 * n + 1] - 1;
        double* outs[3] = {tab, tabx, tabxx};
        for (int kk = 0; kk < n; ++kk) {
            int lo = u[2 * kk] - 1, hi = u[2 * kk + 1] - 1;
            const bool colon = kk >= cols[2 * slot] && kk <= cols[2 * slot + 1];
            if (!colon) { lo = 0; hi = -1; }
            band[2 * (slot * n + kk)] = lo;
            band[2 * (slot * n + kk) + 1] = hi;
            for (int kind = 0; kind < 3; ++kind) {
                if (!outs[kind]) continue;
                const double* src = F(kind, k);
                for (int j = 0; j < n; ++j)
                    outs[kind][(static_cast<size_t>(slot) * n + kk) * n + j] =
                        (j >= lo && j <= hi) ? src[j + kk * n] : 0.0;
            }
        }
    }
}

namespace {

void set_err(std::string* err, const std::string& m) {
    if (err) *err = m;
}

}  // namespace

// Modified list 1 of leaf b (gnd_compute_modified_list1,
// src/common/tree_routs_nd.f:864-1055): the boxes whose near field b takes
// directly -- leaf colleagues (b itself unless b forms a plane-wave
// expansion), leaf children of non-leaf colleagues below the cutoff level,
// and leaf colleagues of the parent that touch b.  Empty for non-leaves and
// for boxes below the cutoff level.  The relation is symmetric, so the list
// serves b as a target.
void modified_list1(const Plan& P, int b, std::vector<int>& out) {
    const TreeView& T = P.tree;
    const int ndim = T.ndim, mc = T.mc, mnb = T.mnbors;
    const int l = T.level[b];
    out.clear();
    if (l > std::min(P.npwlevel, T.nlevels)) return;
    const double bs0 = P.bs0;
    auto wrapdis = [&](double d) {
        double dis = std::fabs(d);
        if (P.iper_eff == 1) dis = std::min(dis, bs0 - dis);
        return dis;
    };
    if (l == 0) {
        if (T.nchild[0] == 0) out.push_back(0);
        return;
    }
    if (T.nchild[b] != 0 || P.all_pw) return;
    const int dad = T.parent[b];
    for (int i = 0; i < T.ncoll[b]; ++i) {
        const int jb = T.coll[static_cast<size_t>(mnb) * b + i];
        if (T.nchild[jb] == 0) {
            if (P.ifpwexp[b] != 1 || jb != b) out.push_back(jb);
        } else if (l < P.npwlevel) {
            const double dt = 1.05 * (T.boxsize[l] + T.boxsize[l + 1]) / 2.0;
            for (int j = 0; j < mc; ++j) {
                const int kb = T.child[static_cast<size_t>(mc) * jb + j];
                if (kb < 0 || T.nchild[kb] != 0) continue;
                bool ok = true;
                for (int a = 0; a < ndim; ++a)
                    if (wrapdis(T.centers[ndim * kb + a] - T.centers[ndim * b + a]) >= dt) { ok = false; break; }
                if (ok) out.push_back(kb);
            }
        }
    }
    for (int i = 0; i < T.ncoll[dad]; ++i) {
        const int jb = T.coll[static_cast<size_t>(mnb) * dad + i];
        if (T.nchild[jb] != 0) continue;
        const double dt = 1.05 * (T.boxsize[l] + T.boxsize[l - 1]) / 2.0;
        bool ok = true;
        for (int a = 0; a < ndim; ++a)
            if (wrapdis(T.centers[ndim * jb + a] - T.centers[ndim * b + a]) >= dt) { ok = false; break; }
        if (ok) out.push_back(jb);
    }
}

int make_plan(const PlanInput& in, Plan& P, std::string* err) {
    // ---- validation (the reference has none; we return a status instead) ----
    if (in.ndim < 1 || in.ndim > 3) { set_err(err, "ndim must be 1, 2 or 3"); return 2; }
    if (in.nd < 1) { set_err(err, "nd must be >= 1"); return 2; }
    if (in.norder < 2 || in.norder > 16) { set_err(err, "norder must be in [2,16]"); return 2; }
    if (in.ipoly != 0 && in.ipoly != 1) { set_err(err, "ipoly must be 0 or 1"); return 2; }
    if (in.iperiod != 0 && in.iperiod != 1) { set_err(err, "iperiod must be 0 or 1"); return 2; }
    if (!(in.delta > 0)) { set_err(err, "delta must be > 0"); return 2; }
    if (!(in.eps > 0 && in.eps < 1)) { set_err(err, "eps must be in (0,1)"); return 2; }
    if (in.nboxes < 1 || in.nlevels < 0 || in.nlevels > 60) { set_err(err, "bad nboxes/nlevels"); return 2; }
    const int ndim = in.ndim, n = in.norder, mc = 1 << ndim, mnb = ipow(3, ndim);
    if (in.npbox != ipow(n, ndim)) { set_err(err, "npbox != norder**ndim"); return 2; }
    const int nb = in.nboxes, nl = in.nlevels;
    const int64_t need = static_cast<int64_t>(4 + mc + mnb) * nb + 2 * (nl + 1);
    if (in.ltree < need) { set_err(err, "ltree too small for nboxes"); return 2; }
    const int expect[8] = {1, 2 * (nl + 1) + 1, 0, 0, 0, 0, 0, 0};
    if (in.iptr[0] != expect[0] || in.iptr[1] != expect[1] || in.iptr[2] != in.iptr[1] + nb ||
        in.iptr[3] != in.iptr[2] + nb || in.iptr[4] != in.iptr[3] + nb ||
        in.iptr[5] != in.iptr[4] + mc * nb || in.iptr[6] != in.iptr[5] + nb) {
        set_err(err, "iptr does not describe a vol_tree_build layout");
        return 2;
    }

    P = Plan();
    P.nd = in.nd; P.ndim = ndim; P.n = n; P.ipoly = in.ipoly; P.iperiod = in.iperiod;
    P.npbox = in.npbox; P.nlevels = nl; P.delta = in.delta; P.eps = in.eps;
    P.bs0 = in.boxsize[0];

    // ---- tree arrays (1-based Fortran -> 0-based) ----
    TreeView& T = P.tree;
    T.ndim = ndim; T.nlevels = nl; T.mc = mc; T.mnbors = mnb;
    const int* it = in.itree;
    T.lfirst.resize(nl + 1); T.llast.resize(nl + 1);
    for (int l = 0; l <= nl; ++l) {
        T.lfirst[l] = it[2 * l] - 1;
        T.llast[l] = it[2 * l + 1] - 1;
        if (T.llast[l] >= nb || (T.lfirst[l] < 0 && T.llast[l] >= 0)) {
            set_err(err, "laddr out of range"); return 2;
        }
    }
    if (T.lfirst[0] != 0 || T.llast[0] != 0) { set_err(err, "level 0 must hold only box 1"); return 2; }
    // the tree's boxes are 1..laddr(2,nlevels); nboxes may be larger (room
    // a caller keeps for refinement, vol_tree_mem with ifnewtree = 1): the
    // sections of itree are laid out for nboxes, boxes past the tree's own
    // are not read
    const int nba = T.llast[nl] + 1;
    if (nba < 1) { set_err(err, "laddr out of range"); return 2; }
    P.nboxes = nba;
    T.nboxes = nba;
    T.level.resize(nba); T.parent.resize(nba); T.nchild.resize(nba);
    T.child.resize(static_cast<size_t>(mc) * nba); T.ncoll.resize(nba);
    T.coll.resize(static_cast<size_t>(mnb) * nba);
    auto bad_id = [&](int v) { return v < -1 || v == 0 || v > nba; };
    for (int b = 0; b < nba; ++b) {
        T.level[b] = it[in.iptr[1] - 1 + b];
        T.parent[b] = it[in.iptr[2] - 1 + b] - (it[in.iptr[2] - 1 + b] > 0 ? 1 : 0);
        T.nchild[b] = it[in.iptr[3] - 1 + b];
        T.ncoll[b] = it[in.iptr[5] - 1 + b];
        if (T.level[b] < 0 || T.level[b] > nl || T.nchild[b] < 0 || T.nchild[b] > mc ||
            T.ncoll[b] < 0 || T.ncoll[b] > mnb) {
            set_err(err, "invalid level/nchild/ncoll entry"); return 2;
        }
        for (int j = 0; j < mc; ++j) {
            const int v = it[in.iptr[4] - 1 + mc * b + j];
            if (bad_id(v) && j < T.nchild[b]) { set_err(err, "invalid ichild entry"); return 2; }
            T.child[static_cast<size_t>(mc) * b + j] = (v > 0 && v <= nba) ? v - 1 : -1;
        }
        for (int j = 0; j < mnb; ++j) {
            const int v = it[in.iptr[6] - 1 + mnb * b + j];
            if (j < T.ncoll[b] && (v < 1 || v > nba)) { set_err(err, "invalid coll entry"); return 2; }
            T.coll[static_cast<size_t>(mnb) * b + j] = (v > 0 && v <= nba) ? v - 1 : -1;
        }
        for (int j = 0; j < T.nchild[b]; ++j)
            if (T.child[static_cast<size_t>(mc) * b + j] < 0) { set_err(err, "missing child id"); return 2; }
    }
    // levels agree with laddr; every box below the root has an in-range
    // parent one level up, and every child sits one level down
    for (int l = 0; l <= nl; ++l)
        for (int b = T.lfirst[l]; b >= 0 && b <= T.llast[l]; ++b)
            if (T.level[b] != l) { set_err(err, "box level disagrees with laddr"); return 2; }
    for (int b = 0; b < nba; ++b) {
        const int l = T.level[b];
        if (b < T.lfirst[l] || b > T.llast[l]) { set_err(err, "box outside its level's laddr range"); return 2; }
        if (l >= 1) {
            const int p = T.parent[b];
            if (p < 0 || p >= nba || T.level[p] != l - 1) { set_err(err, "invalid iparent entry"); return 2; }
        }
        for (int j = 0; j < T.nchild[b]; ++j)
            if (T.level[T.child[static_cast<size_t>(mc) * b + j]] != l + 1) {
                set_err(err, "child not one level below its parent"); return 2;
            }
    }
    T.centers.assign(in.centers, in.centers + static_cast<size_t>(ndim) * nba);
    T.boxsize.assign(in.boxsize, in.boxsize + nl + 1);

    // ---- cutoff level and plane-wave length (boxfgt.f:132-162) ----
    const double dcut = std::sqrt(in.delta * std::log(1.0 / in.eps));
    int npwlevel = nl + 1;
    for (int i = nl; i >= 0; --i)
        if (T.boxsize[i] >= dcut) { npwlevel = i; break; }
    if (T.boxsize[nl] > dcut) npwlevel = nl + 1;
    if (T.boxsize[0] <= dcut) npwlevel = 0;
    int npw = 0;
    double pmax = 0;
    if (npwlevel <= nl) {
        if (npwlevel <= 1) pw_terms(2 * T.boxsize[std::max(npwlevel, 0)], in.delta, in.eps, in.iperiod, &pmax, &npw);
        else pw_terms(2 * T.boxsize[npwlevel], in.delta, in.eps, 0, &pmax, &npw);
    }
    if (npw > 200) { set_err(err, "npw exceeds 200 (delta too small for the tree's coarse levels)"); return 3; }

    // ---- STEP 0: nodes, periodic branch (boxfgt.f:475-533) ----
    P.delta0 = in.delta;
    P.iper_eff = in.iperiod;
    P.ws.assign(std::max(npw, 1), 0.0);
    P.ts.assign(std::max(npw, 1), 0.0);
    if (npwlevel <= 1 && in.iperiod == 1) {
        // Fourier series of the periodic kernel over the whole box
        // (get_periodic_pwnodes, fgtterms.f:191-211); colleagues are then
        // free-space ones (computecoll with iperiod = 0), the tables use delta = 1.
        P.periodic_small = true;
        const double h = 2 * kPi / P.bs0, w = std::sqrt(kPi * in.delta) / P.bs0, hw = h * std::sqrt(in.delta) / 2;
        for (int j = -npw / 2; j <= (npw - 1) / 2; ++j) {
            P.ts[j + npw / 2] = j * h;
            P.ws[j + npw / 2] = w * std::exp(-(hw * j) * (hw * j));
        }
        P.iper_eff = 0;
        npwlevel = 0;
        P.delta0 = 1.0;
        // recompute colleagues without periodic wrap (computecoll,
        // tree_routs_nd.f:106-242) -- into our own copy, the caller's tree is untouched
        std::fill(T.ncoll.begin(), T.ncoll.end(), 0);
        std::fill(T.coll.begin(), T.coll.end(), -1);
        T.ncoll[0] = 1;
        T.coll[0] = 0;
        for (int l = 1; l <= nl; ++l)
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                const int dad = T.parent[b];
                for (int i = 0; i < T.ncoll[dad]; ++i) {
                    const int jb = T.coll[static_cast<size_t>(mnb) * dad + i];
                    for (int j = 0; j < mc; ++j) {
                        const int kb = T.child[static_cast<size_t>(mc) * jb + j];
                        if (kb < 0) continue;
                        bool ok = true;
                        for (int a = 0; a < ndim; ++a)
                            if (std::fabs(T.centers[ndim * kb + a] - T.centers[ndim * b + a]) > 1.05 * T.boxsize[l]) { ok = false; break; }
                        if (ok) T.coll[static_cast<size_t>(mnb) * b + T.ncoll[b]++] = kb;
                    }
                }
            }
    } else if (npw > 0) {
        // midpoint rule (get_pwnodes itype=0, fgtterms.f:151-179)
        const int h2 = npw / 2;
        const double h = pmax / h2, w = h / (2.0 * std::sqrt(kPi));
        for (int j = -h2; j <= h2 - 1; ++j) {
            const double t = (j + 0.5) * h;
            P.ts[j + h2] = t;
            P.ws[j + h2] = w * std::exp(-t * t / 4);
        }
    }
    P.npwlevel = npwlevel;
    P.npw = npw;
    P.pmax = pmax;
    P.nh = (npw + 1) / 2;
    P.nexp = npw > 0 ? P.nh * ipow(npw, ndim - 1) : 0;
    P.nlevstart = std::max(npwlevel, 0);
    P.do_pw = npwlevel < nl;
    P.xnodes.resize(n);
    {
        std::vector<double> w(n), u(n * n);
        poly_exps(in.ipoly, n, P.xnodes.data(), w.data(), u.data());
    }

    // ---- which boxes carry PW expansions (gnd_find_pwexp_boxes,
    //      tree_routs_nd.f:1266-1366) ----
    P.ifpwexp.assign(nba, 0);
    if (npwlevel <= 0) {
        P.ifpwexp[0] = 1;
        // Every box forms and evaluates its PW expansion on this path, as
        // described at boxfgt.f:485-497 (the reference flags only the root,
        // so its leaf potentials stay zero; see DESIGN notes on this choice).
        if (nl >= 1) {
            P.all_pw = true;
            std::fill(P.ifpwexp.begin(), P.ifpwexp.end(), 1);
        }
    } else if (npwlevel <= nl) {
        for (int b = T.lfirst[npwlevel]; b <= T.llast[npwlevel]; ++b) {
            if (T.nchild[b] == 0) {
                for (int j = 0; j < T.ncoll[b]; ++j)
                    if (T.nchild[T.coll[static_cast<size_t>(mnb) * b + j]] > 0) { P.ifpwexp[b] = 1; break; }
            } else {
                P.ifpwexp[b] = 1;
            }
        }
        for (int l = npwlevel + 1; l <= nl; ++l)
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) P.ifpwexp[b] = 1;
    }

    // ---- leaves ----
    for (int b = 0; b < nba; ++b) if (T.nchild[b] == 0) P.nleaves++;

    // ---- near-field tables per level ----
    const size_t nn = static_cast<size_t>(n) * n;
    P.loctab_ks = nn * kLocSlots * (nl + 1);
    P.loctab.assign(3 * P.loctab_ks, 0.0);
    P.locband.assign(static_cast<size_t>(2) * n * kLocSlots * (nl + 1), 0);
    P.loccols.assign(static_cast<size_t>(2) * kLocSlots * (nl + 1), 0);
    const int nlevend = std::min(npwlevel, nl);
    // levels are independent (each writes its own slices)
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l <= nlevend; ++l)
        make_loctab_level(in.eps, in.ipoly, n, in.delta, T.boxsize[l], &P.loctab[nn * kLocSlots * l],
                          &P.locband[static_cast<size_t>(2) * n * kLocSlots * l], &P.loccols[2 * kLocSlots * l],
                          &P.loctab[P.loctab_ks + nn * kLocSlots * l],
                          &P.loctab[2 * P.loctab_ks + nn * kLocSlots * l]);

    // ---- direct interactions: modified list1 (gnd_compute_modified_list1,
    //      tree_routs_nd.f:864-1055) flattened per target, with table slots ----
    {
        const double bs0 = P.bs0;
        auto wrapdis = [&](double d) {
            double dis = std::fabs(d);
            if (P.iper_eff == 1) dis = std::min(dis, bs0 - dis);
            return dis;
        };
        auto add_pair = [&](int tgt, int src, std::vector<int>& srcs, std::vector<int>& codes) {
            // gnd_find_loctab_ind (bfgt_local.f:1624-1660), mrefinelev = 0
            const int jl = T.level[src];
            const double bsq = T.boxsize[jl] / 4;
            int code = 0, sh = 0;
            bool empty = false;
            for (int a = 0; a < ndim; ++a) {
                double dx = (T.centers[ndim * tgt + a] - T.centers[ndim * src + a]) / bsq;
                if (P.iper_eff == 1) {
                    const double dp = dx - bs0 / bsq, dm = dx + bs0 / bsq;
                    if (std::fabs(dx) > std::fabs(dp)) dx = dp;
                    if (std::fabs(dx) > std::fabs(dm)) dx = dm;
                }
                const int ix = static_cast<int>(dx);
                if (ix < -6 || ix > 6) empty = true;
                const int slot = std::min(std::max(ix + 6, 0), kLocSlots - 1);
                const int tl = T.level[tgt];
                const int* cr = &P.loccols[2 * (kLocSlots * tl + slot)];
                if (cr[0] > cr[1]) empty = true;
                code |= slot << sh;
                sh += 8;
            }
            if (empty) return;  // the reference returns early for an empty band
            srcs.push_back(src);
            codes.push_back(code);
        };
        // boxes are independent: each fills its own lists (in parallel), and
        // the lists are appended in level order, box order within a level
        std::vector<int> order_box, order_lev;
        for (int l = 0; l <= nlevend; ++l)
            for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
                order_box.push_back(b);
                order_lev.push_back(l);
            }
        const int nq = static_cast<int>(order_box.size());
        std::vector<std::vector<int>> qsrc(nq), qcode(nq);
#pragma omp parallel for schedule(dynamic, 64)
        for (int q = 0; q < nq; ++q) {
            const int b = order_box[q], l = order_lev[q];
            std::vector<int>& srcs = qsrc[q];
            std::vector<int>& codes = qcode[q];
            std::vector<int> raw;
            modified_list1(P, b, raw);
            for (int jb : raw) add_pair(b, jb, srcs, codes);
            (void)l;
        }
        for (int q = 0; q < nq; ++q) {
            if (qsrc[q].empty()) continue;
            P.dir_tgt.push_back(order_box[q]);
            P.dir_lev.push_back(order_lev[q]);
            P.dir_off.push_back(static_cast<int>(P.dir_src.size()));
            P.dir_src.insert(P.dir_src.end(), qsrc[q].begin(), qsrc[q].end());
            P.dir_code.insert(P.dir_code.end(), qcode[q].begin(), qcode[q].end());
        }
        P.dir_off.push_back(static_cast<int>(P.dir_src.size()));
        P.ndirect_pairs = static_cast<int64_t>(P.dir_src.size());
    }

    if (!P.do_pw) return 0;

    // ---- PW tables (boxfgt.f:629-664) ----
    const int nls = P.nlevstart;
    P.p2pw.assign(static_cast<size_t>(n) * npw * (nl + 1), cplx(0));
    P.pw2pot_ks = static_cast<size_t>(n) * npw * (nl + 1);
    P.pw2pot.assign(3 * P.pw2pot_ks, cplx(0));
    for (int l = nls; l <= nl; ++l) {
        make_p2pw(in.ipoly, n, npw, P.ws.data(), P.ts.data(), P.delta0, T.boxsize[l],
                  &P.p2pw[static_cast<size_t>(n) * npw * l]);
        const size_t o = static_cast<size_t>(n) * npw * l;
        make_pw2pgh(n, npw, P.ts.data(), P.xnodes.data(), P.delta0, T.boxsize[l], &P.pw2pot[o],
                    &P.pw2pot[P.pw2pot_ks + o], &P.pw2pot[2 * P.pw2pot_ks + o]);
    }
    // merge/split phases by repeated squaring (g?d_mk_merge_split_pw_matrices,
    // bfgt_pwrouts.f:1344-1543); kept as 1-D factors, the kernels form the
    // separable products
    P.nmax = nl - nls;
    P.wms.assign(static_cast<size_t>(std::max(P.nmax, 1)) * npw, cplx(1.0));
    {
        const double xmin2 = T.boxsize[nl] / std::sqrt(P.delta0) / 2;
        for (int j = 0; j < npw; ++j) {
            cplx z = std::exp(cplx(0.0, P.ts[j] * xmin2));
            for (int k = 0; k < P.nmax; ++k) {
                P.wms[static_cast<size_t>(k) * npw + j] = z;
                z = z * z;
            }
        }
    }
    // M2L phases for offsets -1, 0, +1 at the cutoff level
    // (g?d_mk_pw_translation_matrices, bfgt_pwrouts.f:1076-1293)
    P.wm2l.assign(static_cast<size_t>(3) * npw, cplx(1.0));
    {
        const double xmin = T.boxsize[nls] / std::sqrt(P.delta0);
        for (int j = 0; j < npw; ++j) {
            const cplx z = std::exp(cplx(0.0, P.ts[j] * xmin));
            P.wm2l[j] = std::conj(z);
            P.wm2l[npw + j] = 1.0;
            P.wm2l[2 * npw + j] = z;
        }
    }

    // ---- expansion slots: every box at levels >= nlevstart that forms or
    //      receives an expansion (non-leaf, or PW leaf) ----
    P.slot_of_box.assign(nba, -1);
    for (int l = nls; l <= nl; ++l)
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b)
            if (T.nchild[b] > 0 || P.ifpwexp[b]) P.slot_of_box[b] = P.nslots++;

    // S1/S5 leaves, fine to coarse as STEP 1 walks them
    for (int l = nl; l >= nls; --l)
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b)
            if (P.ifpwexp[b] && T.nchild[b] == 0) {
                P.leaf_box.push_back(b);
                P.leaf_lev.push_back(l);
                P.leaf_slot.push_back(P.slot_of_box[b]);
            }
    P.npw_leaves = static_cast<int64_t>(P.leaf_box.size());

    // S2/S4 parent-child entries per parent level; sign bit a set when the
    // parent's center is below the child's along axis a (gnd_find_msshift_ind,
    // bfgt_pwrouts.f:1687-1717; the split uses the opposite sense).
    P.pc_level_off.assign(nl + 2, 0);
    P.par_level_off.assign(nl + 2, 0);
    for (int l = 0; l <= nl; ++l) {
        P.pc_level_off[l] = static_cast<int>(P.pc_parent.size());
        P.par_level_off[l] = static_cast<int>(P.par_slot.size());
        if (l < nls || l >= nl) continue;
        for (int b = T.lfirst[l]; b <= T.llast[l]; ++b) {
            if (T.nchild[b] == 0) continue;
            P.par_slot.push_back(P.slot_of_box[b]);
            P.par_first.push_back(static_cast<int>(P.pc_parent.size()));
            P.par_count.push_back(T.nchild[b]);
            for (int i = 0; i < T.nchild[b]; ++i) {
                const int c = T.child[static_cast<size_t>(mc) * b + i];
                int k = 0;
                for (int a = 0; a < ndim; ++a)
                    if (T.centers[ndim * b + a] - T.centers[ndim * c + a] < 0) k |= 1 << a;
                P.pc_parent.push_back(P.slot_of_box[b]);
                P.pc_child.push_back(P.slot_of_box[c]);
                P.pc_sign.push_back(k);
            }
        }
    }
    P.pc_level_off[nl + 1] = static_cast<int>(P.pc_parent.size());
    P.par_level_off[nl + 1] = static_cast<int>(P.par_slot.size());

    // S3 at the cutoff level (gnd_compute_listpw, tree_routs_nd.f:1562-1628 and
    // gnd_find_pwshift_ind, bfgt_pwrouts.f:1592-1636)
    for (int b = T.lfirst[nls]; b <= T.llast[nls]; ++b) {
        if (P.slot_of_box[b] < 0) continue;
        P.m2l_tgt.push_back(P.slot_of_box[b]);
        P.m2l_off.push_back(static_cast<int>(P.m2l_src.size()));
        for (int i = 0; i < T.ncoll[b]; ++i) {
            const int jb = T.coll[static_cast<size_t>(mnb) * b + i];
            if (jb == b || (T.nchild[jb] == 0 && T.nchild[b] == 0)) continue;
            int idx = 0, inc = 1;
            const double xmin = T.boxsize[nls];
            for (int a = 0; a < ndim; ++a) {
                const double d = T.centers[ndim * b + a] - T.centers[ndim * jb + a];
                int jx = static_cast<int>(std::lround(d / xmin));
                if (P.iper_eff == 1) {
                    const int jp = static_cast<int>(std::lround((d - P.bs0) / xmin));
                    const int jm = static_cast<int>(std::lround((d + P.bs0) / xmin));
                    if (std::abs(jx) > std::abs(jp)) jx = jp;
                    if (std::abs(jx) > std::abs(jm)) jx = jm;
                }
                if (jx < -1 || jx > 1) { set_err(err, "M2L offset outside the 3^d stencil"); return 4; }
                idx += (jx + 1) * inc;
                inc *= 3;
            }
            if (P.slot_of_box[jb] < 0) { set_err(err, "M2L source without expansion"); return 4; }
            P.m2l_src.push_back(P.slot_of_box[jb]);
            P.m2l_idx.push_back(idx);
        }
    }
    P.m2l_off.push_back(static_cast<int>(P.m2l_src.size()));
    return 0;
}

}  // namespace fgt
