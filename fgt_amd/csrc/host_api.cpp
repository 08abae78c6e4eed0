// host_api.cpp -- exported host-only planner pieces (include/fgt_host.h).
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fgt_host.h"
#include "plan.h"

extern "C" {

void fgt_host_pwterms(double bsize, double delta, double eps, int iperiod, double* pmax, int* npw) {
    fgt::pw_terms(bsize, delta, eps, iperiod, pmax, npw);
}

void fgt_host_polyexps(int ipoly, int n, double* x, double* w, double* u) { fgt::poly_exps(ipoly, n, x, w, u); }

void fgt_host_sph_bessel(int nterms, double z, double* fjs) { fgt::sph_bessel(nterms, z, fjs); }

void fgt_host_poly2pw(int ipoly, int n, int npw, const double* ws, const double* ts, double delta,
                      double boxdim, double* tab) {
    std::vector<fgt::cplx> t(static_cast<size_t>(n) * npw);
    fgt::make_p2pw(ipoly, n, npw, ws, ts, delta, boxdim, t.data());
    // out[k*n + m] is Fortran tab(m,k): same memory order
    std::memcpy(tab, t.data(), sizeof(double) * 2 * t.size());
}

void fgt_host_pw2pot(int ipoly, int n, int npw, const double* ts, double delta, double boxdim, double* tab) {
    std::vector<double> x(n), w(n), u(static_cast<size_t>(n) * n);
    fgt::poly_exps(ipoly, n, x.data(), w.data(), u.data());
    std::vector<fgt::cplx> t(static_cast<size_t>(n) * npw);
    fgt::make_pw2pot(n, npw, ts, x.data(), delta, boxdim, t.data());
    // out[i*npw + k] is Fortran tab(k,i): same memory order
    std::memcpy(tab, t.data(), sizeof(double) * 2 * t.size());
}

void fgt_host_loctab_mref(double eps, int ipoly, int n, double delta, double boxdim, int mref, double* tab,
                          int* ind) {
    const size_t nn = static_cast<size_t>(n) * n;
    const int ns = 2 * fgt::loc_nloctab(mref) + 1;
    std::vector<double> t(nn * ns);
    std::vector<int> band(static_cast<size_t>(2) * n * ns), cols(2 * ns);
    fgt::make_loctab_level(eps, ipoly, n, delta, boxdim, t.data(), band.data(), cols.data(), nullptr, nullptr,
                           mref);
    for (int s = 0; s < ns; ++s) {
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) tab[s * nn + j + static_cast<size_t>(k) * n] = t[(s * n + k) * n + j];
        int* o = ind + static_cast<size_t>(s) * 2 * (n + 1);
        for (int k = 0; k < n; ++k) {
            const int lo = band[2 * (s * n + k)], hi = band[2 * (s * n + k) + 1];
            if (lo > hi) { o[2 * k] = 0; o[2 * k + 1] = -1; }
            else { o[2 * k] = lo + 1; o[2 * k + 1] = hi + 1; }
        }
        o[2 * n] = cols[2 * s] + 1;
        o[2 * n + 1] = cols[2 * s + 1] + 1;
    }
}

void fgt_host_loctab(double eps, int ipoly, int n, double delta, double boxdim, double* tab, int* ind) {
    fgt_host_loctab_mref(eps, ipoly, n, delta, boxdim, 0, tab, ind);
}

void fgt_host_pw2pgh(int ipoly, int n, int npw, const double* ts, double delta, double boxdim, double* tab,
                     double* tabx, double* tabxx) {
    std::vector<double> x(n), w(n), u(static_cast<size_t>(n) * n);
    fgt::poly_exps(ipoly, n, x.data(), w.data(), u.data());
    const size_t m = static_cast<size_t>(n) * npw;
    std::vector<fgt::cplx> t(3 * m);
    fgt::make_pw2pgh(n, npw, ts, x.data(), delta, boxdim, t.data(), t.data() + m, t.data() + 2 * m);
    std::memcpy(tab, t.data(), sizeof(double) * 2 * m);
    std::memcpy(tabx, t.data() + m, sizeof(double) * 2 * m);
    std::memcpy(tabxx, t.data() + 2 * m, sizeof(double) * 2 * m);
}

void fgt_host_loctab_der(double eps, int ipoly, int n, double delta, double boxdim, double* tabx,
                         double* tabxx) {
    const size_t nn = static_cast<size_t>(n) * n;
    std::vector<double> t(nn * fgt::kLocSlots), tx(nn * fgt::kLocSlots), txx(nn * fgt::kLocSlots);
    std::vector<int> band(static_cast<size_t>(2) * n * fgt::kLocSlots), cols(2 * fgt::kLocSlots);
    fgt::make_loctab_level(eps, ipoly, n, delta, boxdim, t.data(), band.data(), cols.data(), tx.data(),
                           txx.data());
    for (int s = 0; s < fgt::kLocSlots; ++s)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j) {
                tabx[s * nn + j + static_cast<size_t>(k) * n] = tx[(s * n + k) * n + j];
                tabxx[s * nn + j + static_cast<size_t>(k) * n] = txx[(s * n + k) * n + j];
            }
}

static int plan_for(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int nboxes,
                    int nlevels, int ltree, const int* itree, const int* iptr, const double* centers,
                    const double* boxsize, fgt::Plan& P) {
    int npbox = 1;
    for (int i = 0; i < ndim && i < 3; ++i) npbox *= norder;
    fgt::PlanInput in{nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree,
                      itree, iptr, centers, boxsize};
    std::string err;
    return fgt::make_plan(in, P, &err);
}

int fgt_plan_summary(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder, int nboxes,
                     int nlevels, int ltree, const int* itree, const int* iptr, const double* centers,
                     const double* boxsize, fgt_plan_info* out) {
    if (!out) return 2;
    fgt::Plan P;
    const int rc = plan_for(nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels, ltree, itree, iptr,
                            centers, boxsize, P);
    if (rc) return rc;
    std::memset(out, 0, sizeof *out);
    out->npwlevel = P.npwlevel;
    out->npw = P.npw;
    out->nexp = P.nexp;
    out->nlevstart = P.nlevstart;
    out->nslots = P.nslots;
    out->periodic_small = P.periodic_small;
    out->all_pw = P.all_pw;
    out->nleaves = P.nleaves;
    out->npw_leaves = P.npw_leaves;
    out->ndirect_pairs = P.ndirect_pairs;
    out->n_m2l_pairs = static_cast<long long>(P.m2l_src.size());
    out->n_parent_child = static_cast<long long>(P.pc_child.size());
    out->n_m2l_targets = static_cast<long long>(P.m2l_tgt.size());
    return 0;
}

long long fgt_plan_direct_pairs(int nd, int ndim, double delta, double eps, int ipoly, int iperiod, int norder,
                                int nboxes, int nlevels, int ltree, const int* itree, const int* iptr,
                                const double* centers, const double* boxsize, long long cap, int* tgt, int* src,
                                int* ixyz) {
    fgt::Plan P;
    const int rc = plan_for(nd, ndim, delta, eps, ipoly, iperiod, norder, nboxes, nlevels, ltree, itree, iptr,
                            centers, boxsize, P);
    if (rc) return -rc;
    const long long total = static_cast<long long>(P.dir_src.size());
    long long p = 0;
    for (size_t t = 0; t + 1 < P.dir_off.size(); ++t)
        for (int j = P.dir_off[t]; j < P.dir_off[t + 1]; ++j, ++p) {
            if (p >= cap) continue;
            tgt[p] = P.dir_tgt[t] + 1;
            src[p] = P.dir_src[j] + 1;
            for (int a = 0; a < 3; ++a) ixyz[3 * p + a] = a < ndim ? ((P.dir_code[j] >> (8 * a)) & 255) - 6 : 0;
        }
    return total;
}

}  // extern "C"
