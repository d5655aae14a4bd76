/*
 * tp_oracle.cpp -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see tp_oracle.h).
 *
 * Restates the reference's classed_version CPU path ("CV" below =
 * classed_version/) function by function, in FP64, with the same loop
 * structure and convergence tests so that adaptive decisions agree:
 *   - cosmology tables:     CV/bispectrum.cpp:91-164, 676-912, 951-987
 *   - BiHalofit bispectrum:  CV/bispectrum.cpp:166-236, 609-633, 785-841
 *   - Limber B_kappa:        CV/bispectrum.cpp:238-286, 471-541, 1010-1020
 *   - <Map^3> integrand:     CV/apertureStatistics.cpp:10-28, 254-279, 321-353
 *   - adaptive cubature:     CV/hcubature.cpp:115-1068, CV/converged.h
 *   - Gamma_i (Ogata):       CV/gamma.cpp:6-81, 121-219, 291-377, 462-563, 624-643
 *   - P_kappa, Halofit:      CV/bispectrum.cpp:288-417, 543-577, 1022-1059
 *   - <Map^2>, T17 factor:   GV/apertureStatistics.cu:29-136, GV/bispectrum.cu:951-982
 *   - Gaussian covariance:   CV/apertureStatistics.cpp:50-159, 355-396, CV/bispectrum.cpp:580-607;
 *                            T1 (infinite survey) GV/apertureStatisticsCovariance.cu:1065-1106,
 *                            1958-1985; T1 and T2 (circle, square, rectangle) :117-177, 328-538,
 *                            1065-1163, 1886-1955, 1984-2110 (J1 of the circle window: the C
 *                            library's j1)
 *                            ("GV" = cuda_version/)
 *   - file readers:          CV/helper.cpp:31-96, CV/cosmology.cpp:10-72
 * GSL routines the reference calls are replaced by:
 *   - orc_qag15: QUADPACK QAG with the 15-point Gauss-Kronrod rule (Piessens et
 *     al. 1983), the algorithm behind gsl_integration_qag(..., key=1, ...);
 *   - orc_bessel_jn: power series / Miller backward recurrence / Hankel
 *     asymptotic expansion (Abramowitz & Stegun 9.1.10, 9.12, 9.2.5-9.2.10);
 *   - orc_bessel_zero_jn: McMahon expansion (A&S 9.5.12) + Newton refinement.
 */
#include "tp_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes (own Newton iteration on P_n; used in place of the   */
/* 15-digit GL-96 literals of CV/bispectrum.hpp:157-183).                     */
/* ------------------------------------------------------------------------ */
void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[n - 1 - i] = w[i];
    }
}

/* 48 positive half-nodes of the 96-point rule, ascending (the CV A96/W96 layout) */
struct GL96 {
    double A[48], W[48];
    GL96()
    {
        double x[96], w[96];
        gauss_legendre(96, x, w);
        for (int i = 0; i < 48; ++i) {
            A[i] = x[48 + i];
            W[i] = w[48 + i];
        }
    }
};
const GL96 &gl96()
{
    static GL96 g;
    return g;
}

/* ------------------------------------------------------------------------ */
/* QUADPACK QAG / GK15 (replaces gsl_integration_qag, CV/bispectrum.cpp:961) */
/* ------------------------------------------------------------------------ */
/* 15-point Kronrod abscissae / weights and 7-point Gauss weights
   (published QUADPACK constants, Piessens et al. 1983). */
const double XGK15[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                         0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                         0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
                         0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
const double WG7[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                       0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
const double WGK15[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                         0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                         0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                         0.204432940075298892414161999234649, 0.209482141084727828012999174891714};

double rescale_error(double err, double result_abs, double result_asc)
{
    err = std::fabs(err);
    if (result_asc != 0 && err != 0) {
        double scale = std::pow(200 * err / result_asc, 1.5);
        err = (scale < 1) ? result_asc * scale : result_asc;
    }
    if (result_abs > DBL_MIN / (50 * DBL_EPSILON)) {
        double min_err = 50 * DBL_EPSILON * result_abs;
        if (min_err > err) err = min_err;
    }
    return err;
}

template <class F>
void qk15(const F &f, double a, double b, double *result, double *abserr, double *resabs,
          double *resasc)
{
    const double center = 0.5 * (a + b), half = 0.5 * (b - a), abs_half = std::fabs(half);
    const double f_center = f(center);
    double result_gauss = f_center * WG7[3];
    double result_kronrod = f_center * WGK15[7];
    double result_abs = std::fabs(result_kronrod);
    double fv1[8], fv2[8];
    for (int j = 0; j < 3; ++j) {
        int jtw = 2 * j + 1;
        double ab = half * XGK15[jtw];
        double f1 = f(center - ab), f2 = f(center + ab);
        fv1[jtw] = f1;
        fv2[jtw] = f2;
        result_gauss += WG7[j] * (f1 + f2);
        result_kronrod += WGK15[jtw] * (f1 + f2);
        result_abs += WGK15[jtw] * (std::fabs(f1) + std::fabs(f2));
    }
    for (int j = 0; j < 4; ++j) {
        int jtwm1 = 2 * j;
        double ab = half * XGK15[jtwm1];
        double f1 = f(center - ab), f2 = f(center + ab);
        fv1[jtwm1] = f1;
        fv2[jtwm1] = f2;
        result_kronrod += WGK15[jtwm1] * (f1 + f2);
        result_abs += WGK15[jtwm1] * (std::fabs(f1) + std::fabs(f2));
    }
    double mean = result_kronrod * 0.5;
    double result_asc = WGK15[7] * std::fabs(f_center - mean);
    for (int j = 0; j < 7; ++j)
        result_asc += WGK15[j] * (std::fabs(fv1[j] - mean) + std::fabs(fv2[j] - mean));
    double err = (result_kronrod - result_gauss) * half;
    result_kronrod *= half;
    result_abs *= abs_half;
    result_asc *= abs_half;
    *result = result_kronrod;
    *resabs = result_abs;
    *resasc = result_asc;
    *abserr = rescale_error(err, result_abs, result_asc);
}

template <class F>
double qag15(const F &f, double a, double b, double epsabs, double epsrel, int limit)
{
    double result0, abserr0, resabs0, resasc0;
    qk15(f, a, b, &result0, &abserr0, &resabs0, &resasc0);
    double tolerance = std::max(epsabs, epsrel * std::fabs(result0));
    double round_off = 50 * DBL_EPSILON * resabs0;
    if (abserr0 <= round_off && abserr0 > tolerance) return result0;
    if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) return result0;
    if (limit == 1) return result0;
    struct Iv { double a, b, r, e; };
    std::vector<Iv> iv;
    iv.push_back({a, b, result0, abserr0});
    double area = result0, errsum = abserr0;
    int roundoff1 = 0, roundoff2 = 0, iteration = 1, error_type = 0;
    do {
        size_t imax = 0;
        for (size_t i = 1; i < iv.size(); ++i)
            if (iv[i].e > iv[imax].e) imax = i;
        Iv cur = iv[imax];
        double a1 = cur.a, b1 = 0.5 * (cur.a + cur.b), a2 = b1, b2 = cur.b;
        double area1, error1, resabs1, resasc1, area2, error2, resabs2, resasc2;
        qk15(f, a1, b1, &area1, &error1, &resabs1, &resasc1);
        qk15(f, a2, b2, &area2, &error2, &resabs2, &resasc2);
        double area12 = area1 + area2, error12 = error1 + error2;
        errsum += (error12 - cur.e);
        area += area12 - cur.r;
        if (resasc1 != error1 && resasc2 != error2) {
            double delta = cur.r - area12;
            if (std::fabs(delta) <= 1.0e-5 * std::fabs(area12) && error12 >= 0.99 * cur.e) roundoff1++;
            if (iteration >= 10 && error12 > cur.e) roundoff2++;
        }
        tolerance = std::max(epsabs, epsrel * std::fabs(area));
        if (errsum > tolerance) {
            if (roundoff1 >= 6 || roundoff2 >= 20) error_type = 2;
            double tmp = (1 + 100 * DBL_EPSILON) * (std::fabs(a2) + 1000 * DBL_MIN);
            if (std::fabs(a1) <= tmp && std::fabs(b2) <= tmp) error_type = 3;
        }
        iv[imax] = {a1, b1, area1, error1};
        iv.push_back({a2, b2, area2, error2});
        iteration++;
    } while (iteration < limit && !error_type && errsum > tolerance);
    double s = 0;
    for (auto &x : iv) s += x.r;
    return s;
}

/* ------------------------------------------------------------------------ */
/* Bessel J_n (replaces gsl_sf_bessel_Jn, CV/gamma.cpp:158-171)              */
/* ------------------------------------------------------------------------ */
double bessel_jn(int n, double x)
{
    if (n < 0) return ((n & 1) ? -1.0 : 1.0) * bessel_jn(-n, x);
    if (x < 0) return ((n & 1) ? -1.0 : 1.0) * bessel_jn(n, -x);
    if (x == 0) return n == 0 ? 1.0 : 0.0;
    if (x < 2.0) { /* power series, A&S 9.1.10 */
        double t = std::pow(0.5 * x, n);
        for (int k = 1; k <= n; ++k) t /= k;
        double s = t, y = -0.25 * x * x;
        for (int m = 1; m < 60; ++m) {
            t *= y / (m * (double)(m + n));
            s += t;
            if (std::fabs(t) < 1e-18 * std::fabs(s)) break;
        }
        return s;
    }
    if (x < 30.0) { /* Miller backward recurrence normalised by J0 + 2 sum J_2k = 1 */
        int M = 2 * ((std::max(n, (int)x) + 40) / 2);
        double jp1 = 0.0, j = 1e-300, sum = 0.0, res = 0.0;
        for (int k = M; k > 0; --k) {
            double jm1 = (2.0 * k / x) * j - jp1;
            jp1 = j;
            j = jm1;
            if (k - 1 == n) res = j;
            if (((k - 1) & 1) == 0 && k - 1 > 0) sum += 2.0 * j;
            if (std::fabs(j) > 1e250) {
                j *= 1e-250;
                jp1 *= 1e-250;
                sum *= 1e-250;
                res *= 1e-250;
            }
        }
        sum += j; /* J_0 */
        if (n == 0) res = j;
        return res / sum;
    }
    /* Hankel asymptotic expansion, A&S 9.2.5-9.2.10 */
    double mu = 4.0 * n * n, ex = 8.0 * x;
    double P = 1.0, Q = 0.0, term = 1.0;
    for (int k = 1; k < 60; ++k) {
        double odd = 2.0 * k - 1.0;
        term *= (mu - odd * odd) / (k * ex);
        if (k > 2 && std::fabs(term) < 1e-17) break;
        if (k & 1)
            Q += ((k / 2) & 1 ? -1.0 : 1.0) * term;
        else
            P += ((k / 2) & 1 ? -1.0 : 1.0) * term;
    }
    double c = (0.5 * n + 0.25) * M_PI;
    double cx = std::cos(x), sx = std::sin(x), cc = std::cos(c), sc = std::sin(c);
    double cchi = cx * cc + sx * sc, schi = sx * cc - cx * sc;
    return std::sqrt(2.0 / (M_PI * x)) * (P * cchi - Q * schi);
}

double bessel_zero_jn(int n, int s)
{
    if (s == 0) return 0.0; /* the "zeroth zero" convention of gsl_sf_bessel_zero_Jnu */
    double mu = 4.0 * n * n;
    double beta = (s + 0.5 * n - 0.25) * M_PI;
    double b8 = 8.0 * beta;
    double x = beta - (mu - 1) / b8 - 4 * (mu - 1) * (7 * mu - 31) / (3 * b8 * b8 * b8) -
               32 * (mu - 1) * (83 * mu * mu - 982 * mu + 3779) / (15 * std::pow(b8, 5));
    for (int it = 0; it < 30; ++it) {
        double J = bessel_jn(n, x);
        double dJ = bessel_jn(n - 1, x) - n / x * J;
        double dx = J / dJ;
        x -= dx;
        if (std::fabs(dx) < 1e-15 * x) break;
    }
    return x;
}

/* ------------------------------------------------------------------------ */
/* Adaptive cubature (restates CV/hcubature.cpp:115-1068 + CV/converged.h)  */
/* ------------------------------------------------------------------------ */
struct Region {
    std::vector<double> c, hw; /* center, half-widths */
    double vol;
    unsigned split;
    std::vector<double> val, err;
    double errmax;
};

double compute_vol(const std::vector<double> &hw)
{ /* CV/hcubature.cpp:134-141 */
    double vol = 1;
    for (double h : hw) vol *= 2 * h;
    return vol;
}

bool converged_check(unsigned fdim, const double *val, const double *err, double reqAbs,
                     double reqRel, int norm)
{ /* CV/converged.h:6-73 */
    switch (norm) {
    case ORC_ERROR_INDIVIDUAL:
        for (unsigned j = 0; j < fdim; ++j)
            if (err[j] > reqAbs && err[j] > std::fabs(val[j]) * reqRel) return false;
        return true;
    case ORC_ERROR_PAIRED: {
        unsigned j;
        for (j = 0; j + 1 < fdim; j += 2) {
            double maxerr = err[j] > err[j + 1] ? err[j] : err[j + 1];
            double maxval = val[j] > val[j + 1] ? val[j] : val[j + 1];
            double serr = maxerr > 0 ? 1 / maxerr : 1, sval = maxval > 0 ? 1 / maxval : 1;
            double e = std::sqrt((err[j] * serr) * (err[j] * serr) + (err[j + 1] * serr) * (err[j + 1] * serr)) * maxerr;
            double v = std::sqrt((val[j] * sval) * (val[j] * sval) + (val[j + 1] * sval) * (val[j + 1] * sval)) * maxval;
            if (e > reqAbs && e > v * reqRel) return false;
        }
        if (j < fdim)
            if (err[j] > reqAbs && err[j] > std::fabs(val[j]) * reqRel) return false;
        return true;
    }
    case ORC_ERROR_L1: {
        double e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            e += err[j];
            v += std::fabs(val[j]);
        }
        return e <= reqAbs || e <= v * reqRel;
    }
    case ORC_ERROR_LINF: {
        double e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            double av = std::fabs(val[j]);
            if (err[j] > e) e = err[j];
            if (av > v) v = av;
        }
        return e <= reqAbs || e <= v * reqRel;
    }
    case ORC_ERROR_L2: {
        double maxerr = 0, maxval = 0, e = 0, v = 0;
        for (unsigned j = 0; j < fdim; ++j) {
            double av = std::fabs(val[j]);
            if (err[j] > maxerr) maxerr = err[j];
            if (av > maxval) maxval = av;
        }
        double serr = maxerr > 0 ? 1 / maxerr : 1, sval = maxval > 0 ? 1 / maxval : 1;
        for (unsigned j = 0; j < fdim; ++j) {
            e += (err[j] * serr) * (err[j] * serr);
            v += (std::fabs(val[j]) * sval) * (std::fabs(val[j]) * sval);
        }
        e = std::sqrt(e) * maxerr;
        v = std::sqrt(v) * maxval;
        return e <= reqAbs || e <= v * reqRel;
    }
    }
    return true;
}

/* Genz-Malik degree-7/5 rule over nR regions (CV/hcubature.cpp:456-614).
   Point order per region: center; +-lambda2 / +-lambda4 along each axis;
   (lambda4,lambda4) pairs; 2^dim lambda5 corners in Gray-code order. */
struct GMRule {
    unsigned dim, fdim, npts;
    double w1, w3, w5, wE1, wE3;
    std::vector<double> pts, vals;
    GMRule(unsigned d, unsigned f) : dim(d), fdim(f)
    {
        npts = 1 + 4 * d + 2 * d * (d - 1) + (1u << d);
        int id = (int)d;
        w1 = (double)(12824 - 9120 * id + 400 * id * id) / 19683.0;
        w3 = (double)(1820 - 400 * id) / 19683.0;
        w5 = 6859.0 / 19683.0 / (double)(1u << d);
        wE1 = (double)(729 - 950 * id + 50 * id * id) / 729.0;
        wE3 = (double)(265 - 100 * id) / 1458.0;
    }
    int eval(orc_integrand_v f, void *fdata, std::vector<Region> &R, size_t nR)
    {
        const double l2 = 0.3585685828003180919906451539079374954541; /* sqrt(9/70) */
        const double l4 = 0.9486832980505137995996680633298155601160; /* sqrt(9/10) */
        const double l5 = 0.6882472016116852977216287342936235251269; /* sqrt(9/19) */
        const double wt2 = 980. / 6561., wt4 = 200. / 19683.;
        const double wtE2 = 245. / 486., wtE4 = 25. / 729.;
        const double ratio = (l2 * l2) / (l4 * l4);
        pts.assign(nR * npts * dim, 0.0);
        vals.assign(nR * npts * fdim, 0.0);
        double *P = pts.data();
        std::vector<double> p(dim);
        auto emit = [&]() { for (unsigned i = 0; i < dim; ++i) *P++ = p[i]; };
        for (size_t iR = 0; iR < nR; ++iR) {
            const std::vector<double> &c = R[iR].c, &h = R[iR].hw;
            for (unsigned i = 0; i < dim; ++i) p[i] = c[i];
            emit();
            for (unsigned i = 0; i < dim; ++i) {
                double a2 = h[i] * l2, a4 = h[i] * l4;
                p[i] = c[i] - a2; emit();
                p[i] = c[i] + a2; emit();
                p[i] = c[i] - a4; emit();
                p[i] = c[i] + a4; emit();
                p[i] = c[i];
            }
            for (unsigned i = 0; i + 1 < dim; ++i) {
                double ai = h[i] * l4;
                p[i] = c[i] - ai;
                for (unsigned j = i + 1; j < dim; ++j) {
                    double aj = h[j] * l4;
                    p[j] = c[j] - aj; emit();
                    p[i] = c[i] + ai; emit();
                    p[j] = c[j] + aj; emit();
                    p[i] = c[i] - ai; emit();
                    p[j] = c[j];
                }
                p[i] = c[i];
            }
            unsigned signs = 0;
            for (unsigned i = 0; i < dim; ++i) p[i] = c[i] + h[i] * l5;
            for (unsigned i = 0;; ++i) {
                emit();
                unsigned d = (unsigned)__builtin_ctz(~i);
                if (d >= dim) break;
                signs ^= (1u << d);
                p[d] = (signs & (1u << d)) ? c[d] - h[d] * l5 : c[d] + h[d] * l5;
            }
        }
        if (f(dim, nR * npts, pts.data(), fdata, fdim, vals.data())) return 1;
        std::vector<double> diff(nR * dim, 0.0);
        for (unsigned j = 0; j < fdim; ++j) {
            for (size_t iR = 0; iR < nR; ++iR) {
                const double *v = vals.data() + j + iR * npts * fdim;
#define VALS(i) v[fdim * (i)]
                double val0 = VALS(0), sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0;
                unsigned k0 = 1, k;
                for (k = 0; k < dim; ++k) {
                    double v0 = VALS(k0 + 4 * k), v1 = VALS(k0 + 4 * k + 1);
                    double v2 = VALS(k0 + 4 * k + 2), v3 = VALS(k0 + 4 * k + 3);
                    sum2 += v0 + v1;
                    sum3 += v2 + v3;
                    diff[iR * dim + k] += std::fabs(v0 + v1 - 2 * val0 - ratio * (v2 + v3 - 2 * val0));
                }
                k0 += 4 * k;
                for (k = 0; k < 2 * dim * (dim - 1); ++k) sum4 += VALS(k0 + k);
                k0 += k;
                for (k = 0; k < (1u << dim); ++k) sum5 += VALS(k0 + k);
#undef VALS
                double result = R[iR].vol * (w1 * val0 + wt2 * sum2 + w3 * sum3 + wt4 * sum4 + w5 * sum5);
                double res5th = R[iR].vol * (wE1 * val0 + wtE2 * sum2 + wE3 * sum3 + wtE4 * sum4);
                R[iR].val[j] = result;
                R[iR].err[j] = std::fabs(res5th - result);
            }
        }
        for (size_t iR = 0; iR < nR; ++iR) {
            double maxdiff = 0;
            unsigned dd = 0;
            for (unsigned i = 0; i < dim; ++i)
                if (diff[iR * dim + i] > maxdiff) {
                    maxdiff = diff[iR * dim + i];
                    dd = i;
                }
            R[iR].split = dd;
            double em = 0;
            for (unsigned k = 0; k < fdim; ++k)
                if (R[iR].err[k] > em) em = R[iR].err[k];
            R[iR].errmax = em;
        }
        return 0;
    }
};

/* 15-point Gauss-Kronrod rule of one-dimensional regions (CV/hcubature.cpp:620-765, the rule the
   driver takes for dim = 1).  Point order per region: center, the three Gauss pairs (+-xgk[1, 3, 5]),
   the four Kronrod pairs (+-xgk[0, 2, 4, 6]); sums accumulate in that order.  The error estimate is
   QUADPACK's rescaled |K15 - G7|. */
struct GK15Rule {
    unsigned dim = 1, fdim, npts = 15;
    std::vector<double> pts, vals;
    explicit GK15Rule(unsigned f) : fdim(f) {}
    int eval(orc_integrand_v f, void *fdata, std::vector<Region> &R, size_t nR)
    {
        pts.assign(nR * npts, 0.0);
        vals.assign(nR * npts * fdim, 0.0);
        size_t n = 0;
        for (size_t iR = 0; iR < nR; ++iR) {
            const double c = R[iR].c[0], hw = R[iR].hw[0];
            pts[n++] = c;
            for (int j = 1; j < 7; j += 2) {
                pts[n++] = c - hw * XGK15[j];
                pts[n++] = c + hw * XGK15[j];
            }
            for (int j = 0; j < 8; j += 2) {
                pts[n++] = c - hw * XGK15[j];
                pts[n++] = c + hw * XGK15[j];
            }
            R[iR].split = 0;
        }
        if (f(1, nR * npts, pts.data(), fdata, fdim, vals.data())) return 1;
        for (size_t iR = 0; iR < nR; ++iR) {
            const double hw = R[iR].hw[0];
            for (unsigned k = 0; k < fdim; ++k) {
                const double *v = vals.data() + iR * npts * fdim + k;
                auto V = [&](unsigned i) { return v[fdim * i]; };
                double g = V(0) * WG7[3], kr = V(0) * WGK15[7], ra = std::fabs(kr);
                unsigned i = 1;
                for (int j = 0; j < 3; ++j, i += 2) {
                    double s = V(i) + V(i + 1);
                    g += WG7[j] * s;
                    kr += WGK15[2 * j + 1] * s;
                    ra += WGK15[2 * j + 1] * (std::fabs(V(i)) + std::fabs(V(i + 1)));
                }
                for (int j = 0; j < 4; ++j, i += 2) {
                    kr += WGK15[2 * j] * (V(i) + V(i + 1));
                    ra += WGK15[2 * j] * (std::fabs(V(i)) + std::fabs(V(i + 1)));
                }
                R[iR].val[k] = kr * hw;
                double mean = kr * 0.5, rasc = WGK15[7] * std::fabs(V(0) - mean);
                i = 1;
                for (int j = 0; j < 3; ++j, i += 2)
                    rasc += WGK15[2 * j + 1] * (std::fabs(V(i) - mean) + std::fabs(V(i + 1) - mean));
                for (int j = 0; j < 4; ++j, i += 2)
                    rasc += WGK15[2 * j] * (std::fabs(V(i) - mean) + std::fabs(V(i + 1) - mean));
                R[iR].err[k] = rescale_error((kr - g) * hw, ra * hw, rasc * hw);
            }
            double em = 0;
            for (unsigned k = 0; k < fdim; ++k)
                if (R[iR].err[k] > em) em = R[iR].err[k];
            R[iR].errmax = em;
        }
        return 0;
    }
};

/* binary max-heap keyed on errmax, with running sums (CV/hcubature.cpp:772-890) */
struct RegionHeap {
    std::vector<Region> items;
    std::vector<double> sval, serr;
    explicit RegionHeap(unsigned fdim) : sval(fdim, 0.0), serr(fdim, 0.0) {}
    void push(const Region &r)
    {
        for (size_t i = 0; i < sval.size(); ++i) {
            sval[i] += r.val[i];
            serr[i] += r.err[i];
        }
        size_t ins = items.size();
        items.push_back(r);
        while (ins) {
            size_t parent = (ins - 1) / 2;
            if (r.errmax <= items[parent].errmax) break;
            items[ins] = items[parent];
            ins = parent;
        }
        items[ins] = r;
    }
    Region pop()
    {
        Region ret = items[0];
        size_t n = items.size() - 1;
        items[0] = items[n];
        items.pop_back();
        size_t i = 0, child;
        while ((child = i * 2 + 1) < n) {
            size_t largest = (items[child].errmax <= items[i].errmax) ? i : child;
            if (++child < n && items[largest].errmax < items[child].errmax) largest = child;
            if (largest == i) break;
            std::swap(items[i], items[largest]);
            i = largest;
        }
        for (size_t k = 0; k < sval.size(); ++k) {
            sval[k] -= ret.val[k];
            serr[k] -= ret.err[k];
        }
        return ret;
    }
};

thread_local long long g_last_neval = 0, g_last_nregions = 0;

template <class Rule>
int hcubature_run(Rule &rule, unsigned fdim, orc_integrand_v f, void *fdata, unsigned dim, const double *xmin,
                  const double *xmax, size_t maxEval, double reqAbs, double reqRel, int norm, double *val,
                  double *err)
{ /* CV/hcubature.cpp:904-1068 (parallel / Gladwell variant used by hcubature_v) */
    RegionHeap heap(fdim);
    std::vector<Region> R(1);
    R[0].c.resize(dim);
    R[0].hw.resize(dim);
    for (unsigned i = 0; i < dim; ++i) {
        R[0].c[i] = 0.5 * (xmin[i] + xmax[i]);
        R[0].hw[i] = 0.5 * (xmax[i] - xmin[i]);
    }
    R[0].vol = compute_vol(R[0].hw);
    R[0].split = 0;
    R[0].val.assign(fdim, 0.0);
    R[0].err.assign(fdim, 0.0);
    if (rule.eval(f, fdata, R, 1)) return 1;
    heap.push(R[0]);
    size_t numEval = rule.npts;
    long long nreg = 1;
    std::vector<double> eval_(fdim), eerr(fdim);
    while (numEval < maxEval || !maxEval) {
        if (converged_check(fdim, heap.sval.data(), heap.serr.data(), reqAbs, reqRel, norm)) break;
        size_t nR = 0;
        eval_ = heap.sval;
        eerr = heap.serr;
        R.clear();
        do {
            R.push_back(heap.pop());
            for (unsigned j = 0; j < fdim; ++j) eerr[j] -= R[nR].err[j];
            /* cut_region (CV/hcubature.cpp:206-218): the kept half updates vol
               incrementally, the new half recomputes it from its widths. */
            Region &A = R[nR];
            unsigned d = A.split;
            A.hw[d] *= 0.5;
            A.vol *= 0.5;
            Region B = A;
            B.vol = compute_vol(B.hw);
            A.c[d] -= A.hw[d];
            B.c[d] += A.hw[d];
            R.push_back(B);
            numEval += rule.npts * 2;
            nR += 2;
            if (converged_check(fdim, eval_.data(), eerr.data(), reqAbs, reqRel, norm)) break;
        } while (!heap.items.empty() && (numEval < maxEval || !maxEval));
        if (rule.eval(f, fdata, R, nR)) return 1;
        for (size_t i = 0; i < nR; ++i) heap.push(R[i]);
        nreg += (long long)nR;
    }
    for (unsigned j = 0; j < fdim; ++j) val[j] = err[j] = 0;
    for (auto &it : heap.items)
        for (unsigned j = 0; j < fdim; ++j) {
            val[j] += it.val[j];
            err[j] += it.err[j];
        }
    g_last_neval = (long long)numEval;
    g_last_nregions = nreg;
    return 0;
}

int hcubature_v_impl(unsigned fdim, orc_integrand_v f, void *fdata, unsigned dim,
                     const double *xmin, const double *xmax, size_t maxEval, double reqAbs,
                     double reqRel, int norm, double *val, double *err)
{ /* CV/hcubature.cpp:1030-1047: Gauss-Kronrod for dim = 1, Genz-Malik otherwise */
    g_last_neval = 0;
    g_last_nregions = 0;
    if (fdim == 0) return 0;
    if (dim < 1) return 1;
    if (fdim <= 1) norm = ORC_ERROR_INDIVIDUAL;
    if (norm < 0 || norm > ORC_ERROR_LINF) return 1;
    if (dim == 1) {
        GK15Rule rule(fdim);
        return hcubature_run(rule, fdim, f, fdata, dim, xmin, xmax, maxEval, reqAbs, reqRel, norm, val, err);
    }
    GMRule rule(dim, fdim);
    return hcubature_run(rule, fdim, f, fdata, dim, xmin, xmax, maxEval, reqAbs, reqRel, norm, val, err);
}

/* ------------------------------------------------------------------------ */
/* Bispectrum calculator (restates CV/bispectrum.cpp)                         */
/* ------------------------------------------------------------------------ */
struct Cosmo {
    double h, sigma8, omb, omc, ns, w, om, ow, zmax;
};

struct Calc {
    Cosmo cosmo;
    double norm = 1.0;
    double eps = 1.e-4;           /* CV/bispectrum.hpp:123 */
    int n;                        /* n_redshift_bins */
    double z_max, dz;
    bool nz_from_file = true;
    bool analytic = false;        /* CV/bispectrum.hpp:18 test_analytical */
    std::vector<double> nz_file, fK, g, D1, rs, neff, ncur;
    const double H0_over_c = 100. / 299792.; /* CV/bispectrum.hpp:148 */
    const double c_over_H0 = 2997.92;        /* CV/bispectrum.hpp:149 */
    /* Gamma (Ogata) arrays, CV/gamma.cpp:133-176 */
    double prec_h = 0;
    unsigned prec_k = 0;
    std::vector<double> psi6, prod6, psi2, prod2;

    double E(double z) const { return std::sqrt(cosmo.om * std::pow(1 + z, 3) + cosmo.ow); } /* CV:967-970 */

    double f_K_at_z(double z) const
    { /* CV/bispectrum.cpp:951-965 */
        auto Einv = [this](double x) { return 1. / E(x); };
        return c_over_H0 * qag15(Einv, 0, z, 0, 1.0e-6, 10000);
    }

    double n_of_z(double z) const
    { /* CV/bispectrum.cpp:922-949 (file branch); the read of element [n] that the
         reference performs at the last bin carries weight 0 and is dropped here */
        if (z <= 0 || z >= z_max) return 0;
        if (nz_from_file) {
            double diff_z = z / z_max * (n);
            int ix = (int)diff_z;
            diff_z = diff_z - ix;
            double hi = (ix + 1 < n) ? nz_file[ix + 1] : 0.0;
            return nz_file[ix] * (1. - diff_z) + hi * diff_z;
        }
        return (1.7865 * (std::pow(z, 0.4710) + std::pow(z, 0.4710 * 5.1843)) /
                (std::pow(z, 5.1843) + 0.7259)) / 2.97653;
    }

    bool Pk_given = false;                 /* CV/bispectrum.cpp:81-89 */
    std::map<double, double> linearPk_given;

    double linear_pk(double k) const
    { /* CV/bispectrum.cpp:785-809: the given table (linear in k, end values outside) ... */
        if (Pk_given) {
            auto ix = linearPk_given.upper_bound(k);
            if (ix == linearPk_given.end()) return (--ix)->second;
            if (ix == linearPk_given.begin()) return ix->second;
            auto lo = ix;
            --lo;
            double diff = (k - lo->first) / (ix->first - lo->first);
            return diff * ix->second + (1 - diff) * lo->second;
        }
        return linear_pk_eh(k);
    }

    double linear_pk_eh(double k) const
    { /* ... or Eisenstein & Hu no-wiggle, CV/bispectrum.cpp:811-841 */
        double pk, delk, alnu, geff, qeff, L, C;
        k *= cosmo.h;
        double fc = cosmo.omc / cosmo.om;
        double fb = cosmo.omb / cosmo.om;
        double theta = 2.728 / 2.7;
        double pc = 0.25 * (5.0 - std::sqrt(1.0 + 24.0 * fc));
        double omh2 = cosmo.om * cosmo.h * cosmo.h;
        double ombh2 = cosmo.omb * cosmo.h * cosmo.h;
        double zeq = 2.5e+4 * omh2 / std::pow(theta, 4);
        double b1 = 0.313 * std::pow(omh2, -0.419) * (1.0 + 0.607 * std::pow(omh2, 0.674));
        double b2 = 0.238 * std::pow(omh2, 0.223);
        double zd = 1291.0 * std::pow(omh2, 0.251) / (1.0 + 0.659 * std::pow(omh2, 0.828)) *
                    (1.0 + b1 * std::pow(ombh2, b2));
        double yd = (1.0 + zeq) / (1.0 + zd);
        double sh = 44.5 * std::log(9.83 / (omh2)) / std::sqrt(1.0 + 10.0 * std::pow(ombh2, 0.75));
        alnu = fc * (5.0 - 2.0 * pc) / 5.0 * (1.0 - 0.553 * fb + 0.126 * fb * fb * fb) *
               std::pow(1.0 + yd, -pc) *
               (1.0 + 0.5 * pc * (1.0 + 1.0 / (7.0 * (3.0 - 4.0 * pc))) / (1.0 + yd));
        geff = omh2 * (std::sqrt(alnu) + (1.0 - std::sqrt(alnu)) / (1.0 + std::pow(0.43 * k * sh, 4)));
        qeff = k / geff * theta * theta;
        L = std::log(2.718281828 + 1.84 * std::sqrt(alnu) * qeff / (1.0 - 0.949 * fb));
        C = 14.4 + 325.0 / (1.0 + 60.5 * std::pow(qeff, 1.11));
        delk = std::pow(norm, 2) * std::pow(k * 2997.9 / cosmo.h, 3. + cosmo.ns) *
               std::pow(L / (L + C * qeff * qeff), 2);
        pk = 2.0 * M_PI * M_PI / (k * k * k) * delk;
        return std::pow(cosmo.h, 3) * pk;
    }

    static double window(double x, int i)
    { /* CV/bispectrum.cpp:769-783 */
        if (i == 0) return 3.0 / std::pow(x, 3) * (std::sin(x) - x * std::cos(x));
        if (i == 1) return std::exp(-0.5 * x * x);
        if (i == 2) return x * std::exp(-0.5 * x * x);
        return x * x * (1 - x * x) * std::exp(-x * x);
    }

    double sigmam(double r, int j) const
    { /* CV/bispectrum.cpp:709-767: nested doubling trapezoid in ln k with range growth */
        int nn, i;
        double k1, k2, xx = 0, xxp, xxpp, k, a, b, hh;
        k1 = 2. * M_PI / r;
        k2 = 2. * M_PI / r;
        xxpp = -1.0;
        for (;;) {
            k1 = k1 / 10.0;
            k2 = k2 * 2.0;
            a = std::log(k1), b = std::log(k2);
            xxp = -1.0;
            nn = 2;
            for (;;) {
                nn = nn * 2;
                hh = (b - a) / (double)nn;
                xx = 0.;
                for (i = 1; i < nn; i++) {
                    k = std::exp(a + hh * i);
                    if (j < 3)
                        xx += k * k * k * linear_pk(k) * std::pow(window(k * r, j), 2);
                    else
                        xx += k * k * k * linear_pk(k) * window(k * r, j);
                }
                if (j < 3)
                    xx += 0.5 * (k1 * k1 * k1 * linear_pk(k1) * std::pow(window(k1 * r, j), 2) +
                                 k2 * k2 * k2 * linear_pk(k2) * std::pow(window(k2 * r, j), 2));
                else
                    xx += 0.5 * (k1 * k1 * k1 * linear_pk(k1) * window(k1 * r, j) +
                                 k2 * k2 * k2 * linear_pk(k2) * window(k2 * r, j));
                xx *= hh;
                if (std::fabs((xx - xxp) / xx) < eps) break;
                xxp = xx;
            }
            if (std::fabs((xx - xxpp) / xx) < eps) break;
            xxpp = xx;
        }
        if (j < 3) return std::sqrt(xx / (2.0 * M_PI * M_PI));
        return xx / (2.0 * M_PI * M_PI);
    }

    double calc_r_sigma(double D1v) const
    { /* CV/bispectrum.cpp:676-707 (bracket, then bisection on D1*sigma_Gauss(1/k) = 1) */
        double k, k1, k2;
        k1 = k2 = 1.;
        for (;;) {
            if (D1v * sigmam(1. / k1, 1) < 1.) break;
            k1 *= 0.5;
        }
        for (;;) {
            if (D1v * sigmam(1. / k2, 1) > 1.) break;
            k2 *= 2.;
        }
        for (;;) {
            k = 0.5 * (k1 + k2);
            double s = D1v * sigmam(1. / k, 1);
            if (s < 1.)
                k1 = k;
            else if (s > 1.)
                k2 = k;
            if (s == 1. || std::fabs(k2 / k1 - 1.) < eps * 0.1) break;
        }
        return 1. / k;
    }

    double lgr_func(int j, double la, const double y[2]) const
    { /* CV/bispectrum.cpp:895-912 */
        if (j == 0) return y[1];
        double a = std::exp(la);
        double g_ = -0.5 * (5. * cosmo.om + (5. - 3 * cosmo.w) * cosmo.ow * std::pow(a, -3. * cosmo.w)) * y[1] -
                    1.5 * (1. - cosmo.w) * cosmo.ow * std::pow(a, -3. * cosmo.w) * y[0];
        g_ = g_ / (cosmo.om + cosmo.ow * std::pow(a, -3. * cosmo.w));
        return g_;
    }

    double lgr(double z) const
    { /* CV/bispectrum.cpp:843-893: RK4 in ln a from a0 = 1/1100, steps doubling */
        int i, j, nn;
        double a, a0, x, hs, yp;
        double k1[2], k2[2], k3[2], k4[2], y[2], y2[2], y3[2], y4[2];
        a = 1. / (1. + z);
        a0 = 1. / 1100.;
        yp = -1.;
        nn = 10;
        for (;;) {
            nn *= 2;
            hs = (std::log(a) - std::log(a0)) / nn;
            x = std::log(a0);
            y[0] = 1., y[1] = 0.;
            for (i = 0; i < nn; i++) {
                for (j = 0; j < 2; j++) k1[j] = hs * lgr_func(j, x, y);
                for (j = 0; j < 2; j++) y2[j] = y[j] + 0.5 * k1[j];
                for (j = 0; j < 2; j++) k2[j] = hs * lgr_func(j, x + 0.5 * hs, y2);
                for (j = 0; j < 2; j++) y3[j] = y[j] + 0.5 * k2[j];
                for (j = 0; j < 2; j++) k3[j] = hs * lgr_func(j, x + 0.5 * hs, y3);
                for (j = 0; j < 2; j++) y4[j] = y[j] + k3[j];
                for (j = 0; j < 2; j++) k4[j] = hs * lgr_func(j, x + hs, y4);
                for (j = 0; j < 2; j++) y[j] += (k1[j] + k4[j]) / 6. + (k2[j] + k3[j]) / 3.;
                x += hs;
            }
            if (std::fabs(y[0] / yp - 1.) < 0.1 * eps) break;
            yp = y[0];
        }
        return a * y[0];
    }

    void set_cosmology()
    { /* CV/bispectrum.cpp:111-164 */
        fK.assign(n, 0);
        g.assign(n, 0);
        D1.assign(n, 0);
        rs.assign(n, 0);
        neff.assign(n, 0);
        ncur.assign(n, 0);
        for (int i = 0; i < n; i++) fK[i] = f_K_at_z(i * dz);
        for (int i = 0; i < n; i++) {
            g[i] = 0;
            for (int j = i; j < n; j++) {
                double z_now = j * dz;
                if (j == i || j == n - 1)
                    g[i] += n_of_z(z_now) * (fK[j] - fK[i]) / fK[j] / 2;
                else
                    g[i] += n_of_z(z_now) * (fK[j] - fK[i]) / fK[j];
            }
            g[i] = g[i] * dz;
        }
        g[0] = 1.;
        norm = 1.;
        norm *= cosmo.sigma8 / sigmam(8., 0);
        for (int i = 0; i < n; i++) {
            double z_now = i * dz;
            D1[i] = lgr(z_now) / lgr(0.);
            rs[i] = calc_r_sigma(D1[i]);
            double s2 = sigmam(rs[i], 2);
            double d1 = -2. * std::pow(D1[i] * s2, 2);
            neff[i] = -3. + 2. * std::pow(D1[i] * s2, 2);
            ncur[i] = d1 * d1 + 4. * sigmam(rs[i], 3) * std::pow(D1[i], 2);
        }
    }

    static double F2_tree(double k1, double k2, double k3)
    { /* CV/bispectrum.cpp:629-633 */
        double c12 = 0.5 * (k3 * k3 - k1 * k1 - k2 * k2) / (k1 * k2);
        return (5. / 7.) + 0.5 * c12 * (k1 / k2 + k2 / k1) + (2. / 7.) * c12 * c12;
    }

    double F2(double k1, double k2, double k3, double z, double D1v, double r_sigma) const
    { /* CV/bispectrum.cpp:614-627 */
        double q3 = k3 * r_sigma;
        double logsigma8z = std::log10(D1v * cosmo.sigma8);
        double a = 1. / (1. + z);
        double omz = cosmo.om / (cosmo.om + cosmo.ow * std::pow(a, -3. * cosmo.w));
        double dn = std::pow(10., -0.483 + 0.892 * logsigma8z - 0.086 * omz);
        return F2_tree(k1, k2, k3) + dn * q3;
    }

    void coeffs(int idx, double didx, double *D1v, double *r_sigma, double *n_eff) const
    { /* CV/bispectrum.cpp:914-920 */
        *D1v = D1[idx] * (1 - didx) + D1[idx + 1] * didx;
        *r_sigma = rs[idx] * (1 - didx) + rs[idx + 1] * didx;
        *n_eff = neff[idx] * (1 - didx) + neff[idx + 1] * didx;
    }

    double bispec_tree(double k1, double k2, double k3, double D1v) const
    { /* CV/bispectrum.cpp:609-612 */
        return std::pow(D1v, 4) * 2. *
               (F2_tree(k1, k2, k3) * linear_pk(k1) * linear_pk(k2) +
                F2_tree(k2, k3, k1) * linear_pk(k2) * linear_pk(k3) +
                F2_tree(k3, k1, k2) * linear_pk(k3) * linear_pk(k1));
    }

    double bispec(double k1, double k2, double k3, double z, int idx, double didx) const
    { /* CV/bispectrum.cpp:166-236 (BiHalofit, Takahashi et al. 2020 eqs. B4-B9) */
        int i, j;
        double q[4], qt, logsigma8z, r1, r2;
        double an, bn, cn, en, fn, gn, hn, mn, nn_, pn, alphan, betan, mun, nun, BS1h, BS3h, PSE[4];
        double r_sigma, n_eff, D1v;
        coeffs(idx, didx, &D1v, &r_sigma, &n_eff);
        if (z > 10.) return bispec_tree(k1, k2, k3, D1v);
        q[1] = k1 * r_sigma, q[2] = k2 * r_sigma, q[3] = k3 * r_sigma;
        for (i = 1; i <= 3; i++)
            for (j = i + 1; j <= 3; j++)
                if (q[i] < q[j]) {
                    qt = q[j];
                    q[j] = q[i];
                    q[i] = qt;
                }
        r1 = q[3] / q[1], r2 = (q[2] + q[3] - q[1]) / q[1];
        q[1] = k1 * r_sigma, q[2] = k2 * r_sigma, q[3] = k3 * r_sigma;
        logsigma8z = std::log10(D1v * cosmo.sigma8);
        an = std::pow(10., -2.167 - 2.944 * logsigma8z - 1.106 * std::pow(logsigma8z, 2) -
                               2.865 * std::pow(logsigma8z, 3) -
                               0.310 * std::pow(r1, std::pow(10., 0.182 + 0.57 * n_eff)));
        bn = std::pow(10., -3.428 - 2.681 * logsigma8z + 1.624 * std::pow(logsigma8z, 2) -
                               0.095 * std::pow(logsigma8z, 3));
        cn = std::pow(10., 0.159 - 1.107 * n_eff);
        alphan = std::pow(10., -4.348 - 3.006 * n_eff - 0.5745 * std::pow(n_eff, 2) +
                                   std::pow(10., -0.9 + 0.2 * n_eff) * std::pow(r2, 2));
        if (alphan > 1. - (2. / 3.) * cosmo.ns) alphan = 1. - (2. / 3.) * cosmo.ns;
        betan = std::pow(10., -1.731 - 2.845 * n_eff - 1.4995 * std::pow(n_eff, 2) -
                                  0.2811 * std::pow(n_eff, 3) + 0.007 * r2);
        BS1h = 1.;
        for (i = 1; i <= 3; i++)
            BS1h *= 1. / (an * std::pow(q[i], alphan) + bn * std::pow(q[i], betan)) / (1. + 1. / (cn * q[i]));
        fn = std::pow(10., -10.533 - 16.838 * n_eff - 9.3048 * std::pow(n_eff, 2) - 1.8263 * std::pow(n_eff, 3));
        gn = std::pow(10., 2.787 + 2.405 * n_eff + 0.4577 * std::pow(n_eff, 2));
        hn = std::pow(10., -1.118 - 0.394 * n_eff);
        mn = std::pow(10., -2.605 - 2.434 * logsigma8z + 5.71 * std::pow(logsigma8z, 2));
        nn_ = std::pow(10., -4.468 - 3.08 * logsigma8z + 1.035 * std::pow(logsigma8z, 2));
        mun = std::pow(10., 15.312 + 22.977 * n_eff + 10.9579 * std::pow(n_eff, 2) + 1.6586 * std::pow(n_eff, 3));
        nun = std::pow(10., 1.347 + 1.246 * n_eff + 0.4525 * std::pow(n_eff, 2));
        pn = std::pow(10., 0.071 - 0.433 * n_eff);
        en = std::pow(10., -0.632 + 0.646 * n_eff);
        for (i = 1; i <= 3; i++)
            PSE[i] = (1. + fn * std::pow(q[i], 2)) / (1. + gn * q[i] + hn * std::pow(q[i], 2)) *
                         std::pow(D1v, 2) * linear_pk(q[i] / r_sigma) +
                     1. / (mn * std::pow(q[i], mun) + nn_ * std::pow(q[i], nun)) / (1. + std::pow(pn * q[i], -3));
        BS3h = 2. * (F2(k1, k2, k3, z, D1v, r_sigma) * PSE[1] * PSE[2] +
                     F2(k2, k3, k1, z, D1v, r_sigma) * PSE[2] * PSE[3] +
                     F2(k3, k1, k2, z, D1v, r_sigma) * PSE[3] * PSE[1]);
        for (i = 1; i <= 3; i++) BS3h *= 1. / (1. + en * q[i]);
        return BS1h + BS3h;
    }

    void z_index(double z, int *idx, double *didx) const
    { /* CV/bispectrum.cpp:506-513 */
        double d = z / z_max * (n - 1);
        int i = (int)d;
        d = d - i;
        if (i == n - 1) {
            i = n - 2;
            d = 1.;
        }
        *idx = i;
        *didx = d;
    }

    static double bispectrum_analytic_single_a(double l1, double l2, double phi, double a)
    { /* CV/bispectrum.cpp:989-996 */
        double bis_pref = -std::pow(M_PI, 2) / (8 * 72 * std::pow(a, 8));
        double pref = std::pow(l1, 2) * std::pow(l2, 2) * (std::pow(l1, 2) + std::pow(l2, 2) + 2 * l1 * l2 * std::cos(phi));
        double expon = std::exp(-(std::pow(l1, 2) + std::pow(l2, 2) + l1 * l2 * std::cos(phi)) / (6 * a));
        return bis_pref * pref * expon;
    }

    double integrand_bkappa(double z, double ell1, double ell2, double ell3, bool *bad = nullptr) const
    { /* CV/bispectrum.cpp:471-541 */
        if (analytic) {
            if (z > 1) return 0;
            double phi = M_PI - std::acos((ell1 * ell1 + ell2 * ell2 - ell3 * ell3) / (2 * ell1 * ell2));
            if (std::isnan(phi)) phi = 0;
            static const double wts[9] = {0, 1.0e+8, 1.0e+6, 1.0e+4, 1.0e+2, 1.0e+0, 5.0e-3, 1.0e-5, 1.0e-6};
            static const double avs[9] = {3.0e+3, 1.0e+4, 3.0e+4, 1.0e+5, 3.0e+5, 1.0e+6, 3.0e+6, 1.0e+7, 3.0e+7};
            double temp = 0;
            for (int i = 0; i < 9; i++) temp += wts[i] * bispectrum_analytic_single_a(ell1, ell2, phi, avs[i]);
            temp /= (27. / 8. * std::pow(cosmo.om, 3) * std::pow(100. / 299792., 5));
            return temp * 3.;
        }
        if (z == 0) return 0.;
        if (ell3 == 0) return 0.;
        int idx;
        double didx;
        z_index(z, &idx, &didx);
        double g_value = g[idx] * (1 - didx) + g[idx + 1] * didx;
        double f_K_value = fK[idx] * (1 - didx) + fK[idx + 1] * didx;
        double result = std::pow(g_value * (1. + z), 3) *
                        bispec(ell1 / f_K_value, ell2 / f_K_value, ell3 / f_K_value, z, idx, didx) /
                        f_K_value / E(z);
        if (!std::isfinite(result) && bad) *bad = true;
        return result;
    }

    double bkappa(double ell1, double ell2, double ell3, bool *bad = nullptr) const
    { /* CV/bispectrum.cpp:238-286 with GQ96_of_bdelta CV:1010-1020 */
        if (ell1 == 0 || ell2 == 0 || ell3 == 0) return 0;
        if (analytic) {
            /* the argument is clamped to [-1, 1]; the reference's bkappa branch has no guard and
               returns NaN where rounding pushes it past 1 (phi = 0, pi) */
            double carg = (ell1 * ell1 + ell2 * ell2 - ell3 * ell3) / (2 * ell1 * ell2);
            carg = std::min(1.0, std::max(-1.0, carg));
            double phi = std::acos(carg);
            static const double wts[9] = {0, 1.0e+8, 1.0e+6, 1.0e+4, 1.0e+2, 1.0e+0, 5.0e-3, 1.0e-5, 1.0e-6};
            static const double avs[9] = {3.0e+3, 1.0e+4, 3.0e+4, 1.0e+5, 3.0e+5, 1.0e+6, 3.0e+6, 1.0e+7, 3.0e+7};
            double temp = 0;
            for (int i = 0; i < 9; i++) temp += wts[i] * bispectrum_analytic_single_a(ell1, ell2, phi, avs[i]);
            return temp * (3. * std::pow(2 * M_PI, 2));
        }
        double prefactor = 27. / 8. * std::pow(cosmo.om, 3) * std::pow(H0_over_c, 5);
        const GL96 &G = gl96();
        double a = 0, b = z_max;
        double cx = (a + b) / 2, dx = (b - a) / 2, q = 0;
        for (int i = 0; i < 48; i++)
            q += G.W[i] * (integrand_bkappa(cx - dx * G.A[i], ell1, ell2, ell3, bad) +
                           integrand_bkappa(cx + dx * G.A[i], ell1, ell2, ell3, bad));
        return prefactor * (q * dx);
    }

    static double uHat(double eta)
    { /* CV/apertureStatistics.cpp:10-14 */
        double temp = 0.5 * eta * eta;
        return temp * std::exp(-temp);
    }

    double map3_integrand(double l1, double l2, double phi, const double *th, bool *bad = nullptr) const
    { /* CV/apertureStatistics.cpp:16-28 */
        double l3 = std::sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * std::cos(phi));
        if (!std::isfinite(l1) || !std::isfinite(l2) || !std::isfinite(l3)) {
            if (bad) *bad = true;
            return NAN;
        }
        return l1 * l2 * bkappa(l1, l2, l3, bad) * uHat(l1 * th[0]) * uHat(l2 * th[1]) * uHat(l3 * th[2]);
    }

    /* ---------------- Gamma_i (CV/gamma.cpp) ---------------- */
    static double psi_f(double t) { return t * std::tanh(M_PI * std::sinh(t) / 2); } /* CV:178-181 */
    static double psip_f(double t)
    { /* CV/gamma.cpp:183-188 */
        double zahler = std::sinh(M_PI * std::sinh(t)) + M_PI * t * std::cosh(t);
        double nenner = std::cosh(M_PI * std::sinh(t)) + 1;
        return zahler / nenner;
    }

    void initialize_bessel(double ph, double pk_arg)
    { /* CV/gamma.cpp:133-176 */
        prec_h = ph;
        prec_k = (unsigned)(pk_arg / prec_h);
        psi6.assign(prec_k, 0);
        prod6.assign(prec_k, 0);
        psi2.assign(prec_k, 0);
        prod2.assign(prec_k, 0);
        for (int nu : {6, 2}) {
            std::vector<double> &ps = (nu == 6) ? psi6 : psi2;
            std::vector<double> &pr = (nu == 6) ? prod6 : prod2;
            for (unsigned i = 0; i < prec_k; i++) {
                double bz = bessel_zero_jn(nu, (int)i);
                double pbz = bz / M_PI;
                ps[i] = M_PI * psi_f(pbz * prec_h) / prec_h;
                double jb = bessel_jn(nu, ps[i]);
                double pp = psip_f(prec_h * pbz);
                double w = 2 / (M_PI * bz * std::pow(bessel_jn(nu + 1, bz), 2));
                pr[i] = w * std::pow(ps[i], 3) * jb * pp;
            }
        }
    }

    static double interior_angle(double an1, double an2, double opp)
    { /* CV/gamma.cpp:121-124 */
        return std::acos((std::pow(an1, 2) + std::pow(an2, 2) - std::pow(opp, 2)) / (2.0 * an1 * an2));
    }
    static double height_of_triangle(double x1, double x2, double x3)
    {
        return 0.5 * std::sqrt(2 * std::pow(x1, 2) + 2 * std::pow(x2, 2) - std::pow(x3, 2));
    }
    static double alpha_f(double psi, double x1, double x2, double phi, double varpsi)
    { /* CV/gamma.cpp:192-197 */
        double zahler = (std::cos(psi) * x2 - std::sin(psi) * x1) * std::sin((phi + varpsi) / 2);
        double nenner = (std::cos(psi) * x2 + std::sin(psi) * x1) * std::cos((phi + varpsi) / 2);
        return std::atan2(zahler, nenner);
    }
    static double betabar(double psi, double phi)
    { /* CV/gamma.cpp:199-204 */
        double zahler = std::cos(2 * psi) * std::sin(phi);
        double nenner = std::cos(phi) + std::sin(2 * psi);
        return 0.5 * std::atan2(zahler, nenner);
    }
    static double A_f(double psi, double x1, double x2, double phi, double varpsi)
    { /* CV/gamma.cpp:216-219 */
        return std::sqrt(std::pow(std::cos(psi) * x2, 2) + std::pow(std::sin(psi) * x1, 2) +
                         std::sin(2 * psi) * x1 * x2 * std::cos(phi + varpsi));
    }

    double ogata_sum(int nu, double z, double phi, double psi, double A3) const
    { /* CV/gamma.cpp:291-318 (J6) and 462-488 (J2) */
        const std::vector<double> &ps = (nu == 6) ? psi6 : psi2;
        const std::vector<double> &pr = (nu == 6) ? prod6 : prod2;
        double bis = 0, cpsi = std::cos(psi), spsi = std::sin(psi);
        for (unsigned k = 1; k < prec_k; k++) {
            double ell1 = ps[k] / A3 * cpsi;
            double ell2 = ps[k] / A3 * spsi;
            double ell3 = ell1 * ell1 + ell2 * ell2 + 2 * ell1 * ell2 * std::cos(phi);
            if (ell3 <= 0)
                ell3 = 0;
            else
                ell3 = std::sqrt(ell3);
            bis += integrand_bkappa(z, ell1, ell2, ell3) * pr[k];
        }
        return bis * M_PI / std::pow(A3, 4);
    }

    std::complex<double> gamma0_one_x(double z, double phi, double psi, double x1, double x2, double x3) const
    { /* CV/gamma.cpp:320-331 with exponential_prefactor / exponential_of_betabar CV:206-214 */
        double varpsi = interior_angle(x1, x2, x3);
        double A3 = A_f(psi, x1, x2, phi, varpsi);
        std::complex<double> ep = std::exp(std::complex<double>(
            0, (interior_angle(x2, x3, x1) - interior_angle(x1, x3, x2) - 6 * alpha_f(psi, x1, x2, phi, varpsi))));
        std::complex<double> eb = std::exp(std::complex<double>(0, betabar(psi, phi) * 2.));
        std::complex<double> prefactor = std::sin(2 * psi) * ep * eb;
        return prefactor * ogata_sum(6, z, phi, psi, A3);
    }

    std::complex<double> gamma_integrand(int which, double z, double phi, double psi, double x1, double x2,
                                         double x3) const
    {
        if (which == 0) /* CV/gamma.cpp:333-336 */
            return gamma0_one_x(z, phi, psi, x1, x2, x3) + gamma0_one_x(z, phi, psi, x2, x3, x1) +
                   gamma0_one_x(z, phi, psi, x3, x1, x2);
        /* CV/gamma.cpp:490-520 */
        double vb = betabar(psi, phi);
        double vp1 = interior_angle(x2, x3, x1), vp2 = interior_angle(x3, x1, x2), vp3 = interior_angle(x1, x2, x3);
        double A1 = A_f(psi, x2, x3, phi, vp1), A2 = A_f(psi, x3, x1, phi, vp2), A3 = A_f(psi, x1, x2, phi, vp3);
        double al1 = alpha_f(psi, x2, x3, phi, vp1), al2 = alpha_f(psi, x3, x1, phi, vp2),
               al3 = alpha_f(psi, x1, x2, phi, vp3);
        std::complex<double> i3 = std::exp(std::complex<double>(0, vp1 - vp2 + 2 * vp3 + 2 * (vb - phi - al3)));
        i3 *= ogata_sum(2, z, phi, psi, A3);
        std::complex<double> i1 = std::exp(std::complex<double>(0, vp3 - vp2 - 2 * (vb + al1)));
        i1 *= ogata_sum(2, z, phi, psi, A1);
        std::complex<double> i2 = std::exp(std::complex<double>(0, vp3 - vp1 - 2 * vp2 + 2 * (vb + phi - al2)));
        i2 *= ogata_sum(2, z, phi, psi, A2);
        return std::sin(2 * psi) * (i1 + i2 + i3);
    }

    /* ---------------- convergence power spectrum and <Map^2> ---------------- */
    double om_m_of_z(double z) const
    { /* CV/bispectrum.cpp:1022-1026 (no w: the flat-universe form) */
        double aa = 1. / (1 + z);
        return cosmo.om / (cosmo.om + aa * (aa * aa * cosmo.ow + (1 - cosmo.om - cosmo.ow)));
    }
    double om_v_of_z(double z) const
    { /* CV/bispectrum.cpp:1028-1032 */
        double aa = 1. / (1 + z);
        return cosmo.ow * std::pow(aa, 3) / (cosmo.om + aa * (aa * aa * cosmo.ow + (1 - cosmo.om - cosmo.ow)));
    }

    double P_k_nonlinear(double k, double z) const
    { /* CV/bispectrum.cpp:288-367: revised Halofit (Takahashi et al. 2012) with the tabulated
         D1, r_sigma, n_eff and n_cur; the reference's flatness check cannot fail here (ow = 1 - om) */
        int idx;
        double didx;
        z_index(z, &idx, &didx);
        double r_sigma, n_eff, D1v;
        coeffs(idx, didx, &D1v, &r_sigma, &n_eff);
        double nsqr = n_eff * n_eff;
        double nc = ncur[idx] * (1 - didx) + ncur[idx + 1] * didx;
        double om_m = om_m_of_z(z), om_v = om_v_of_z(z);
        double f1a = std::pow(om_m, -0.0732), f2a = std::pow(om_m, -0.1423), f3a = std::pow(om_m, 0.0725);
        double f1b = std::pow(om_m, -0.0307), f2b = std::pow(om_m, -0.0585), f3b = std::pow(om_m, 0.0743);
        double frac = om_v / (1. - om_m);
        double f1 = frac * f1b + (1 - frac) * f1a;
        double f2 = frac * f2b + (1 - frac) * f2a;
        double f3 = frac * f3b + (1 - frac) * f3a;
        double a = 1.5222 + 2.8553 * n_eff + 2.3706 * nsqr + 0.9903 * n_eff * nsqr + 0.2250 * nsqr * nsqr -
                   0.6038 * nc + 0.1749 * om_v * (1.0 + cosmo.w);
        a = std::pow(10.0, a);
        double b = std::pow(10.0, -0.5642 + 0.5864 * n_eff + 0.5716 * nsqr - 1.5474 * nc +
                                      0.2279 * om_v * (1.0 + cosmo.w));
        double c = std::pow(10.0, 0.3698 + 2.0404 * n_eff + 0.8161 * nsqr + 0.5869 * nc);
        double gam = 0.1971 - 0.0843 * n_eff + 0.8460 * nc;
        double alpha = std::fabs(6.0835 + 1.3373 * n_eff - 0.1959 * nsqr - 5.5274 * nc);
        double beta = 2.0379 - 0.7354 * n_eff + 0.3157 * nsqr + 1.2490 * n_eff * nsqr + 0.3980 * nsqr * nsqr -
                      0.1682 * nc;
        double xnu = std::pow(10.0, 5.2105 + 3.6902 * n_eff);
        double plin = linear_pk(k) * D1v * D1v * k * k * k / (2 * M_PI * M_PI);
        double y = k * r_sigma, ysqr = y * y;
        double ph = a * std::pow(y, f1 * 3) / (1 + b * std::pow(y, f2) + std::pow(f3 * c * y, 3 - gam));
        ph = ph / (1 + xnu / ysqr);
        double pq = plin * std::pow(1 + plin, beta) / (1 + plin * alpha) * std::exp(-y / 4.0 - ysqr / 8.0);
        return 2 * M_PI * M_PI * (pq + ph) / (k * k * k);
    }

    double linear_pk_at_z(double k, double z) const
    { /* CV/bispectrum.cpp:369-380 */
        int idx;
        double didx;
        z_index(z, &idx, &didx);
        double r_sigma, n_eff, D1v;
        coeffs(idx, didx, &D1v, &r_sigma, &n_eff);
        return D1v * D1v * linear_pk(k);
    }

    static double t17_correction(double k)
    { /* GV/bispectrum.cu:971-979: the ratio fitted to the T17 simulations */
        const double c1 = 9.5171e-4, c2 = 5.1543e-3, a1 = 1.3063, a2 = 1.1475, a3 = 0.62793;
        return std::pow(1 + c1 * std::pow(k, -1. * a1), a1) / std::pow(1 + c2 * std::pow(k, -1. * a2), a3);
    }

    double limber_integrand_power_spectrum(double ell, double z, bool nonlinear, bool t17 = false) const
    { /* CV/bispectrum.cpp:543-577 (GV/bispectrum.cu:951-982 for the T17 factor); the reference's
         f_K = 0 exit cannot trigger for z >= 1e-5 */
        if (z < 1e-5) return 0;
        int idx;
        double didx;
        z_index(z, &idx, &didx);
        double g_value = g[idx] * (1 - didx) + g[idx + 1] * didx;
        double f_K_value = fK[idx] * (1 - didx) + fK[idx + 1] * didx;
        double prefactor = 9. / 4. * H0_over_c * H0_over_c * H0_over_c * cosmo.om * cosmo.om * (1. + z) * (1. + z) *
                           g_value * g_value / E(z);
        double k = ell / f_K_value;
        double corr = t17 ? t17_correction(k) : 1.0;
        return corr * prefactor * (nonlinear ? P_k_nonlinear(k, z) : linear_pk_at_z(k, z));
    }

    double GQ96_of_Pk(double a, double b, double ell, bool nonlinear = true, bool t17 = false) const
    { /* CV/bispectrum.cpp:1034-1046 */
        const GL96 &G = gl96();
        double cx = (a + b) / 2, dx = (b - a) / 2, q = 0;
        for (int i = 0; i < 48; i++)
            q += G.W[i] * (limber_integrand_power_spectrum(ell, cx - dx * G.A[i], nonlinear, t17) +
                           limber_integrand_power_spectrum(ell, cx + dx * G.A[i], nonlinear, t17));
        return q * dx;
    }

    double Pell(double ell, double sigma, double n_gal, bool t17 = false) const
    { /* CV/bispectrum.cpp:1048-1059: Limber-integrated P(k) plus the shape-noise power 0.5 sigma^2 / n */
        double P_shapenoise = 0.5 * sigma * sigma / n_gal;
        if (ell == 0) return P_shapenoise;
        return GQ96_of_Pk(0, z_max, ell, true, t17) + P_shapenoise;
    }

    /* ---------------- Gaussian <Map^3> covariance ---------------- */
    double limber_integrand_triple_power_spectrum(double ell1, double ell2, double ell3, double z) const
    { /* CV/bispectrum.cpp:580-607 (the f_K = 0 exit cannot trigger for z >= 1e-5) */
        if (z < 1e-5) return 0;
        int idx;
        double didx;
        z_index(z, &idx, &didx);
        double g_value = g[idx] * (1 - didx) + g[idx + 1] * didx;
        double f_K_value = fK[idx] * (1 - didx) + fK[idx + 1] * didx;
        double prefactor = 9. / 4. * H0_over_c * H0_over_c * H0_over_c * cosmo.om * cosmo.om * (1. + z) * (1. + z) *
                           g_value * g_value / E(z);
        return std::pow(prefactor, 3) * P_k_nonlinear(ell1 / f_K_value, z) * P_k_nonlinear(ell2 / f_K_value, z) *
               P_k_nonlinear(ell3 / f_K_value, z);
    }

    static double uhat_product(double l1, double l2, double l3, const double *th)
    { /* CV/apertureStatistics.cpp:50-63 */
        return uHat(l1 * th[0]) * uHat(l2 * th[1]) * uHat(l3 * th[2]);
    }

    double cov_gauss_integrand(double l1, double l2, double phi, double z, const double *th6) const
    { /* CV/apertureStatistics.cpp:66-106: uHat_product(123) x the six permutations of (456), the
         single-z triple power spectrum, l1 l2 */
        double l3 = std::sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * std::cos(phi));
        double result = uhat_product(l1, l2, l3, th6);
        double perm = uhat_product(l1, l2, l3, th6 + 3);
        perm += uhat_product(l2, l3, l1, th6 + 3);
        perm += uhat_product(l3, l1, l2, th6 + 3);
        perm += uhat_product(l1, l3, l2, th6 + 3);
        perm += uhat_product(l3, l2, l1, th6 + 3);
        perm += uhat_product(l2, l1, l3, th6 + 3);
        result *= perm;
        result *= limber_integrand_triple_power_spectrum(l1, l2, l3, z);
        result *= l1 * l2;
        return result;
    }

    double pell_flags(double ell, double sigma, double n_gal, int flags) const
    { /* the cuda_version's device Pell (GV/bispectrum.cu:511-529, __CUDA_ARCH__ branch): the
         96-node Limber sum + shape noise, shape noise alone with the constant spectrum (flags & 4) */
        double P_shapenoise = 0.5 * sigma * sigma / n_gal;
        if (flags & 4) return P_shapenoise;
        if (ell == 0) return P_shapenoise;
        return GQ96_of_Pk(0, z_max, ell, !(flags & 1), (flags & 2) != 0) + P_shapenoise;
    }

    double t1_integrand(double l1, double l2, double phi, const double *th6, double lmin, double lmax, double sigma,
                        double n_gal, int flags) const
    { /* GV/apertureStatisticsCovariance.cu:1958-1985 (integrand_T1_infinite) */
        double l3 = std::sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * std::cos(phi));
        if (l1 <= lmin || l2 <= lmin || l3 <= lmin || l3 > lmax) return 0;
        double result = pell_flags(l1, sigma, n_gal, flags) * pell_flags(l2, sigma, n_gal, flags) *
                        pell_flags(l3, sigma, n_gal, flags);
        result *= uHat(l1 * th6[0]) * uHat(l2 * th6[1]) * uHat(l3 * th6[2]);
        result *= uHat(l1 * th6[3]) * uHat(l2 * th6[4]) * uHat(l3 * th6[5]);
        result *= l1 * l2;
        return result;
    }

    double map2_integrand(double ell, double z, double theta, bool nonlinear, bool t17, double sigma,
                          double ngal_arcmin2) const
    { /* GV/apertureStatistics.cu:29-57 for one source bin; sigma = 0 drops the shape-noise term */
        double u = uHat(ell * theta);
        double v = ell * std::pow(u, 2) * limber_integrand_power_spectrum(ell, z, nonlinear, t17);
        if (t17) v /= 1 + std::pow(ell / (1.6 * 8192.), 2);
        if (sigma != 0)
            v += ell * std::pow(u * sigma, 2) * 0.5 / ngal_arcmin2 * std::pow(2.9088820866e-4, 2) / z_max;
        return v;
    }

    static double one_rotation_angle_otc(double x1, double x2, double x3)
    { /* CV/gamma.cpp:624-633 */
        double psi3 = interior_angle(x1, x2, x3);
        double h3 = height_of_triangle(x1, x2, x3);
        double cos_2 = std::pow(std::pow(x2, 2) - std::pow(x1, 2), 2) - 4 * std::pow(x1 * x2 * std::sin(psi3), 2);
        cos_2 /= std::pow(2 * h3 * x3, 2);
        double sin_2 = (std::pow(x2, 2) - std::pow(x1, 2)) * x1 * x2 * std::sin(psi3);
        sin_2 /= std::pow(h3 * x3, 2);
        return std::atan2(sin_2, cos_2) / 2;
    }
};

struct Map3Ctx {
    const Calc *c;
    double th[3];
    int nthreads;
    bool bad;
};

int map3_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* CV/apertureStatistics.cpp:254-279 */
    Map3Ctx *m = (Map3Ctx *)fdata;
    if (fdim != 1 || ndim != 3) return 1;
    bool bad = false;
#pragma omp parallel for schedule(static) num_threads(m->nthreads) reduction(|| : bad)
    for (long long i = 0; i < (long long)npts; i++) {
        bool b = false;
        fval[i] = m->c->map3_integrand(x[i * 3], x[i * 3 + 1], x[i * 3 + 2], m->th, &b);
        bad = bad || b;
    }
    if (bad) m->bad = true;
    return bad ? 1 : 0;
}

struct GammaCtx {
    const Calc *c;
    int which;
    double x1, x2, x3;
    int nthreads;
};

int gamma_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* CV/gamma.cpp:338-359, 522-543 */
    GammaCtx *g = (GammaCtx *)fdata;
    if (fdim != 2 || ndim != 3) return 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g->nthreads)
    for (long long i = 0; i < (long long)npts; i++) {
        std::complex<double> t =
            g->c->gamma_integrand(g->which == 0 ? 0 : 1, x[i * 3], x[i * 3 + 1], x[i * 3 + 2], g->x1, g->x2, g->x3);
        fval[i * 2] = t.real();
        fval[i * 2 + 1] = t.imag();
    }
    return 0;
}

struct PowerCtx {
    const Calc *c;
    double ell;   // P_kappa: the multipole; <Map^2>: the aperture radius [rad]
    bool nonlinear, t17;
    double sigma, ngal;
};

int pkappa_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* CV/bispectrum.cpp:390-417 */
    PowerCtx *p = (PowerCtx *)fdata;
    if (fdim != 1 || ndim != 1) return 1;
    for (size_t i = 0; i < npts; i++)
        fval[i] = p->c->limber_integrand_power_spectrum(p->ell, x[i], p->nonlinear, p->t17);
    return 0;
}

int map2_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* GV/apertureStatistics.cu:59-108: points (ell, z) */
    PowerCtx *p = (PowerCtx *)fdata;
    if (fdim != 1 || ndim != 2) return 1;
    for (size_t i = 0; i < npts; i++)
        fval[i] = p->c->map2_integrand(x[2 * i], x[2 * i + 1], p->ell, p->nonlinear, p->t17, p->sigma, p->ngal);
    return 0;
}

struct CovCtx {
    const Calc *c;
    double th[6];
    double lmin, lmax, sigma, ngal;
    int flags, nthreads;
};

int cov_gauss_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* CV/apertureStatistics.cpp:108-159 */
    CovCtx *p = (CovCtx *)fdata;
    if (fdim != 1 || ndim != 4) return 1;
#pragma omp parallel for schedule(static) num_threads(p->nthreads)
    for (long long i = 0; i < (long long)npts; i++)
        fval[i] = p->c->cov_gauss_integrand(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3], p->th);
    return 0;
}

int t1_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{ /* GV/apertureStatisticsCovariance.cu integrand_T1 with the infinite-survey kernel */
    CovCtx *p = (CovCtx *)fdata;
    if (fdim != 1 || ndim != 3) return 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(p->nthreads)
    for (long long i = 0; i < (long long)npts; i++)
        fval[i] = p->c->t1_integrand(x[3 * i], x[3 * i + 1], x[3 * i + 2], p->th, p->lmin, p->lmax, p->sigma, p->ngal,
                                     p->flags);
    return 0;
}

/* Survey window factors of the finite geometries (GV/apertureStatisticsCovariance.cu:117-177):
   circle 4 J1(theta_max l)^2 / (theta_max l)^2; square and rectangle the product of the two squared
   sinc factors of the half side lengths, 1 below |argument| 1e-6 */
double g_circle(double ell, double tmax)
{
    double t = tmax * ell;
    double r = j1(t);
    r *= r;
    r *= 4 / t / t;
    return r;
}

double sinc_cut(double t) { return std::fabs(t) <= 1e-6 ? 1 : std::sin(t) / t; }

double g_box(double lx, double ly, double tx, double ty)
{
    double a = sinc_cut(0.5 * lx * tx), b = sinc_cut(0.5 * ly * ty);
    return a * a * b * b;
}

/* The four T2 integrands (GV/apertureStatisticsCovariance.cu:2014-2110): kind 0 integrand_T2_part1
   (ell Pell uHat uHat, 1-D), 1 integrand_T2_part2_circle (1-D, x G_circle), 2 the square and 3 the
   rectangle (2-D over (ell_x, ell_y), x G) */
struct T2Ctx {
    const Calc *c;
    int kind;
    double ta, tb, tmax, tmax2, lmin, sigma, ngal;
    int flags, nthreads;
};

double t2_point(const T2Ctx &p, const double *x)
{
    if (p.kind <= 1) {
        double ell = x[0];
        if (ell < p.lmin) return 0;
        if (p.kind == 0) {
            double r = ell * p.c->pell_flags(ell, p.sigma, p.ngal, p.flags);
            r *= Calc::uHat(ell * p.ta) * Calc::uHat(ell * p.tb);
            return r;
        }
        double G = g_circle(ell, p.tmax);
        double r = p.c->pell_flags(ell, p.sigma, p.ngal, p.flags);
        r *= ell * Calc::uHat(ell * p.ta) * Calc::uHat(ell * p.tb);
        r *= G;
        return r;
    }
    double lx = x[0], ly = x[1];
    double G = g_box(lx, ly, p.tmax, p.kind == 2 ? p.tmax : p.tmax2);
    double ell = std::sqrt(lx * lx + ly * ly);
    if (ell < p.lmin) return 0;
    double r = p.c->pell_flags(ell, p.sigma, p.ngal, p.flags);
    r *= Calc::uHat(ell * p.ta) * Calc::uHat(ell * p.tb);
    r *= G;
    return r;
}

int t2_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{
    T2Ctx *p = (T2Ctx *)fdata;
    if (fdim != 1 || ndim != (p->kind <= 1 ? 1u : 2u)) return 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(p->nthreads)
    for (long long i = 0; i < (long long)npts; i++) fval[i] = t2_point(*p, x + ndim * i);
    return 0;
}

/* The pairings of T2_total (GV/apertureStatisticsCovariance.cu:328-538), one side at a time: each
   entry is (multiplicity, three indices into the side's radii).  For theta_123 the indices are
   T2's arguments 1-3 (the two radii of one A factor, then the radius of B); for theta_456 its
   arguments 4-6 (the radius of B, then the two of the other A factor).  The reference's table is
   the product of the two lists, with the equality tests in its order.  The theta_1 == theta_3 list is
   the reference's (:451-488): (theta_2, theta_1, theta_1) where the pairing sum has (theta_1,
   theta_1, theta_2). */
int t2_side_123(const double *t, int mult[3], int idx[3][3])
{
    static const int all[1][3] = {{0, 0, 0}}, e01[2][3] = {{0, 0, 2}, {0, 2, 0}},
                     e12[2][3] = {{0, 1, 1}, {1, 1, 0}}, e02[2][3] = {{0, 1, 0}, {1, 0, 0}},
                     dif[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    const int(*L)[3];
    int n;
    if (t[0] == t[1] && t[1] == t[2]) {
        L = all; n = 1; mult[0] = 3;
    } else if (t[0] == t[1]) {
        L = e01; n = 2; mult[0] = 1; mult[1] = 2;
    } else if (t[1] == t[2]) {
        L = e12; n = 2; mult[0] = 2; mult[1] = 1;
    } else if (t[0] == t[2]) {
        L = e02; n = 2; mult[0] = 2; mult[1] = 1;
    } else {
        L = dif; n = 3; mult[0] = mult[1] = mult[2] = 1;
    }
    for (int k = 0; k < n; ++k)
        for (int d = 0; d < 3; ++d) idx[k][d] = L[k][d];
    return n;
}

int t2_side_456(const double *t, int mult[3], int idx[3][3])
{
    static const int all[1][3] = {{0, 0, 0}}, e01[2][3] = {{0, 0, 2}, {2, 0, 0}},
                     e02[2][3] = {{0, 1, 0}, {1, 0, 0}}, e12[2][3] = {{1, 0, 1}, {0, 1, 1}},
                     dif[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
    const int(*L)[3];
    int n;
    if (t[0] == t[1] && t[1] == t[2]) {
        L = all; n = 1; mult[0] = 3;
    } else if (t[0] == t[1]) {
        L = e01; n = 2; mult[0] = 2; mult[1] = 1;
    } else if (t[0] == t[2]) {
        L = e02; n = 2; mult[0] = 2; mult[1] = 1;
    } else if (t[1] == t[2]) {
        L = e12; n = 2; mult[0] = 2; mult[1] = 1;
    } else {
        L = dif; n = 3; mult[0] = mult[1] = mult[2] = 1;
    }
    for (int k = 0; k < n; ++k)
        for (int d = 0; d < 3; ++d) idx[k][d] = L[k][d];
    return n;
}

/* T1 of a finite survey (GV/apertureStatisticsCovariance.cu:1886-1955, 1984-2012): 6-D over
   (a, b) the window's wave vector and (c, d), (e, f) two of the three Pell wave vectors,
   l1 = |(a - c - e, b - d - f)|, l2 = |(c, d)|, l3 = |(e, f)|; the circle cuts l_i <= lMin, the square
   and the rectangle l_i <= 0 and l3 > lMax */
struct T1FinCtx {
    const Calc *c;
    int geometry;
    double th[6], tmax, tmax2, lmin, lmax, sigma, ngal;
    int flags, nthreads;
};

double t1fin_point(const T1FinCtx &p, const double *x)
{
    double a = x[0], b = x[1], c = x[2], d = x[3], e = x[4], f = x[5];
    double G;
    if (p.geometry == 0) G = g_circle(std::sqrt(a * a + b * b), p.tmax);
    else G = g_box(a, b, p.tmax, p.geometry == 1 ? p.tmax : p.tmax2);
    double ell1 = std::sqrt((a - c - e) * (a - c - e) + (b - d - f) * (b - d - f));
    double ell2 = std::sqrt(c * c + d * d);
    double ell3 = std::sqrt(e * e + f * f);
    if (p.geometry == 0) {
        if (ell1 <= p.lmin || ell2 <= p.lmin || ell3 <= p.lmin) return 0;
    } else if (ell1 <= 0 || ell2 <= 0 || ell3 <= 0 || ell3 > p.lmax) {
        return 0;
    }
    const Calc &C = *p.c;
    double r = C.pell_flags(ell1, p.sigma, p.ngal, p.flags) * C.pell_flags(ell2, p.sigma, p.ngal, p.flags) *
               C.pell_flags(ell3, p.sigma, p.ngal, p.flags);
    r *= Calc::uHat(ell1 * p.th[0]) * Calc::uHat(ell2 * p.th[1]) * Calc::uHat(ell3 * p.th[2]);
    r *= Calc::uHat(ell1 * p.th[3]) * Calc::uHat(ell2 * p.th[4]) * Calc::uHat(ell3 * p.th[5]);
    r *= G;
    return r;
}

int t1fin_cb(unsigned ndim, size_t npts, const double *x, void *fdata, unsigned fdim, double *fval)
{
    T1FinCtx *p = (T1FinCtx *)fdata;
    if (fdim != 1 || ndim != 6) return 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(p->nthreads)
    for (long long i = 0; i < (long long)npts; i++) fval[i] = t1fin_point(*p, x + 6 * i);
    return 0;
}

int nthreads_or_default(int nt)
{
#ifdef _OPENMP
    return nt > 0 ? nt : omp_get_max_threads();
#else
    (void)nt;
    return 1;
#endif
}

} // namespace

/* ======================================================================== */
extern "C" {

int orc_hcubature_v(unsigned fdim, orc_integrand_v f, void *fdata, unsigned dim, const double *xmin,
                    const double *xmax, size_t maxEval, double reqAbsError, double reqRelError, int norm,
                    double *val, double *err)
{
    return hcubature_v_impl(fdim, f, fdata, dim, xmin, xmax, maxEval, reqAbsError, reqRelError, norm, val, err);
}
long long orc_last_neval(void) { return g_last_neval; }
long long orc_last_nregions(void) { return g_last_nregions; }

double orc_qag15(orc_f1 f, void *p, double a, double b, double epsabs, double epsrel, int limit)
{
    auto F = [&](double x) { return f(x, p); };
    return qag15(F, a, b, epsabs, epsrel, limit);
}

void orc_gauss_legendre(int n, double *x, double *w) { gauss_legendre(n, x, w); }

int orc_read_cosmology(const char *fn, double *out)
{ /* CV/cosmology.cpp:10-73; out = {h, sigma8, Omega_b, n_s, w, Omega_m, z_max} */
    std::ifstream in(fn);
    if (in.fail()) return 1;
    std::string line;
    for (int i = 0; i < 7; ++i) out[i] = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string name;
        double v;
        iss >> name >> v;
        if (name == "h") out[0] = v;
        else if (name == "sigma_8") out[1] = v;
        else if (name == "Omega_b") out[2] = v;
        else if (name == "n_s") out[3] = v;
        else if (name == "w") out[4] = v;
        else if (name == "Omega_m") out[5] = v;
        else if (name == "z_max") out[6] = v;
        else return 2;
    }
    return 0;
}

int orc_read_n_of_z(const char *fn, int n_bins, double zMax, double *nz_out)
{ /* CV/helper.cpp:31-96 */
    std::ifstream in(fn);
    if (in.fail()) return 1;
    std::vector<double> zs, ns;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double z, v;
        iss >> z >> v;
        zs.push_back(z);
        ns.push_back(v);
    }
    if (zs.size() < 2) return 2;
    int nf = (int)zs.size();
    double zmin_f = zs.front(), zmax_f = zs.back();
    double dz_f = (zmax_f - zmin_f) / (nf - 1);
    double dz = zMax / n_bins;
    std::vector<double> nz;
    for (int i = 0; i < n_bins; i++) {
        double z = i * dz;
        int ix = int((z - zmin_f) / dz_f);
        double dix = (z - zmin_f) / dz_f - ix;
        double v = 0;
        if (ix >= 0 && ix < nf - 1) v = ns[ix + 1] * dix + ns[ix] * (1 - dix);
        if (ix == nf - 1) v = ns[nf - 1];
        nz.push_back(v);
    }
    double norm = 0.0;
    for (double v : nz) norm += v;
    if (norm == 0) return 3;
    norm *= dz;
    for (int i = 0; i < n_bins; i++) nz_out[i] = nz[i] / norm;
    return 0;
}

void *orc_create(const double *p, const double *nz, int n_z, double z_max)
{ /* CV/bispectrum.cpp:73-79, 91-109 (n(z) constructor) */
    Calc *c = new Calc();
    c->cosmo.h = p[0];
    c->cosmo.sigma8 = p[1];
    c->cosmo.omb = p[2];
    c->cosmo.ns = p[3];
    c->cosmo.w = p[4];
    c->cosmo.om = p[5];
    c->cosmo.zmax = p[6];
    c->cosmo.ow = 1 - c->cosmo.om;
    c->cosmo.omc = c->cosmo.om - c->cosmo.omb;
    c->n = n_z;
    c->z_max = z_max;
    c->dz = z_max / ((double)n_z);
    c->nz_file.assign(nz, nz + n_z);
    c->set_cosmology();
    return c;
}

void *orc_create_pk(const double *p, const double *nz, int n_z, double z_max, const double *k, const double *pk,
                    int nk)
{ /* CV/bispectrum.cpp:81-89 (n(z) and given linear P(k)) */
    Calc *c = new Calc();
    c->cosmo.h = p[0];
    c->cosmo.sigma8 = p[1];
    c->cosmo.omb = p[2];
    c->cosmo.ns = p[3];
    c->cosmo.w = p[4];
    c->cosmo.om = p[5];
    c->cosmo.zmax = p[6];
    c->cosmo.ow = 1 - c->cosmo.om;
    c->cosmo.omc = c->cosmo.om - c->cosmo.omb;
    c->n = n_z;
    c->z_max = z_max;
    c->dz = z_max / ((double)n_z);
    c->nz_file.assign(nz, nz + n_z);
    c->Pk_given = true;
    for (int i = 0; i < nk; ++i) c->linearPk_given[k[i]] = pk[i];
    c->set_cosmology();
    return c;
}

void orc_destroy(void *h) { delete (Calc *)h; }
void orc_set_analytic(void *h, int on) { ((Calc *)h)->analytic = on != 0; }

void orc_get_tables(void *h, double *fK, double *g, double *D1, double *rs, double *neff, double *ncur, double *norm)
{
    Calc *c = (Calc *)h;
    for (int i = 0; i < c->n; ++i) {
        fK[i] = c->fK[i];
        g[i] = c->g[i];
        D1[i] = c->D1[i];
        rs[i] = c->rs[i];
        neff[i] = c->neff[i];
        ncur[i] = c->ncur[i];
    }
    *norm = c->norm;
}

double orc_linear_pk(void *h, double k) { return ((Calc *)h)->linear_pk(k); }
double orc_sigmam(void *h, double r, int j) { return ((Calc *)h)->sigmam(r, j); }
double orc_lgr(void *h, double z) { return ((Calc *)h)->lgr(z); }
double orc_f_K_at_z(void *h, double z) { return ((Calc *)h)->f_K_at_z(z); }
double orc_bispec(void *h, double k1, double k2, double k3, double z)
{
    Calc *c = (Calc *)h;
    int idx;
    double didx;
    c->z_index(z, &idx, &didx);
    return c->bispec(k1, k2, k3, z, idx, didx);
}
double orc_bispec_idx(void *h, double k1, double k2, double k3, double z, int idx, double didx)
{
    return ((Calc *)h)->bispec(k1, k2, k3, z, idx, didx);  // CV/bispectrum.cpp:166 (the 6-argument form)
}
double orc_integrand_bkappa(void *h, double z, double l1, double l2, double l3)
{
    return ((Calc *)h)->integrand_bkappa(z, l1, l2, l3);
}
double orc_bkappa(void *h, double l1, double l2, double l3) { return ((Calc *)h)->bkappa(l1, l2, l3); }

void orc_bkappa_batch(void *h, long long n, const double *l, double *out, int nthreads)
{
    Calc *c = (Calc *)h;
    int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long long i = 0; i < n; ++i) out[i] = c->bkappa(l[3 * i], l[3 * i + 1], l[3 * i + 2]);
}

double orc_map3_integrand(void *h, double l1, double l2, double phi, const double *th)
{
    return ((Calc *)h)->map3_integrand(l1, l2, phi, th);
}

void orc_map3_integrand_batch(void *h, long long n, const double *x, const double *th, double *out, int nthreads)
{
    Calc *c = (Calc *)h;
    int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long long i = 0; i < n; ++i) out[i] = c->map3_integrand(x[3 * i], x[3 * i + 1], x[3 * i + 2], th);
}

int orc_map3(void *h, const double *th, double reltol, int nthreads, double *val, double *err, long long *neval)
{ /* CV/apertureStatistics.cpp:321-353, INTEGRATE4D=false branch: 3-D hcubature over (l1, l2, phi) */
    Calc *c = (Calc *)h;
    double thetaMin = std::min(th[0], std::min(th[1], th[2]));
    double lMax = 10. / thetaMin;
    double lMin = 1, phiMin = 0, phiMax = 2 * M_PI;
    double vmin[3] = {lMin, lMin, phiMin}, vmax[3] = {lMax, lMax, phiMax};
    Map3Ctx m{c, {th[0], th[1], th[2]}, nthreads_or_default(nthreads), false};
    double result = 0, error = 0;
    int rc = hcubature_v_impl(1, map3_cb, &m, 3, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, &result, &error);
    *val = result / 8 / M_PI / M_PI / M_PI;
    *err = error / 8 / M_PI / M_PI / M_PI;
    if (neval) *neval = g_last_neval;
    return rc;
}

double orc_bessel_jn(int n, double x) { return bessel_jn(n, x); }
double orc_bessel_zero_jn(int n, int s) { return bessel_zero_jn(n, s); }

int orc_gamma_init(void *h, double prec_h, double prec_k_arg)
{
    ((Calc *)h)->initialize_bessel(prec_h, prec_k_arg);
    return 0;
}

int orc_gamma_integrand(void *h, int which, double z, double phi, double psi, double x1, double x2, double x3,
                        double *re, double *im)
{
    std::complex<double> t = ((Calc *)h)->gamma_integrand(which, z, phi, psi, x1, x2, x3);
    *re = t.real();
    *im = t.imag();
    return 0;
}

int orc_gamma_integrand_batch(void *h, int which, double x1, double x2, double x3, size_t npts, const double *x,
                              double *fval, int nthreads)
{ /* the integrand_v adapters of CV/gamma.cpp:338-359, 522-543 on npts points (z, phi, psi) */
    GammaCtx g{(Calc *)h, which == 0 ? 0 : 1, x1, x2, x3, nthreads_or_default(nthreads)};
    return gamma_cb(3, npts, x, &g, 2, fval);
}

int orc_gamma_center(void *h, int which, double x1, double x2, double x3, double reltol, int nthreads, int center,
                     double *re, double *im, long long *neval)
{ /* CV/gamma.cpp:27-81 (gamma0..3), 361-377 / 545-563 (cubature + Ogata); center 0 = centroid (rotation
     624-643), 1 = orthocenter (no rotation, gamma.cpp:12-13) */
    Calc *c = (Calc *)h;
    bool conj = which != 0;
    if (which == 2) { double t = x1; x1 = x2; x2 = x3; x3 = t; }           /* gamma1(x2, x3, x1) */
    else if (which == 3) { double t = x3; x3 = x2; x2 = x1; x1 = t; }      /* gamma1(x3, x1, x2) */
    GammaCtx g{c, which == 0 ? 0 : 1, x1, x2, x3, nthreads_or_default(nthreads)};
    double vmin[3] = {0, 0, 0}, vmax[3] = {c->z_max, 2 * M_PI, M_PI / 2};
    double result[2] = {0, 0}, error[2] = {0, 0};
    int rc = hcubature_v_impl(2, gamma_cb, &g, 3, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, result, error);
    std::complex<double> r(result[0], result[1]);
    r *= 27. / 8. * std::pow(c->cosmo.om, 3) * std::pow(100. / 299792., 5);
    r /= 3.;
    r /= (2 * std::pow(2 * M_PI, 3));
    if (center == 0) {
        double a1 = Calc::one_rotation_angle_otc(x2, x3, x1);
        double a2 = Calc::one_rotation_angle_otc(x3, x1, x2);
        double a3 = Calc::one_rotation_angle_otc(x1, x2, x3);
        if (conj) a1 *= -1.;
        r = r * std::exp(std::complex<double>(0, -2.) * (a1 + a2 + a3));
    }
    *re = r.real();
    *im = r.imag();
    if (neval) *neval = g_last_neval;
    return rc;
}

int orc_gamma(void *h, int which, double x1, double x2, double x3, double reltol, int nthreads, double *re,
              double *im, long long *neval)
{
    return orc_gamma_center(h, which, x1, x2, x3, reltol, nthreads, 0, re, im, neval);
}

double orc_pk_nonlinear(void *h, double k, double z) { return ((Calc *)h)->P_k_nonlinear(k, z); }
double orc_linear_pk_at_z(void *h, double k, double z) { return ((Calc *)h)->linear_pk_at_z(k, z); }
double orc_om_m_of_z(void *h, double z) { return ((Calc *)h)->om_m_of_z(z); }
double orc_om_v_of_z(void *h, double z) { return ((Calc *)h)->om_v_of_z(z); }
double orc_limber_pkappa(void *h, double ell, double z, int flags)
{
    return ((Calc *)h)->limber_integrand_power_spectrum(ell, z, !(flags & 1), (flags & 2) != 0);
}
double orc_gq96_pkappa(void *h, double ell, int flags)
{
    Calc *c = (Calc *)h;
    return c->GQ96_of_Pk(0, c->z_max, ell, !(flags & 1), (flags & 2) != 0);
}
double orc_pell(void *h, double ell, double sigma, double n_gal, int flags)
{ /* CV's Pell is nonlinear only; flags & 1 gives the linear spectrum for comparison */
    Calc *c = (Calc *)h;
    double P_shapenoise = 0.5 * sigma * sigma / n_gal;
    if (!(flags & 1)) return c->Pell(ell, sigma, n_gal, (flags & 2) != 0);
    if (ell == 0) return P_shapenoise;
    return c->GQ96_of_Pk(0, c->z_max, ell, false, (flags & 2) != 0) + P_shapenoise;
}
double orc_map2_integrand(void *h, double ell, double z, double theta, int flags, double sigma, double ngal)
{
    return ((Calc *)h)->map2_integrand(ell, z, theta, !(flags & 1), (flags & 2) != 0, sigma, ngal);
}

int orc_convergence_power_spectrum(void *h, double ell, double reltol, int flags, double *val, double *err,
                                   long long *neval)
{ /* CV/bispectrum.cpp:382-388: 1-D hcubature over z in [0, z_max] (GK15 rule); flags as orc_pell */
    Calc *c = (Calc *)h;
    PowerCtx p{c, ell, !(flags & 1), (flags & 2) != 0, 0, 1};
    double vmin[1] = {0}, vmax[1] = {c->z_max};
    int rc = hcubature_v_impl(1, pkappa_cb, &p, 1, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, val, err);
    if (neval) *neval = g_last_neval;
    return rc;
}

int orc_map2(void *h, double theta, double reltol, int flags, double sigma, double ngal_arcmin2, double *val,
             double *err, long long *neval)
{ /* GV/apertureStatistics.cu:110-136: 2-D hcubature over ell in [1, 10/theta], z in [0, z_max], / 2 pi */
    Calc *c = (Calc *)h;
    PowerCtx p{c, theta, !(flags & 1), (flags & 2) != 0, sigma, ngal_arcmin2};
    double vmin[2] = {1., 0}, vmax[2] = {10. / theta, c->z_max};
    double result = 0, error = 0;
    int rc = hcubature_v_impl(1, map2_cb, &p, 2, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, &result, &error);
    *val = result / 2. / M_PI;
    *err = error / 2. / M_PI;
    if (neval) *neval = g_last_neval;
    return rc;
}

double orc_limber_triple_pk(void *h, double l1, double l2, double l3, double z)
{
    return ((Calc *)h)->limber_integrand_triple_power_spectrum(l1, l2, l3, z);
}
double orc_cov_gauss_integrand(void *h, double l1, double l2, double phi, double z, const double *th6)
{
    return ((Calc *)h)->cov_gauss_integrand(l1, l2, phi, z, th6);
}
double orc_t1_integrand(void *h, double l1, double l2, double phi, const double *th6, double lmin, double lmax,
                        double sigma, double n_gal, int flags)
{
    return ((Calc *)h)->t1_integrand(l1, l2, phi, th6, lmin, lmax, sigma, n_gal, flags);
}

int orc_map3_cov_gauss(void *h, const double *th6, double survey_area, double reltol, int nthreads, double *val,
                       double *err, long long *neval)
{ /* CV/apertureStatistics.cpp:355-396 (INTEGRATE4D): 4-D hcubature over (l1, l2, phi in [0, pi], z),
     lMax = 10 / max(min theta_123, min theta_456), result 2 / area / (2 pi)^3 */
    Calc *c = (Calc *)h;
    double m123 = std::min({th6[0], th6[1], th6[2]}), m456 = std::min({th6[3], th6[4], th6[5]});
    double lMax = 10. / std::max(m123, m456);
    CovCtx p{c, {th6[0], th6[1], th6[2], th6[3], th6[4], th6[5]}, 1, lMax, 0, 1, 0, nthreads_or_default(nthreads)};
    double vmin[4] = {1, 1, 0, 0}, vmax[4] = {lMax, lMax, 2 * M_PI / 2., c->z_max};
    double result = 0, error = 0;
    int rc = hcubature_v_impl(1, cov_gauss_cb, &p, 4, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, &result, &error);
    *val = 2. * result / survey_area / 8 / M_PI / M_PI / M_PI;
    *err = 2. * error / survey_area / 8 / M_PI / M_PI / M_PI;
    if (neval) *neval = g_last_neval;
    return rc;
}

int orc_t1(void *h, const double *th6, double area, double lmin, double sigma, double n_gal, int flags,
           double reltol, int nthreads, double *val, double *err, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:1065-1106, infinite survey (type 2): 3-D hcubature over
     [lMin, lMax]^2 x [0, pi], lMax = 10 / max(min theta_123, min theta_456), x 2 / area / (2 pi)^3 */
    Calc *c = (Calc *)h;
    double m123 = std::min({th6[0], th6[1], th6[2]}), m456 = std::min({th6[3], th6[4], th6[5]});
    double lMax = 10. / std::max(m123, m456);
    CovCtx p{c, {th6[0], th6[1], th6[2], th6[3], th6[4], th6[5]}, lmin, lMax, sigma, n_gal, flags,
             nthreads_or_default(nthreads)};
    double vmin[3] = {lmin, lmin, 0}, vmax[3] = {lMax, lMax, M_PI};
    double result = 0, error = 0;
    int rc = hcubature_v_impl(1, t1_cb, &p, 3, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, &result, &error);
    result *= 2 / area / std::pow(2 * M_PI, 3);
    error *= 2 / area / std::pow(2 * M_PI, 3);
    *val = result;
    *err = error;
    if (neval) *neval = g_last_neval;
    return rc;
}

int orc_t1_total(void *h, const double *th123, const double *th456, double area, double lmin, double sigma,
                 double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:272-326: T1 summed over the distinct permutations of theta_456 */
    const double th0 = th123[0], th1 = th123[1], th2 = th123[2], th3 = th456[0], th4 = th456[1], th5 = th456[2];
    long long ntot = 0;
    int rc = 0;
    auto T1 = [&](double a, double b, double c) {
        double th6[6] = {th0, th1, th2, a, b, c}, v = 0, e = 0;
        long long n = 0;
        rc |= orc_t1(h, th6, area, lmin, sigma, n_gal, flags, reltol, nthreads, &v, &e, &n);
        ntot += n;
        return v;
    };
    double result;
    if (th3 == th4) {
        if (th4 == th5) {
            result = 6 * T1(th3, th3, th3);
        } else {
            result = 2 * T1(th3, th3, th5);
            result += 2 * T1(th3, th5, th3);
            result += 2 * T1(th5, th3, th3);
        }
    } else if (th3 == th5) {
        result = 2 * T1(th3, th4, th3);
        result += 2 * T1(th3, th3, th4);
        result += 2 * T1(th4, th3, th3);
    } else if (th4 == th5) {
        result = 2 * T1(th3, th4, th4);
        result += 2 * T1(th4, th3, th4);
        result += 2 * T1(th4, th4, th3);
    } else {
        result = T1(th3, th4, th5);
        result += T1(th3, th5, th4);
        result += T1(th4, th3, th5);
        result += T1(th4, th5, th3);
        result += T1(th5, th3, th4);
        result += T1(th5, th4, th3);
    }
    *val = result;
    if (neval) *neval = ntot;
    return rc;
}

int orc_t1_finite(void *h, const double *th6, int geometry, double theta_max, double theta_max2, double lmin,
                  double sigma, double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:1065-1106, types 0, 1, 3: 6-D hcubature over
     [-lMax, 1.02 lMax]^6, lMax = 10 / max(min theta_123, min theta_456), result / (2 pi)^6 */
    Calc *c = (Calc *)h;
    *val = 0;
    if (neval) *neval = 0;
    if (geometry != 0 && geometry != 1 && geometry != 3) return 1;
    double m123 = std::min({th6[0], th6[1], th6[2]}), m456 = std::min({th6[3], th6[4], th6[5]});
    double lMax = 10. / std::max(m123, m456);
    double vMax = lMax;
    T1FinCtx p{c, geometry, {th6[0], th6[1], th6[2], th6[3], th6[4], th6[5]}, theta_max, theta_max2, lmin, lMax,
               sigma, n_gal, flags, nthreads_or_default(nthreads)};
    double vmin[6] = {-vMax, -vMax, -lMax, -lMax, -lMax, -lMax};
    double vmax[6] = {1.02 * vMax, 1.02 * vMax, 1.02 * lMax, 1.02 * lMax, 1.02 * lMax, 1.02 * lMax};
    double result = 0, error = 0;
    int rc = hcubature_v_impl(1, t1fin_cb, &p, 6, vmin, vmax, 0, 0, reltol, ORC_ERROR_L1, &result, &error);
    result /= std::pow(2 * M_PI, 6);
    *val = result;
    if (neval) *neval = g_last_neval;
    return rc;
}

int orc_t1_total_finite(void *h, const double *th123, const double *th456, int geometry, double theta_max,
                        double theta_max2, double lmin, double sigma, double n_gal, int flags, double reltol,
                        int nthreads, double *val, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:272-326 with the finite-survey T1 */
    const double th3 = th456[0], th4 = th456[1], th5 = th456[2];
    long long ntot = 0;
    int rc = 0;
    auto T1 = [&](double a, double b, double d) {
        double th6[6] = {th123[0], th123[1], th123[2], a, b, d}, v = 0;
        long long n = 0;
        rc |= orc_t1_finite(h, th6, geometry, theta_max, theta_max2, lmin, sigma, n_gal, flags, reltol, nthreads,
                            &v, &n);
        ntot += n;
        return v;
    };
    double result;
    if (th3 == th4) {
        if (th4 == th5) {
            result = 6 * T1(th3, th3, th3);
        } else {
            result = 2 * T1(th3, th3, th5);
            result += 2 * T1(th3, th5, th3);
            result += 2 * T1(th5, th3, th3);
        }
    } else if (th3 == th5) {
        result = 2 * T1(th3, th4, th3);
        result += 2 * T1(th3, th3, th4);
        result += 2 * T1(th4, th3, th3);
    } else if (th4 == th5) {
        result = 2 * T1(th3, th4, th4);
        result += 2 * T1(th4, th3, th4);
        result += 2 * T1(th4, th4, th3);
    } else {
        result = T1(th3, th4, th5);
        result += T1(th3, th5, th4);
        result += T1(th4, th3, th5);
        result += T1(th4, th5, th3);
        result += T1(th5, th3, th4);
        result += T1(th5, th4, th3);
    }
    *val = result;
    if (neval) *neval = ntot;
    return rc;
}

int orc_t2(void *h, const double *th6, int geometry, double theta_max, double theta_max2, double lmin,
           double sigma, double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:1108-1163: A1 A2 B, each a hcubature at reltol; the second A
     integral and the circle's B reuse the first integral's domain [lMin, 10 / min(theta_1..theta_5)];
     the square's and the rectangle's B run over [-lMax, lMax]^2, lMax = 10 / min(theta_3, theta_4) */
    Calc *c = (Calc *)h;
    *val = 0;
    if (neval) *neval = 0;
    if (geometry == 2) return 0;
    if (geometry != 0 && geometry != 1 && geometry != 3) return 1;
    const int nt = nthreads_or_default(nthreads);
    double lMax = 10. / std::min({th6[0], th6[1], th6[2], th6[3], th6[4]});
    double vmin1[1] = {lmin}, vmax1[1] = {lMax};
    long long ntot = 0;
    int rc = 0;
    auto run = [&](int kind, double ta, double tb, unsigned dim, const double *lo, const double *hi) {
        T2Ctx p{c, kind, ta, tb, theta_max, theta_max2, lmin, sigma, n_gal, flags, nt};
        double v = 0, e = 0;
        rc |= hcubature_v_impl(1, t2_cb, &p, dim, lo, hi, 0, 0, reltol, ORC_ERROR_L1, &v, &e);
        ntot += g_last_neval;
        return v;
    };
    double A1 = run(0, th6[0], th6[1], 1, vmin1, vmax1);
    A1 /= 2 * M_PI;
    double A2 = run(0, th6[4], th6[5], 1, vmin1, vmax1);
    A2 /= 2 * M_PI;
    double B;
    if (geometry == 0) {
        B = run(1, th6[2], th6[3], 1, vmin1, vmax1);
        B /= 2 * M_PI;
    } else {
        double l3 = 10. / std::min(th6[2], th6[3]);
        double vmin2[2] = {-l3, -l3}, vmax2[2] = {l3, l3};
        B = run(geometry == 1 ? 2 : 3, th6[2], th6[3], 2, vmin2, vmax2);
        B /= std::pow(2 * M_PI, 2);
    }
    *val = A1 * A2 * B;
    if (neval) *neval = ntot;
    return rc;
}

int orc_t2_total(void *h, const double *th123, const double *th456, int geometry, double theta_max,
                 double theta_max2, double lmin, double sigma, double n_gal, int flags, double reltol, int nthreads,
                 double *val, long long *neval)
{ /* GV/apertureStatisticsCovariance.cu:328-538 as the product of the two sides' pairing lists */
    *val = 0;
    if (neval) *neval = 0;
    if (geometry == 2) return 0;
    int ma[3], mb[3], ia[3][3], ib[3][3];
    const int na = t2_side_123(th123, ma, ia), nb = t2_side_456(th456, mb, ib);
    double result = 0;
    long long ntot = 0;
    int rc = 0;
    bool first = true;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            double th6[6] = {th123[ia[a][0]], th123[ia[a][1]], th123[ia[a][2]],
                             th456[ib[b][0]], th456[ib[b][1]], th456[ib[b][2]]};
            double v = 0;
            long long n = 0;
            rc |= orc_t2(h, th6, geometry, theta_max, theta_max2, lmin, sigma, n_gal, flags, reltol, nthreads, &v, &n);
            const int m = ma[a] * mb[b];
            const double tv = m == 1 ? v : m * v;
            result = first ? tv : result + tv;
            first = false;
            ntot += n;
        }
    *val = result;
    if (neval) *neval = ntot;
    return rc;
}

void orc_ggg_analytic(double r1, double r2, double r3, double *re, double *im)
{ /* CV/gamma.cpp:83-94, 567-586: closed-form Gamma0 of the analytic test bispectrum */
    static const double wts[9] = {0, 1.0e+8, 1.0e+6, 1.0e+4, 1.0e+2, 1.0e+0, 5.0e-3, 1.0e-5, 1.0e-6};
    static const double avs[9] = {3.0e+3, 1.0e+4, 3.0e+4, 1.0e+5, 3.0e+5, 1.0e+6, 3.0e+6, 1.0e+7, 3.0e+7};
    std::complex<double> tot(0, 0);
    for (int k = 0; k < 9; ++k) {
        double a = avs[k];
        std::complex<double> x(r2, 0);
        double height = (r1 + r2 + r3) * (r2 + r3 - r1) * (r1 - r2 + r3) * (r1 + r2 - r3);
        height = std::sqrt(height) / (2. * r2);
        double rest = std::sqrt(r1 * r1 - height * height);
        std::complex<double> y(-rest, -height);
        std::complex<double> z = -x - y;
        const std::complex<double> I(0, 1);
        double p1 = std::atan2(x.imag(), x.real()), p2 = std::atan2(y.imag(), y.real()),
               p3 = std::atan2(z.imag(), z.real());
        std::complex<double> pref = std::pow((x + 2. * y) * (x - y) * (y + 2. * x), 2);
        double expon = std::norm(x) + std::norm(y) + x.real() * y.real() + x.imag() * y.imag();
        std::complex<double> phase = std::exp(-2. * I * (p1 + p2 + p3));
        tot += wts[k] * (pref * std::exp(-2 * a * expon) * phase);
    }
    *re = tot.real();
    *im = tot.imag();
}

} /* extern "C" */
