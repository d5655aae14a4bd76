// tp_tables.hip -- FP64 construction of the per-cosmology redshift tables on gfx950.
//
// Follows classed_version/bispectrum.cpp:111-164 (set_cosmology) and the helpers
// it calls: f_K_at_z (:951-965, adaptive GK15 quadrature of 1/E), the lensing
// efficiency trapezoid (:126-144, then g[0] = 1), the sigma8 normalisation
// (:146-147), lgr (:843-912, RK4 in ln a with step doubling), calc_r_sigma
// (:676-707, bracket + bisection) and sigmam (:709-767, nested doubling
// trapezoid in ln k).  The trapezoid sums inside sigmam are split over the 64
// lanes of a wavefront and combined by an xor-butterfly, so every lane holds the
// same sum and the bisection's control flow stays wave-uniform.
//
// Kernels (one stream, in order):
//   k_cosmo_const   1 thread / cosmology   : derived constants, EH constants
//   k_norm          1 wave / cosmology     : norm = sigma8 / sigma_TH(8)
//   k_fK            1 thread / (cosmo,bin) : f_K(z_i)
//   k_g             1 thread / (cosmo,bin) : g(z_i)
//   k_lgr0          1 thread / cosmology   : lgr(0)
//   k_d1            1 thread / (cosmo,bin) : D1 = lgr(z_i) / lgr(0)
//   k_growth        1 wave / (cosmo,bin)   : r_sigma, n_eff, ncur
//   k_nodes         1 thread / (cosmo,node): 96-node Limber tables (FP64 + FP32 copy)
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "tp_device.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double window(double x, int i)
{
    // classed_version/bispectrum.cpp:769-783
    if (i == 0) return 3.0 / (x * x * x) * (sin(x) - x * cos(x));
    if (i == 1) return exp(-0.5 * x * x);
    if (i == 2) return x * exp(-0.5 * x * x);
    return x * x * (1 - x * x) * exp(-x * x);
}

__device__ __forceinline__ double sigma_term(const TpCosmoConst &C, double k, double r, int j, double norm)
{
    double w = window(k * r, j);
    double kk = k * k * k * tp_pk_lin(C, k, norm);
    return (j < 3) ? kk * w * w : kk * w;
}

__device__ __forceinline__ double wave_sum(double v)
{
    // xor butterfly over the 64 lanes: every lane ends with the same bits
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// sigmam with lane-split trapezoid sums; all 64 lanes call it with equal arguments.
__device__ double sigmam_wave(const TpCosmoConst &C, double r, int j, double norm, int lane)
{
    const double eps = 1.e-4;
    double k1 = 2. * kPi / r, k2 = 2. * kPi / r;
    double xx = 0, xxpp = -1.0;
    for (int outer = 0; outer < 30; ++outer) {  // bounded; converges after a few range doublings
        k1 = k1 / 10.0;
        k2 = k2 * 2.0;
        double a = log(k1), b = log(k2);
        double xxp = -1.0;
        int n = 2;
        for (int inner = 0; inner < 22; ++inner) {  // n <= 2^23
            n = n * 2;
            double hh = (b - a) / (double)n;
            double part = 0.;
            for (int i = 1 + lane; i < n; i += 64) part += sigma_term(C, exp(a + hh * i), r, j, norm);
            xx = wave_sum(part);
            xx += 0.5 * (sigma_term(C, k1, r, j, norm) + sigma_term(C, k2, r, j, norm));
            xx *= hh;
            if (fabs((xx - xxp) / xx) < eps) break;
            xxp = xx;
        }
        if (fabs((xx - xxpp) / xx) < eps) break;
        xxpp = xx;
    }
    if (j < 3) return sqrt(xx / (2.0 * kPi * kPi));
    return xx / (2.0 * kPi * kPi);
}

__device__ double lgr_func1(const TpCosmoConst &C, double la, const double y[2])
{
    // classed_version/bispectrum.cpp:895-907 (j == 1)
    double a = exp(la);
    double aw = pow(a, -3. * C.w);
    double g = -0.5 * (5. * C.om + (5. - 3 * C.w) * C.ow * aw) * y[1] - 1.5 * (1. - C.w) * C.ow * aw * y[0];
    return g / (C.om + C.ow * aw);
}

__device__ double lgr(const TpCosmoConst &C, double z)
{
    // classed_version/bispectrum.cpp:843-893: RK4 in ln a, n = 20, 40, ... until rel 1e-5
    double a = 1. / (1. + z), a0 = 1. / 1100.;
    double yp = -1., y0 = 0;
    int n = 10;
    for (int it = 0; it < 16; ++it) {  // n <= 655360 steps
        n *= 2;
        double h = (log(a) - log(a0)) / n;
        double x = log(a0);
        double y[2] = {1., 0.};
        for (int i = 0; i < n; i++) {
            double k1[2], k2[2], k3[2], k4[2], yt[2];
            k1[0] = h * y[1];
            k1[1] = h * lgr_func1(C, x, y);
            yt[0] = y[0] + 0.5 * k1[0];
            yt[1] = y[1] + 0.5 * k1[1];
            k2[0] = h * yt[1];
            k2[1] = h * lgr_func1(C, x + 0.5 * h, yt);
            yt[0] = y[0] + 0.5 * k2[0];
            yt[1] = y[1] + 0.5 * k2[1];
            k3[0] = h * yt[1];
            k3[1] = h * lgr_func1(C, x + 0.5 * h, yt);
            yt[0] = y[0] + k3[0];
            yt[1] = y[1] + k3[1];
            k4[0] = h * yt[1];
            k4[1] = h * lgr_func1(C, x + h, yt);
            y[0] += (k1[0] + k4[0]) / 6. + (k2[0] + k3[0]) / 3.;
            y[1] += (k1[1] + k4[1]) / 6. + (k2[1] + k3[1]) / 3.;
            x += h;
        }
        y0 = y[0];
        if (fabs(y0 / yp - 1.) < 0.1 * 1.e-4) break;
        yp = y0;
    }
    return a * y0;
}

// ------------------------------------------------- adaptive GK15 (QUADPACK QAG)
__constant__ double c_xgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                                0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                                0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
                                0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
__constant__ double c_wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                               0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
__constant__ double c_wgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                0.204432940075298892414161999234649, 0.209482141084727828012999174891714};

__device__ __forceinline__ double Einv(const TpCosmoConst &C, double z)
{
    return 1. / sqrt(C.om * (1 + z) * (1 + z) * (1 + z) + C.ow);  // classed_version/bispectrum.cpp:967-975
}

__device__ void qk15_Einv(const TpCosmoConst &C, double a, double b, double *res, double *abserr, double *resabs,
                          double *resasc)
{
    double center = 0.5 * (a + b), half = 0.5 * (b - a), abs_half = fabs(half);
    double fc = Einv(C, center);
    double rg = fc * c_wg[3], rk = fc * c_wgk[7], ra = fabs(rk);
    double fv1[8], fv2[8];
    for (int j = 0; j < 3; ++j) {
        int jt = 2 * j + 1;
        double ab = half * c_xgk[jt];
        double f1 = Einv(C, center - ab), f2 = Einv(C, center + ab);
        fv1[jt] = f1;
        fv2[jt] = f2;
        rg += c_wg[j] * (f1 + f2);
        rk += c_wgk[jt] * (f1 + f2);
        ra += c_wgk[jt] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 4; ++j) {
        int jt = 2 * j;
        double ab = half * c_xgk[jt];
        double f1 = Einv(C, center - ab), f2 = Einv(C, center + ab);
        fv1[jt] = f1;
        fv2[jt] = f2;
        rk += c_wgk[jt] * (f1 + f2);
        ra += c_wgk[jt] * (fabs(f1) + fabs(f2));
    }
    double mean = rk * 0.5, rasc = c_wgk[7] * fabs(fc - mean);
    for (int j = 0; j < 7; ++j) rasc += c_wgk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double err = fabs((rk - rg) * half);
    rk *= half;
    ra *= abs_half;
    rasc *= abs_half;
    if (rasc != 0 && err != 0) {
        double scale = pow(200 * err / rasc, 1.5);
        err = (scale < 1) ? rasc * scale : rasc;
    }
    if (ra > DBL_MIN / (50 * DBL_EPSILON)) {
        double me = 50 * DBL_EPSILON * ra;
        if (me > err) err = me;
    }
    *res = rk;
    *abserr = err;
    *resabs = ra;
    *resasc = rasc;
}

// gsl_integration_qag(..., epsabs=0, epsrel=1e-6, key=GK15) equivalent; capacity 48 intervals
__device__ double qag_Einv(const TpCosmoConst &C, double a, double b, int *overflow)
{
    const double epsrel = 1e-6;
    double r0, e0, ra0, rasc0;
    qk15_Einv(C, a, b, &r0, &e0, &ra0, &rasc0);
    double tol = epsrel * fabs(r0);
    double round_off = 50 * DBL_EPSILON * ra0;
    if (e0 <= round_off && e0 > tol) return r0;
    if ((e0 <= tol && e0 != rasc0) || e0 == 0.0) return r0;
    const int CAP = 48;
    double ia[CAP], ib[CAP], ir[CAP], ie[CAP];
    int n = 1;
    ia[0] = a;
    ib[0] = b;
    ir[0] = r0;
    ie[0] = e0;
    double area = r0, errsum = e0;
    int ro1 = 0, ro2 = 0, iteration = 1, etype = 0;
    do {
        int im = 0;
        for (int i = 1; i < n; ++i)
            if (ie[i] > ie[im]) im = i;
        double a1 = ia[im], b1 = 0.5 * (ia[im] + ib[im]), a2 = b1, b2 = ib[im];
        double r1, e1, x1, s1, r2, e2, x2, s2;
        qk15_Einv(C, a1, b1, &r1, &e1, &x1, &s1);
        qk15_Einv(C, a2, b2, &r2, &e2, &x2, &s2);
        double area12 = r1 + r2, error12 = e1 + e2;
        errsum += error12 - ie[im];
        area += area12 - ir[im];
        if (s1 != e1 && s2 != e2) {
            double delta = ir[im] - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * ie[im]) ro1++;
            if (iteration >= 10 && error12 > ie[im]) ro2++;
        }
        tol = epsrel * fabs(area);
        if (errsum > tol) {
            if (ro1 >= 6 || ro2 >= 20) etype = 2;
            double tmp = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
            if (fabs(a1) <= tmp && fabs(b2) <= tmp) etype = 3;
        }
        if (n >= CAP) {
            *overflow = 1;
            break;
        }
        ia[im] = a1;
        ib[im] = b1;
        ir[im] = r1;
        ie[im] = e1;
        ia[n] = a2;
        ib[n] = b2;
        ir[n] = r2;
        ie[n] = e2;
        ++n;
        iteration++;
    } while (iteration < 10000 && !etype && errsum > tol);
    double s = 0;
    for (int i = 0; i < n; ++i) s += ir[i];
    return s;
}

// ----------------------------------------------------------------- kernels
__global__ void k_cosmo_const(int ncosmo, const tp_cosmology *p, double z_max, int n_z, TpCosmoConst *out)
{
    int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= ncosmo) return;
    TpCosmoConst C;
    C.h = p[ic].h;
    C.sigma8 = p[ic].sigma8;
    C.omb = p[ic].omb;
    C.ns = p[ic].ns;
    C.w = p[ic].w;
    C.om = p[ic].om;
    C.zmax = p[ic].zmax;
    C.ow = 1 - C.om;             // classed_version/cosmology.cpp:71-72
    C.omc = C.om - C.omb;
    // EH constants, classed_version/bispectrum.cpp:816-830
    double fc = C.omc / C.om, fb = C.omb / C.om;
    double theta = 2.728 / 2.7;
    double pc = 0.25 * (5.0 - sqrt(1.0 + 24.0 * fc));
    double omh2 = C.om * C.h * C.h, ombh2 = C.omb * C.h * C.h;
    double zeq = 2.5e+4 * omh2 / pow(theta, 4);
    double b1 = 0.313 * pow(omh2, -0.419) * (1.0 + 0.607 * pow(omh2, 0.674));
    double b2 = 0.238 * pow(omh2, 0.223);
    double zd = 1291.0 * pow(omh2, 0.251) / (1.0 + 0.659 * pow(omh2, 0.828)) * (1.0 + b1 * pow(ombh2, b2));
    double yd = (1.0 + zeq) / (1.0 + zd);
    double sh = 44.5 * log(9.83 / (omh2)) / sqrt(1.0 + 10.0 * pow(ombh2, 0.75));
    double alnu = fc * (5.0 - 2.0 * pc) / 5.0 * (1.0 - 0.553 * fb + 0.126 * fb * fb * fb) * pow(1.0 + yd, -pc) *
                  (1.0 + 0.5 * pc * (1.0 + 1.0 / (7.0 * (3.0 - 4.0 * pc))) / (1.0 + yd));
    C.eh_fb = fb;
    C.eh_sqrt_alnu = sqrt(alnu);
    C.eh_sh = sh;
    C.eh_omh2 = omh2;
    C.eh_theta2 = theta * theta;
    C.eh_cL = 1.84 * sqrt(alnu) / (1.0 - 0.949 * fb);
    C.limber_pref = 27. / 8. * C.om * C.om * C.om * pow(100. / 299792., 5);
    C.norm = 1.0;
    C.n_z = n_z;
    C.pad = 0;
    (void)z_max;
    out[ic] = C;
}

// Points every cosmology's constants at its tabulated P(k) (one table of nk entries per cosmology).
__global__ void k_set_pk(int ncosmo, TpCosmoConst *cc, const double *k, const double *p, const int *idx, int nk,
                         int ng, const double *lnk0, const double *inv_dlnk)
{
    int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= ncosmo) return;
    TpPkTable t;
    t.k = k + (size_t)ic * nk;
    t.p = p + (size_t)ic * nk;
    t.idx = idx + (size_t)ic * ng;
    t.nk = nk;
    t.ng = ng;
    t.lnk0 = lnk0[ic];
    t.inv_dlnk = inv_dlnk[ic];
    cc[ic].tab = t;
}

__global__ void __launch_bounds__(64) k_norm(int ncosmo, TpCosmoConst *cc)
{
    int ic = blockIdx.x;
    if (ic >= ncosmo) return;
    int lane = threadIdx.x;
    TpCosmoConst C = cc[ic];
    double s = sigmam_wave(C, 8., 0, 1.0, lane);   // classed_version/bispectrum.cpp:146-147
    if (lane == 0) cc[ic].norm = C.sigma8 / s;
}

__global__ void k_fK(int ncosmo, int n_z, double dz, const TpCosmoConst *cc, double *tab, int *err)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncosmo * n_z) return;
    int ic = t / n_z, i = t % n_z;
    TpCosmoConst C = cc[ic];
    int ovf = 0;
    double z = i * dz;
    double v = (z == 0.0) ? 0.0 : qag_Einv(C, 0.0, z, &ovf);
    tab[(size_t)ic * 6 * n_z + i] = 2997.92 * v;    // c/H0 [Mpc/h], classed_version/bispectrum.hpp:149
    if (ovf) atomicOr(err, 1);
}

__device__ __forceinline__ double n_of_z_dev(const double *nzf, int n_z, double z_max, double z)
{
    // classed_version/bispectrum.cpp:922-933 (file branch); index n_z carries weight 0
    if (z <= 0 || z >= z_max) return 0;
    double d = z / z_max * (n_z);
    int ix = (int)d;
    d = d - ix;
    double hi = (ix + 1 < n_z) ? nzf[ix + 1] : 0.0;
    return nzf[ix] * (1. - d) + hi * d;
}

__global__ void k_g(int ncosmo, int n_z, double dz, double z_max, const double *nz, double *tab)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncosmo * n_z) return;
    int ic = t / n_z, i = t % n_z;
    const double *fK = tab + (size_t)ic * 6 * n_z;
    const double *nzf = nz + (size_t)ic * n_z;
    // classed_version/bispectrum.cpp:126-144
    double g = 0;
    for (int j = i; j < n_z; j++) {
        double zj = j * dz;
        double term = n_of_z_dev(nzf, n_z, z_max, zj) * (fK[j] - fK[i]) / fK[j];
        g += (j == i || j == n_z - 1) ? term / 2 : term;
    }
    g = g * dz;
    if (i == 0) g = 1.;
    tab[(size_t)ic * 6 * n_z + n_z + i] = g;
}

// lgr(0) of each cosmology, one thread per cosmology (every z-bin's D1 divides by it)
__global__ void k_lgr0(int ncosmo, const TpCosmoConst *cc, double *lgr0)
{
    const int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic < ncosmo) lgr0[ic] = lgr(cc[ic], 0.);
}

// D1 = lgr(z) / lgr(0) (classed_version/bispectrum.cpp:157), one thread per (cosmology, z-bin): the growth
// ODE is sequential, so each lane integrates its own bin instead of a whole wave repeating one
__global__ void k_d1(int ncosmo, int n_z, double dz, const TpCosmoConst *cc, const double *lgr0, double *tab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncosmo * n_z) return;
    const int ic = t / n_z, i = t % n_z;
    tab[(size_t)ic * 6 * n_z + 2 * n_z + i] = lgr(cc[ic], i * dz) / lgr0[ic];
}

// r_sigma, n_eff and ncur of one (cosmology, z-bin) per wave, from the bin's D1
__global__ void __launch_bounds__(64) k_growth(int ncosmo, int n_z, double dz, const TpCosmoConst *cc, double *tab)
{
    int ic = blockIdx.x / n_z, i = blockIdx.x % n_z;
    if (ic >= ncosmo) return;
    int lane = threadIdx.x;
    TpCosmoConst C = cc[ic];
    const double eps = 1.e-4;
    const double D1 = tab[(size_t)ic * 6 * n_z + 2 * n_z + i];
    // calc_r_sigma, classed_version/bispectrum.cpp:676-707
    double k, k1 = 1., k2 = 1.;
    for (int it = 0; it < 200; ++it) {
        if (D1 * sigmam_wave(C, 1. / k1, 1, C.norm, lane) < 1.) break;
        k1 *= 0.5;
    }
    for (int it = 0; it < 200; ++it) {
        if (D1 * sigmam_wave(C, 1. / k2, 1, C.norm, lane) > 1.) break;
        k2 *= 2.;
    }
    for (int it = 0; it < 400; ++it) {
        k = 0.5 * (k1 + k2);
        double s = D1 * sigmam_wave(C, 1. / k, 1, C.norm, lane);
        if (s < 1.)
            k1 = k;
        else if (s > 1.)
            k2 = k;
        if (s == 1. || fabs(k2 / k1 - 1.) < eps * 0.1) break;
    }
    double rs = 1. / k;
    double s2 = sigmam_wave(C, rs, 2, C.norm, lane);
    double s3 = sigmam_wave(C, rs, 3, C.norm, lane);
    double d1 = -2. * (D1 * s2) * (D1 * s2);
    if (lane == 0) {
        double *T = tab + (size_t)ic * 6 * n_z;
        T[3 * n_z + i] = rs;
        T[4 * n_z + i] = -3. + 2. * (D1 * s2) * (D1 * s2);   // classed_version/bispectrum.cpp:160
        T[5 * n_z + i] = d1 * d1 + 4. * s3 * D1 * D1;          // :161
    }
}

// 96-point Gauss-Legendre half nodes and weights, filled by the host (own Newton solve)
__constant__ double c_gl_x[48];
__constant__ double c_gl_w[48];

__device__ TpNode64 make_node(const TpCosmoConst &C, const double *T, int n_z, double z, double wgt)
{
    // interpolation coefficients, classed_version/bispectrum.cpp:506-513
    double d = z / C.zmax * (n_z - 1);
    int idx = (int)d;
    d = d - idx;
    if (idx == n_z - 1) {
        idx = n_z - 2;
        d = 1.;
    }
    double fK = T[idx] * (1 - d) + T[idx + 1] * d;
    double g = T[n_z + idx] * (1 - d) + T[n_z + idx + 1] * d;
    double D1 = T[2 * n_z + idx] * (1 - d) + T[2 * n_z + idx + 1] * d;
    double rs = T[3 * n_z + idx] * (1 - d) + T[3 * n_z + idx + 1] * d;
    double ne = T[4 * n_z + idx] * (1 - d) + T[4 * n_z + idx + 1] * d;
    double E = sqrt(C.om * (1 + z) * (1 + z) * (1 + z) + C.ow);
    double gz = g * (1. + z);
    const double ln10 = 2.302585092994045684;
    TpNode64 N;
    N.wz = (z == 0.0) ? 0.0 : wgt * C.limber_pref * gz * gz * gz / fK / E;
    N.ks = rs / fK;
    N.inv_rs = 1. / rs;
    double L = log10(D1 * C.sigma8);
    N.lnA0 = ln10 * (-2.167 - 2.944 * L - 1.106 * L * L - 2.865 * L * L * L);
    N.c1 = pow(10., 0.182 + 0.57 * ne);
    N.m031ln10 = -0.310 * ln10;
    N.bn = pow(10., -3.428 - 2.681 * L + 1.624 * L * L - 0.095 * L * L * L);
    N.cn = pow(10., 0.159 - 1.107 * ne);
    N.a0ln = ln10 * (-4.348 - 3.006 * ne - 0.5745 * ne * ne);
    N.a1ln = ln10 * pow(10., -0.9 + 0.2 * ne);
    N.acap = 1. - (2. / 3.) * C.ns;
    N.b0ln = ln10 * (-1.731 - 2.845 * ne - 1.4995 * ne * ne - 0.2811 * ne * ne * ne);
    N.b1ln = ln10 * 0.007;
    N.fn = pow(10., -10.533 - 16.838 * ne - 9.3048 * ne * ne - 1.8263 * ne * ne * ne);
    N.gn = pow(10., 2.787 + 2.405 * ne + 0.4577 * ne * ne);
    N.hn = pow(10., -1.118 - 0.394 * ne);
    N.mn = pow(10., -2.605 - 2.434 * L + 5.71 * L * L);
    N.nn = pow(10., -4.468 - 3.08 * L + 1.035 * L * L);
    N.mun = pow(10., 15.312 + 22.977 * ne + 10.9579 * ne * ne + 1.6586 * ne * ne * ne);
    N.nun = pow(10., 1.347 + 1.246 * ne + 0.4525 * ne * ne);
    N.pn = pow(10., 0.071 - 0.433 * ne);
    N.en = pow(10., -0.632 + 0.646 * ne);
    N.D1sq = D1 * D1;
    double a = 1. / (1. + z);
    double omz = C.om / (C.om + C.ow * pow(a, -3. * C.w));
    N.dn = pow(10., -0.483 + 0.892 * L - 0.086 * omz);
    N.lnkoff = log(2997.9 / rs);
    // P(k) = h^3 2 pi^2 / kk^3 * norm^2 (kk 2997.9/h)^(3+ns) T^2 with kk = q h / r_sigma
    double hr = C.h / rs;
    N.pk_pref = C.h * C.h * C.h * 2.0 * kPi * kPi * C.norm * C.norm / (hr * hr * hr);
    N.lnks = log(N.ks);
    return N;
}

__device__ void to_f32(const TpNode64 &N, TpNode32 *M)
{
    const double *src = &N.wz;
    float *dst = &M->wz;
    for (int k = 0; k < (int)(sizeof(TpNode64) / sizeof(double)); ++k) dst[k] = (float)src[k];
}

// 8-point Gauss-Legendre nodes/weights on [-1, 1] for the surrogate's baseline Limber sum
__constant__ double c_gl8_x[8];
__constant__ double c_gl8_w[8];

__global__ void k_nodes(int ncosmo, int n_z, const TpCosmoConst *cc, const double *tab, TpNode64 *n64,
                        TpNode32 *n32, TpEH64 *eh64, TpEH32 *eh32, TpNode32 *base32)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncosmo * (TP_NNODES + TP_NBASE)) return;
    int ic = t / (TP_NNODES + TP_NBASE), nd = t % (TP_NNODES + TP_NBASE);
    const TpCosmoConst C = cc[ic];
    const double *T = tab + (size_t)ic * 6 * n_z;
    double cx = 0.5 * C.zmax, dx = 0.5 * C.zmax;
    if (nd >= TP_NNODES) {  // baseline node
        int j = nd - TP_NNODES;
        TpNode64 N = make_node(C, T, n_z, cx + dx * c_gl8_x[j], c_gl8_w[j] * dx);
        to_f32(N, &base32[(size_t)ic * TP_NBASE + j]);
        return;
    }
    // node ordering: pair i = nd/2, sign - for even nd, + for odd (GQ96_of_bdelta, CV:1010-1020)
    int i = nd >> 1;
    double z = (nd & 1) ? cx + dx * c_gl_x[i] : cx - dx * c_gl_x[i];
    TpNode64 N = make_node(C, T, n_z, z, c_gl_w[i] * dx);
    int o = ic * TP_NNODES + nd;
    n64[o] = N;
    to_f32(N, &n32[o]);
    if (nd == 0) {
        TpEH64 e;
        e.h = C.h;
        e.sqrt_alnu = C.eh_sqrt_alnu;
        e.one_m_sqrt_alnu = 1.0 - C.eh_sqrt_alnu;
        e.sh043 = 0.43 * C.eh_sh;
        e.omh2 = C.eh_omh2;
        e.theta2 = C.eh_theta2;
        e.cL = C.eh_cL;
        e.three_ns = 3. + C.ns;
        e.e_const = 2.718281828;
        e.tab = C.tab;
        eh64[ic] = e;
        TpEH32 f;
        f.h = (float)e.h;
        f.sqrt_alnu = (float)e.sqrt_alnu;
        f.one_m_sqrt_alnu = (float)e.one_m_sqrt_alnu;
        f.sh043 = (float)e.sh043;
        f.omh2 = (float)e.omh2;
        f.theta2 = (float)e.theta2;
        f.cL = (float)e.cL;
        f.three_ns = (float)e.three_ns;
        f.e_const = (float)e.e_const;
        f.tab = C.tab;
        eh32[ic] = f;
    }
}

}  // namespace

void tp_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0;
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
            if (fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[n - 1 - i] = w[i];
    }
}

// Host: upload GL-96 half nodes once per process and device.
static int upload_gl96(int device)
{
    static bool done[64] = {false};
    if (device >= 0 && device < 64 && done[device]) return 0;
    double x[96], w[96], hx[48], hw[48];
    tp_gauss_legendre(96, x, w);
    for (int i = 0; i < 48; ++i) {
        hx[i] = x[48 + i];
        hw[i] = w[48 + i];
    }
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl_x), hx, sizeof(hx)));
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl_w), hw, sizeof(hw)));
    double x8[8], w8[8];
    tp_gauss_legendre(8, x8, w8);
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl8_x), x8, sizeof(x8)));
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gl8_w), w8, sizeof(w8)));
    if (device >= 0 && device < 64) done[device] = true;
    return 0;
}

static int cosmo_build(tp_ctx *ctx, int ncosmo, const tp_cosmology *params, const double *nz, int n_z,
                       size_t nz_stride, double z_max, const double *pk_k, const double *pk_p, int nk,
                       size_t pk_stride, tp_cosmo **out)
{
    if (!ctx || !params || !nz || !out || ncosmo <= 0 || n_z < 2 || n_z > TP_NZ_MAX || !(z_max > 0))
        return tp_set_error(TP_EINVAL, "tp_cosmo_build: bad arguments");
    // tabulated P(k): strictly increasing positive k, finite P (the reference keeps a std::map)
    std::vector<double> hk, hp, hl0, hinv;
    std::vector<int> hidx;
    const int ng = nk > 0 ? std::max(64, 4 * nk) : 0;
    if (nk) {
        if (nk < 2 || !pk_k || !pk_p) return tp_set_error(TP_EINVAL, "tp_cosmo_build_pk: need at least 2 P(k) entries");
        hk.resize((size_t)ncosmo * nk);
        hp.resize((size_t)ncosmo * nk);
        hidx.resize((size_t)ncosmo * ng);
        hl0.resize(ncosmo);
        hinv.resize(ncosmo);
        for (int i = 0; i < ncosmo; ++i) {
            const double *ks = pk_k + pk_stride * i, *ps = pk_p + pk_stride * i;
            for (int j = 0; j < nk; ++j) {
                if (!(ks[j] > 0) || !std::isfinite(ks[j]) || !std::isfinite(ps[j]) || (j && !(ks[j] > ks[j - 1])))
                    return tp_set_error(TP_EINVAL, "tp_cosmo_build_pk: k must be positive and strictly increasing, "
                                                   "P finite (cosmology " + std::to_string(i) + ")");
                hk[(size_t)i * nk + j] = ks[j];
                hp[(size_t)i * nk + j] = ps[j];
            }
            const double l0 = std::log(ks[0]), l1 = std::log(ks[nk - 1]);
            const double inv = ng / (l1 - l0);
            hl0[i] = l0;
            hinv[i] = inv;
            int j = 0;
            for (int c = 0; c < ng; ++c) {  // interval at half a cell before the cell start
                const double kc = std::exp(l0 + (c - 0.5) / inv);
                while (j < nk - 2 && ks[j + 1] <= kc) ++j;
                hidx[(size_t)i * ng + c] = j;
            }
        }
    }
    for (int i = 0; i < ncosmo; ++i) {
        const tp_cosmology &p = params[i];
        if (!(p.om > 0) || !(p.h > 0) || !(p.sigma8 > 0) || !(p.omb >= 0) || p.omb >= p.om)
            return tp_set_error(TP_EINVAL, "tp_cosmo_build: unphysical cosmology " + std::to_string(i));
    }
    TP_HIP(hipSetDevice(ctx->device));
    int rc = upload_gl96(ctx->device);
    if (rc) return rc;
    tp_cosmo *c = new tp_cosmo();
    c->device = ctx->device;
    c->ncosmo = ncosmo;
    c->n_z = n_z;
    c->z_max = z_max;
    hipStream_t s = ctx->stream;
    auto fail = [&](int code) {
        tp_cosmo_destroy(c);
        return code;
    };
#define TPC(call)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(tp_set_error(TP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)
    // one device allocation for the whole build (separate allocations and frees cost ~0.1 ms each, a
    // noticeable share of a one-cosmology build); each array starts on a 256-byte boundary
    const size_t nkt = (size_t)ncosmo * (nk ? nk : 0), ngt = (size_t)ncosmo * (nk ? ng : 0);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t o_const = take(sizeof(TpCosmoConst) * ncosmo);
    const size_t o_tab = take(sizeof(double) * 6 * n_z * (size_t)ncosmo);
    const size_t o_nz = take(sizeof(double) * n_z * (size_t)ncosmo);
    const size_t o_node64 = take(sizeof(TpNode64) * TP_NNODES * (size_t)ncosmo);
    const size_t o_node32 = take(sizeof(TpNode32) * TP_NNODES * (size_t)ncosmo);
    const size_t o_eh64 = take(sizeof(TpEH64) * (size_t)ncosmo);
    const size_t o_eh32 = take(sizeof(TpEH32) * (size_t)ncosmo);
    const size_t o_base32 = take(sizeof(TpNode32) * TP_NBASE * (size_t)ncosmo);
    const size_t o_err = take(sizeof(int));
    const size_t o_params = take(sizeof(tp_cosmology) * ncosmo + sizeof(double) * ncosmo);
    const size_t o_pk = nk ? take((2 * nkt + 2 * (size_t)ncosmo) * sizeof(double) + ngt * sizeof(int)) : 0;
    TPC(hipMalloc(&c->d_arena, off));
    char *arena = (char *)c->d_arena;
    c->d_const = (TpCosmoConst *)(arena + o_const);
    c->d_tab = (double *)(arena + o_tab);
    c->d_nz = (double *)(arena + o_nz);
    c->d_node64 = (TpNode64 *)(arena + o_node64);
    c->d_node32 = (TpNode32 *)(arena + o_node32);
    c->d_eh64 = (TpEH64 *)(arena + o_eh64);
    c->d_eh32 = (TpEH32 *)(arena + o_eh32);
    c->d_base32 = (TpNode32 *)(arena + o_base32);
    c->d_err = (int *)(arena + o_err);
    TPC(hipMemsetAsync(c->d_err, 0, sizeof(int), s));
    double *d_pkk = nullptr, *d_pkp = nullptr, *d_l0 = nullptr, *d_inv = nullptr;
    int *d_idx = nullptr;
    if (nk) {
        c->d_pk = arena + o_pk;
        d_pkk = (double *)c->d_pk;
        d_pkp = d_pkk + nkt;
        d_l0 = d_pkp + nkt;
        d_inv = d_l0 + ncosmo;
        d_idx = (int *)(d_inv + ncosmo);
        c->pk_nk = nk;
        TPC(hipMemcpy(d_pkk, hk.data(), nkt * sizeof(double), hipMemcpyHostToDevice));
        TPC(hipMemcpy(d_pkp, hp.data(), nkt * sizeof(double), hipMemcpyHostToDevice));
        TPC(hipMemcpy(d_l0, hl0.data(), ncosmo * sizeof(double), hipMemcpyHostToDevice));
        TPC(hipMemcpy(d_inv, hinv.data(), ncosmo * sizeof(double), hipMemcpyHostToDevice));
        TPC(hipMemcpy(d_idx, hidx.data(), ngt * sizeof(int), hipMemcpyHostToDevice));
    }
    tp_cosmology *d_params = (tp_cosmology *)(arena + o_params);
    double *d_lgr0 = (double *)(d_params + ncosmo);  // lgr(0) per cosmology
    TPC(hipMemcpyAsync(d_params, params, sizeof(tp_cosmology) * ncosmo, hipMemcpyHostToDevice, s));
    // n(z) of every cosmology in one copy (the stream is synchronised before h_nz goes out of scope)
    std::vector<double> h_nz((size_t)ncosmo * n_z);
    for (int i = 0; i < ncosmo; ++i)
        std::copy(nz + nz_stride * i, nz + nz_stride * i + n_z, h_nz.begin() + (size_t)i * n_z);
    TPC(hipMemcpyAsync(c->d_nz, h_nz.data(), sizeof(double) * h_nz.size(), hipMemcpyHostToDevice, s));
    double dz = z_max / ((double)n_z);   // classed_version/bispectrum.cpp:98
    int nt = ncosmo * n_z;
    hipLaunchKernelGGL(k_cosmo_const, dim3((ncosmo + 63) / 64), dim3(64), 0, s, ncosmo, d_params, z_max, n_z,
                       c->d_const);
    if (nk)
        hipLaunchKernelGGL(k_set_pk, dim3((ncosmo + 63) / 64), dim3(64), 0, s, ncosmo, c->d_const, d_pkk, d_pkp,
                           d_idx, nk, ng, d_l0, d_inv);
    hipLaunchKernelGGL(k_norm, dim3(ncosmo), dim3(64), 0, s, ncosmo, c->d_const);
    hipLaunchKernelGGL(k_fK, dim3((nt + 127) / 128), dim3(128), 0, s, ncosmo, n_z, dz, c->d_const, c->d_tab,
                       c->d_err);
    hipLaunchKernelGGL(k_g, dim3((nt + 127) / 128), dim3(128), 0, s, ncosmo, n_z, dz, z_max, c->d_nz, c->d_tab);
    hipLaunchKernelGGL(k_lgr0, dim3((ncosmo + 63) / 64), dim3(64), 0, s, ncosmo, c->d_const, d_lgr0);
    hipLaunchKernelGGL(k_d1, dim3((nt + 63) / 64), dim3(64), 0, s, ncosmo, n_z, dz, c->d_const, d_lgr0, c->d_tab);
    hipLaunchKernelGGL(k_growth, dim3(nt), dim3(64), 0, s, ncosmo, n_z, dz, c->d_const, c->d_tab);
    int nn = ncosmo * (TP_NNODES + TP_NBASE);
    hipLaunchKernelGGL(k_nodes, dim3((nn + 127) / 128), dim3(128), 0, s, ncosmo, n_z, c->d_const, c->d_tab,
                       c->d_node64, c->d_node32, c->d_eh64, c->d_eh32, c->d_base32);
    TPC(hipGetLastError());
    int herr = 0;
    TPC(hipMemcpyAsync(&herr, c->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    c->h_const.resize(ncosmo);
    TPC(hipMemcpyAsync(c->h_const.data(), c->d_const, sizeof(TpCosmoConst) * ncosmo, hipMemcpyDeviceToHost, s));
    TPC(hipStreamSynchronize(s));
    if (herr) return fail(tp_set_error(TP_ECAPACITY, "tp_cosmo_build: f_K quadrature exceeded its interval capacity"));
    int prc = tp_build_pnodes(ctx, c);
    if (prc) return fail(prc);
    *out = c;
    return 0;
#undef TPC
}

extern "C" int tp_cosmo_build(tp_ctx *ctx, int ncosmo, const tp_cosmology *params, const double *nz, int n_z,
                              size_t nz_stride, double z_max, tp_cosmo **out)
{
    return cosmo_build(ctx, ncosmo, params, nz, n_z, nz_stride, z_max, nullptr, nullptr, 0, 0, out);
}

extern "C" int tp_cosmo_build_pk(tp_ctx *ctx, int ncosmo, const tp_cosmology *params, const double *nz, int n_z,
                                 size_t nz_stride, double z_max, const double *k_h, const double *p_k, int nk,
                                 size_t pk_stride, tp_cosmo **out)
{
    if (nk < 2) return tp_set_error(TP_EINVAL, "tp_cosmo_build_pk: need at least 2 P(k) entries");
    return cosmo_build(ctx, ncosmo, params, nz, n_z, nz_stride, z_max, k_h, p_k, nk, pk_stride, out);
}

extern "C" int tp_cosmo_destroy(tp_cosmo *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    (void)hipFree(c->d_pnode);
    (void)hipFree(c->d_arena);  // every other array lives in it
    delete c;
    return 0;
}

extern "C" int tp_cosmo_count(const tp_cosmo *c, int *ncosmo, int *n_z)
{
    if (!c) return tp_set_error(TP_EINVAL, "tp_cosmo_count: NULL");
    if (ncosmo) *ncosmo = c->ncosmo;
    if (n_z) *n_z = c->n_z;
    return 0;
}

extern "C" int tp_cosmo_tables(const tp_cosmo *c, int icosmo, double *f_K, double *g, double *D1, double *r_sigma,
                               double *n_eff, double *ncur, double *norm)
{
    if (!c || icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_cosmo_tables: bad index");
    TP_HIP(hipSetDevice(c->device));
    std::vector<double> t(6 * (size_t)c->n_z);
    TP_HIP(hipMemcpy(t.data(), c->d_tab + (size_t)icosmo * 6 * c->n_z, sizeof(double) * t.size(),
                     hipMemcpyDeviceToHost));
    double *dst[6] = {f_K, g, D1, r_sigma, n_eff, ncur};
    for (int k = 0; k < 6; ++k)
        if (dst[k])
            for (int i = 0; i < c->n_z; ++i) dst[k][i] = t[(size_t)k * c->n_z + i];
    if (norm) *norm = c->h_const[icosmo].norm;
    return 0;
}
