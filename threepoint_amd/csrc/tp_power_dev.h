// tp_power_dev.h -- per-point functions of the convergence power spectrum and <Map^2>.
//
// Templates on the arithmetic type T: double in the kernels of tp_power.hip, a counting type in the
// operation census (tp_census.cpp).  tp_dbl(x) gives the value of x as a double for table indexing.
#pragma once
#include "tp_device.h"

__host__ __device__ __forceinline__ double tp_dbl(double x) { return x; }

// Table values at redshift z (classed_version/bispectrum.cpp:506-513 + compute_coefficients :914-920).
template <class T>
struct TpZAtT {
    T fK, g, D1, rs, ne, nc;
};

template <class T>
__host__ __device__ __forceinline__ void tp_z_at(const double *Tab, int n_z, double zmax, T z, TpZAtT<T> *o)
{
    double d = tp_dbl(z) / zmax * (n_z - 1);
    int idx;
    if (!(d >= 0)) {  // below the table (or NaN): its first node; the cubature domains start at 0
        idx = 0;
        d = 0;
    } else if (d >= n_z - 1) {  // z = z_max: the reference's (n-2, 1.0); beyond it: clamped the same way
        idx = n_z - 2;
        d = 1.;
    } else {
        idx = (int)d;
        d -= idx;
    }
    const T dd = d, e = T(1.0) - dd;
    o->fK = Tab[idx] * e + Tab[idx + 1] * dd;
    o->g = Tab[n_z + idx] * e + Tab[n_z + idx + 1] * dd;
    o->D1 = Tab[2 * n_z + idx] * e + Tab[2 * n_z + idx + 1] * dd;
    o->rs = Tab[3 * n_z + idx] * e + Tab[3 * n_z + idx + 1] * dd;
    o->ne = Tab[4 * n_z + idx] * e + Tab[4 * n_z + idx + 1] * dd;
    o->nc = Tab[5 * n_z + idx] * e + Tab[5 * n_z + idx + 1] * dd;
}

// Revised Halofit (Takahashi et al. 2012), classed_version/bispectrum.cpp:288-367, split into its
// z-only coefficients (tp_halofit_z) and the k-dependent part (tp_halofit_k), so that fixed z-nodes
// can keep the former.  The reference's flatness check always passes: ow = 1 - om by construction.
template <class T>
struct TpHaloZT {
    T f1, f2, f3, a, b, c, gam, alpha, beta, xnu;
    T lf3c;  // ln(f3 c): the k-part takes (f3 c y)^(3 - gam) as exp((3 - gam)(ln(f3 c) + ln y))
};

template <class T>
__host__ __device__ void tp_halofit_z(const TpCosmoConst &C, const TpZAtT<T> &Z, T z, TpHaloZT<T> *h)
{
    const T n_eff = Z.ne, nsqr = n_eff * n_eff, nc = Z.nc;
    // om_m_of_z, om_v_of_z (bispectrum.cpp:1022-1032): the flat form, without w
    const T aa = 1. / (1 + z);
    const T den = C.om + aa * (aa * aa * C.ow + (1 - C.om - C.ow));
    const T om_m = C.om / den;
    const T om_v = C.ow * (aa * aa * aa) / den;
    const T f1a = pow(om_m, -0.0732), f2a = pow(om_m, -0.1423), f3a = pow(om_m, 0.0725);
    const T f1b = pow(om_m, -0.0307), f2b = pow(om_m, -0.0585), f3b = pow(om_m, 0.0743);
    const T frac = om_v / (1. - om_m);
    h->f1 = frac * f1b + (1 - frac) * f1a;
    h->f2 = frac * f2b + (1 - frac) * f2a;
    h->f3 = frac * f3b + (1 - frac) * f3a;
    T a = 1.5222 + 2.8553 * n_eff + 2.3706 * nsqr + 0.9903 * n_eff * nsqr + 0.2250 * nsqr * nsqr - 0.6038 * nc +
          0.1749 * om_v * (1.0 + C.w);
    h->a = pow(10.0, a);
    h->b = pow(10.0, -0.5642 + 0.5864 * n_eff + 0.5716 * nsqr - 1.5474 * nc + 0.2279 * om_v * (1.0 + C.w));
    h->c = pow(10.0, 0.3698 + 2.0404 * n_eff + 0.8161 * nsqr + 0.5869 * nc);
    h->gam = 0.1971 - 0.0843 * n_eff + 0.8460 * nc;
    h->alpha = fabs(6.0835 + 1.3373 * n_eff - 0.1959 * nsqr - 5.5274 * nc);
    h->beta = 2.0379 - 0.7354 * n_eff + 0.3157 * nsqr + 1.2490 * n_eff * nsqr + 0.3980 * nsqr * nsqr - 0.1682 * nc;
    h->xnu = pow(10.0, 5.2105 + 3.6902 * n_eff);
    h->lf3c = log(h->f3 * h->c);
}

// The k-dependent part in log space: every power of k is exp(e ln k), with ln k supplied by the caller
// (shared by the nodes of a fixed-node sum) -- 7 exp and 3 log per evaluation instead of 6 pow, an exp
// and a log; the rounding differs from the pow forms by ~|e ln k| ulp (< 1e-14 relative).
// Eisenstein & Hu P_lin (tp_pk_eh): (k h 2997.9 / h)^(3 + n_s) = exp((3 + n_s)(ln k + ln 2997.9)).
template <class T>
__host__ __device__ T tp_pk_eh_log(const TpCosmoConst &C, T k, T lk, double norm)
{
    const double ln29979 = 8.005667322535853;  // ln 2997.9
    const T kh = k * C.h;
    const T x = 0.43 * kh * C.eh_sh;
    const T x2 = x * x;
    const T geff = C.eh_omh2 * (C.eh_sqrt_alnu + (1.0 - C.eh_sqrt_alnu) / (1.0 + x2 * x2));
    const T qeff = kh / geff * C.eh_theta2;
    const T L = log(2.718281828 + C.eh_cL * qeff);
    const T Cq = 14.4 + 325.0 / (1.0 + 60.5 * exp(1.11 * log(qeff)));
    const T r = L / (L + Cq * qeff * qeff);
    const T delk = norm * norm * exp((3. + C.ns) * (lk + ln29979)) * r * r;
    const T pk = 2.0 * tpd::kPi * tpd::kPi / (kh * kh * kh) * delk;
    return C.h * C.h * C.h * pk;
}

template <class T>
__host__ __device__ __forceinline__ T tp_pk_lin_log(const TpCosmoConst &C, T k, T lk)
{
    if constexpr (std::is_same<T, double>::value)
        if (C.tab.nk) return tp_pk_table(C.tab, k);
    return tp_pk_eh_log<T>(C, k, lk, C.norm);
}

template <class T>
__host__ __device__ T tp_halofit_k(const TpCosmoConst &C, T D1, T rs, T lrs, const TpHaloZT<T> &h, T k, T lk)
{
    const double twopi2 = 2 * tpd::kPi * tpd::kPi;
    const T plin = tp_pk_lin_log<T>(C, k, lk) * D1 * D1 * k * k * k / twopi2;
    const T y = k * rs, ly = lk + lrs, ysqr = y * y;
    T ph = h.a * exp(h.f1 * 3 * ly) / (1 + h.b * exp(h.f2 * ly) + exp((3 - h.gam) * (h.lf3c + ly)));
    ph = ph / (1 + h.xnu / ysqr);
    const T pq = plin * exp(h.beta * log1p(plin)) / (1 + plin * h.alpha) * exp(-y / 4.0 - ysqr / 8.0);
    return twopi2 * (pq + ph) / (k * k * k);
}

template <class T>
__host__ __device__ T tp_pnl(const TpCosmoConst &C, const TpZAtT<T> &Z, T k, T z)
{
    TpHaloZT<T> h;
    tp_halofit_z<T>(C, Z, z, &h);
    return tp_halofit_k<T>(C, Z.D1, Z.rs, log(Z.rs), h, k, log(k));
}

// P(k, z): Halofit, or D1^2 P_lin with TP_PK_LINEAR (linear_pk_at_z, bispectrum.cpp:369-380).
template <class T>
__host__ __device__ __forceinline__ T tp_pk_at(const TpCosmoConst &C, const TpZAtT<T> &Z, T k, T z, int flags)
{
    return (flags & TP_PK_LINEAR) ? Z.D1 * Z.D1 * tp_pk_lin_log<T>(C, k, log(k)) : tp_pnl<T>(C, Z, k, z);
}

// The ratio fitted to the T17 simulations, cuda_version/bispectrum.cu:971-979.
// (1 + c1 k^-a1)^a1 / (1 + c2 k^-a2)^a3 evaluated as exp(a1 ln(1 + c1 e^(-a1 ln k)) - a3 ln(1 + c2 e^(-a2 ln k))).
template <class T>
__host__ __device__ __forceinline__ T tp_t17_factor(T lk)
{
    const double c1 = 9.5171e-4, c2 = 5.1543e-3, a1 = 1.3063, a2 = 1.1475, a3 = 0.62793;
    return exp(a1 * log1p(c1 * exp(-1. * a1 * lk)) - a3 * log1p(c2 * exp(-1. * a2 * lk)));
}

// The Limber prefactor 9/4 (H0/c)^3 om^2 (1+z)^2 g^2 / E(z) (classed_version/bispectrum.cpp:574-577).
template <class T>
__host__ __device__ __forceinline__ T tp_limber_pref(const TpCosmoConst &C, T z, T g)
{
    const double H = 100. / 299792.;  // H0/c [h/Mpc], classed_version/bispectrum.hpp:148
    const T zp = 1. + z;
    const T E = sqrt(C.om * (zp * zp * zp) + C.ow);
    return 9. / 4. * H * H * H * C.om * C.om * zp * zp * g * g / E;
}

// limber_integrand_power_spectrum, classed_version/bispectrum.cpp:543-577.
template <class T>
__host__ __device__ T tp_limber_pk(const TpCosmoConst &C, const double *Tab, int n_z, T ell, T z, int flags)
{
    if (tp_dbl(z) < 1e-5) return T(0.0);
    TpZAtT<T> Z;
    tp_z_at<T>(Tab, n_z, C.zmax, z, &Z);
    const T pref = tp_limber_pref<T>(C, z, Z.g);
    const T k = ell / Z.fK, lk = log(k);
    T P;
    if (flags & TP_PK_LINEAR) {
        P = Z.D1 * Z.D1 * tp_pk_lin_log<T>(C, k, lk);
    } else {
        TpHaloZT<T> h;
        tp_halofit_z<T>(C, Z, z, &h);
        P = tp_halofit_k<T>(C, Z.D1, Z.rs, log(Z.rs), h, k, lk);
    }
    return (flags & TP_PK_T17) ? tp_t17_factor<T>(lk) * pref * P : T(1.0) * pref * P;
}

// One z-node of a fixed-node Limber sum of P: every z-only factor of tp_limber_pk, evaluated once
// (same expressions, same bits).
struct TpPNode {
    double fK, D1, rs, pref;
    double lfK, lrs;  // ln f_K, ln r_sigma: ln k = ln ell - ln f_K, ln y = ln k + ln r_sigma
    TpHaloZT<double> h;
    int zero;  // z < 1e-5: the integrand is 0 there
    int pad;
};

__host__ __device__ __forceinline__ void tp_pnode_make(const TpCosmoConst &C, const double *Tab, int n_z, double z,
                                                       TpPNode *o)
{
    TpZAtT<double> Z;
    tp_z_at<double>(Tab, n_z, C.zmax, z, &Z);
    o->fK = Z.fK;
    o->D1 = Z.D1;
    o->rs = Z.rs;
    o->lfK = log(Z.fK);
    o->lrs = log(Z.rs);
    o->pref = tp_limber_pref<double>(C, z, Z.g);
    tp_halofit_z<double>(C, Z, z, &o->h);
    o->zero = z < 1e-5 ? 1 : 0;
    o->pad = 0;
}

// FP32 form of a node record, for the FP32 Pell of T1: the node's constants folded so that a node value
// is a few native exp/log.  With Delta^2_lin(k) = D1^2 P_lin k^3 / (2 pi^2) = r_EH(k)^2 exp((3 + n_s) ln k
// + ld), ld = ln(D1^2 norm^2) + (3 + n_s) ln 2997.9, and the Limber factor pref 2 pi^2 / k^3 =
// exp(lp - 3 ln k), lp = ln(2 pi^2 pref), the node value is (pq + ph) exp(lp - 3 ln k) [x T17].
// The Eisenstein & Hu fit only (a tabulated P(k) keeps the FP64 form).
struct TpPNode32 {
    float inv_fK, lfK, lrs, rs;
    float lp, ld, ns3;                       // see above; ns3 = 3 + n_s
    float h, sh043, omh2, sqa, theta2, cL;   // E&H: h, 0.43 s h, om h^2, sqrt(alpha_nu), theta^2, c_L
    float f1x3, f2, a, b, gam3, lf3c, alpha, beta, xnu;  // Halofit z-part (f1x3 = 3 f1, gam3 = 3 - gam)
    int zero;
};

__host__ __device__ inline void tp_pnode32_make(const TpCosmoConst &C, const TpPNode &N, TpPNode32 *o)
{
    const double ln29979 = 8.005667322535853;  // ln 2997.9
    o->inv_fK = (float)(1.0 / N.fK);
    o->lfK = (float)N.lfK;
    o->lrs = (float)N.lrs;
    o->rs = (float)N.rs;
    o->lp = (float)(N.pref > 0 ? log(2 * tpd::kPi * tpd::kPi * N.pref) : -HUGE_VAL);
    o->ld = (float)(log(N.D1 * N.D1 * C.norm * C.norm) + (3. + C.ns) * ln29979);
    o->ns3 = (float)(3. + C.ns);
    o->h = (float)C.h;
    o->sh043 = (float)(0.43 * C.eh_sh);
    o->omh2 = (float)C.eh_omh2;
    o->sqa = (float)C.eh_sqrt_alnu;
    o->theta2 = (float)C.eh_theta2;
    o->cL = (float)C.eh_cL;
    o->f1x3 = (float)(3 * N.h.f1);
    o->f2 = (float)N.h.f2;
    o->a = (float)N.h.a;
    o->b = (float)N.h.b;
    o->gam3 = (float)(3 - N.h.gam);
    o->lf3c = (float)N.h.lf3c;
    o->alpha = (float)N.h.alpha;
    o->beta = (float)N.h.beta;
    o->xnu = (float)N.h.xnu;
    o->zero = N.zero;
}

// The FP32 node value (tp_limber_pk_node's arithmetic with the folded constants; native exp, log and
// reciprocal).  lell = ln ell.
__device__ __forceinline__ float tp_limber_pk_node32(const TpPNode32 &N, float ell, float lell, int flags)
{
    if (N.zero) return 0.f;
    const float k = ell * N.inv_fK, lk = lell - N.lfK;
    // Eisenstein & Hu transfer function r (tp_pk_eh_log)
    const float kh = k * N.h, x = N.sh043 * kh, x2 = x * x;
    const float geff = N.omh2 * (N.sqa + (1.f - N.sqa) * __builtin_amdgcn_rcpf(1.f + x2 * x2));
    const float qeff = kh * __builtin_amdgcn_rcpf(geff) * N.theta2;
    const float L = __logf(2.718281828f + N.cL * qeff);
    const float Cq = 14.4f + 325.f * __builtin_amdgcn_rcpf(1.f + 60.5f * __expf(1.11f * __logf(qeff)));
    const float r = L * __builtin_amdgcn_rcpf(L + Cq * qeff * qeff);
    const float plin = r * r * __expf(N.ns3 * lk + N.ld);  // Delta^2_lin
    float P = plin;
    if (!(flags & TP_PK_LINEAR)) {  // revised Halofit k-part (tp_halofit_k)
        const float y = k * N.rs, ly = lk + N.lrs, ysqr = y * y;
        float ph = N.a * __expf(N.f1x3 * ly) *
                   __builtin_amdgcn_rcpf(1.f + N.b * __expf(N.f2 * ly) + __expf(N.gam3 * (N.lf3c + ly)));
        ph = ph * __builtin_amdgcn_rcpf(1.f + N.xnu * __builtin_amdgcn_rcpf(ysqr));
        const float pq = plin * __expf(N.beta * __logf(1.f + plin)) * __builtin_amdgcn_rcpf(1.f + plin * N.alpha) *
                         __expf(-0.25f * y - 0.125f * ysqr);
        P = pq + ph;
    }
    float v = P * __expf(N.lp - 3.f * lk);
    if (flags & TP_PK_T17) {  // tp_t17_factor
        const float c1 = 9.5171e-4f, c2 = 5.1543e-3f, a1 = 1.3063f, a2 = 1.1475f, a3 = 0.62793f;
        v *= __expf(a1 * __logf(1.f + c1 * __expf(-a1 * lk)) - a3 * __logf(1.f + c2 * __expf(-a2 * lk)));
    }
    return v;
}

// tp_limber_pk at a precomputed node
// (lell = ln ell, computed once per multipole)
__host__ __device__ __forceinline__ double tp_limber_pk_node(const TpCosmoConst &C, const TpPNode &N, double ell,
                                                             double lell, int flags)
{
    if (N.zero) return 0.0;
    const double k = ell / N.fK, lk = lell - N.lfK;
    const double P = (flags & TP_PK_LINEAR) ? N.D1 * N.D1 * tp_pk_lin_log<double>(C, k, lk)
                                            : tp_halofit_k<double>(C, N.D1, N.rs, N.lrs, N.h, k, lk);
    return (flags & TP_PK_T17) ? tp_t17_factor<double>(lk) * N.pref * P : 1.0 * N.pref * P;
}

// <Map^2> integrand at (ell, z) for one source bin, cuda_version/apertureStatistics.cu:29-57:
// theta the aperture radius, sigma/ngal the shape noise (sigma = 0: none; ngal in arcmin^-2).
template <class T>
__host__ __device__ T tp_map2_point(const TpCosmoConst &C, const double *Tab, int n_z, T ell, T z, double theta,
                                    int flags, double sigma, double ngal)
{
    const double arcmin = 2.9088820866e-4;  // classed_version/helper.cpp:139
    const T eta = ell * theta, t = 0.5 * eta * eta;
    const T u = t * exp(-t);  // uHat, classed_version/apertureStatistics.cpp:10-14
    T v = ell * (u * u) * tp_limber_pk<T>(C, Tab, n_z, ell, z, flags);
    if (flags & TP_PK_T17) {
        const T r = ell / (1.6 * 8192.);
        v = v / (1 + r * r);
    }
    if (sigma != 0) {
        const T us = u * sigma;
        v = v + ell * (us * us) * 0.5 / ngal * (arcmin * arcmin) / C.zmax;
    }
    return v;
}
