// tp_gamma.hip -- Gamma_i natural components: Ogata (2005) Hankel quadrature
// inside a 3-D adaptive cubature over (z, phi, psi).
//
// Follows classed_version/gamma.cpp: node tables (:133-188), geometry
// (:121-219), the J6 / J2 Ogata sums (:291-318, :462-488), the Gamma0 / Gamma1
// integrands (:320-336, :490-520), gamma0..3 with their prefactors (:27-81) and
// the orthocenter -> centroid rotation (:624-643) of the centroid convention (the
// orthocenter convention returns the components unrotated, :12-13).
//
// Device layout: one 256-thread workgroup per cubature point.  The z-dependent
// BiHalofit coefficients are computed once per point into LDS; each thread then
// walks the Ogata nodes k = 1 + tid, 1 + tid + 256, ... for the three terms, and
// a wave-shuffle + LDS reduction combines the three partial sums.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>

#include "tp_cubature.h"
#include "tp_device.h"

namespace {

constexpr int kGammaBlock = 256;
constexpr size_t kGammaLaunchPts = (size_t)1 << 22;  // points per launch: 2^22 blocks x 256 threads = 2^30 work-items
// the FP32 form's miss mask holds 3 bits per k slot of a thread in 128 bits: 42 slots of 256 k
constexpr int kGammaMaxK32 = 42 * kGammaBlock + 1;

struct TpGammaDesc {
    double x[3];
    int which;   // 0: Gamma0 integrand, 1: Gamma1 integrand (already permuted for Gamma2/3)
    int icosmo;
};

// z-node record at an arbitrary redshift (the per-point analogue of k_nodes), without
// the GL weight and Limber prefactor: integrand_bkappa = wz * B (classed_version/
// bispectrum.cpp:471-541).  Returns false at z == 0 (the reference's zero guard).
// The record from the table interval idx and the fraction d within it (compute_coefficients and
// g/f_K_interpolated, bispectrum.cpp:914-920, 977-987); z enters E(z), (1 + z) and Omega_m(z).
__device__ void z_record_at(const TpCosmoConst &C, const double *T, int n_z, double z, int idx, double d, TpNode64 *N)
{
    double fK = T[idx] * (1 - d) + T[idx + 1] * d;
    double g = T[n_z + idx] * (1 - d) + T[n_z + idx + 1] * d;
    double D1 = T[2 * n_z + idx] * (1 - d) + T[2 * n_z + idx + 1] * d;
    double rs = T[3 * n_z + idx] * (1 - d) + T[3 * n_z + idx + 1] * d;
    double ne = T[4 * n_z + idx] * (1 - d) + T[4 * n_z + idx + 1] * d;
    double E = sqrt(C.om * (1 + z) * (1 + z) * (1 + z) + C.ow);
    double gz = g * (1. + z);
    const double ln10 = 2.302585092994045684;
    N->wz = gz * gz * gz / fK / E;
    N->ks = rs / fK;
    N->inv_rs = 1. / rs;
    double L = log10(D1 * C.sigma8);
    N->lnA0 = ln10 * (-2.167 - 2.944 * L - 1.106 * L * L - 2.865 * L * L * L);
    N->c1 = pow(10., 0.182 + 0.57 * ne);
    N->m031ln10 = -0.310 * ln10;
    N->bn = pow(10., -3.428 - 2.681 * L + 1.624 * L * L - 0.095 * L * L * L);
    N->cn = pow(10., 0.159 - 1.107 * ne);
    N->a0ln = ln10 * (-4.348 - 3.006 * ne - 0.5745 * ne * ne);
    N->a1ln = ln10 * pow(10., -0.9 + 0.2 * ne);
    N->acap = 1. - (2. / 3.) * C.ns;
    N->b0ln = ln10 * (-1.731 - 2.845 * ne - 1.4995 * ne * ne - 0.2811 * ne * ne * ne);
    N->b1ln = ln10 * 0.007;
    N->fn = pow(10., -10.533 - 16.838 * ne - 9.3048 * ne * ne - 1.8263 * ne * ne * ne);
    N->gn = pow(10., 2.787 + 2.405 * ne + 0.4577 * ne * ne);
    N->hn = pow(10., -1.118 - 0.394 * ne);
    N->mn = pow(10., -2.605 - 2.434 * L + 5.71 * L * L);
    N->nn = pow(10., -4.468 - 3.08 * L + 1.035 * L * L);
    N->mun = pow(10., 15.312 + 22.977 * ne + 10.9579 * ne * ne + 1.6586 * ne * ne * ne);
    N->nun = pow(10., 1.347 + 1.246 * ne + 0.4525 * ne * ne);
    N->pn = pow(10., 0.071 - 0.433 * ne);
    N->en = pow(10., -0.632 + 0.646 * ne);
    N->D1sq = D1 * D1;
    double a = 1. / (1. + z);
    double omz = C.om / (C.om + C.ow * pow(a, -3. * C.w));
    N->dn = pow(10., -0.483 + 0.892 * L - 0.086 * omz);
    N->lnkoff = log(2997.9 / rs);
    double hr = C.h / rs;
    N->pk_pref = C.h * C.h * C.h * 2.0 * tpd::kPi * tpd::kPi * C.norm * C.norm / (hr * hr * hr);
    N->lnks = log(N->ks);
}

// The table interval of z as the reference's integrand_bkappa finds it (:506-513): idx = int(z / z_max *
// (n - 1)), the last interval with fraction 1 at z = z_max.  A z outside [0, z_max], which the cubature
// domains never produce, reads the nearest interval (extrapolating linearly) instead of past the tables.
__device__ void z_interval(const TpCosmoConst &C, int n_z, double z, int *idx_out, double *d_out)
{
    double d = z / C.zmax * (n_z - 1);
    int idx = (int)d;
    d = d - idx;
    if (idx > n_z - 2) {  // z >= z_max (at z = z_max: the last interval, fraction 1)
        d += idx - (n_z - 2);
        idx = n_z - 2;
    }
    if (idx < 0) {
        idx = 0;
        d = 0;
    }
    *idx_out = idx;
    *d_out = d;
}
__device__ bool z_record(const TpCosmoConst &C, const double *T, int n_z, double z, TpNode64 *N)
{
    if (z == 0) return false;
    int idx;
    double d;
    z_interval(C, n_z, z, &idx, &d);
    z_record_at(C, T, n_z, z, idx, d, N);
    return true;
}

// F2 tree kernel from the wavenumbers as the reference forms it (bispectrum.cpp:629-633), contraction
// off.  In squeezed triangles the three F2 P P terms of eq. B5 cancel to ~1/squeeze of their size, and a
// rounding that breaks the consistency of the three cosines (an FMA in one of them, or wavenumbers
// rescaled one by one) is amplified ~squeeze^2; formed this way the single-z entry points keep the
// reference's cosines.
__device__ double f2_tree_k(double k1, double k2, double k3)
{
#pragma clang fp contract(off)
    double c12 = 0.5 * (k3 * k3 - k1 * k1 - k2 * k2) / (k1 * k2);
    return (5. / 7.) + 0.5 * c12 * (k1 / k2 + k2 / k1) + (2. / 7.) * c12 * c12;
}

// BiHalofit at wavenumbers k [h/Mpc] (bispectrum.cpp:166-236): q = k r_sigma with the tables' r_sigma
// (:177), r1 and r2 of eq. B8 from the sorted q (:179-192), the F2 kernels from k.
__device__ double bispec_k(const TpNode64 &N, const TpEH64 &E, double rs, double k1, double k2, double k3)
{
    const double q[3] = {k1 * rs, k2 * rs, k3 * rs};
    double a = q[0], b = q[1], c = q[2], t;
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    const double lq[3] = {log(q[0]), log(q[1]), log(q[2])};
    const double f2t[3] = {f2_tree_k(k1, k2, k3), f2_tree_k(k2, k3, k1), f2_tree_k(k3, k1, k2)};
    return tp_bispec_node<double>(N, E, q, lq, f2t, log(c / a), (b + c - a) / a);
}

// Tree-level bispectrum, bispectrum.cpp:609-612 (its z > 10 branch, :174-175): D1^4 2 sum F2 P P with
// the linear P(k) (the fit or the given table).
__device__ double tree_bispec(const TpCosmoConst &C, double D1, double k1, double k2, double k3)
{
    const double p1 = tp_pk_lin<double>(C, k1, C.norm), p2 = tp_pk_lin<double>(C, k2, C.norm),
                 p3 = tp_pk_lin<double>(C, k3, C.norm);
    return pow(D1, 4) * 2. * (f2_tree_k(k1, k2, k3) * p1 * p2 + f2_tree_k(k2, k3, k1) * p2 * p3 +
                              f2_tree_k(k3, k1, k2) * p3 * p1);
}
__device__ double table_lerp(const double *T, int n_z, int row, int idx, double d)
{
    return T[row * n_z + idx] * (1 - d) + T[row * n_z + idx + 1] * d;
}

__device__ __forceinline__ double interior_angle(double an1, double an2, double opp)
{
    return acos((an1 * an1 + an2 * an2 - opp * opp) / (2.0 * an1 * an2));  // gamma.cpp:121-124
}
__device__ __forceinline__ double alpha_f(double psi, double x1, double x2, double phi, double vp)
{
    double zz = (cos(psi) * x2 - sin(psi) * x1) * sin((phi + vp) / 2);  // gamma.cpp:192-197
    double nn = (cos(psi) * x2 + sin(psi) * x1) * cos((phi + vp) / 2);
    return atan2(zz, nn);
}
__device__ __forceinline__ double betabar(double psi, double phi)
{
    return 0.5 * atan2(cos(2 * psi) * sin(phi), cos(phi) + sin(2 * psi));  // gamma.cpp:199-204
}
__device__ __forceinline__ double A_f(double psi, double x1, double x2, double phi, double vp)
{
    double a = cos(psi) * x2, b = sin(psi) * x1;  // gamma.cpp:216-219
    return sqrt(a * a + b * b + sin(2 * psi) * x1 * x2 * cos(phi + vp));
}

// integrand_bkappa of the reference's analytic test bispectrum (bispectrum.cpp:477-499, 989-996): zero above
// z = 1; phi = pi - acos(x) of the triangle, 0 where acos gives NaN (as the reference sets it)
__device__ double integrand_bkappa_analytic(double z, double l1, double l2, double l3, double limber_pref)
{
    if (z > 1) return 0.0;
    double phi = tpd::kPi - acos((l1 * l1 + l2 * l2 - l3 * l3) / (2 * l1 * l2));
    if (isnan(phi)) phi = 0;
    const double wts[9] = {0, 1.0e+8, 1.0e+6, 1.0e+4, 1.0e+2, 1.0e+0, 5.0e-3, 1.0e-5, 1.0e-6};
    const double avs[9] = {3.0e+3, 1.0e+4, 3.0e+4, 1.0e+5, 3.0e+5, 1.0e+6, 3.0e+6, 1.0e+7, 3.0e+7};
    const double cphi = cos(phi);
    double temp = 0;
    for (int i = 0; i < 9; i++) {
        const double a = avs[i];
        const double bis_pref = -pow(tpd::kPi, 2) / (8 * 72 * pow(a, 8));
        const double pref = pow(l1, 2) * pow(l2, 2) * (pow(l1, 2) + pow(l2, 2) + 2 * l1 * l2 * cphi);
        const double expon = exp(-(pow(l1, 2) + pow(l2, 2) + l1 * l2 * cphi) / (6 * a));
        temp += wts[i] * (bis_pref * pref * expon);
    }
    return temp / limber_pref * 3.;
}

__device__ __forceinline__ double ogata_term(const TpNode64 &N, const TpEH64 &E, double psik, double prodk,
                                             double invA, double cpsi, double spsi, double cphi)
{
    // one node of integrated_bdelta_times_rcubed_J6/J2 (gamma.cpp:298-314)
    double l1 = psik * invA * cpsi, l2 = psik * invA * spsi;
    double s = l1 * l1 + l2 * l2 + 2 * l1 * l2 * cphi;
    if (s <= 0) return 0.0;  // ell3 = 0 -> integrand_bkappa = 0
    double l3 = sqrt(s);
    double b = tp_bispec_q<double>(N, E, l1 * N.ks, l2 * N.ks, l3 * N.ks);
    return N.wz * b * prodk;
}

// FP32 form: an Ogata term whose wavenumbers q = s ks l_i all lie in the window [q32lo, q32hi]
// (tp_ctx::gamma_q32, by default [1e-5, 1e6]: on the C2 triangles a wider window gains nothing and
// a narrower one leaves speed, at the same 1e-5 agreement with FP64) is evaluated in FP32; the rest (small z, where f_K -> 0, the first Ogata nodes, where psi_k -> 0, the far nodes)
// and any FP32 term that leaves the float range stay FP64.  Accumulation is FP64.

template <bool F32, bool ANA = false>
__global__ void __launch_bounds__(kGammaBlock) __attribute__((amdgpu_waves_per_eu(F32 ? 3 : 1))) k_gamma_points(const TpCosmoConst *__restrict__ cc,
                                                              const double *__restrict__ tab, const TpEH64 *__restrict__ eh,
                                                              const TpEH32 *__restrict__ eh32,
                                                              int n_z, const double *__restrict__ ogata, int prec_k,
                                                              const TpGammaDesc *__restrict__ desc,
                                                              const int *__restrict__ pdesc, size_t p0, size_t npts,
                                                              const double *__restrict__ x, double *__restrict__ f,
                                                              int *flag, double q32lo, double q32hi)
{
    __shared__ TpNode64 sN;
    __shared__ TpNode32 sN32;
    __shared__ TpPointPre sP;              // the point's shape (FP32 form: read back by the FP64 pass)
    __shared__ TpShapeNodeT<double> sS;
    __shared__ int sok;
    __shared__ double red[3][kGammaBlock / 64];
    size_t p = p0 + blockIdx.x;  // one block per point; a launch covers points [p0, p0 + gridDim.x)
    if (p >= npts) return;
    const TpGammaDesc D = desc[pdesc[p]];
    const double z = x[3 * p], phi = x[3 * p + 1], psi = x[3 * p + 2];
    const TpCosmoConst &C = cc[D.icosmo];
    if (threadIdx.x == 0) {
        sok = z_record(C, tab + (size_t)D.icosmo * 6 * n_z, n_z, z, &sN) ? 1 : 0;
        if (F32) {  // the record's FP32 copy (TpNodeT<double> and TpNodeT<float> list the same fields)
            const double *src = &sN.wz;
            float *dst = &sN32.wz;
            for (int k = 0; k < (int)(sizeof(TpNode64) / sizeof(double)); ++k) dst[k] = (float)src[k];
        }
    }
    __syncthreads();
    const double x1 = D.x[0], x2 = D.x[1], x3 = D.x[2];
    // term geometry: (A_j, phase_j) for j = 0, 1, 2
    double A[3];
    if (D.which == 0) {  // cyclic (x1,x2,x3), (x2,x3,x1), (x3,x1,x2): gamma.cpp:333-336
        A[0] = A_f(psi, x1, x2, phi, interior_angle(x1, x2, x3));
        A[1] = A_f(psi, x2, x3, phi, interior_angle(x2, x3, x1));
        A[2] = A_f(psi, x3, x1, phi, interior_angle(x3, x1, x2));
    } else {  // A1, A2, A3 of gamma.cpp:499-501
        A[0] = A_f(psi, x2, x3, phi, interior_angle(x2, x3, x1));
        A[1] = A_f(psi, x3, x1, phi, interior_angle(x3, x1, x2));
        A[2] = A_f(psi, x1, x2, phi, interior_angle(x1, x2, x3));
    }
    const double *psik = ogata + (D.which == 0 ? 0 : 2) * (size_t)prec_k;
    const double *prodk = psik + prec_k;
    const double *lnpsik = ogata + (D.which == 0 ? 4 : 5) * (size_t)prec_k;
    double s[3] = {0, 0, 0};
    if (ANA) {  // the analytic test bispectrum (no tables): every Ogata term of integrated_bdelta_times_rcubed
        const double cpsi = cos(psi), spsi = sin(psi), cphi = cos(phi);
        for (int k = 1 + threadIdx.x; k < prec_k; k += kGammaBlock) {
            const double pk = psik[k], wk = prodk[k];
#pragma unroll 1
            for (int j = 0; j < 3; ++j) {
                const double l1 = pk / A[j] * cpsi, l2 = pk / A[j] * spsi;  // gamma.cpp:300-307
                double l3s = l1 * l1 + l2 * l2 + 2 * l1 * l2 * cphi;
                const double l3 = l3s > 0 ? sqrt(l3s) : 0.0;
                s[j] += integrand_bkappa_analytic(z, l1, l2, l3, C.limber_pref) * wk;
            }
        }
    } else if (sok) {
        const TpEH64 &E = eh[D.icosmo];
        const double cpsi = cos(psi), spsi = sin(psi), cphi = cos(phi);
        double invA[3] = {1.0 / A[0], 1.0 / A[1], 1.0 / A[2]};
        // every term's triangle has the shape (cos psi, sin psi, u3); away from the degenerate
        // shape u3 -> 0 (where ell3 = 0 decides the value) the shape is computed once per point
        const double u3s = cpsi * cpsi + spsi * spsi + 2 * cpsi * spsi * cphi;
        if (u3s > 1e-6) {
            TpPointPre P;
            tp_point_pre<double>(cpsi, spsi, sqrt(u3s), &P);
            const double lninvA[3] = {-log(A[0]), -log(A[1]), -log(A[2])};
            // the node (one z per point) and the shape are those of every term: a_n, alpha_n, beta_n once
            TpShapeNodeT<double> S;
            tp_shape_node<double>(sN, P.lr1, P.r2, &S);
            if (F32) {
                // every thread of the block evaluates the same point: each stores the same shape
                sP = P;
                sS = S;
                const TpEH32 E32 = eh32[D.icosmo];
                TpPointPreT<float> P32;
                for (int d = 0; d < 3; ++d) {
                    P32.l[d] = (float)P.l[d];
                    P32.ll[d] = (float)P.ll[d];
                    P32.f2t[d] = (float)P.f2t[d];
                }
                P32.lr1 = (float)P.lr1;
                P32.r2 = (float)P.r2;
                TpShapeNodeT<float> S32;
                tp_shape_node<float>(sN32, P32.lr1, P32.r2, &S32);
                const double lmax = fmax(P.l[0], fmax(P.l[1], P.l[2]));
                const double lmin = fmin(P.l[0], fmin(P.l[1], P.l[2]));
                // FP32 term for s in [smin32, smax32): every wavenumber of the term in [1e-3, 1e4], the
                // range of the FP32 <Map^3> path (below it the one-halo denominators underflow in FP32)
                const double smax32 = q32hi / (sN.ks * lmax);
                const double smin32 = q32lo / (sN.ks * lmin);
                // Two passes, so that the rare FP64 terms do not set the register budget of the FP32 loop:
                // 1. the FP32 terms inside the window; a term whose FP32 form leaves the float range sets
                //    its bit in `miss` (slot m of this thread's k = 1 + tid + 256 m, component j: bit
                //    3 m + j of the two words; prec_k <= kGammaMaxK32, checked by the host);
                // 2. the FP64 terms: outside the window, and the missed ones.
                // Each component is the FP64 sum of its FP32 terms, then of its FP64 terms, in k order.
                uint64_t miss0 = 0, miss1 = 0;
                int m = 0;
                for (int k = 1 + threadIdx.x; k < prec_k; k += kGammaBlock, ++m) {
                    const double pk = psik[k], wk = prodk[k], lpk = lnpsik[k];
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const double sj = pk * invA[j];
                        if (sj < smax32 && sj >= smin32) {
                            const float t32 = tp_ogata_term<float>(sN32, E32, P32, S32, (float)sj,
                                                                   (float)(lpk + lninvA[j]), (float)wk);
                            const int bit = 3 * m + j;
                            if (isfinite(t32))
                                s[j] += (double)t32;
                            else if (bit < 64)
                                miss0 |= 1ull << bit;
                            else
                                miss1 |= 1ull << (bit - 64);
                        }
                    }
                }
                // the FP64 pass reads the point's shape and the E&H constants from memory, so they do not
                // stay in registers through the FP32 loop (the empty asm: memory may have changed)
                __asm__ __volatile__("" ::: "memory");
                const TpPointPre &Pm = sP;
                const TpShapeNodeT<double> &Sm = sS;
                const TpEH64 &Em = eh[D.icosmo];
                m = 0;
                double s0 = 0, s1 = 0, s2 = 0;
                for (int k = 1 + threadIdx.x; k < prec_k; k += kGammaBlock, ++m) {
                    const double pk = psik[k], wk = prodk[k], lpk = lnpsik[k];
#pragma unroll 1
                    for (int j = 0; j < 3; ++j) {
                        const double ia = j == 0 ? invA[0] : (j == 1 ? invA[1] : invA[2]);
                        const double la = j == 0 ? lninvA[0] : (j == 1 ? lninvA[1] : lninvA[2]);
                        const double sj = pk * ia;
                        const int bit = 3 * m + j;
                        const bool missed = bit < 64 ? ((miss0 >> bit) & 1) : ((miss1 >> (bit - 64)) & 1);
                        if (sj < smax32 && sj >= smin32 && !missed) continue;
                        const double t = tp_ogata_term<double>(sN, Em, Pm, Sm, sj, lpk + la, wk);
                        if (j == 0)
                            s0 += t;
                        else if (j == 1)
                            s1 += t;
                        else
                            s2 += t;
                    }
                }
                s[0] += s0;
                s[1] += s1;
                s[2] += s2;
            } else {
                for (int k = 1 + threadIdx.x; k < prec_k; k += kGammaBlock) {
                    const double pk = psik[k], wk = prodk[k], lpk = lnpsik[k];
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        s[j] += tp_ogata_term<double>(sN, E, P, S, pk * invA[j], lpk + lninvA[j], wk);
                }
            }
        } else {  // the degenerate shape (rare): one term at a time
            for (int k = 1 + threadIdx.x; k < prec_k; k += kGammaBlock) {
                const double pk = psik[k], wk = prodk[k];
#pragma unroll 1
                for (int j = 0; j < 3; ++j) s[j] += ogata_term(sN, E, pk, wk, invA[j], cpsi, spsi, cphi);
            }
        }
    }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double v = s[j];
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        if (lane == 0) red[j][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double S[3];
    for (int j = 0; j < 3; ++j) {
        double v = 0;
        for (int w = 0; w < kGammaBlock / 64; ++w) v += red[j][w];
        double A2 = A[j] * A[j];
        S[j] = v * tpd::kPi / (A2 * A2);  // bis * M_PI / pow(A3, 4)
    }
    double re = 0, im = 0;
    const double s2p = sin(2 * psi);
    if (D.which == 0) {
        // gamma0_one_x (gamma.cpp:320-331): sin(2psi) e^{i(int(x2,x3,x1) - int(x1,x3,x2) - 6 alpha)} e^{2i betabar}
        const double xs[3][3] = {{x1, x2, x3}, {x2, x3, x1}, {x3, x1, x2}};
        double bb = betabar(psi, phi);
        for (int j = 0; j < 3; ++j) {
            double a = xs[j][0], b = xs[j][1], cc3 = xs[j][2];
            double vp = interior_angle(a, b, cc3);
            double ang = interior_angle(b, cc3, a) - interior_angle(a, cc3, b) - 6 * alpha_f(psi, a, b, phi, vp) +
                         2. * bb;
            re += s2p * cos(ang) * S[j];
            im += s2p * sin(ang) * S[j];
        }
    } else {
        // integrand_z_phi_psi_gamma1 (gamma.cpp:490-520)
        double vb = betabar(psi, phi);
        double vp1 = interior_angle(x2, x3, x1), vp2 = interior_angle(x3, x1, x2), vp3 = interior_angle(x1, x2, x3);
        double al1 = alpha_f(psi, x2, x3, phi, vp1), al2 = alpha_f(psi, x3, x1, phi, vp2),
               al3 = alpha_f(psi, x1, x2, phi, vp3);
        double ang1 = vp3 - vp2 - 2 * (vb + al1);
        double ang2 = vp3 - vp1 - 2 * vp2 + 2 * (vb + phi - al2);
        double ang3 = vp1 - vp2 + 2 * vp3 + 2 * (vb - phi - al3);
        re = s2p * (cos(ang1) * S[0] + cos(ang2) * S[1] + cos(ang3) * S[2]);
        im = s2p * (sin(ang1) * S[0] + sin(ang2) * S[1] + sin(ang3) * S[2]);
    }
    if (!isfinite(re) || !isfinite(im)) atomicOr(flag, 4);
    f[2 * p] = re;
    f[2 * p + 1] = im;
}

// single-z bispectrum / Limber integrand (one thread per point)
__global__ void __launch_bounds__(128) k_zpoint(const TpCosmoConst *__restrict__ cc, const double *__restrict__ tab,
                                                const TpEH64 *__restrict__ eh, int n_z, int icosmo, int what,
                                                size_t npts, const double *__restrict__ in, double *__restrict__ out,
                                                int *flag)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const double *p = in + 4 * i;
    const TpCosmoConst &C = cc[icosmo];
    double v = 0.0;
    TpNode64 N;
    const double *T = tab + (size_t)icosmo * 6 * n_z;
    if (what == 0 || what == 2) {  // bispec(k1, k2, k3, z[, idx, didx]): B at wavenumbers k [h/Mpc]
        const double *q = what == 2 ? in + 6 * i : p;
        const double z = q[3];
        int idx;
        double d;
        if (what == 2) {  // the caller's table interval (bispectrum.cpp:166-175)
            idx = (int)q[4];
            d = q[5];
        } else {          // the interval of z (:506-513); the reference has no z guard in bispec
            z_interval(C, n_z, z, &idx, &d);
        }
        if (z > 10.) {
            v = tree_bispec(C, table_lerp(T, n_z, 2, idx, d), q[0], q[1], q[2]);
        } else {
            z_record_at(C, T, n_z, z, idx, d, &N);
            v = bispec_k(N, eh[icosmo], table_lerp(T, n_z, 3, idx, d), q[0], q[1], q[2]);
        }
    } else {  // integrand_bkappa(z, ell1, ell2, ell3), bispectrum.cpp:502-524: 0 at z = 0 or ell3 = 0
        const double z = p[0];
        if (z != 0.0 && p[3] != 0.0) {
            int idx;
            double d;
            z_interval(C, n_z, z, &idx, &d);
            z_record_at(C, T, n_z, z, idx, d, &N);
            const double gv = table_lerp(T, n_z, 1, idx, d), fK = table_lerp(T, n_z, 0, idx, d);
            const double E = sqrt(C.om * pow(1 + z, 3) + C.ow);  // E(z), bispectrum.cpp:967-970
            const double b = bispec_k(N, eh[icosmo], table_lerp(T, n_z, 3, idx, d), p[1] / fK, p[2] / fK, p[3] / fK);
            v = pow(gv * (1. + z), 3) * b / fK / E;
        }
    }
    if (!isfinite(v)) atomicOr(flag, 8);
    out[i] = v;
}

// ---------------------------------------------------------- host Bessel
// J_n(x): power series (x < 2), Miller backward recurrence normalised by the
// Neumann sum J0 + 2 sum J_2k = 1 (x < 30), Hankel's asymptotic expansion (x >= 30);
// Abramowitz & Stegun 9.1.10, 9.1.46, 9.2.5-9.2.10.
double jn_host(int n, double x)
{
    if (x == 0) return n == 0 ? 1.0 : 0.0;
    if (x < 2.0) {
        double term = 1.0;
        for (int k = 1; k <= n; ++k) term *= 0.5 * x / k;
        double sum = term, q = -0.25 * x * x;
        for (int m = 1; m < 80; ++m) {
            term *= q / ((double)m * (double)(m + n));
            sum += term;
            if (std::fabs(term) <= 1e-17 * std::fabs(sum)) break;
        }
        return sum;
    }
    if (x < 30.0) {
        int top = 2 * ((std::max(n, (int)x) + 44) / 2);
        double bj1 = 0.0, bj = 1e-280, neumann = 0.0, keep = 0.0;
        for (int k = top; k >= 1; --k) {
            double bjm = 2.0 * k / x * bj - bj1;
            bj1 = bj;
            bj = bjm;  // now bj = J_{k-1} (unnormalised)
            if (k - 1 == n) keep = bj;
            if (k - 1 > 0 && ((k - 1) % 2 == 0)) neumann += 2.0 * bj;
            if (std::fabs(bj) > 1e200) {
                bj *= 1e-200;
                bj1 *= 1e-200;
                neumann *= 1e-200;
                keep *= 1e-200;
            }
        }
        neumann += bj;
        if (n == 0) keep = bj;
        return keep / neumann;
    }
    double mu = 4.0 * n * n, P = 1.0, Q = 0.0, t = 1.0;
    for (int k = 1; k < 64; ++k) {
        double o = 2.0 * k - 1.0;
        t *= (mu - o * o) / (k * 8.0 * x);
        if (k > 2 && std::fabs(t) < 1e-18) break;
        double sgn = ((k / 2) % 2) ? -1.0 : 1.0;
        if (k % 2) Q += sgn * t;
        else P += sgn * t;
    }
    double chi_c = (0.5 * n + 0.25) * M_PI;
    double cx = std::cos(x), sx = std::sin(x), cc = std::cos(chi_c), sc = std::sin(chi_c);
    return std::sqrt(2.0 / (M_PI * x)) * (P * (cx * cc + sx * sc) - Q * (sx * cc - cx * sc));
}

double jn_zero_host(int n, int s)
{
    if (s == 0) return 0.0;  // zeroth zero = 0 (the gsl_sf_bessel_zero_Jnu convention)
    double mu = 4.0 * n * n, beta = (s + 0.5 * n - 0.25) * M_PI, b8 = 8.0 * beta;
    double x = beta - (mu - 1) / b8 - 4 * (mu - 1) * (7 * mu - 31) / (3 * b8 * b8 * b8);  // McMahon
    for (int it = 0; it < 40; ++it) {
        double J = jn_host(n, x);
        double dJ = jn_host(n - 1, x) - n / x * J;
        double dx = J / dJ;
        x -= dx;
        if (std::fabs(dx) <= 1e-15 * x) break;
    }
    return x;
}

struct GammaIntegral {
    tpc::AdaptiveIntegral32 I;
    size_t tri;
    int comp;
};

}  // namespace

static int zpoint_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, int what, size_t npts, const double *in,
                       double *out)
{
    if (!ctx || !c || (npts && (!in || !out))) return tp_set_error(TP_EINVAL, "NULL argument");
    if (icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "cosmology index out of range");
    if (npts == 0) return 0;
    TP_HIP(hipSetDevice(ctx->device));
    int rc;
    const size_t stride = what == 2 ? 6 : 4;
    if (what == 2)
        for (size_t i = 0; i < npts; ++i) {
            const double ix = in[6 * i + 4];
            if (!(ix >= 0 && ix <= c->n_z - 2) || ix != (double)(int)ix)
                return tp_set_error(TP_EINVAL, "tp_bispec_eval_idx: idx must be an integer in [0, n_z - 2]");
        }
    size_t bi = npts * stride * sizeof(double), bo = npts * sizeof(double);
    if ((rc = ctx->d_in.ensure(bi)) || (rc = ctx->d_out.ensure(bo))) return rc;
    TP_HIP(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(int), ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_in.p, in, bi, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_zpoint, dim3((unsigned)((npts + 127) / 128)), dim3(128), 0, ctx->stream, c->d_const,
                       c->d_tab, c->d_eh64, c->n_z, icosmo, what, npts, (const double *)ctx->d_in.p,
                       (double *)ctx->d_out.p, (int *)ctx->d_flag.p);
    TP_HIP(hipGetLastError());
    TP_HIP(hipMemcpyAsync(out, ctx->d_out.p, bo, hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    if (flag) return tp_set_error(TP_ENONFINITE, "non-finite bispectrum value");
    return 0;
}

extern "C" int tp_bispec_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *k123z, double *out)
{
    return zpoint_eval(ctx, c, icosmo, 0, npts, k123z, out);
}

extern "C" int tp_bispec_eval_idx(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *k123z_idx,
                                  double *out)
{
    return zpoint_eval(ctx, c, icosmo, 2, npts, k123z_idx, out);
}

extern "C" int tp_integrand_bkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *zell,
                                        double *out)
{
    return zpoint_eval(ctx, c, icosmo, 1, npts, zell, out);
}

extern "C" int tp_gamma_setup(tp_ctx *ctx, double prec_h, double prec_k_arg)
{
    // classed_version/gamma.cpp:133-176
    if (!ctx || !(prec_h > 0) || !(prec_k_arg > 0)) return tp_set_error(TP_EINVAL, "tp_gamma_setup: bad arguments");
    unsigned pk = (unsigned)(prec_k_arg / prec_h);
    if (pk < 2 || pk > (1u << 24)) return tp_set_error(TP_EINVAL, "tp_gamma_setup: prec_k out of range");
    std::vector<double> h(6 * (size_t)pk, 0.0);  // psi6, prod6, psi2, prod2, ln psi6, ln psi2
    auto psi = [](double t) { return t * std::tanh(M_PI * std::sinh(t) / 2); };
    auto psip = [](double t) {
        return (std::sinh(M_PI * std::sinh(t)) + M_PI * t * std::cosh(t)) / (std::cosh(M_PI * std::sinh(t)) + 1);
    };
#pragma omp parallel for schedule(dynamic, 64) num_threads(tp_host_threads(ctx, kTpHostThreadsSync))
    for (long long i = 1; i < (long long)pk; ++i) {
        for (int which = 0; which < 2; ++which) {
            int nu = which == 0 ? 6 : 2;
            double bz = jn_zero_host(nu, (int)i);
            double pbz = bz / M_PI;
            double ps = M_PI * psi(pbz * prec_h) / prec_h;
            double jn1 = jn_host(nu + 1, bz);
            double w = 2 / (M_PI * bz * jn1 * jn1);
            double prod = w * ps * ps * ps * jn_host(nu, ps) * psip(prec_h * pbz);
            h[(size_t)(2 * which) * pk + i] = ps;
            h[(size_t)(2 * which + 1) * pk + i] = prod;
            h[(size_t)(4 + which) * pk + i] = std::log(ps);
        }
    }
    TP_HIP(hipSetDevice(ctx->device));
    if (ctx->d_ogata) hipFree(ctx->d_ogata);
    ctx->d_ogata = nullptr;
    TP_HIP(hipMalloc(&ctx->d_ogata, sizeof(double) * h.size()));
    TP_HIP(hipMemcpy(ctx->d_ogata, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    ctx->gamma_prec_h = prec_h;
    ctx->gamma_prec_k = pk;
    return 0;
}

static int gamma_launch(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpGammaDesc> &desc,
                        const std::vector<int> &pdesc, size_t npts, const double *x, double *fval,
                        int mode = TP_MODE_FP64)
{
    if (npts == 0) return 0;
    if (mode == TP_MODE_FP32 && ctx->gamma_prec_k > (unsigned)kGammaMaxK32)
        return tp_set_error(TP_EINVAL, "Gamma FP32 form: prec_k above " + std::to_string(kGammaMaxK32) +
                                           " (set up with tp_gamma_setup); use the FP64 form");
    int rc;
    size_t bx = npts * 3 * sizeof(double), bf = npts * 2 * sizeof(double);
    size_t bd = desc.size() * sizeof(TpGammaDesc), bp = npts * sizeof(int);
    if ((rc = ctx->d_in.ensure(bx)) || (rc = ctx->d_out.ensure(bf)) || (rc = ctx->d_desc.ensure(bd)) ||
        (rc = ctx->d_misc.ensure(bp)) || (rc = ctx->h_in.ensure(bx)) || (rc = ctx->h_out.ensure(std::max(bf, (size_t)64))))
        return rc;
    std::memcpy(ctx->h_in.p, x, bx);
    TP_HIP(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(int), ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_in.p, ctx->h_in.p, bx, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_desc.p, desc.data(), bd, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_misc.p, pdesc.data(), bp, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    // One block of kGammaBlock threads per point.  A dispatch's grid size is a 32-bit count of work-items,
    // so one launch covers at most kGammaLaunchPts points; larger rounds run as several launches.
    for (size_t p0 = 0; p0 < npts; p0 += kGammaLaunchPts) {
        const dim3 gr((unsigned)std::min(npts - p0, kGammaLaunchPts));
        if (mode == TP_MODE_FP32)
            hipLaunchKernelGGL(k_gamma_points<true>, gr, dim3(kGammaBlock), 0, ctx->stream, c->d_const, c->d_tab,
                               c->d_eh64, c->d_eh32, c->n_z, ctx->d_ogata, (int)ctx->gamma_prec_k,
                               (const TpGammaDesc *)ctx->d_desc.p, (const int *)ctx->d_misc.p, p0, npts,
                               (const double *)ctx->d_in.p, (double *)ctx->d_out.p, (int *)ctx->d_flag.p,
                               ctx->gamma_q32[0], ctx->gamma_q32[1]);
        else if (mode == TP_MODE_ANALYTIC)
            hipLaunchKernelGGL((k_gamma_points<false, true>), gr, dim3(kGammaBlock), 0, ctx->stream, c->d_const,
                               c->d_tab, c->d_eh64, c->d_eh32, c->n_z, ctx->d_ogata, (int)ctx->gamma_prec_k,
                               (const TpGammaDesc *)ctx->d_desc.p, (const int *)ctx->d_misc.p, p0, npts,
                               (const double *)ctx->d_in.p, (double *)ctx->d_out.p, (int *)ctx->d_flag.p, 0.0, 0.0);
        else
            hipLaunchKernelGGL(k_gamma_points<false>, gr, dim3(kGammaBlock), 0, ctx->stream, c->d_const, c->d_tab,
                               c->d_eh64, c->d_eh32, c->n_z, ctx->d_ogata, (int)ctx->gamma_prec_k,
                               (const TpGammaDesc *)ctx->d_desc.p, (const int *)ctx->d_misc.p, p0, npts,
                               (const double *)ctx->d_in.p, (double *)ctx->d_out.p, (int *)ctx->d_flag.p, 0.0, 0.0);
        TP_HIP(hipGetLastError());
    }
    TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, bf, hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(fval, ctx->h_out.p, bf);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats.kernel_ms += ms;
    ctx->stats.integrand_ms += ms;
    ctx->stats.integrand_sum_ms += ms;
    ctx->stats.launches += (long long)((npts + kGammaLaunchPts - 1) / kGammaLaunchPts);
    int flag = 0;
    TP_HIP(hipMemcpy(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) return tp_set_error(TP_ENONFINITE, "Gamma integrand: non-finite value");
    return 0;
}

extern "C" int tp_gamma_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, int which, double x1, double x2, double x3,
                             size_t npts, const double *x, double *fval)
{
    if (!ctx || !c || (npts && (!x || !fval))) return tp_set_error(TP_EINVAL, "tp_gamma_eval: NULL");
    if (!ctx->d_ogata) return tp_set_error(TP_EINVAL, "tp_gamma_eval: call tp_gamma_setup first");
    if (icosmo < 0 || icosmo >= c->ncosmo || (which != 0 && which != 1))
        return tp_set_error(TP_EINVAL, "tp_gamma_eval: bad icosmo or which");
    TP_HIP(hipSetDevice(ctx->device));
    std::vector<TpGammaDesc> desc(1, TpGammaDesc{{x1, x2, x3}, which, icosmo});
    std::vector<int> pdesc(npts, 0);
    return gamma_launch(ctx, c, desc, pdesc, npts, x, fval);
}

static double one_rotation_angle_otc(double x1, double x2, double x3)
{
    // classed_version/gamma.cpp:624-633
    double psi3 = std::acos((x1 * x1 + x2 * x2 - x3 * x3) / (2.0 * x1 * x2));
    double h3 = 0.5 * std::sqrt(2 * x1 * x1 + 2 * x2 * x2 - x3 * x3);
    double c2 = std::pow(x2 * x2 - x1 * x1, 2) - 4 * std::pow(x1 * x2 * std::sin(psi3), 2);
    c2 /= std::pow(2 * h3 * x3, 2);
    double s2 = (x2 * x2 - x1 * x1) * x1 * x2 * std::sin(psi3);
    s2 /= std::pow(h3 * x3, 2);
    return std::atan2(s2, c2) / 2;
}

extern "C" int tp_gamma_batch(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                              double reltol, double *out_re, double *out_im, long long *neval)
{
    return tp_gamma_batch_mode(ctx, c, icosmo, ntri, x123, reltol, TP_MODE_FP64, out_re, out_im, neval);
}

extern "C" int tp_gamma_batch_mode(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                                   double reltol, int mode, double *out_re, double *out_im, long long *neval)
{
    return tp_gamma_batch_center(ctx, c, icosmo, ntri, x123, reltol, mode, TP_CENTER_CENTROID, out_re, out_im, neval);
}

extern "C" int tp_gamma_batch_center(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                                     double reltol, int mode, int center, double *out_re, double *out_im,
                                     long long *neval)
{
    if (!ctx || !c || !x123 || !out_re || !out_im) return tp_set_error(TP_EINVAL, "tp_gamma_batch: NULL");
    if (center != TP_CENTER_CENTROID && center != TP_CENTER_ORTHOCENTER)
        return tp_set_error(TP_EINVAL, "tp_gamma_batch: center must be TP_CENTER_CENTROID or TP_CENTER_ORTHOCENTER");
    if (mode != TP_MODE_FP64 && mode != TP_MODE_FP32 && mode != TP_MODE_ANALYTIC)
        return tp_set_error(TP_EINVAL, "tp_gamma_batch: mode must be TP_MODE_FP64, TP_MODE_FP32 or TP_MODE_ANALYTIC");
    if (!ctx->d_ogata) return tp_set_error(TP_EINVAL, "tp_gamma_batch: call tp_gamma_setup first");
    if (icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_gamma_batch: bad icosmo");
    TP_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->stats = tp_batch_stats{};
    const int nt = tp_host_threads(ctx, kTpHostThreadsSync);
    const size_t nint = 4 * ntri;
    std::vector<GammaIntegral> G(nint);
    std::vector<TpGammaDesc> desc(nint);
    const double zmax = c->z_max;
    for (size_t t = 0; t < ntri; ++t) {
        double x1 = x123[3 * t], x2 = x123[3 * t + 1], x3 = x123[3 * t + 2];
        for (int i = 0; i < 4; ++i) {
            // gamma2(x1,x2,x3) = gamma1(x2,x3,x1), gamma3 = gamma1(x3,x1,x2) (gamma.cpp:73-81)
            double a = x1, b = x2, d = x3;
            if (i == 2) { a = x2; b = x3; d = x1; }
            if (i == 3) { a = x3; b = x1; d = x2; }
            desc[4 * t + i] = TpGammaDesc{{a, b, d}, i == 0 ? 0 : 1, icosmo};
            double vmin[3] = {0, 0, 0}, vmax[3] = {zmax, 2 * M_PI, M_PI / 2};
            G[4 * t + i].I.init(3, 2, vmin, vmax, 0, 0, reltol, tpc::L1);
            G[4 * t + i].tri = t;
            G[4 * t + i].comp = i;
        }
    }
    const unsigned np = tpc::rule_points(3);
    std::vector<size_t> active(nint);
    for (size_t k = 0; k < nint; ++k) active[k] = k;
    std::vector<double> pts, vals;
    std::vector<int> pdesc;
    std::vector<size_t> offs;
    long long regions = 0;
    while (!active.empty()) {
        if (ctx->stats.rounds >= ctx->max_rounds)
            return tp_set_error(TP_ECAPACITY, "tp_gamma_batch: round limit reached (integrals not converging)");
        size_t na = active.size();
        offs.assign(na + 1, 0);
        for (size_t a = 0; a < na; ++a) offs[a + 1] = offs[a] + G[active[a]].I.pending().size() * np;
        size_t npts = offs[na];
        regions += (long long)(npts / np);
        if (regions > ctx->max_regions)
            return tp_set_error(TP_ECAPACITY, "tp_gamma_batch: region limit reached (integrals not converging)");
        pts.resize(npts * 3);
        vals.resize(npts * 2);
        pdesc.resize(npts);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) if (na > 1 && npts > 16384)
        for (size_t a = 0; a < na; ++a) {
            auto &A = G[active[a]].I;
            const auto &pend = A.pending();
            for (size_t r = 0; r < pend.size(); ++r) {
                size_t o = offs[a] + r * np;
                tpc::rule_points_of(3, A.regions()[pend[r]], pts.data() + 3 * o);
                for (unsigned q = 0; q < np; ++q) pdesc[o + q] = (int)active[a];
            }
        }
        int rc = gamma_launch(ctx, c, desc, pdesc, npts, pts.data(), vals.data(), mode);
        if (rc) return rc;
        ctx->stats.rounds += 1;
        ctx->stats.points += (long long)npts;
        std::vector<char> fin(na, 0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) if (na > 1 && npts > 16384)
        for (size_t a = 0; a < na; ++a) {
            auto &A = G[active[a]].I;
            const auto &pend = A.pending();
            for (size_t r = 0; r < pend.size(); ++r)
                tpc::rule_apply(3, 2, A.regions()[pend[r]], vals.data() + 2 * (offs[a] + r * np));
            A.absorb();
            fin[a] = A.advance() ? 1 : 0;
        }
        std::vector<size_t> next;
        for (size_t a = 0; a < na; ++a)
            if (!fin[a]) next.push_back(active[a]);
        active.swap(next);
    }
    for (size_t k = 0; k < nint; ++k) {
        double v[2], e[2];
        G[k].I.result(v, e);
        std::complex<double> r(v[0], v[1]);
        const TpCosmoConst &C = c->h_const[icosmo];
        r *= 27. / 8. * std::pow(C.om, 3) * std::pow(100. / 299792., 5);  // gamma.cpp:40-42
        r /= 3.;
        r /= (2 * std::pow(2 * M_PI, 3));
        // the integrand's natural components refer to the orthocenter; the centroid convention rotates them
        // (gamma.cpp:44-45, 67-68), the orthocenter convention keeps them (gamma.cpp:12-13)
        if (center == TP_CENTER_CENTROID) {
            const TpGammaDesc &D = desc[k];
            double a1 = one_rotation_angle_otc(D.x[1], D.x[2], D.x[0]);
            double a2 = one_rotation_angle_otc(D.x[2], D.x[0], D.x[1]);
            double a3 = one_rotation_angle_otc(D.x[0], D.x[1], D.x[2]);
            if (G[k].comp != 0) a1 *= -1.;  // conjugate_phi1 for gamma1..3 (gamma.cpp:635-643)
            r = r * std::exp(std::complex<double>(0, -2.) * (a1 + a2 + a3));
        }
        out_re[k] = r.real();
        out_im[k] = r.imag();
        if (neval) neval[k] = (long long)G[k].I.num_eval();
    }
    ctx->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

extern "C" int tp_gamma_fp32_window(tp_ctx *ctx, double qmin, double qmax)
{
    if (!ctx || !(qmin > 0) || !(qmax > qmin)) return tp_set_error(TP_EINVAL, "tp_gamma_fp32_window: need 0 < qmin < qmax");
    ctx->gamma_q32[0] = qmin;
    ctx->gamma_q32[1] = qmax;
    return 0;
}
