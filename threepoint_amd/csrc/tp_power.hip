// tp_power.hip -- convergence power spectrum P_kappa(ell) and <Map^2>(theta) on gfx950.
//
// The second-order statistics of the same model: revised Halofit (Takahashi et al. 2012) P_NL(k, z)
// from the per-cosmology FP64 tables (D1, r_sigma, n_eff, n_cur), Limber-projected with the lensing
// efficiency g(z), then
//   * Pell(ell)                : 96-point Gauss-Legendre z-sum + shape noise
//                                (classed_version/bispectrum.cpp:1034-1059), one wave per ell;
//   * convergence_power_spectrum(ell): adaptive 1-D cubature over z (bispectrum.cpp:382-417);
//   * <Map^2>(theta)           : adaptive 2-D cubature over (ell, z) (cuda_version/apertureStatistics.cu:29-136).
// and the Gaussian <Map^3> covariance integrands (the 4-D single-z form of the classed_version and
// T1 of the cuda_version for an infinite survey).
// The adaptive integrals of one call run together: every round gathers the pending regions of all
// unfinished integrals and one launch evaluates all their rule points (15 Gauss-Kronrod points in 1-D,
// 17 / 33 / 57 Genz-Malik points in 2-D / 3-D / 4-D); each integral's heap advances with the decision
// sequence of the reference's one-at-a-time hcubature.  P_kappa and the 4-D covariance apply the rules
// on the host; <Map^2> and T1 apply them on the device (bit-identical arithmetic order).
//
// Per point the integrand is one table interpolation, the Halofit z-coefficients (pow), the k-part in
// log space and the filter: transcendental-bound FP64 work on a few doubles of input, so a launch is
// bounded by its FP64 issue rate, and the rounds by the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "tp_cubature.h"
#include "tp_power_dev.h"

namespace {

// What one integral (or one evaluation call) computes at its points.
enum PowerKind {
    kPkappa = 0, kMap2 = 1, kPkAtZ = 2, kLimber = 3, kCovGauss = 4, kT1 = 5,
    // the T2 integrals of a finite survey (1-D: A and the circle's B; 2-D: the square's and rectangle's B)
    kT2A = 6, kT2Circle = 7, kT2Square = 8, kT2Rect = 9,
    // T1 of a finite survey (6-D)
    kT1Circle = 10, kT1Square = 11, kT1Rect = 12
};

struct TpPowerDesc {
    double par;          // kPkappa: ell; kMap2: theta [rad]
    double sigma, ngal;  // kMap2 shape noise (sigma = 0: none), ngal [arcmin^-2]; kT1: Pell's shape noise
    int icosmo;
    int kind;
    int flags;           // TP_PK_LINEAR | TP_PK_T17 (| TP_PK_SHAPENOISE_ONLY for kT1)
    int pad;
    double th[6];        // kCovGauss, kT1: theta_123, theta_456 [rad]; kT2*: the two radii, the survey's
                         // side lengths or radius (theta_max, theta_max_smaller) [rad]
    double lmin, lmax;   // kT1, kT2*, kT1Circle..: the integrand's ell cut
    double wx, wy;       // kT1Circle..kT1Rect: the survey's radius or side lengths [rad]
};

// 96-point Gauss-Legendre half nodes/weights (own Newton solve, uploaded once per device)
__constant__ double c_pgl_x[48];
__constant__ double c_pgl_w[48];

__device__ __forceinline__ double uhat(double eta)
{
    const double t = 0.5 * eta * eta;  // classed_version/apertureStatistics.cpp:10-14
    return t * exp(-t);
}

// Pell from the z-node records N[96] (pairs 2j, 2j+1 at cx -+ dx x_j): the cuda_version's device Pell
// (bispectrum.cu:511-529), node pairs added in the reference's order (GQ96_of_Pk, bispectrum.cpp:1034-1046).
__device__ double pell_nodes(const TpCosmoConst &C, const TpPNode *__restrict__ N, double ell, int flags, double dx,
                             double shot)
{
#pragma clang fp contract(off)
    if ((flags & TP_PK_SHAPENOISE_ONLY) || ell == 0) return shot;
    const double lell = log(ell);
    double q = 0;
    for (int j = 0; j < 48; ++j)
        q += c_pgl_w[j] * (tp_limber_pk_node(C, N[2 * j], ell, lell, flags) +
                           tp_limber_pk_node(C, N[2 * j + 1], ell, lell, flags));
    return q * dx + shot;
}

// FP32 form of pell_nodes: FP32 node values (tp_limber_pk_node32), added in FP64 in the same order.
__device__ double pell_nodes32(const TpPNode32 *__restrict__ N, double ell, int flags, double dx, double shot)
{
    if ((flags & TP_PK_SHAPENOISE_ONLY) || ell == 0) return shot;
    const float e = (float)ell, le = (float)log(ell);
    double q = 0;
    for (int j = 0; j < 48; ++j)
        q += c_pgl_w[j] * ((double)tp_limber_pk_node32(N[2 * j], e, le, flags) +
                           (double)tp_limber_pk_node32(N[2 * j + 1], e, le, flags));
    return q * dx + shot;
}

// l3 = |l1 + l2| of the T1 integrand at (l1, l2, phi), cuda_version/apertureStatisticsCovariance.cu:1962
__device__ __forceinline__ double t1_l3(double l1, double l2, double phi)
{
#pragma clang fp contract(off)
    return sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
}

// The classed_version's Gaussian covariance integrand at (l1, l2, phi, z),
// classed_version/apertureStatistics.cpp:66-106, with the single-z triple spectrum of bispectrum.cpp:580-607.
__device__ double covg_point(const TpCosmoConst &C, const double *T, int n_z, double l1, double l2, double phi,
                             double z, const double *th)
{
    double l3;
    {
#pragma clang fp contract(off)
        l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
    }
    auto up = [&](double a, double b, double c, const double *t) { return uhat(a * t[0]) * uhat(b * t[1]) * uhat(c * t[2]); };
    double result = up(l1, l2, l3, th);
    double perm = up(l1, l2, l3, th + 3);
    perm += up(l2, l3, l1, th + 3);
    perm += up(l3, l1, l2, th + 3);
    perm += up(l1, l3, l2, th + 3);
    perm += up(l3, l2, l1, th + 3);
    perm += up(l2, l1, l3, th + 3);
    result *= perm;
    double tri = 0;
    if (!(z < 1e-5)) {
        TpZAtT<double> Z;
        tp_z_at<double>(T, n_z, C.zmax, z, &Z);
        const double pref = tp_limber_pref<double>(C, z, Z.g);
        TpHaloZT<double> h;
        tp_halofit_z<double>(C, Z, z, &h);
        const double lrs = log(Z.rs), k1 = l1 / Z.fK, k2 = l2 / Z.fK, k3 = l3 / Z.fK;
        tri = pow(pref, 3) * tp_halofit_k<double>(C, Z.D1, Z.rs, lrs, h, k1, log(k1)) *
              tp_halofit_k<double>(C, Z.D1, Z.rs, lrs, h, k2, log(k2)) *
              tp_halofit_k<double>(C, Z.D1, Z.rs, lrs, h, k3, log(k3));
    }
    result *= tri;
    result *= l1 * l2;
    return result;
}


// Survey window factors (cuda_version/apertureStatisticsCovariance.cu:117-177): the circle's
// 4 J1(theta_max l)^2 / (theta_max l)^2, and the squared sinc factors of the half side lengths of the
// square and the rectangle (1 below |argument| 1e-6).
__device__ __forceinline__ double g_circle(double ell, double tmax)
{
#pragma clang fp contract(off)
    const double t = tmax * ell;
    double r = j1(t);
    r *= r;
    r *= 4 / t / t;
    return r;
}

__device__ __forceinline__ double sinc_cut(double t) { return fabs(t) <= 1e-6 ? 1.0 : sin(t) / t; }

// The T2 integrands at one point (apertureStatisticsCovariance.cu:2014-2110), in the reference's
// product order: kT2A integrand_T2_part1 (ell Pell uHat uHat), kT2Circle integrand_T2_part2_circle,
// kT2Square / kT2Rect integrand_T2_part2_square / _rectangle over (ell_x, ell_y).
__device__ double t2_point(const TpCosmoConst &C, const TpPNode *__restrict__ N, const TpPowerDesc &D,
                           const double *x, double dx)
{
#pragma clang fp contract(off)
    const double shot = 0.5 * D.sigma * D.sigma / D.ngal;
    if (D.kind == kT2A || D.kind == kT2Circle) {
        const double ell = x[0];
        if (ell < D.lmin) return 0;
        if (D.kind == kT2A) {
            double r = ell * pell_nodes(C, N, ell, D.flags, dx, shot);
            r *= uhat(ell * D.th[0]) * uhat(ell * D.th[1]);
            return r;
        }
        const double G = g_circle(ell, D.th[2]);
        double r = pell_nodes(C, N, ell, D.flags, dx, shot);
        r *= ell * uhat(ell * D.th[0]) * uhat(ell * D.th[1]);
        r *= G;
        return r;
    }
    const double lx = x[0], ly = x[1];
    const double a = sinc_cut(0.5 * lx * D.th[2]), b = sinc_cut(0.5 * ly * (D.kind == kT2Square ? D.th[2] : D.th[3]));
    const double G = a * a * b * b;
    const double ell = sqrt(lx * lx + ly * ly);
    if (ell < D.lmin) return 0;
    double r = pell_nodes(C, N, ell, D.flags, dx, shot);
    r *= uhat(ell * D.th[0]) * uhat(ell * D.th[1]);
    r *= G;
    return r;
}

// T1 of a finite survey at one point (apertureStatisticsCovariance.cu:1886-1955, 1984-2012): x = (a, b)
// the window's wave vector, (c, d) and (e, f) two of the Pell wave vectors, l1 = |(a - c - e, b - d - f)|;
// the circle cuts l_i <= lMin, the square and the rectangle l_i <= 0 and l3 > lMax; the reference's
// product order.
__device__ double t1fin_point(const TpCosmoConst &C, const TpPNode *__restrict__ N, const TpPowerDesc &D,
                              const double *x, double dx)
{
#pragma clang fp contract(off)
    const double a = x[0], b = x[1], c = x[2], d = x[3], e = x[4], f = x[5];
    double G;
    if (D.kind == kT1Circle) {
        G = g_circle(sqrt(a * a + b * b), D.wx);
    } else {
        const double p = sinc_cut(0.5 * a * D.wx), q = sinc_cut(0.5 * b * (D.kind == kT1Square ? D.wx : D.wy));
        G = p * p * q * q;
    }
    const double ell1 = sqrt((a - c - e) * (a - c - e) + (b - d - f) * (b - d - f));
    const double ell2 = sqrt(c * c + d * d), ell3 = sqrt(e * e + f * f);
    if (D.kind == kT1Circle) {
        if (ell1 <= D.lmin || ell2 <= D.lmin || ell3 <= D.lmin) return 0;
    } else if (ell1 <= 0 || ell2 <= 0 || ell3 <= 0 || ell3 > D.lmax) {
        return 0;
    }
    const double shot = 0.5 * D.sigma * D.sigma / D.ngal;
    double r = pell_nodes(C, N, ell1, D.flags, dx, shot) * pell_nodes(C, N, ell2, D.flags, dx, shot) *
               pell_nodes(C, N, ell3, D.flags, dx, shot);
    r *= uhat(ell1 * D.th[0]) * uhat(ell2 * D.th[1]) * uhat(ell3 * D.th[2]);
    r *= uhat(ell1 * D.th[3]) * uhat(ell2 * D.th[4]) * uhat(ell3 * D.th[5]);
    r *= G;
    return r;
}

// One thread per point; pdesc[i] names the integral (desc entry) of point i.
__global__ void k_power_points(const TpCosmoConst *__restrict__ cc, const double *__restrict__ tab, int n_z,
                               const TpPowerDesc *__restrict__ desc,
                               const int *__restrict__ pdesc, size_t npts, int dim, const double *__restrict__ x,
                               double *__restrict__ f, int *flag, const TpPNode *__restrict__ pnode, double dx)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const TpPowerDesc D = desc[pdesc ? pdesc[i] : 0];
    const TpCosmoConst &C = cc[D.icosmo];
    const double *T = tab + (size_t)D.icosmo * 6 * n_z;
    const double *xi = x + (size_t)dim * i;
    double v;
    if (D.kind == kPkappa) {
        v = tp_limber_pk<double>(C, T, n_z, D.par, xi[0], D.flags);
    } else if (D.kind == kMap2) {
        v = tp_map2_point<double>(C, T, n_z, xi[0], xi[1], D.par, D.flags, D.sigma, D.ngal);
    } else if (D.kind == kLimber) {
        v = tp_limber_pk<double>(C, T, n_z, xi[0], xi[1], D.flags);
    } else if (D.kind == kPkAtZ) {
        TpZAtT<double> Z;
        tp_z_at<double>(T, n_z, C.zmax, xi[1], &Z);
        v = tp_pk_at<double>(C, Z, xi[0], xi[1], D.flags);
    } else if (D.kind >= kT1Circle) {
        v = t1fin_point(C, pnode + (size_t)D.icosmo * TP_NNODES, D, xi, dx);
    } else if (D.kind >= kT2A) {
        v = t2_point(C, pnode + (size_t)D.icosmo * TP_NNODES, D, xi, dx);
    } else {
        v = covg_point(C, T, n_z, xi[0], xi[1], xi[2], xi[3], D.th);
    }
    if (!isfinite(v)) atomicOr(flag, 1);
    f[i] = v;
}

// T1 of the Gaussian <Map^3> covariance for an infinite survey (cuda_version/
// apertureStatisticsCovariance.cu:1958-1985) over whole 3-D regions (l1, l2, phi): the rule points are
// generated and the Genz-Malik rule applied on the device, so a round moves one region record in and
// one result out per region.  Of a region's 33 points only 7 distinct l1 (the axis-0 codes -3..3) and
// 7 distinct l2 occur, so a region needs 7 + 7 + 33 = 47 Pell, not 99.  A workgroup takes kT1Reg
// regions in three phases:
//   1. one thread per Pell (7 l1, 7 l2, 33 l3 per region; those of cut points skipped), the node order
//      of pell_nodes, so each value has the bits of a per-point evaluation;
//   2. one thread per point: the integrand from the cached Pell, in the reference's product order;
//   3. one thread per region: the rule (gm_rule3, the host rule's arithmetic order).
constexpr int kT1Reg = 5;                     // regions per workgroup: 235 Pell on 256 threads
constexpr int kT1Pell = 14 + TP_NPTS_GM3;     // Pell per region

template <bool F32>  // F32: the Pell from the FP32 node records (pell_nodes32)
__global__ void __launch_bounds__(256) k_t1_regions(const TpCosmoConst *__restrict__ cc,
                                                    const TpPNode *__restrict__ pnode,
                                                    const TpPNode32 *__restrict__ pnode32, double dx,
                                                    const TpPowerDesc *__restrict__ desc,
                                                    const TpRegion *__restrict__ reg, size_t nreg,
                                                    TpRegionOut *__restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ double pl[kT1Reg][kT1Pell];
    __shared__ double fv[kT1Reg][TP_NPTS_GM3];
    const size_t r0 = (size_t)blockIdx.x * kT1Reg;
    const int t = threadIdx.x;
    if (t < kT1Reg * kT1Pell) {
        const int rr = t / kT1Pell, k = t - rr * kT1Pell;
        const size_t r = r0 + rr;
        double v = 0;
        if (r < nreg) {
            const TpRegion R = reg[r];
            const TpPowerDesc &D = desc[R.integral];
            double ell;
            bool need;
            if (k < 14) {  // k = 7 axis + code + 3
                const int ax = k / 7;  // selected, not indexed: the region record stays in registers
                ell = tp_gm_coord(ax ? R.c[1] : R.c[0], ax ? R.hw[1] : R.hw[0], k - 7 * ax - 3);
                need = ell > D.lmin;
            } else {
                double l1, l2, phi;
                gm_point3(R, k - 14, &l1, &l2, &phi);
                ell = t1_l3(l1, l2, phi);
                need = !(l1 <= D.lmin || l2 <= D.lmin || ell <= D.lmin || ell > D.lmax);
            }
            if (need) {
                const double shot = 0.5 * D.sigma * D.sigma / D.ngal;
                if (F32)
                    v = pell_nodes32(pnode32 + (size_t)D.icosmo * TP_NNODES, ell, D.flags, dx, shot);
                else
                    v = pell_nodes(cc[D.icosmo], pnode + (size_t)D.icosmo * TP_NNODES, ell, D.flags, dx, shot);
            }
        }
        pl[rr][k] = v;
    }
    __syncthreads();
    if (t < kT1Reg * TP_NPTS_GM3) {
        const int rr = t / TP_NPTS_GM3, p = t - rr * TP_NPTS_GM3;
        const size_t r = r0 + rr;
        double v = 0;
        if (r < nreg) {
            const TpRegion R = reg[r];
            const TpPowerDesc &D = desc[R.integral];
            double l1, l2, phi;
            gm_point3(R, p, &l1, &l2, &phi);
            const double l3 = t1_l3(l1, l2, phi);
            if (!(l1 <= D.lmin || l2 <= D.lmin || l3 <= D.lmin || l3 > D.lmax)) {
                v = pl[rr][c_gm3_code[p][0] + 3] * pl[rr][c_gm3_code[p][1] + 10] * pl[rr][14 + p];
                v *= uhat(l1 * D.th[0]) * uhat(l2 * D.th[1]) * uhat(l3 * D.th[2]);
                v *= uhat(l1 * D.th[3]) * uhat(l2 * D.th[4]) * uhat(l3 * D.th[5]);
                v *= l1 * l2;
            }
        }
        fv[rr][p] = v;
    }
    __syncthreads();
    if (t < kT1Reg && r0 + t < nreg) {
        TpRegionOut o;
        gm_rule3(reg[r0 + t].vol, fv[t], &o);
        out[r0 + t] = o;
    }
}

// ---------------------------------------------------------------- 2-D regions on the device
// The 17 Genz-Malik points of a 2-D region in the reference's order (classed_version/hcubature.cpp:
// 456-513): centre; -+lambda2, -+lambda4 along each axis; the four (lambda4, lambda4) sign pairs;
// the four lambda5 corners in Gray-code order from (+, +).  Codes: 0 = centre, +-1 = lambda2,
// +-2 = lambda4, +-3 = lambda5 (sign = side).
__constant__ signed char c_gm2_code[17][2] = {{0, 0},   {-1, 0}, {1, 0},  {-2, 0}, {2, 0},  {0, -1},
                                              {0, 1},   {0, -2}, {0, 2},  {-2, -2}, {2, -2}, {2, 2},
                                              {-2, 2},  {3, 3},  {-3, 3}, {-3, -3}, {3, -3}};
constexpr int kGm2Pts = 17;

struct TpRegion2 {
    double c[2], hw[2];
    double vol;
    int integral;  // index into the integral descriptors of the call
    int pad;
};
struct TpRegion2Out {
    double val, err;
    int split;
    int flag;  // nonzero: non-finite rule estimate
};

// One thread per (region, rule point): the integrand at the point.
__global__ void __launch_bounds__(256) k_map2_region_points(const TpCosmoConst *__restrict__ cc,
                                                            const double *__restrict__ tab, int n_z,
                                                            const TpPowerDesc *__restrict__ desc,
                                                            const TpRegion2 *__restrict__ reg, size_t nreg,
                                                            double *__restrict__ f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nreg * kGm2Pts) return;
    const size_t r = i / kGm2Pts;
    const int p = (int)(i - r * kGm2Pts);
    const TpRegion2 R = reg[r];
    const TpPowerDesc D = desc[R.integral];
    const double ell = tp_gm_coord(R.c[0], R.hw[0], c_gm2_code[p][0]);
    const double z = tp_gm_coord(R.c[1], R.hw[1], c_gm2_code[p][1]);
    f[i] = tp_map2_point<double>(cc[D.icosmo], tab + (size_t)D.icosmo * 6 * n_z, n_z, ell, z, D.par, D.flags,
                                 D.sigma, D.ngal);
}

// One thread per region: value, error and split axis from its 17 values, in the arithmetic order of
// the host rule (classed_version/hcubature.cpp:515-575 with dim = 2, fdim = 1), contraction off.
__global__ void __launch_bounds__(256) k_map2_region_rule(size_t nreg, const TpRegion2 *__restrict__ reg,
                                                          const double *__restrict__ fv,
                                                          TpRegion2Out *__restrict__ out)
{
#pragma clang fp contract(off)
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    const double *v = fv + r * kGm2Pts;
    const double l2 = 0.3585685828003180919906451539079374954541;  // sqrt(9/70)
    const double l4 = 0.9486832980505137995996680633298155601160;  // sqrt(9/10)
    const double ratio = (l2 * l2) / (l4 * l4);
    const double w1 = (double)(12824 - 9120 * 2 + 400 * 4) / 19683.0;
    const double w3 = (double)(1820 - 400 * 2) / 19683.0;
    const double w5 = 6859.0 / 19683.0 / 4.0;
    const double wE1 = (double)(729 - 950 * 2 + 50 * 4) / 729.0;
    const double wE3 = (double)(265 - 100 * 2) / 1458.0;
    const double wt2 = 980. / 6561., wt4 = 200. / 19683., wtE2 = 245. / 486., wtE4 = 25. / 729.;
    const double val0 = v[0];
    double sum2 = 0, sum3 = 0, sum4 = 0, sum5 = 0, diff[2];
    for (int k = 0; k < 2; ++k) {
        const double a0 = v[1 + 4 * k], a1 = v[2 + 4 * k], a2 = v[3 + 4 * k], a3 = v[4 + 4 * k];
        sum2 += a0 + a1;
        sum3 += a2 + a3;
        diff[k] = fabs(a0 + a1 - 2 * val0 - ratio * (a2 + a3 - 2 * val0));
    }
    for (int k = 0; k < 4; ++k) sum4 += v[9 + k];
    for (int k = 0; k < 4; ++k) sum5 += v[13 + k];
    const double vol = reg[r].vol;
    const double result = vol * (w1 * val0 + wt2 * sum2 + w3 * sum3 + wt4 * sum4 + w5 * sum5);
    const double res5th = vol * (wE1 * val0 + wtE2 * sum2 + wE3 * sum3 + wtE4 * sum4);
    TpRegion2Out o;
    o.val = result;
    o.err = fabs(res5th - result);
    double md = 0;  // the first axis of largest difference, the host rule's scan
    o.split = 0;
    for (int i = 0; i < 2; ++i)
        if (diff[i] > md) {
            md = diff[i];
            o.split = i;
        }
    o.flag = (isfinite(result) && isfinite(res5th)) ? 0 : 1;
    out[r] = o;
}

// Pell: one wave per multipole.  Lane j < 48 evaluates the node pair j (z = cx -+ dx x_j); lane 0
// then adds the 48 pair sums in node order (the reference's loop order, GQ96_of_Pk).
__global__ void __launch_bounds__(64) k_pell(const TpCosmoConst *__restrict__ cc, const TpPNode *__restrict__ pnode,
                                             int icosmo, int flags, size_t n, const double *__restrict__ ell,
                                             double zmax, double shot, double *__restrict__ out, int *flag)
{
    __shared__ double part[48];
    const size_t i = blockIdx.x;
    const int j = threadIdx.x;
    if (i >= n) return;
    const TpCosmoConst &C = cc[icosmo];
    const TpPNode *N = pnode + (size_t)icosmo * TP_NNODES;
    const double e = ell[i];
    const double dx = zmax / 2;
    if (j < 48 && e != 0) {
#pragma clang fp contract(off)
        const double le = log(e);
        part[j] = c_pgl_w[j] * (tp_limber_pk_node(C, N[2 * j], e, le, flags) +
                                tp_limber_pk_node(C, N[2 * j + 1], e, le, flags));
    }
    __syncthreads();
    if (j == 0) {
#pragma clang fp contract(off)
        double r = shot;
        if (e != 0) {
            double q = 0;
            for (int k = 0; k < 48; ++k) q += part[k];
            r = q * dx + shot;
        }
        if (!isfinite(r)) atomicOr(flag, 1);
        out[i] = r;
    }
}

// The P Limber sum's z-nodes of every cosmology: node 2j at cx - dx x_j, node 2j+1 at cx + dx x_j.
__global__ void k_pnodes(int ncosmo, int n_z, const TpCosmoConst *__restrict__ cc, const double *__restrict__ tab,
                         double zmax, TpPNode *__restrict__ out, TpPNode32 *__restrict__ out32)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncosmo * TP_NNODES) return;
    const int ic = i / TP_NNODES, nd = i - ic * TP_NNODES, j = nd >> 1;
    const double cx = zmax / 2, dx = zmax / 2;
    const double z = (nd & 1) ? cx + dx * c_pgl_x[j] : cx - dx * c_pgl_x[j];
    tp_pnode_make(cc[ic], tab + (size_t)ic * 6 * n_z, n_z, z, &out[i]);
    tp_pnode32_make(cc[ic], out[i], &out32[i]);
}

int upload_pgl(int device)
{
    static bool done[64] = {false};
    if (device >= 0 && device < 64 && done[device]) return 0;
    double x[96], w[96], hx[48], hw[48];
    tp_gauss_legendre(96, x, w);
    // the reference's half-node order: x_j ascending in (0, 1) with their weights
    for (int j = 0; j < 48; ++j) {
        hx[j] = x[48 + j];
        hw[j] = w[48 + j];
    }
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_pgl_x), hx, sizeof(hx)));
    TP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_pgl_w), hw, sizeof(hw)));
    if (device >= 0 && device < 64) done[device] = true;
    return 0;
}

int check_args(tp_ctx *ctx, const tp_cosmo *c, int flags)
{
    if (!ctx || !c) return tp_set_error(TP_EINVAL, "NULL context or tables");
    if (flags & ~(TP_PK_LINEAR | TP_PK_T17 | TP_PK_SHAPENOISE_ONLY))
        return tp_set_error(TP_EINVAL, "unknown power-spectrum flags");
    return 0;
}

// Evaluates npts points x (dim doubles each) of the integrals desc (pdesc[i] per point, or all of
// desc[0] when pdesc is null); fval receives npts values.  Synchronous.
int power_launch(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpPowerDesc> &desc, const int *pdesc, size_t npts,
                 int dim, const double *x, double *fval)
{
    if (npts == 0) return 0;
    int rc;
    const size_t bx = npts * dim * sizeof(double), bf = npts * sizeof(double);
    const size_t bd = desc.size() * sizeof(TpPowerDesc), bp = pdesc ? npts * sizeof(int) : 0;
    if ((rc = ctx->d_in.ensure(bx)) || (rc = ctx->d_out.ensure(bf)) || (rc = ctx->d_desc.ensure(bd)) ||
        (rc = ctx->d_misc.ensure(std::max(bp, (size_t)4))) || (rc = ctx->h_in.ensure(bx + bp)) ||
        (rc = ctx->h_out.ensure(bf)))
        return rc;
    std::memcpy(ctx->h_in.p, x, bx);
    if (pdesc) std::memcpy((char *)ctx->h_in.p + bx, pdesc, bp);
    TP_HIP(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(int), ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_in.p, ctx->h_in.p, bx, hipMemcpyHostToDevice, ctx->stream));
    if (pdesc)
        TP_HIP(hipMemcpyAsync(ctx->d_misc.p, (char *)ctx->h_in.p + bx, bp, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_desc.p, desc.data(), bd, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const unsigned blk = 256;
    hipLaunchKernelGGL(k_power_points, dim3((unsigned)((npts + blk - 1) / blk)), dim3(blk), 0, ctx->stream,
                       c->d_const, c->d_tab, c->n_z, (const TpPowerDesc *)ctx->d_desc.p,
                       pdesc ? (const int *)ctx->d_misc.p : nullptr, npts, dim, (const double *)ctx->d_in.p,
                       (double *)ctx->d_out.p, (int *)ctx->d_flag.p, (const TpPNode *)c->d_pnode, c->z_max / 2);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, bf, hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(fval, ctx->h_out.p, bf);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats.kernel_ms += ms;
    ctx->stats.integrand_ms += ms;
    ctx->stats.integrand_sum_ms += ms;
    ctx->stats.launches += 1;
    if (flag) return tp_set_error(TP_ENONFINITE, "power spectrum integrand: non-finite value");
    return 0;
}

// All integrals of one call advance together, one launch per round (see the file comment).
template <unsigned D, class Post>
int power_batch(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpPowerDesc> &desc, const double *lo,
                const double *hi, double reltol, Post post, double *out, double *err, long long *neval)
{
    auto t0 = std::chrono::steady_clock::now();
    ctx->stats = tp_batch_stats{};
    const int nt = tp_host_threads(ctx, kTpHostThreadsSync);
    const size_t nint = desc.size();
    std::vector<tpc::AdaptiveIntegralT<D, 1>> I(nint);
    for (size_t k = 0; k < nint; ++k)
        if (!I[k].init(D, 1, lo + D * k, hi + D * k, 0, 0, reltol, tpc::L1))
            return tp_set_error(TP_EINVAL, "bad integration domain");
    const unsigned np = tpc::rule_points(D);
    std::vector<size_t> active(nint), offs;
    for (size_t k = 0; k < nint; ++k) active[k] = k;
    std::vector<double> pts, vals;
    std::vector<int> pdesc;
    long long regions = 0;
    const size_t budget = (size_t)std::max(1LL, ctx->round_points);
    while (!active.empty()) {
        if (ctx->stats.rounds >= ctx->max_rounds)
            return tp_set_error(TP_ECAPACITY, "batched cubature: round limit reached (integrals not converging)");
        const size_t na = active.size();
        offs.assign(na + 1, 0);
        for (size_t a = 0; a < na; ++a) offs[a + 1] = offs[a] + I[active[a]].pending().size() * np;
        regions += (long long)(offs[na] / np);
        if (regions > ctx->max_regions)
            return tp_set_error(TP_ECAPACITY, "batched cubature: region limit reached (integrals not converging)");
        std::vector<char> fin(na, 0);
        // the round in slices of whole integrals of at most `budget` points (at least one integral), so
        // the point and value buffers stay bounded however many regions the integrals hold
        for (size_t a0 = 0; a0 < na;) {
            size_t a1 = a0 + 1;
            while (a1 < na && offs[a1 + 1] - offs[a0] <= budget) ++a1;
            const size_t base = offs[a0], npts = offs[a1] - base;
            pts.resize(npts * D);
            vals.resize(npts);
            pdesc.resize(npts);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) if (a1 - a0 > 64)
            for (size_t a = a0; a < a1; ++a) {
                auto &A = I[active[a]];
                const auto &pend = A.pending();
                for (size_t r = 0; r < pend.size(); ++r) {
                    const size_t o = offs[a] - base + r * np;
                    tpc::rule_points_of(D, A.regions()[pend[r]], pts.data() + D * o);
                    for (unsigned q = 0; q < np; ++q) pdesc[o + q] = (int)active[a];
                }
            }
            int rc = power_launch(ctx, c, desc, pdesc.data(), npts, (int)D, pts.data(), vals.data());
            if (rc) return rc;
            ctx->stats.points += (long long)npts;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) if (a1 - a0 > 64)
            for (size_t a = a0; a < a1; ++a) {
                auto &A = I[active[a]];
                const auto &pend = A.pending();
                for (size_t r = 0; r < pend.size(); ++r)
                    tpc::rule_apply(D, 1, A.regions()[pend[r]], vals.data() + offs[a] - base + r * np);
                A.absorb();
                fin[a] = A.advance() ? 1 : 0;
            }
            a0 = a1;
        }
        ctx->stats.rounds += 1;
        std::vector<size_t> next;
        for (size_t a = 0; a < na; ++a)
            if (!fin[a]) next.push_back(active[a]);
        active.swap(next);
    }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) if (nint > 256)
    for (size_t k = 0; k < nint; ++k) {
        double v, e;
        I[k].result(&v, &e);
        out[k] = post(v);
        if (err) err[k] = post(e);
        if (neval) neval[k] = (long long)I[k].num_eval();
    }
    ctx->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// The context's kept integrators of a dimension (their pools keep capacity across calls).
template <unsigned D>
std::vector<tpc::AdaptiveIntegralT<D, 1>> &kept_integrals(tp_ctx *ctx);
template <>
std::vector<tpc::AdaptiveIntegralT<2, 1>> &kept_integrals<2>(tp_ctx *ctx)
{
    return ctx->map2_keep;
}
template <>
std::vector<tpc::AdaptiveIntegralT<3, 1>> &kept_integrals<3>(tp_ctx *ctx)
{
    return ctx->map3_keep;
}

// Region records of the device-rule batches: the 2-D form for <Map^2>, the engine's 3-D form for T1.
template <unsigned D>
struct RegionIO;
template <>
struct RegionIO<2> {
    using In = TpRegion2;
    using Out = TpRegion2Out;
    static constexpr int kPts = kGm2Pts;
};
template <>
struct RegionIO<3> {
    using In = TpRegion;
    using Out = TpRegionOut;
    static constexpr int kPts = TP_NPTS_GM3;
};

// As power_batch, with the rule points generated and the rule applied on the device, so a round moves
// one region record in and one result out per region (not the rule's points and values) and the host
// only keeps the heaps.  The integrals are split into two groups, each with its own stream and buffers
// (the context's lanes), that alternate as in tp_map3_batch: while the device evaluates one group's
// regions and copies them over, the host absorbs the other group's results.  Each integral's decision
// sequence is unchanged.  launch(nreg, d_regions, d_out, d_fvals, stream, evm) queues a round's kernels
// on `stream` and records evm after the integrand; fv_per_region doubles of d_fvals per region are
// reserved for kernels that pass the point values between launches.
template <unsigned D, class Launch, class Post>
int regions_batch(tp_ctx *ctx, const std::vector<TpPowerDesc> &desc, const double *lo, const double *hi,
                  double reltol, size_t fv_per_region, Launch launch, Post post, const char *what, double *out,
                  double *err, long long *neval)
{
    using In = typename RegionIO<D>::In;
    using Out = typename RegionIO<D>::Out;
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    ctx->stats = tp_batch_stats{};
    const int nt = tp_host_threads(ctx, kTpHostThreadsSync);
    const size_t nint = desc.size();
    std::vector<tpc::AdaptiveIntegralT<D, 1>> I;
    I.swap(kept_integrals<D>(ctx));  // the last call's integrators, if kept
    if (I.size() < nint) I.resize(nint);
    for (size_t k = 0; k < nint; ++k)
        if (!I[k].init(D, 1, lo + D * k, hi + D * k, 0, 0, reltol, tpc::L1))
            return tp_set_error(TP_EINVAL, "bad integration domain");
    int rc;
    const size_t bd = nint * sizeof(TpPowerDesc);
    if ((rc = ctx->d_desc.ensure(std::max(bd, (size_t)64)))) return rc;
    TP_HIP(hipMemcpyAsync(ctx->d_desc.p, desc.data(), bd, hipMemcpyHostToDevice, ctx->stream));
    const int nl = nint > 1 ? 2 : 1;
    struct Group {
        std::vector<size_t> active, offs;
        size_t nreg = 0;
        bool inflight = false;
    };
    Group G[2];
    for (size_t k = 0; k < nint; ++k) G[nl == 1 ? 0 : (int)(k & 1)].active.push_back(k);
    hipStream_t gst[2];
    TP_HIP(hipEventRecord(ctx->ev_batch, ctx->stream));
    for (int g = 0; g < nl; ++g) {
        if ((rc = ctx->lane[g].create())) return rc;
        gst[g] = nl == 1 ? ctx->stream : ctx->lane[g].stream;
        if (nl > 1) TP_HIP(hipStreamWaitEvent(gst[g], ctx->ev_batch, 0));
    }
    std::vector<std::pair<float, float>> iv_pts, iv_ker;  // device intervals of the rounds [ms]

    long long total_regions = 0;
    auto issue = [&](int g) -> int {
        Group &Gr = G[g];
        TpLane &L = ctx->lane[g];
        auto th0 = clk::now();
        const std::vector<size_t> &act = Gr.active;
        const size_t na = act.size();
        Gr.offs.assign(na + 1, 0);
        for (size_t a = 0; a < na; ++a) Gr.offs[a + 1] = Gr.offs[a] + I[act[a]].pending().size();
        const size_t nreg = Gr.offs[na];
        Gr.nreg = nreg;
        total_regions += (long long)nreg;
        if (total_regions > ctx->max_regions)
            return tp_set_error(TP_ECAPACITY, std::string(what) + ": region limit reached (integrals not converging)");
        const size_t br = nreg * sizeof(In), bo = nreg * sizeof(Out);
        int e;
        if ((e = L.d_regions.ensure(br)) || (e = L.d_rout.ensure(bo)) ||
            (e = L.d_fvals.ensure(std::max(nreg * fv_per_region * sizeof(double), (size_t)64))) ||
            (e = L.h_regions.ensure(br)) || (e = L.h_rout.ensure(bo)))
            return e;
        In *hr = (In *)L.h_regions.p;
        const std::vector<size_t> &offs = Gr.offs;
        // static schedules in both loops: a thread keeps the same integrals from loop to loop and round to
        // round, so their region pools stay in its caches (dynamic chunks: ~28% slower on 5 120 <Map^2>)
#pragma omp parallel for schedule(static) num_threads(nt) if (na > 1 && (na > 256 || nreg > 2048))
        for (size_t a = 0; a < na; ++a) {
            auto &A = I[act[a]];
            const auto &pend = A.pending();
            for (size_t q = 0; q < pend.size(); ++q) {
                const auto &R = A.regions()[pend[q]];
                In &o = hr[offs[a] + q];
                for (unsigned d = 0; d < D; ++d) {
                    o.c[d] = R.c[d];
                    o.hw[d] = R.hw[d];
                }
                o.vol = R.vol;
                o.integral = (int)act[a];
                o.pad = 0;
            }
        }
        ctx->stats.host_ms += ms_since(th0);
        TP_HIP(hipMemcpyAsync(L.d_regions.p, hr, br, hipMemcpyHostToDevice, gst[g]));
        TP_HIP(hipEventRecord(L.ev0, gst[g]));
        if ((e = launch(nreg, (const In *)L.d_regions.p, (Out *)L.d_rout.p, (double *)L.d_fvals.p, gst[g], L.evm)))
            return e;
        TP_HIP(hipEventRecord(L.ev1, gst[g]));
        TP_HIP(hipMemcpyAsync(L.h_rout.p, L.d_rout.p, bo, hipMemcpyDeviceToHost, gst[g]));
        TP_HIP(hipEventRecord(L.done, gst[g]));
        Gr.inflight = true;
        ctx->stats.launches += 1;
        ctx->stats.regions += (long long)nreg;
        ctx->stats.points += (long long)nreg * RegionIO<D>::kPts;
        return 0;
    };
    auto complete = [&](int g) -> int {
        Group &Gr = G[g];
        TpLane &L = ctx->lane[g];
        auto tw = clk::now();
        TP_HIP(hipEventSynchronize(L.done));
        ctx->stats.wait_ms += ms_since(tw);
        Gr.inflight = false;
        float ta = 0, tm = 0, tb = 0;
        TP_HIP(hipEventElapsedTime(&ta, ctx->ev_batch, L.ev0));
        TP_HIP(hipEventElapsedTime(&tm, ctx->ev_batch, L.evm));
        TP_HIP(hipEventElapsedTime(&tb, ctx->ev_batch, L.ev1));
        iv_pts.push_back({ta, tm});
        iv_ker.push_back({ta, tb});
        ctx->stats.integrand_sum_ms += tm - ta;
        auto th0 = clk::now();
        const Out *ho = (const Out *)L.h_rout.p;
        const std::vector<size_t> &act = Gr.active;
        const std::vector<size_t> &offs = Gr.offs;
        const size_t na = act.size();
        std::vector<char> fin(na, 0);
        int bad = 0;
#pragma omp parallel for schedule(static) num_threads(nt) if (na > 1 && (na > 256 || Gr.nreg > 2048)) reduction(| : bad)
        for (size_t a = 0; a < na; ++a) {
            auto &A = I[act[a]];
            const auto &pend = A.pending();
            for (size_t q = 0; q < pend.size(); ++q) {
                auto &R = A.regions()[pend[q]];
                const Out &o = ho[offs[a] + q];
                R.val[0] = o.val;
                R.err[0] = o.err;
                R.split = (unsigned)o.split;
                R.errmax = o.err;
                R.corr = 0.0;
                R.cvar = 0.0;
                bad |= o.flag;
            }
            A.absorb();
            fin[a] = A.advance() ? 1 : 0;
        }
        if (bad) return tp_set_error(TP_ENONFINITE, std::string(what) + ": non-finite value");
        std::vector<size_t> next;
        next.reserve(na);
        for (size_t a = 0; a < na; ++a)
            if (!fin[a]) next.push_back(act[a]);
        Gr.active.swap(next);
        ctx->stats.host_ms += ms_since(th0);
        return 0;
    };
    // on an error, let queued work drain before the buffers can be reused
    auto fail = [&](int e) {
        for (int g = 0; g < nl; ++g) (void)hipStreamSynchronize(gst[g]);
        return e;
    };
    long long rounds[2] = {0, 0};
    for (int g = 0; g < nl; ++g)
        if (!G[g].active.empty()) {
            if ((rc = issue(g))) return fail(rc);
            ++rounds[g];
        }
    for (;;) {
        bool any = false;
        for (int g = 0; g < nl; ++g) {
            if (!G[g].inflight) continue;
            any = true;
            if ((rc = complete(g))) return fail(rc);
            if (!G[g].active.empty()) {
                if (rounds[g] >= ctx->max_rounds)
                    return fail(tp_set_error(TP_ECAPACITY, std::string(what) +
                                                               ": round limit reached (integrals not converging)"));
                if ((rc = issue(g))) return fail(rc);
                ++rounds[g];
            }
        }
        if (!any) break;
    }
    ctx->stats.rounds = std::max(rounds[0], rounds[1]);
    auto union_ms = [](std::vector<std::pair<float, float>> &iv) {
        std::sort(iv.begin(), iv.end());
        double tot = 0, ua = 0, ub = -1;
        for (const auto &x : iv) {
            if (x.first > ub) {
                if (ub > ua) tot += ub - ua;
                ua = x.first;
                ub = x.second;
            } else if (x.second > ub) {
                ub = x.second;
            }
        }
        if (ub > ua) tot += ub - ua;
        return tot;
    };
    ctx->stats.kernel_ms = union_ms(iv_ker);
    ctx->stats.integrand_ms = union_ms(iv_pts);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt) if (nint > 256)
    for (size_t k = 0; k < nint; ++k) {
        double v, e;
        I[k].result(&v, &e);
        out[k] = post(v);
        if (err) err[k] = post(e);
        if (neval) neval[k] = (long long)I[k].num_eval();
    }
    {
        size_t held = 0;
        for (const auto &A : I) held += A.bytes_reserved();
        if (held <= ctx->keep_pool_bytes) kept_integrals<D>(ctx).swap(I);
    }
    ctx->stats.wall_ms = ms_since(t0);
    return 0;
}

// <Map^2> rounds: one thread per (region, point), then one thread per region for the rule.
int map2_regions(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpPowerDesc> &desc, const double *lo,
                 const double *hi, double reltol, double *out, double *err, long long *neval)
{
    auto launch = [&](size_t nreg, const TpRegion2 *d_reg, TpRegion2Out *d_out, double *d_fv, hipStream_t st,
                      hipEvent_t evm) -> int {
        const size_t npts = nreg * kGm2Pts;
        hipLaunchKernelGGL(k_map2_region_points, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, c->d_const,
                           c->d_tab, c->n_z, (const TpPowerDesc *)ctx->d_desc.p, d_reg, nreg, d_fv);
        TP_HIP(hipGetLastError());
        TP_HIP(hipEventRecord(evm, st));
        hipLaunchKernelGGL(k_map2_region_rule, dim3((unsigned)((nreg + 255) / 256)), dim3(256), 0, st, nreg, d_reg,
                           (const double *)d_fv, d_out);
        TP_HIP(hipGetLastError());
        return 0;
    };
    return regions_batch<2>(ctx, desc, lo, hi, reltol, kGm2Pts, launch,
                            [](double v) { return v / 2. / tpd::kPi; },  // apertureStatistics.cu:135
                            "<Map^2> integrand", out, err, neval);
}

// T1 rounds: one launch of k_t1_regions (points, cached Pell and rule in one kernel).
int t1_regions(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpPowerDesc> &desc, const double *lo,
               const double *hi, double reltol, double area, int mode, double *out, long long *neval)
{
    auto launch = [&](size_t nreg, const TpRegion *d_reg, TpRegionOut *d_out, double *, hipStream_t st,
                      hipEvent_t evm) -> int {
        const size_t nblk = (nreg + kT1Reg - 1) / kT1Reg;
        if (nblk > 0x7fffffffu) return tp_set_error(TP_ECAPACITY, "T1: too many regions for one launch");
        hipLaunchKernelGGL(mode == TP_MODE_FP32 ? k_t1_regions<true> : k_t1_regions<false>, dim3((unsigned)nblk),
                           dim3(256), 0, st, c->d_const, (const TpPNode *)c->d_pnode,
                           (const TpPNode32 *)c->d_pnode32, c->z_max / 2, (const TpPowerDesc *)ctx->d_desc.p, d_reg,
                           nreg, d_out);
        TP_HIP(hipGetLastError());
        TP_HIP(hipEventRecord(evm, st));
        return 0;
    };
    std::vector<double> e(desc.size());
    return regions_batch<3>(
        ctx, desc, lo, hi, reltol, 0, launch,
        [area](double x) { return x * (2 / area / pow(2 * tpd::kPi, 3)); },  // apertureStatisticsCovariance.cu:1089
        "T1 integrand", out, e.data(), neval);
}

int point_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, int kind, size_t npts, const double *in, int flags,
               double *out)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (npts && (!in || !out)) return tp_set_error(TP_EINVAL, "NULL point or output array");
    if (icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "cosmology index out of range");
    TP_HIP(hipSetDevice(ctx->device));
    std::vector<TpPowerDesc> desc(1, TpPowerDesc{0, 0, 1, icosmo, kind, flags, 0});
    return power_launch(ctx, c, desc, nullptr, npts, 2, in, out);
}

}  // namespace

extern "C" int tp_pk_at_z_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *kz, int flags,
                               double *out)
{
    return point_eval(ctx, c, icosmo, kPkAtZ, npts, kz, flags, out);
}

extern "C" int tp_limber_pkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ellz,
                                     int flags, double *out)
{
    return point_eval(ctx, c, icosmo, kLimber, npts, ellz, flags, out);
}

extern "C" int tp_pell(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t n, const double *ell, int flags,
                       double sigma, double n_gal, double *out)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (n && (!ell || !out)) return tp_set_error(TP_EINVAL, "tp_pell: NULL array");
    if (icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_pell: cosmology index out of range");
    if (n == 0) return 0;
    if (n > 0x7fffffffu) return tp_set_error(TP_EINVAL, "tp_pell: too many multipoles for one call");
    TP_HIP(hipSetDevice(ctx->device));
    if ((rc = upload_pgl(ctx->device))) return rc;
    const double shot = 0.5 * sigma * sigma / n_gal;  // bispectrum.cpp:1050
    const size_t b = n * sizeof(double);
    if ((rc = ctx->d_in.ensure(b)) || (rc = ctx->d_out.ensure(b))) return rc;
    TP_HIP(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(int), ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_in.p, ell, b, hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(k_pell, dim3((unsigned)n), dim3(64), 0, ctx->stream, c->d_const, (const TpPNode *)c->d_pnode,
                       icosmo, flags, n, (const double *)ctx->d_in.p, c->z_max, shot, (double *)ctx->d_out.p,
                       (int *)ctx->d_flag.p);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    TP_HIP(hipMemcpyAsync(out, ctx->d_out.p, b, hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats = tp_batch_stats{};
    ctx->stats.kernel_ms = ctx->stats.integrand_ms = ctx->stats.integrand_sum_ms = ms;
    ctx->stats.launches = 1;
    ctx->stats.points = (long long)n * 96;
    if (flag) return tp_set_error(TP_ENONFINITE, "tp_pell: non-finite value");
    return 0;
}

extern "C" int tp_pkappa_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *ell,
                               double reltol, int flags, double *out, double *err, long long *neval)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (n && (!ell || !out)) return tp_set_error(TP_EINVAL, "tp_pkappa_batch: NULL array");
    if (!(reltol > 0)) return tp_set_error(TP_EINVAL, "tp_pkappa_batch: reltol must be positive");
    if (n > 0x7fffffffu) return tp_set_error(TP_EINVAL, "tp_pkappa_batch: too many integrals for one call");
    TP_HIP(hipSetDevice(ctx->device));
    std::vector<TpPowerDesc> desc(n);
    std::vector<double> lo(n, 0.0), hi(n, c->z_max);
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_pkappa_batch: cosmology index out of range");
        desc[i] = TpPowerDesc{ell[i], 0, 1, ic, kPkappa, flags, 0};
    }
    return power_batch<1>(ctx, c, desc, lo.data(), hi.data(), reltol, [](double v) { return v; }, out, err, neval);
}

extern "C" int tp_map2_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta,
                             double reltol, int flags, double sigma, double ngal_arcmin2, double *out, double *err,
                             long long *neval)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (n && (!theta || !out)) return tp_set_error(TP_EINVAL, "tp_map2_batch: NULL array");
    if (!(reltol > 0)) return tp_set_error(TP_EINVAL, "tp_map2_batch: reltol must be positive");
    if (sigma != 0 && !(ngal_arcmin2 > 0)) return tp_set_error(TP_EINVAL, "tp_map2_batch: ngal must be positive");
    if (n > 0x7fffffffu) return tp_set_error(TP_EINVAL, "tp_map2_batch: too many integrals for one call");
    TP_HIP(hipSetDevice(ctx->device));
    std::vector<TpPowerDesc> desc(n);
    std::vector<double> lo(2 * n), hi(2 * n);
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_map2_batch: cosmology index out of range");
        if (!(theta[i] > 0)) return tp_set_error(TP_EINVAL, "tp_map2_batch: aperture radii must be positive");
        desc[i] = TpPowerDesc{theta[i], sigma, sigma != 0 ? ngal_arcmin2 : 1.0, ic, kMap2, flags, 0};
        lo[2 * i] = 1.;                // ell range [1, 10 / theta] (apertureStatistics.cu:112-113)
        hi[2 * i] = 10. / theta[i];
        lo[2 * i + 1] = 0;             // z range [0, z_max]
        hi[2 * i + 1] = c->z_max;
    }
    return map2_regions(ctx, c, desc, lo.data(), hi.data(), reltol, out, err, neval);
}

int tp_build_pnodes(tp_ctx *ctx, tp_cosmo *c)
{
    int rc = upload_pgl(ctx->device);
    if (rc) return rc;
    const size_t n = (size_t)c->ncosmo * TP_NNODES;
    const size_t b64 = (n * sizeof(TpPNode) + 255) / 256 * 256;
    TP_HIP(hipMalloc(&c->d_pnode, b64 + n * sizeof(TpPNode32)));
    c->d_pnode32 = (char *)c->d_pnode + b64;
    hipLaunchKernelGGL(k_pnodes, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream, c->ncosmo, c->n_z,
                       c->d_const, c->d_tab, c->z_max, (TpPNode *)c->d_pnode, (TpPNode32 *)c->d_pnode32);
    TP_HIP(hipGetLastError());
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

extern "C" int tp_map3_cov_gauss_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo,
                                       const double *theta6, double survey_area, double reltol, double *out,
                                       double *err, long long *neval)
{
    int rc = check_args(ctx, c, 0);
    if (rc) return rc;
    if (n && (!theta6 || !out)) return tp_set_error(TP_EINVAL, "tp_map3_cov_gauss_batch: NULL array");
    if (!(reltol > 0) || !(survey_area > 0))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_gauss_batch: reltol and survey_area must be positive");
    if (n > 0x7fffffffu) return tp_set_error(TP_EINVAL, "tp_map3_cov_gauss_batch: too many integrals for one call");
    TP_HIP(hipSetDevice(ctx->device));
    std::vector<TpPowerDesc> desc(n);
    std::vector<double> lo(4 * n), hi(4 * n);
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo)
            return tp_set_error(TP_EINVAL, "tp_map3_cov_gauss_batch: cosmology index out of range");
        TpPowerDesc D{};
        D.icosmo = ic;
        D.kind = kCovGauss;
        const double *t = theta6 + 6 * i;
        for (int k = 0; k < 6; ++k) {
            if (!(t[k] > 0)) return tp_set_error(TP_EINVAL, "tp_map3_cov_gauss_batch: aperture radii must be positive");
            D.th[k] = t[k];
        }
        desc[i] = D;
        // apertureStatistics.cpp:358-380: lMax = 10 / max(min theta_123, min theta_456), phi in [0, pi]
        const double lmax = 10. / std::max(std::min({t[0], t[1], t[2]}), std::min({t[3], t[4], t[5]}));
        const double l[4] = {1., 1., 0., 0.}, h[4] = {lmax, lmax, 2 * tpd::kPi / 2., c->z_max};
        for (int k = 0; k < 4; ++k) {
            lo[4 * i + k] = l[k];
            hi[4 * i + k] = h[k];
        }
    }
    const double A = survey_area;
    return power_batch<4>(
        ctx, c, desc, lo.data(), hi.data(), reltol,
        [A](double v) { return 2. * v / A / 8 / tpd::kPi / tpd::kPi / tpd::kPi; },  // apertureStatistics.cpp:395
        out, err, neval);
}

extern "C" int tp_map3_cov_t1_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo,
                                    const double *theta6, double area, double lmin, double sigma, double n_gal,
                                    int flags, double reltol, double *out, long long *neval)
{
    return tp_map3_cov_t1_batch_mode(ctx, c, n, icosmo, theta6, area, lmin, sigma, n_gal, flags, reltol,
                                     TP_MODE_FP64, out, neval);
}

extern "C" int tp_map3_cov_t1_batch_mode(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo,
                                         const double *theta6, double area, double lmin, double sigma, double n_gal,
                                         int flags, double reltol, int mode, double *out, long long *neval)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (mode != TP_MODE_FP64 && mode != TP_MODE_FP32)
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_mode: mode must be TP_MODE_FP64 or TP_MODE_FP32");
    if (mode == TP_MODE_FP32 && c->pk_nk)
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_mode: the FP32 form uses the Eisenstein & Hu P(k); "
                                       "tables built from a tabulated P(k) need TP_MODE_FP64");
    if (n && (!theta6 || !out)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch: NULL array");
    if (!(reltol > 0) || !(area > 0) || !(n_gal > 0) || !(lmin >= 0))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch: reltol, area and n_gal must be positive, lmin >= 0");
    if (n > 0x7fffffffu / 6) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch: too many pairs for one call");
    TP_HIP(hipSetDevice(ctx->device));
    // T1_total (apertureStatisticsCovariance.cu:272-326): each pair is a sum of T1 integrals over the
    // distinct permutations of theta_456, with multiplicities; terms[i] lists (coefficient, integral).
    std::vector<TpPowerDesc> desc;
    std::vector<double> lo, hi;
    std::vector<std::vector<std::pair<int, size_t>>> terms(n);
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch: cosmology index out of range");
        const double *t = theta6 + 6 * i;
        for (int k = 0; k < 6; ++k)
            if (!(t[k] > 0)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch: aperture radii must be positive");
        const double th3 = t[3], th4 = t[4], th5 = t[5];
        auto add = [&](int coef, double a, double b, double d) {
            TpPowerDesc D{};
            D.icosmo = ic;
            D.kind = kT1;
            D.flags = flags;
            D.sigma = sigma;
            D.ngal = n_gal;
            D.th[0] = t[0];
            D.th[1] = t[1];
            D.th[2] = t[2];
            D.th[3] = a;
            D.th[4] = b;
            D.th[5] = d;
            D.lmin = lmin;
            // T1 (apertureStatisticsCovariance.cu:1065-1093): lMax = 10 / max(min theta_123, min theta_456)
            D.lmax = 10. / std::max(std::min({t[0], t[1], t[2]}), std::min({a, b, d}));
            terms[i].push_back({coef, desc.size()});
            desc.push_back(D);
            lo.insert(lo.end(), {lmin, lmin, 0.});
            hi.insert(hi.end(), {D.lmax, D.lmax, tpd::kPi});
        };
        if (th3 == th4) {
            if (th4 == th5) {
                add(6, th3, th3, th3);
            } else {
                add(2, th3, th3, th5);
                add(2, th3, th5, th3);
                add(2, th5, th3, th3);
            }
        } else if (th3 == th5) {
            add(2, th3, th4, th3);
            add(2, th3, th3, th4);
            add(2, th4, th3, th3);
        } else if (th4 == th5) {
            add(2, th3, th4, th4);
            add(2, th4, th3, th4);
            add(2, th4, th4, th3);
        } else {
            add(1, th3, th4, th5);
            add(1, th3, th5, th4);
            add(1, th4, th3, th5);
            add(1, th4, th5, th3);
            add(1, th5, th3, th4);
            add(1, th5, th4, th3);
        }
    }
    std::vector<double> v(desc.size());
    std::vector<long long> ne(desc.size());
    rc = t1_regions(ctx, c, desc, lo.data(), hi.data(), reltol, area, mode, v.data(), ne.data());
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        double r = 0;
        long long np = 0;
        for (size_t q = 0; q < terms[i].size(); ++q) {
            const double tq = terms[i][q].first == 1 ? v[terms[i][q].second] : terms[i][q].first * v[terms[i][q].second];
            r = q == 0 ? tq : r + tq;
            np += ne[terms[i][q].second];
        }
        out[i] = r;
        if (neval) neval[i] = np;
    }
    return 0;
}

extern "C" int tp_map3_cov_t1_batch_geometry(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo,
                                             const double *theta6, int geometry, double theta_max,
                                             double theta_max2, double area, double lmin, double sigma,
                                             double n_gal, int flags, double reltol, double *out,
                                             long long *neval)
{
    if (geometry == 2)
        return tp_map3_cov_t1_batch(ctx, c, n, icosmo, theta6, area, lmin, sigma, n_gal, flags, reltol, out, neval);
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (geometry != 0 && geometry != 1 && geometry != 3)
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: geometry must be 0 (circle), 1 (square), "
                                       "2 (infinite) or 3 (rectangle)");
    if (n && (!theta6 || !out)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: NULL array");
    if (!(reltol > 0) || !(n_gal > 0) || !(lmin >= 0))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: reltol and n_gal must be positive, lmin >= 0");
    if (!(theta_max > 0) || (geometry == 3 && !(theta_max2 > 0)))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: the survey's side lengths or radius must "
                                       "be positive");
    if (n > 0x7fffffffu / 6) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: too many pairs");
    TP_HIP(hipSetDevice(ctx->device));
    const int kind = geometry == 0 ? kT1Circle : geometry == 1 ? kT1Square : kT1Rect;
    std::vector<TpPowerDesc> desc;
    std::vector<double> lo, hi;
    std::vector<std::vector<std::pair<int, size_t>>> terms(n);
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo)
            return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: cosmology index out of range");
        const double *t = theta6 + 6 * i;
        for (int k = 0; k < 6; ++k)
            if (!(t[k] > 0)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t1_batch_geometry: aperture radii must be positive");
        const double th3 = t[3], th4 = t[4], th5 = t[5];
        auto add = [&](int coef, double a, double b, double d) {
            TpPowerDesc D{};
            D.icosmo = ic;
            D.kind = kind;
            D.flags = flags;
            D.sigma = sigma;
            D.ngal = n_gal;
            D.th[0] = t[0];
            D.th[1] = t[1];
            D.th[2] = t[2];
            D.th[3] = a;
            D.th[4] = b;
            D.th[5] = d;
            D.lmin = lmin;
            D.wx = theta_max;
            D.wy = theta_max2;
            // T1 (apertureStatisticsCovariance.cu:1065-1102): lMax = 10 / max(min theta_123, min theta_456),
            // every axis over [-lMax, 1.02 lMax]
            D.lmax = 10. / std::max(std::min({t[0], t[1], t[2]}), std::min({a, b, d}));
            terms[i].push_back({coef, desc.size()});
            desc.push_back(D);
            for (int q = 0; q < 6; ++q) {
                lo.push_back(-D.lmax);
                hi.push_back(1.02 * D.lmax);
            }
        };
        // T1_total's permutations of theta_456 (apertureStatisticsCovariance.cu:272-326)
        if (th3 == th4) {
            if (th4 == th5) {
                add(6, th3, th3, th3);
            } else {
                add(2, th3, th3, th5);
                add(2, th3, th5, th3);
                add(2, th5, th3, th3);
            }
        } else if (th3 == th5) {
            add(2, th3, th4, th3);
            add(2, th3, th3, th4);
            add(2, th4, th3, th3);
        } else if (th4 == th5) {
            add(2, th3, th4, th4);
            add(2, th4, th3, th4);
            add(2, th4, th4, th3);
        } else {
            add(1, th3, th4, th5);
            add(1, th3, th5, th4);
            add(1, th4, th3, th5);
            add(1, th4, th5, th3);
            add(1, th5, th3, th4);
            add(1, th5, th4, th3);
        }
    }
    std::vector<double> v(desc.size());
    std::vector<long long> ne(desc.size());
    const double norm = std::pow(2 * tpd::kPi, 6);
    // the 6-D integrals hold ~1e4-1e5 regions each: a chunk of kChunk integrals at a time bounds the
    // host heaps (each chunk's integrators are released before the next one starts)
    const size_t kChunk = 256;
    tp_batch_stats tot{};
    for (size_t k0 = 0; k0 < desc.size(); k0 += kChunk) {
        const size_t k1 = std::min(desc.size(), k0 + kChunk);
        const std::vector<TpPowerDesc> dk(desc.begin() + k0, desc.begin() + k1);
        rc = power_batch<6>(ctx, c, dk, lo.data() + 6 * k0, hi.data() + 6 * k0, reltol,
                            [norm](double x) { return x / norm; }, v.data() + k0, nullptr, ne.data() + k0);
        if (rc) return rc;
        tot.rounds += ctx->stats.rounds;
        tot.points += ctx->stats.points;
        tot.launches += ctx->stats.launches;
        tot.kernel_ms += ctx->stats.kernel_ms;
        tot.integrand_ms += ctx->stats.integrand_ms;
        tot.integrand_sum_ms += ctx->stats.integrand_sum_ms;
        tot.wall_ms += ctx->stats.wall_ms;
    }
    ctx->stats = tot;
    for (size_t i = 0; i < n; ++i) {
        double r = 0;
        long long np = 0;
        for (size_t q = 0; q < terms[i].size(); ++q) {
            const double tq = terms[i][q].first == 1 ? v[terms[i][q].second] : terms[i][q].first * v[terms[i][q].second];
            r = q == 0 ? tq : r + tq;
            np += ne[terms[i][q].second];
        }
        out[i] = r;
        if (neval) neval[i] = np;
    }
    return 0;
}

// The pairings of T2_total (apertureStatisticsCovariance.cu:328-538), one side at a time: (multiplicity,
// three indices into the side's radii).  For theta_123 the indices give T2's arguments 1-3 (the radii of
// one A factor, then the radius of B), for theta_456 its arguments 4-6 (the radius of B, then the radii
// of the other A factor); the reference's table is the product of the two lists, with the equality
// tests in its order.  Its theta_1 == theta_3 list (:451-488) is kept as it is: (theta_1, theta_2,
// theta_1) twice and (theta_2, theta_1, theta_1) once, so theta_2 never enters B, where the pairing
// sum has (theta_1, theta_1, theta_2) for the second entry.
static int t2_side(const double *t, bool first, int mult[3], int idx[3][3])
{
    static const int all[1][3] = {{0, 0, 0}};
    static const int a01[2][3] = {{0, 0, 2}, {0, 2, 0}}, a12[2][3] = {{0, 1, 1}, {1, 1, 0}},
                     a02[2][3] = {{0, 1, 0}, {1, 0, 0}}, adif[3][3] = {{0, 1, 2}, {0, 2, 1}, {1, 2, 0}};
    static const int b01[2][3] = {{0, 0, 2}, {2, 0, 0}}, b02[2][3] = {{0, 1, 0}, {1, 0, 0}},
                     b12[2][3] = {{1, 0, 1}, {0, 1, 1}}, bdif[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
    const int(*L)[3] = first ? adif : bdif;
    int n = 3;
    mult[0] = mult[1] = mult[2] = 1;
    if (t[0] == t[1] && t[1] == t[2]) {
        L = all;
        n = 1;
        mult[0] = 3;
    } else if (t[0] == t[1]) {
        L = first ? a01 : b01;
        n = 2;
        mult[0] = first ? 1 : 2;
        mult[1] = first ? 2 : 1;
    } else if (first ? t[1] == t[2] : t[0] == t[2]) {
        L = first ? a12 : b02;
        n = 2;
        mult[0] = 2;
    } else if (first ? t[0] == t[2] : t[1] == t[2]) {
        L = first ? a02 : b12;
        n = 2;
        mult[0] = 2;
    }
    for (int k = 0; k < n; ++k)
        for (int d = 0; d < 3; ++d) idx[k][d] = L[k][d];
    return n;
}

extern "C" int tp_map3_cov_t2_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo,
                                    const double *theta6, int geometry, double theta_max, double theta_max2,
                                    double lmin, double sigma, double n_gal, int flags, double reltol, double *out,
                                    long long *neval)
{
    int rc = check_args(ctx, c, flags);
    if (rc) return rc;
    if (n && (!theta6 || !out)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: NULL array");
    if (geometry < 0 || geometry > 3)
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: geometry must be 0 (circle), 1 (square), 2 (infinite) "
                                       "or 3 (rectangle)");
    if (!(reltol > 0) || !(n_gal > 0) || !(lmin >= 0))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: reltol and n_gal must be positive, lmin >= 0");
    if (geometry != 2 && (!(theta_max > 0) || (geometry == 3 && !(theta_max2 > 0))))
        return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: the survey's side lengths or radius must be positive");
    if (n > 0x7fffffffu / 81) return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: too many pairs for one call");
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        if (ic < 0 || ic >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: cosmology index out of range");
        for (int k = 0; k < 6; ++k)
            if (!(theta6[6 * i + k] > 0)) return tp_set_error(TP_EINVAL, "tp_map3_cov_t2_batch: aperture radii must be positive");
    }
    if (geometry == 2) {  // T2 vanishes for an infinite survey (apertureStatisticsCovariance.cu:330-331)
        for (size_t i = 0; i < n; ++i) {
            out[i] = 0;
            if (neval) neval[i] = 0;
        }
        return 0;
    }
    TP_HIP(hipSetDevice(ctx->device));
    // Every T2 call of the reference (apertureStatisticsCovariance.cu:1108-1163) is A1 A2 B: A1 and A2 over
    // [lMin, lMax], lMax = 10 / min(theta_1..theta_5) (the second A reuses the first one's domain), B over
    // the same domain for the circle, over [-lMax', lMax']^2, lMax' = 10 / min(theta_3, theta_4), for the
    // square and the rectangle.  Equal integrals (same kind, radii and domain) are evaluated once.
    using Key = std::tuple<int, int, double, double, double, double>;
    std::map<Key, size_t> ids[2];
    std::vector<TpPowerDesc> desc[2];
    std::vector<double> lo[2], hi[2];
    auto integral = [&](int kind, int ic, double ta, double tb, double l, double h) -> size_t {
        const int d = kind == kT2Square || kind == kT2Rect;
        const Key k{kind, ic, ta, tb, l, h};
        auto it = ids[d].find(k);
        if (it != ids[d].end()) return it->second;
        TpPowerDesc D{};
        D.icosmo = ic;
        D.kind = kind;
        D.flags = flags;
        D.sigma = sigma;
        D.ngal = n_gal;
        D.th[0] = ta;
        D.th[1] = tb;
        D.th[2] = theta_max;
        D.th[3] = theta_max2;
        D.lmin = lmin;
        const size_t id = desc[d].size();
        desc[d].push_back(D);
        for (int q = 0; q <= d; ++q) {
            lo[d].push_back(l);
            hi[d].push_back(h);
        }
        ids[d].emplace(k, id);
        return id;
    };
    struct Term {
        int mult;
        size_t a1, a2, b;  // a1, a2 index the 1-D integrals; b the 1-D (circle) or the 2-D ones
    };
    std::vector<std::vector<Term>> terms(n);
    const int kb = geometry == 0 ? kT2Circle : geometry == 1 ? kT2Square : kT2Rect;
    for (size_t i = 0; i < n; ++i) {
        const int ic = icosmo ? icosmo[i] : 0;
        const double *t = theta6 + 6 * i;
        int ma[3], mb[3], ia[3][3], ib[3][3];
        const int na = t2_side(t, true, ma, ia), nb = t2_side(t + 3, false, mb, ib);
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
                const double u[6] = {t[ia[a][0]], t[ia[a][1]], t[ia[a][2]], t[3 + ib[b][0]], t[3 + ib[b][1]],
                                     t[3 + ib[b][2]]};
                const double lmax = 10. / std::min({u[0], u[1], u[2], u[3], u[4]});
                Term T;
                T.mult = ma[a] * mb[b];
                T.a1 = integral(kT2A, ic, u[0], u[1], lmin, lmax);
                T.a2 = integral(kT2A, ic, u[4], u[5], lmin, lmax);
                if (kb == kT2Circle) {
                    T.b = integral(kb, ic, u[2], u[3], lmin, lmax);
                } else {
                    const double l3 = 10. / std::min(u[2], u[3]);
                    T.b = integral(kb, ic, u[2], u[3], -l3, l3);
                }
                terms[i].push_back(T);
            }
    }
    std::vector<double> v1(desc[0].size()), v2(desc[1].size());
    std::vector<long long> n1(desc[0].size()), n2(desc[1].size());
    auto ident = [](double v) { return v; };
    if ((rc = power_batch<1>(ctx, c, desc[0], lo[0].data(), hi[0].data(), reltol, ident, v1.data(), nullptr,
                             n1.data())))
        return rc;
    if (!desc[1].empty() && (rc = power_batch<2>(ctx, c, desc[1], lo[1].data(), hi[1].data(), reltol, ident,
                                                 v2.data(), nullptr, n2.data())))
        return rc;
    for (size_t i = 0; i < n; ++i) {
        double r = 0;
        long long np = 0;
        for (size_t q = 0; q < terms[i].size(); ++q) {
            const Term &T = terms[i][q];
            double A1 = v1[T.a1], A2 = v1[T.a2], B;
            A1 /= 2 * tpd::kPi;
            A2 /= 2 * tpd::kPi;
            if (kb == kT2Circle) {
                B = v1[T.b];
                B /= 2 * tpd::kPi;
            } else {
                B = v2[T.b];
                B /= std::pow(2 * tpd::kPi, 2);
            }
            const double t2 = A1 * A2 * B;
            const double tq = T.mult == 1 ? t2 : T.mult * t2;
            r = q == 0 ? tq : r + tq;
            np += n1[T.a1] + n1[T.a2] + (kb == kT2Circle ? n1[T.b] : n2[T.b]);
        }
        out[i] = r;
        if (neval) neval[i] = np;
    }
    return 0;
}
