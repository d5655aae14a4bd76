// tp_power.hip -- convergence power spectrum P_kappa(ell) and <Map^2>(theta) on gfx950.
//
// The second-order statistics of the same model: revised Halofit (Takahashi et al. 2012) P_NL(k, z)
// from the per-cosmology FP64 tables (D1, r_sigma, n_eff, n_cur), Limber-projected with the lensing
// efficiency g(z), then
//   * Pell(ell)                : 96-point Gauss-Legendre z-sum + shape noise
//                                (classed_version/bispectrum.cpp:1034-1059), one wave per ell;
//   * convergence_power_spectrum(ell): adaptive 1-D cubature over z (bispectrum.cpp:382-417);
//   * <Map^2>(theta)           : adaptive 2-D cubature over (ell, z) (cuda_version/apertureStatistics.cu:29-136).
// The adaptive integrals of one call run together: every round gathers the pending regions of all
// unfinished integrals, one launch evaluates all their rule points (15 Gauss-Kronrod points in 1-D,
// 17 Genz-Malik points in 2-D), and the host applies the rules and advances each integral's heap
// with the decision sequence of the reference's one-at-a-time hcubature.
//
// Per point the integrand is one table interpolation, the Halofit coefficients (13 pow), the
// linear P(k) (E&H fit: 2 pow, 1 log) and the filter: it is transcendental-bound FP64 work on a
// few doubles of input, so a launch is bounded by its FP64 issue rate, and the rounds by the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "tp_cubature.h"
#include "tp_device.h"

namespace {

constexpr double kArcmin = 2.9088820866e-4;  // classed_version/helper.cpp:139, the shape-noise term's unit

// What one integral (or one evaluation call) computes at its points.
enum PowerKind { kPkappa = 0, kMap2 = 1, kPkAtZ = 2, kLimber = 3 };

struct TpPowerDesc {
    double par;          // kPkappa: ell; kMap2: theta [rad]
    double sigma, ngal;  // kMap2 shape noise (sigma = 0: none), ngal [arcmin^-2]
    int icosmo;
    int kind;
    int flags;           // TP_PK_LINEAR | TP_PK_T17
    int pad;
};

// Table values at redshift z (classed_version/bispectrum.cpp:506-513 + compute_coefficients :914-920).
struct TpZAt {
    double fK, g, D1, rs, ne, nc;
};

__device__ __forceinline__ void z_at(const double *T, int n_z, double zmax, double z, TpZAt *o)
{
    double d = z / zmax * (n_z - 1);
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
    const double e = 1 - d;
    o->fK = T[idx] * e + T[idx + 1] * d;
    o->g = T[n_z + idx] * e + T[n_z + idx + 1] * d;
    o->D1 = T[2 * n_z + idx] * e + T[2 * n_z + idx + 1] * d;
    o->rs = T[3 * n_z + idx] * e + T[3 * n_z + idx + 1] * d;
    o->ne = T[4 * n_z + idx] * e + T[4 * n_z + idx + 1] * d;
    o->nc = T[5 * n_z + idx] * e + T[5 * n_z + idx + 1] * d;
}

// Revised Halofit, classed_version/bispectrum.cpp:288-367.  The reference's flatness check always
// passes: ow = 1 - om by construction.
__device__ double pnl(const TpCosmoConst &C, const TpZAt &Z, double k, double z)
{
    const double n_eff = Z.ne, nsqr = n_eff * n_eff, nc = Z.nc;
    // om_m_of_z, om_v_of_z (bispectrum.cpp:1022-1032): the flat form, without w
    const double aa = 1. / (1 + z);
    const double den = C.om + aa * (aa * aa * C.ow + (1 - C.om - C.ow));
    const double om_m = C.om / den;
    const double om_v = C.ow * (aa * aa * aa) / den;
    const double f1a = pow(om_m, -0.0732), f2a = pow(om_m, -0.1423), f3a = pow(om_m, 0.0725);
    const double f1b = pow(om_m, -0.0307), f2b = pow(om_m, -0.0585), f3b = pow(om_m, 0.0743);
    const double frac = om_v / (1. - om_m);
    const double f1 = frac * f1b + (1 - frac) * f1a;
    const double f2 = frac * f2b + (1 - frac) * f2a;
    const double f3 = frac * f3b + (1 - frac) * f3a;
    double a = 1.5222 + 2.8553 * n_eff + 2.3706 * nsqr + 0.9903 * n_eff * nsqr + 0.2250 * nsqr * nsqr -
               0.6038 * nc + 0.1749 * om_v * (1.0 + C.w);
    a = pow(10.0, a);
    const double b = pow(10.0, -0.5642 + 0.5864 * n_eff + 0.5716 * nsqr - 1.5474 * nc + 0.2279 * om_v * (1.0 + C.w));
    const double c = pow(10.0, 0.3698 + 2.0404 * n_eff + 0.8161 * nsqr + 0.5869 * nc);
    const double gam = 0.1971 - 0.0843 * n_eff + 0.8460 * nc;
    const double alpha = fabs(6.0835 + 1.3373 * n_eff - 0.1959 * nsqr - 5.5274 * nc);
    const double beta = 2.0379 - 0.7354 * n_eff + 0.3157 * nsqr + 1.2490 * n_eff * nsqr + 0.3980 * nsqr * nsqr -
                        0.1682 * nc;
    const double xnu = pow(10.0, 5.2105 + 3.6902 * n_eff);
    const double twopi2 = 2 * tpd::kPi * tpd::kPi;
    const double plin = tp_pk_lin(C, k, C.norm) * Z.D1 * Z.D1 * k * k * k / twopi2;
    const double y = k * Z.rs, ysqr = y * y;
    double ph = a * pow(y, f1 * 3) / (1 + b * pow(y, f2) + pow(f3 * c * y, 3 - gam));
    ph = ph / (1 + xnu / ysqr);
    const double pq = plin * pow(1 + plin, beta) / (1 + plin * alpha) * exp(-y / 4.0 - ysqr / 8.0);
    return twopi2 * (pq + ph) / (k * k * k);
}

// P(k, z): Halofit, or D1^2 P_lin with TP_PK_LINEAR (linear_pk_at_z, bispectrum.cpp:369-380).
__device__ __forceinline__ double pk_at(const TpCosmoConst &C, const TpZAt &Z, double k, double z, int flags)
{
    return (flags & TP_PK_LINEAR) ? Z.D1 * Z.D1 * tp_pk_lin(C, k, C.norm) : pnl(C, Z, k, z);
}

// The ratio fitted to the T17 simulations, cuda_version/bispectrum.cu:971-979.
__device__ __forceinline__ double t17_factor(double k)
{
    const double c1 = 9.5171e-4, c2 = 5.1543e-3, a1 = 1.3063, a2 = 1.1475, a3 = 0.62793;
    return pow(1 + c1 * pow(k, -1. * a1), a1) / pow(1 + c2 * pow(k, -1. * a2), a3);
}

// limber_integrand_power_spectrum, classed_version/bispectrum.cpp:543-577.
__device__ double limber_pk(const TpCosmoConst &C, const double *T, int n_z, double ell, double z, int flags)
{
    if (z < 1e-5) return 0;
    TpZAt Z;
    z_at(T, n_z, C.zmax, z, &Z);
    const double H = 100. / 299792.;  // H0/c [h/Mpc], classed_version/bispectrum.hpp:148
    const double zp = 1. + z;
    const double E = sqrt(C.om * (zp * zp * zp) + C.ow);
    const double pref = 9. / 4. * H * H * H * C.om * C.om * zp * zp * Z.g * Z.g / E;
    const double k = ell / Z.fK;
    const double corr = (flags & TP_PK_T17) ? t17_factor(k) : 1.0;
    return corr * pref * pk_at(C, Z, k, z, flags);
}

// <Map^2> integrand at (ell, z) for one source bin, cuda_version/apertureStatistics.cu:29-57.
__device__ double map2_f(const TpCosmoConst &C, const double *T, int n_z, double ell, double z, const TpPowerDesc &D)
{
    const double eta = ell * D.par, t = 0.5 * eta * eta;
    const double u = t * exp(-t);  // uHat, classed_version/apertureStatistics.cpp:10-14
    double v = ell * (u * u) * limber_pk(C, T, n_z, ell, z, D.flags);
    if (D.flags & TP_PK_T17) {
        const double r = ell / (1.6 * 8192.);
        v /= 1 + r * r;
    }
    if (D.sigma != 0) {
        const double us = u * D.sigma;
        v += ell * (us * us) * 0.5 / D.ngal * (kArcmin * kArcmin) / C.zmax;
    }
    return v;
}

// One thread per point; pdesc[i] names the integral (desc entry) of point i.
__global__ void k_power_points(const TpCosmoConst *__restrict__ cc, const double *__restrict__ tab, int n_z,
                               const TpPowerDesc *__restrict__ desc, const int *__restrict__ pdesc, size_t npts,
                               int dim, const double *__restrict__ x, double *__restrict__ f, int *flag)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const TpPowerDesc D = desc[pdesc ? pdesc[i] : 0];
    const TpCosmoConst &C = cc[D.icosmo];
    const double *T = tab + (size_t)D.icosmo * 6 * n_z;
    double v;
    if (D.kind == kPkappa) {
        v = limber_pk(C, T, n_z, D.par, x[i], D.flags);
    } else {
        const double x0 = x[2 * i], x1 = x[2 * i + 1];
        if (D.kind == kMap2) {
            v = map2_f(C, T, n_z, x0, x1, D);
        } else if (D.kind == kLimber) {
            v = limber_pk(C, T, n_z, x0, x1, D.flags);
        } else {
            TpZAt Z;
            z_at(T, n_z, C.zmax, x1, &Z);
            v = pk_at(C, Z, x0, x1, D.flags);
        }
    }
    if (!isfinite(v)) atomicOr(flag, 1);
    f[i] = v;
}

// 96-point Gauss-Legendre half nodes/weights (own Newton solve, uploaded once per device)
__constant__ double c_pgl_x[48];
__constant__ double c_pgl_w[48];

// Pell: one wave per multipole.  Lane j < 48 evaluates the node pair j (z = cx -+ dx x_j); lane 0
// then adds the 48 pair sums in node order (the reference's loop order, GQ96_of_Pk).
__global__ void __launch_bounds__(64) k_pell(const TpCosmoConst *__restrict__ cc, const double *__restrict__ tab,
                                             int n_z, int icosmo, int flags, size_t n, const double *__restrict__ ell,
                                             double zmax, double shot, double *__restrict__ out, int *flag)
{
    __shared__ double part[48];
    const size_t i = blockIdx.x;
    const int j = threadIdx.x;
    if (i >= n) return;
    const TpCosmoConst &C = cc[icosmo];
    const double *T = tab + (size_t)icosmo * 6 * n_z;
    const double e = ell[i];
    const double cx = zmax / 2, dx = zmax / 2;
    if (j < 48 && e != 0)
        part[j] = c_pgl_w[j] * (limber_pk(C, T, n_z, e, cx - dx * c_pgl_x[j], flags) +
                                limber_pk(C, T, n_z, e, cx + dx * c_pgl_x[j], flags));
    __syncthreads();
    if (j == 0) {
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
    if (flags & ~(TP_PK_LINEAR | TP_PK_T17)) return tp_set_error(TP_EINVAL, "unknown power-spectrum flags");
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
                       (double *)ctx->d_out.p, (int *)ctx->d_flag.p);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, bf, hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(fval, ctx->h_out.p, bf);
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats.kernel_ms += ms;
    ctx->stats.integrand_ms += ms;
    ctx->stats.launches += 1;
    if (flag) return tp_set_error(TP_ENONFINITE, "power spectrum integrand: non-finite value");
    return 0;
}

// All integrals of one call advance together, one launch per round (see the file comment).
template <unsigned D>
int power_batch(tp_ctx *ctx, const tp_cosmo *c, const std::vector<TpPowerDesc> &desc, const double *lo,
                const double *hi, double reltol, bool per_2pi, double *out, double *err, long long *neval)
{
    auto t0 = std::chrono::steady_clock::now();
    ctx->stats = tp_batch_stats{};
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
    while (!active.empty()) {
        const size_t na = active.size();
        offs.assign(na + 1, 0);
        for (size_t a = 0; a < na; ++a) offs[a + 1] = offs[a] + I[active[a]].pending().size() * np;
        const size_t npts = offs[na];
        pts.resize(npts * D);
        vals.resize(npts);
        pdesc.resize(npts);
#pragma omp parallel for schedule(dynamic, 16) if (na > 64)
        for (size_t a = 0; a < na; ++a) {
            auto &A = I[active[a]];
            const auto &pend = A.pending();
            for (size_t r = 0; r < pend.size(); ++r) {
                const size_t o = offs[a] + r * np;
                tpc::rule_points_of(D, A.regions()[pend[r]], pts.data() + D * o);
                for (unsigned q = 0; q < np; ++q) pdesc[o + q] = (int)active[a];
            }
        }
        int rc = power_launch(ctx, c, desc, pdesc.data(), npts, (int)D, pts.data(), vals.data());
        if (rc) return rc;
        ctx->stats.rounds += 1;
        ctx->stats.points += (long long)npts;
        std::vector<char> fin(na, 0);
#pragma omp parallel for schedule(dynamic, 16) if (na > 64)
        for (size_t a = 0; a < na; ++a) {
            auto &A = I[active[a]];
            const auto &pend = A.pending();
            for (size_t r = 0; r < pend.size(); ++r)
                tpc::rule_apply(D, 1, A.regions()[pend[r]], vals.data() + offs[a] + r * np);
            A.absorb();
            fin[a] = A.advance() ? 1 : 0;
        }
        std::vector<size_t> next;
        for (size_t a = 0; a < na; ++a)
            if (!fin[a]) next.push_back(active[a]);
        active.swap(next);
    }
    for (size_t k = 0; k < nint; ++k) {
        double v, e;
        I[k].result(&v, &e);
        out[k] = per_2pi ? v / 2. / tpd::kPi : v;  // apertureStatistics.cu:135
        if (err) err[k] = per_2pi ? e / 2. / tpd::kPi : e;
        if (neval) neval[k] = (long long)I[k].num_eval();
    }
    ctx->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
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
    hipLaunchKernelGGL(k_pell, dim3((unsigned)n), dim3(64), 0, ctx->stream, c->d_const, c->d_tab, c->n_z, icosmo,
                       flags, n, (const double *)ctx->d_in.p, c->z_max, shot, (double *)ctx->d_out.p,
                       (int *)ctx->d_flag.p);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    TP_HIP(hipMemcpyAsync(out, ctx->d_out.p, b, hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->stats = tp_batch_stats{};
    ctx->stats.kernel_ms = ctx->stats.integrand_ms = ms;
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
    return power_batch<1>(ctx, c, desc, lo.data(), hi.data(), reltol, false, out, err, neval);
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
    return power_batch<2>(ctx, c, desc, lo.data(), hi.data(), reltol, true, out, err, neval);
}
