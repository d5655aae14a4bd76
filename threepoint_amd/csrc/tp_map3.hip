// tp_map3.hip -- gfx950 kernels of the <Map^3> integrand path.
//
//  * bispec_q<T>: BiHalofit (Takahashi et al. 2020, eqs. B4-B9) at one Limber
//    z-node, restating classed_version/bispectrum.cpp:166-236 (with F2 :614-633
//    and the Eisenstein & Hu P(k) :811-841) in terms of the dimensionless
//    wavenumbers q = k r_sigma.  Everything that depends on z only comes from a
//    per-node record (TpNodeT) that the table builder evaluated once in FP64.
//  * bkappa_sum<T>: the 96-node Limber sum of classed_version/bispectrum.cpp:
//    238-286 / 1010-1020, accumulated in FP64.  T = double is the plain path;
//    T = float runs the bispectrum in FP32 (native v_exp_f32 / v_log_f32).
//    Node records are wave-uniform, so they are read through scalar loads.
//  * map3 point kernel: classed_version/apertureStatistics.cpp:16-28.
//  * region kernels: Genz-Malik degree-7/5 rule (classed_version/hcubature.cpp:
//    456-577) for many 3-D regions of many integrals per launch -- one thread
//    per (region, point), then one thread per region for the rule sums, in the
//    reference's summation order.
#include "tp_device.h"

namespace {

// --------------------------------------------------------------- kernels
template <class T>
__global__ void __launch_bounds__(256) k_map3_points(const TpNodeT<T> *__restrict__ nodes,
                                                     const TpEHT<T> *__restrict__ eh, double th1, double th2,
                                                     double th3, size_t npts, const double *__restrict__ x,
                                                     double *__restrict__ f, int mode, int *flag)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double l1 = x[3 * i], l2 = x[3 * i + 1], phi = x[3 * i + 2];
    double v = tp_map3_integrand<T>(nodes, *eh, l1, l2, phi, th1, th2, th3, mode, flag);
    f[i] = v;
}

template <class T>
__global__ void __launch_bounds__(256) k_bkappa_points(const TpNodeT<T> *__restrict__ nodes,
                                                       const TpEHT<T> *__restrict__ eh, size_t npts,
                                                       const double *__restrict__ l, double *__restrict__ out,
                                                       int mode, int *flag)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double l1 = l[3 * i], l2 = l[3 * i + 1], l3 = l[3 * i + 2];
    double v;
    if (mode == TP_MODE_ANALYTIC)
        v = tp_bkappa_analytic(l1, l2, l3);
    else
        v = (l1 == 0 || l2 == 0 || l3 == 0) ? 0.0 : tp_bkappa_sum<T>(nodes, *eh, l1, l2, l3);
    if (!isfinite(v)) atomicOr(flag, 1);
    out[i] = v;
}

// One thread per (region, point).  nodes/eh are indexed by the integral's cosmology.
template <class T>
__global__ void __launch_bounds__(256) k_region_points(const TpNodeT<T> *__restrict__ nodes,
                                                       const TpEHT<T> *__restrict__ eh, size_t nreg,
                                                       const TpRegion *__restrict__ reg,
                                                       const TpIntegralDesc *__restrict__ desc,
                                                       double *__restrict__ fvals, int mode, int *flag)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nreg * TP_NPTS_GM3) return;
    size_t r = t / TP_NPTS_GM3;
    int p = (int)(t - r * TP_NPTS_GM3);
    const TpRegion R = reg[r];
    const TpIntegralDesc D = desc[R.integral];
    double x0, x1, x2;
    gm_point3(R, p, &x0, &x1, &x2);
    const TpNodeT<T> *nd = nodes + (size_t)D.icosmo * TP_NNODES;
    fvals[t] = tp_map3_integrand<T>(nd, eh[D.icosmo], x0, x1, x2, D.th[0], D.th[1], D.th[2], mode, flag);
}

// Region r's flag: bit 0 = non-finite rule estimate.  Region 0 also carries the point
// kernel's flag word (non-finite ell: bit 1, non-finite B_kappa: bit 2), so one copy of
// the region outputs returns every error of the round.
__global__ void __launch_bounds__(256) k_region_rule(size_t nreg, const TpRegion *__restrict__ reg,
                                                     const double *__restrict__ fvals,
                                                     TpRegionOut *__restrict__ out, const int *__restrict__ flag)
{
    size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nreg) return;
    gm_rule3(reg[r].vol, fvals + r * TP_NPTS_GM3, &out[r]);
    if (r == 0) out[0].flag |= *flag << 1;
}

int grid_for(size_t n, int block) { return (int)((n + block - 1) / block); }

}  // namespace

int tp_launch_map3_points(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double *th, size_t npts,
                          const double *d_x, double *d_f, int mode, int *d_flag)
{
    if (npts == 0) return 0;
    const int B = 256;
    if (mode == TP_MODE_FP32)
        hipLaunchKernelGGL(k_map3_points<float>, dim3(grid_for(npts, B)), dim3(B), 0, ctx->stream,
                           c->d_node32 + (size_t)icosmo * TP_NNODES, c->d_eh32 + icosmo, th[0], th[1], th[2], npts,
                           d_x, d_f, mode, d_flag);
    else
        hipLaunchKernelGGL(k_map3_points<double>, dim3(grid_for(npts, B)), dim3(B), 0, ctx->stream,
                           c->d_node64 + (size_t)icosmo * TP_NNODES, c->d_eh64 + icosmo, th[0], th[1], th[2], npts,
                           d_x, d_f, mode, d_flag);
    TP_HIP(hipGetLastError());
    return 0;
}

int tp_launch_bkappa_points(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *d_l,
                            double *d_out, int mode, int *d_flag)
{
    if (npts == 0) return 0;
    const int B = 256;
    if (mode == TP_MODE_FP32)
        hipLaunchKernelGGL(k_bkappa_points<float>, dim3(grid_for(npts, B)), dim3(B), 0, ctx->stream,
                           c->d_node32 + (size_t)icosmo * TP_NNODES, c->d_eh32 + icosmo, npts, d_l, d_out, mode,
                           d_flag);
    else
        hipLaunchKernelGGL(k_bkappa_points<double>, dim3(grid_for(npts, B)), dim3(B), 0, ctx->stream,
                           c->d_node64 + (size_t)icosmo * TP_NNODES, c->d_eh64 + icosmo, npts, d_l, d_out, mode,
                           d_flag);
    TP_HIP(hipGetLastError());
    return 0;
}

int tp_launch_regions(tp_ctx *ctx, const tp_cosmo *c, size_t nreg, const TpRegion *d_reg,
                      const TpIntegralDesc *d_desc, double *d_fvals, TpRegionOut *d_out, int mode, int *d_flag,
                      hipEvent_t evm)
{
    if (nreg == 0) return 0;
    const int B = 256;
    size_t np = nreg * TP_NPTS_GM3;
    if (mode == TP_MODE_FP32)
        hipLaunchKernelGGL(k_region_points<float>, dim3(grid_for(np, B)), dim3(B), 0, ctx->stream, c->d_node32,
                           c->d_eh32, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
    else
        hipLaunchKernelGGL(k_region_points<double>, dim3(grid_for(np, B)), dim3(B), 0, ctx->stream, c->d_node64,
                           c->d_eh64, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
    TP_HIP(hipEventRecord(evm, ctx->stream));
    hipLaunchKernelGGL(k_region_rule, dim3(grid_for(nreg, B)), dim3(B), 0, ctx->stream, nreg, d_reg, d_fvals, d_out,
                       d_flag);
    TP_HIP(hipGetLastError());
    return 0;
}
