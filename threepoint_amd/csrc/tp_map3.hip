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
#include <cmath>
#include <vector>

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

// Integrand values of the 33 Genz-Malik points of many regions of many integrals.  A point's 96
// z-nodes are split over LPP waves (1, 2, 4 or 8): wave w of a point group sums the partials
// [w 8/LPP, (w+1) 8/LPP) of the fixed summation tree with wave-uniform (scalar) node reads, and
// the group's first wave combines the waves' sums in tree order -- the same bits for every LPP.
// A workgroup holds max(1, 4/LPP) groups of 64 points.  nodes/eh are indexed by the integral's
// cosmology.
template <class T, int LPP>
__global__ void __launch_bounds__(64 * (LPP < 4 ? 4 : LPP)) k_region_points(
    const TpNodeT<T> *__restrict__ nodes, const TpEHT<T> *__restrict__ eh, size_t nreg,
    const TpRegion *__restrict__ reg, const TpIntegralDesc *__restrict__ desc, double *__restrict__ fvals, int mode,
    int *flag)
{
    constexpr int GPB = LPP < 4 ? 4 / LPP : 1;  // point groups per workgroup
    __shared__ double part[GPB][LPP][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = wv / LPP, w = wv % LPP;
    const size_t t = ((size_t)blockIdx.x * GPB + g) * 64 + lane;
    const bool live = t < nreg * TP_NPTS_GM3;
    TpPointPre P;
    int kind = TP_PT_DONE;
    double v = 0.0;
    TpIntegralDesc D{};
    if (live) {
        const size_t r = t / TP_NPTS_GM3;
        const int p = (int)(t - r * TP_NPTS_GM3);
        const TpRegion R = reg[r];
        D = desc[R.integral];
        double x0, x1, x2;
        gm_point3(R, p, &x0, &x1, &x2);
        kind = tp_map3_begin(x0, x1, x2, mode, flag, &P);
        if (kind == TP_PT_SUM)
            v = tp_bkappa_parts<T, TP_NPART / LPP>(nodes + (size_t)D.icosmo * TP_NNODES, eh[D.icosmo], P,
                                                   w * (TP_NPART / LPP));
    }
    if constexpr (LPP > 1) {
        part[g][w][lane] = v;
        __syncthreads();
        if (w != 0) return;
        double a[LPP];
#pragma unroll
        for (int k = 0; k < LPP; ++k) a[k] = part[g][k][lane];
#pragma unroll
        for (int s = 1; s < LPP; s <<= 1)
#pragma unroll
            for (int k = 0; k < LPP; k += 2 * s) a[k] = a[k] + a[k + s];
        v = a[0];
    }
    if (live) fvals[t] = tp_map3_end(P, kind, v, D.th[0], D.th[1], D.th[2], flag);
}

// Region-cached form of k_region_points (launch shapes 16, 32, 64).  Of a region's 33 points only 7
// distinct l1 (axis-0 codes -3..3) and 7 distinct l2 occur, and the P_E term of BiHalofit
// (tp_pse_node: the linear P(k), its exp/log and reciprocals) depends on one wavenumber only.  A
// workgroup takes kCacheReg regions with G groups of 256 threads; group g owns the partials
// [g 8/G, (g+1) 8/G) of the summation tree and, for each partial (12 z-nodes),
//   A. evaluates the P_E terms of the 47 multipoles per region (7 l1, 7 l2, 33 l3) at the 12 nodes
//      into its LDS block (work items of 64 consecutive (node, multipole) entries, whose at most two
//      node records are read through wave-uniform loads);
//   B. lets each point add its 12 node values, the shape-dependent rest (tp_bispec_joint) with the
//      three P_E terms read from LDS, and push the partial into the group's subtree.
// The groups' subtrees are then joined in tree order.  Every value is formed by the same operations
// as in k_region_points, so every shape gives the same bits.  More groups shorten a thread's share
// of the work, which keeps launches of a few thousand regions from ending on a half-empty wave of
// workgroups.
constexpr int kCacheReg = 7;                       // regions per workgroup: 231 points per group
constexpr int kCacheEll = 14 + TP_NPTS_GM3;        // multipoles per region: 7 l1, 7 l2, 33 l3
constexpr int kCacheN = kCacheReg * kCacheEll;     // 329

template <class T, int G>
__global__ void __launch_bounds__(256 * G) k_region_cached(const TpNodeT<T> *__restrict__ nodes,
                                                           const TpEHT<T> *__restrict__ eh, size_t nreg,
                                                           const TpRegion *__restrict__ reg,
                                                           const TpIntegralDesc *__restrict__ desc,
                                                           double *__restrict__ fvals, int mode, int *flag)
{
    constexpr int PPG = TP_NPART / G;                 // partials per group
    __shared__ T lv[kCacheN], llv[kCacheN];           // multipole and its log, as the points see them
    __shared__ T pse[G][TP_PART_NODES][kCacheN];      // P_E at (group, node of the partial, multipole)
    __shared__ double sub[G > 1 ? G : 1][kCacheReg * TP_NPTS_GM3];  // the groups' subtree values
    __shared__ int ics[kCacheReg];                    // cosmology of each region
    __shared__ int one_cosmo;                         // all regions of the workgroup share ics[0]
    const int t = threadIdx.x & 255, lane = t & 63;
    const int g = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 8);
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const size_t r0 = (size_t)blockIdx.x * kCacheReg;
    const int nr = (int)((nreg - r0) < (size_t)kCacheReg ? (nreg - r0) : (size_t)kCacheReg);  // >= 1
    const int ne = nr * kCacheEll;
    // set-up: every group's point threads form their point; group 0 stores the l3 entries and
    // its threads < 14 nr the l1 / l2 entries
    const bool pt = t < nr * TP_NPTS_GM3;
    const int prr = t / TP_NPTS_GM3, pp = t - prr * TP_NPTS_GM3;
    TpPointPre P;
    int kind = TP_PT_DONE;
    TpIntegralDesc D{};
    if (pt) {
        const TpRegion R = reg[r0 + prr];
        D = desc[R.integral];
        double x0, x1, x2;
        gm_point3(R, pp, &x0, &x1, &x2);
        kind = tp_map3_begin(x0, x1, x2, mode, flag, &P);  // every group flags the same points: same bit
        if (g == 0) {
            const int e = prr * kCacheEll + 14 + pp;
            lv[e] = kind == TP_PT_SUM ? (T)P.l[2] : T(1);
            llv[e] = kind == TP_PT_SUM ? (T)P.ll[2] : T(0);
            if (pp == 0) ics[prr] = D.icosmo;
        }
    }
    if (g == 0 && t < nr * 14) {
        const int rr = t / 14, k = t - rr * 14, ax = k / 7;  // k = 7 axis + code + 3
        const TpRegion R = reg[r0 + rr];
        const double x = tp_gm_coord(ax ? R.c[1] : R.c[0], ax ? R.hw[1] : R.hw[0], k - 7 * ax - 3);
        lv[rr * kCacheEll + k] = (T)x;
        llv[rr * kCacheEll + k] = x > 0 ? (T)TpMath<double>::lg(x) : T(0);  // non-positive: no summed point uses it
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int u = 1;
        for (int rr = 1; rr < nr; ++rr) u &= ics[rr] == ics[0];
        one_cosmo = u;
    }
    __syncthreads();
    // block-uniform: with one cosmology the node and E&H records are read through scalar loads
    const int uni = __builtin_amdgcn_readfirstlane(one_cosmo);
    const int ic0 = __builtin_amdgcn_readfirstlane(ics[0]);
    const TpNodeT<T> *__restrict__ nodes0 = nodes + (size_t)ic0 * TP_NNODES;
    const bool sum = pt && kind == TP_PT_SUM;
    T L[3] = {T(1), T(1), T(1)}, LL[3] = {T(0), T(0), T(0)}, F2T[3] = {T(0), T(0), T(0)}, LR1 = T(0), R2 = T(0);
    int e0 = 0, e1 = 0, e2 = 0;
    const TpNodeT<T> *Np = nodes;
    if (sum) {
        for (int i = 0; i < 3; ++i) {
            L[i] = (T)P.l[i];
            LL[i] = (T)P.ll[i];
            F2T[i] = (T)P.f2t[i];
        }
        LR1 = (T)P.lr1;
        R2 = (T)P.r2;
        const int base = prr * kCacheEll;
        e0 = base + c_gm3_code[pp][0] + 3;
        e1 = base + c_gm3_code[pp][1] + 10;
        e2 = base + 14 + pp;
        Np = nodes + (size_t)D.icosmo * TP_NNODES;
    }
    T(*ps_g)[kCacheN] = pse[g];
    // phase B body: the point's 12 node values of the partial starting at node n0, from records Nb
    auto node_block = [&](const TpNodeT<T> *__restrict__ Nb, int n0) {
        double acc = 0.0;
        for (int j = 0; j < TP_PART_NODES; ++j) {
            const TpNodeT<T> &N = Nb[n0 + j];
            const T ks = N.ks, lks = N.lnks;
            const T q[3] = {L[0] * ks, L[1] * ks, L[2] * ks};
            const T lq[3] = {LL[0] + lks, LL[1] + lks, LL[2] + lks};
            const T ps[3] = {ps_g[j][e0], ps_g[j][e1], ps_g[j][e2]};
            const T b = tp_bispec_joint<T>(N, q, lq, F2T, LR1, R2, ps);
            acc += (double)(N.wz * b);
        }
        return acc;
    };
    // phase A work items.  With one cosmology and at least two regions (ne >= 94): 64 consecutive
    // entries of the flattened (node j, multipole e) space of the partial, so that every lane has
    // work; such an item spans at most two nodes, read as two wave-uniform records and selected per
    // lane.  Otherwise: one node and 64 consecutive multipoles.
    const bool flat = uni && ne >= 64;
    const int nflat = TP_PART_NODES * ne;
    const int nblk = (ne + 63) >> 6, items = flat ? (nflat + 63) >> 6 : TP_PART_NODES * nblk;
    double s1 = 0.0, s2 = 0.0, s4 = 0.0, v = 0.0;
    for (int k = 0; k < PPG; ++k) {
        const int n0 = (g * PPG + k) * TP_PART_NODES;
        if (flat) {
            for (int it = wv; it < items; it += 4) {
                const int f = it * 64 + lane;
                const int ja = (it * 64) / ne, jb = min(ja + 1, TP_PART_NODES - 1);  // wave-uniform
                const TpNodeT<T> &Na = nodes0[n0 + ja], &Nb = nodes0[n0 + jb];
                if (f < nflat) {
                    const int j = f / ne, e = f - j * ne;
                    const TpNodeT<T> N = j == ja ? Na : Nb;
                    ps_g[j][e] = tp_pse_node<T>(N, eh[ic0], lv[e] * N.ks, llv[e] + N.lnks);
                }
            }
        } else {
            for (int it = wv; it < items; it += 4) {
                const int j = it / nblk, e = (it - j * nblk) * 64 + lane;
                if (e < ne) {
                    if (uni) {
                        const TpNodeT<T> &N = nodes0[n0 + j];
                        ps_g[j][e] = tp_pse_node<T>(N, eh[ic0], lv[e] * N.ks, llv[e] + N.lnks);
                    } else {
                        const int ic = ics[e / kCacheEll];
                        const TpNodeT<T> &N = nodes[(size_t)ic * TP_NNODES + n0 + j];
                        ps_g[j][e] = tp_pse_node<T>(N, eh[ic], lv[e] * N.ks, llv[e] + N.lnks);
                    }
                }
            }
        }
        __syncthreads();
        if (sum) v = tp_tree8_push(k, uni ? node_block(nodes0, n0) : node_block(Np, n0), s1, s2, s4);
        __syncthreads();
    }
    if constexpr (G > 1) {  // join the groups' subtrees: v0 + v1, or (v0 + v1) + (v2 + v3)
        if (pt) sub[g][t] = v;
        __syncthreads();
        if (g != 0) return;
        if (pt) {
            double a[G];
#pragma unroll
            for (int i = 0; i < G; ++i) a[i] = sub[i][t];
#pragma unroll
            for (int st = 1; st < G; st <<= 1)
#pragma unroll
                for (int i = 0; i < G; i += 2 * st) a[i] = a[i] + a[i + st];
            v = a[0];
        }
    }
    if (pt) fvals[(r0 + prr) * TP_NPTS_GM3 + pp] = tp_map3_end(P, kind, v, D.th[0], D.th[1], D.th[2], flag);
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

// one launch of the region-point kernel with LPP waves per point group
template <class T>
static void launch_region_points(hipStream_t st, int lpp, const TpNodeT<T> *nodes, const TpEHT<T> *eh, size_t nreg,
                                 const TpRegion *d_reg, const TpIntegralDesc *d_desc, double *d_fvals, int mode,
                                 int *d_flag)
{
    const size_t np = nreg * TP_NPTS_GM3;
    const size_t groups = (np + 63) / 64;
    switch (lpp) {
    case 2:
        hipLaunchKernelGGL((k_region_points<T, 2>), dim3((unsigned)((groups + 1) / 2)), dim3(256), 0, st, nodes, eh,
                           nreg, d_reg, d_desc, d_fvals, mode, d_flag);
        break;
    case 4:
        hipLaunchKernelGGL((k_region_points<T, 4>), dim3((unsigned)groups), dim3(256), 0, st, nodes, eh, nreg, d_reg,
                           d_desc, d_fvals, mode, d_flag);
        break;
    case 8:
        hipLaunchKernelGGL((k_region_points<T, 8>), dim3((unsigned)groups), dim3(512), 0, st, nodes, eh, nreg, d_reg,
                           d_desc, d_fvals, mode, d_flag);
        break;
    case 16:
        hipLaunchKernelGGL((k_region_cached<T, 1>), dim3((unsigned)((nreg + kCacheReg - 1) / kCacheReg)), dim3(256),
                           0, st, nodes, eh, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
        break;
    case 32:
        hipLaunchKernelGGL((k_region_cached<T, 2>), dim3((unsigned)((nreg + kCacheReg - 1) / kCacheReg)), dim3(512),
                           0, st, nodes, eh, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
        break;
    case 64:
        hipLaunchKernelGGL((k_region_cached<T, 4>), dim3((unsigned)((nreg + kCacheReg - 1) / kCacheReg)), dim3(1024),
                           0, st, nodes, eh, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
        break;
    default:
        hipLaunchKernelGGL((k_region_points<T, 1>), dim3((unsigned)((groups + 3) / 4)), dim3(256), 0, st, nodes, eh,
                           nreg, d_reg, d_desc, d_fvals, mode, d_flag);
    }
}

// size bucket of a launch for the launch-shape table (points: < 32k, < 160k, < 640k, more)
static int region_bucket(size_t np) { return np < 32768 ? 0 : np < 163840 ? 1 : np < 655360 ? 2 : 3; }

int tp_region_lanes(tp_ctx *ctx, const tp_cosmo *c, int mode, size_t np);  // tp_tune below

int tp_launch_regions(tp_ctx *ctx, const tp_cosmo *c, size_t nreg, const TpRegion *d_reg,
                      const TpIntegralDesc *d_desc, double *d_fvals, TpRegionOut *d_out, int mode, int *d_flag,
                      hipEvent_t evm, hipStream_t st)
{
    if (nreg == 0) return 0;
    const int B = 256;
    size_t np = nreg * TP_NPTS_GM3;
    const int lpp = tp_region_lanes(ctx, c, mode, np);
    if (lpp < 0) return TP_EHIP;
    if (mode == TP_MODE_FP32)
        launch_region_points<float>(st, lpp, c->d_node32, c->d_eh32, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
    else
        launch_region_points<double>(st, lpp, c->d_node64, c->d_eh64, nreg, d_reg, d_desc, d_fvals, mode, d_flag);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(evm, st));
    hipLaunchKernelGGL(k_region_rule, dim3(grid_for(nreg, B)), dim3(B), 0, st, nreg, d_reg, d_fvals, d_out,
                       d_flag);
    TP_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ graph rounds
// The round's kernels as (function, grid, block) for graph kernel nodes, the same choice as
// launch_region_points makes.
template <class T>
static void region_points_shape(int lpp, size_t nreg, void **func, dim3 *grid, dim3 *block)
{
    const size_t groups = (nreg * TP_NPTS_GM3 + 63) / 64;
    switch (lpp) {
    case 2:
        *func = (void *)k_region_points<T, 2>;
        *grid = dim3((unsigned)((groups + 1) / 2));
        *block = dim3(256);
        break;
    case 4:
        *func = (void *)k_region_points<T, 4>;
        *grid = dim3((unsigned)groups);
        *block = dim3(256);
        break;
    case 8:
        *func = (void *)k_region_points<T, 8>;
        *grid = dim3((unsigned)groups);
        *block = dim3(512);
        break;
    case 16:
    case 32:
    case 64:
        *func = lpp == 16 ? (void *)k_region_cached<T, 1> : lpp == 32 ? (void *)k_region_cached<T, 2>
                                                                       : (void *)k_region_cached<T, 4>;
        *grid = dim3((unsigned)((nreg + kCacheReg - 1) / kCacheReg));
        *block = dim3(16 * lpp);
        break;
    default:
        *func = (void *)k_region_points<T, 1>;
        *grid = dim3((unsigned)((groups + 3) / 4));
        *block = dim3(256);
    }
}

int tp_graph_round(tp_ctx *ctx, const tp_cosmo *c, TpLane &L, size_t nreg, int mode, hipStream_t st)
{
    if (nreg == 0) return 0;
    const int lpp = tp_region_lanes(ctx, c, mode, nreg * TP_NPTS_GM3);
    if (lpp < 0) return TP_EHIP;
    const bool f32 = mode == TP_MODE_FP32;
    // kernel arguments (copied into the node at set-up / update)
    const void *nodes = f32 ? (const void *)c->d_node32 : (const void *)c->d_node64;
    const void *eh = f32 ? (const void *)c->d_eh32 : (const void *)c->d_eh64;
    const TpRegion *reg = (const TpRegion *)L.d_regions.p;
    const TpIntegralDesc *desc = (const TpIntegralDesc *)ctx->d_desc.p;
    double *fv = (double *)L.d_fvals.p;
    TpRegionOut *ro = (TpRegionOut *)L.d_rout.p;
    int *flag = (int *)L.d_flag.p;
    int md = mode;
    size_t n = nreg;
    void *pargs[] = {(void *)&nodes, (void *)&eh, (void *)&n, (void *)&reg, (void *)&desc, (void *)&fv, (void *)&md,
                     (void *)&flag};
    const int *cflag = flag;
    void *rargs[] = {(void *)&n, (void *)&reg, (void *)&fv, (void *)&ro, (void *)&cflag};
    hipKernelNodeParams kp{}, kr{};
    if (f32) region_points_shape<float>(lpp, nreg, &kp.func, &kp.gridDim, &kp.blockDim);
    else region_points_shape<double>(lpp, nreg, &kp.func, &kp.gridDim, &kp.blockDim);
    kp.kernelParams = pargs;
    kr.func = (void *)k_region_rule;
    kr.gridDim = dim3(grid_for(nreg, 256));
    kr.blockDim = dim3(256);
    kr.kernelParams = rargs;
    const size_t bin = sizeof(TpRegion) * nreg, bout = sizeof(TpRegionOut) * nreg;
    bool fresh = false;
    if (L.exec) {  // update the instantiated chain; an update the runtime refuses rebuilds it
        hipError_t u = hipGraphExecMemcpyNodeSetParams1D(L.exec, L.n_in, L.d_regions.p, L.h_regions.p, bin,
                                                         hipMemcpyHostToDevice);
        if (u == hipSuccess) u = hipGraphExecKernelNodeSetParams(L.exec, L.n_pts, &kp);
        if (u == hipSuccess) u = hipGraphExecKernelNodeSetParams(L.exec, L.n_rule, &kr);
        if (u == hipSuccess)
            u = hipGraphExecMemcpyNodeSetParams1D(L.exec, L.n_out, L.h_rout.p, L.d_rout.p, bout, hipMemcpyDeviceToHost);
        if (u != hipSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(L.exec);
            (void)hipGraphDestroy(L.graph);
            L.exec = nullptr;
            L.graph = nullptr;
            ctx->stats.graph_rebuilds += 1;
        }
    }
    if (!L.exec) {  // build the chain
        TP_HIP(hipGraphCreate(&L.graph, 0));
        hipGraphNode_t e0, em, e1, ed;
        TP_HIP(hipGraphAddMemcpyNode1D(&L.n_in, L.graph, nullptr, 0, L.d_regions.p, L.h_regions.p, bin,
                                       hipMemcpyHostToDevice));
        TP_HIP(hipGraphAddEventRecordNode(&e0, L.graph, &L.n_in, 1, L.ev0));
        TP_HIP(hipGraphAddKernelNode(&L.n_pts, L.graph, &e0, 1, &kp));
        TP_HIP(hipGraphAddEventRecordNode(&em, L.graph, &L.n_pts, 1, L.evm));
        TP_HIP(hipGraphAddKernelNode(&L.n_rule, L.graph, &em, 1, &kr));
        TP_HIP(hipGraphAddEventRecordNode(&e1, L.graph, &L.n_rule, 1, L.ev1));
        TP_HIP(hipGraphAddMemcpyNode1D(&L.n_out, L.graph, &e1, 1, L.h_rout.p, L.d_rout.p, bout, hipMemcpyDeviceToHost));
        TP_HIP(hipGraphAddEventRecordNode(&ed, L.graph, &L.n_out, 1, L.done));
        TP_HIP(hipGraphInstantiate(&L.exec, L.graph, nullptr, nullptr, 0));
        fresh = true;
    }
    (void)fresh;
    TP_HIP(hipGraphLaunch(L.exec, st));
    return 0;
}

// ------------------------------------------------------------------ launch-shape autotuner
// The waves-per-point split of k_region_points is chosen per launch-size bucket by timing every
// shape on a synthetic batch of that size, once per context and precision, on the first batched
// call (a few tens of ms).  Small launches (the tail rounds of a batch) favour more waves per
// point, large ones one wave.  Every shape computes the same bits, so the choice never changes a
// result.  tp_set_region_lanes forces one shape (0 restores tuning).
int tp_region_lanes(tp_ctx *ctx, const tp_cosmo *c, int mode, size_t np)
{
    if (ctx->region_lpp_force) return ctx->region_lpp_force;
    const int im = mode == TP_MODE_FP32 ? 1 : 0;
    if (mode == TP_MODE_ANALYTIC) return 1;
    if (!ctx->region_tuned[im]) {
        ctx->region_tuned[im] = true;  // also when tuning fails: then the default shape stays
        const size_t sizes[4] = {256, 2048, 8192, 32768};  // regions per bucket probe
        const size_t nmax = sizes[3];
        TpRegion *d_reg = nullptr;
        TpIntegralDesc *d_desc = nullptr;
        double *d_f = nullptr;
        int *d_fl = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        bool ok = hipMalloc(&d_reg, nmax * sizeof(TpRegion)) == hipSuccess &&
                  hipMalloc(&d_desc, sizeof(TpIntegralDesc)) == hipSuccess &&
                  hipMalloc(&d_f, nmax * TP_NPTS_GM3 * sizeof(double)) == hipSuccess &&
                  hipMalloc(&d_fl, sizeof(int)) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
                  hipEventCreate(&e1) == hipSuccess;
        if (ok) {
            // regions spread in log ell over [1, 7000]^2 x [0, 2 pi] for radii (2', 4', 8'), cosmology 0
            std::vector<TpRegion> hr(nmax);
            for (size_t k = 0; k < nmax; ++k) {
                const double u = (double)((k * 2654435761u) % 1000003) / 1000003.0;
                const double v = (double)((k * 40503u + 17) % 999983) / 999983.0;
                const double ph = (double)((k * 97u + 5) % 1009) / 1009.0;
                TpRegion &R = hr[k];
                R.c[0] = std::exp(8.85 * u);
                R.c[1] = std::exp(8.85 * v);
                R.c[2] = 2 * M_PI * ph;
                R.hw[0] = 0.05 * R.c[0];
                R.hw[1] = 0.05 * R.c[1];
                R.hw[2] = 0.05;
                R.vol = 8 * R.hw[0] * R.hw[1] * R.hw[2];
                R.integral = 0;
                R.pad = 0;
            }
            const double arcmin = 2.9088820866e-4;
            TpIntegralDesc hd{{2 * arcmin, 4 * arcmin, 8 * arcmin}, 0, 0};
            ok = hipMemcpy(d_reg, hr.data(), nmax * sizeof(TpRegion), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(d_desc, &hd, sizeof(hd), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemset(d_fl, 0, sizeof(int)) == hipSuccess;
        }
        const int shapes[7] = {1, 2, 4, 8, 16, 32, 64};
        for (int b = 0; ok && b < 4; ++b) {
            float best = 0;
            for (int s = 0; ok && s < 7; ++s) {
                float ms = 0;
                for (int rep = 0; rep < 3; ++rep) {  // first launch warms up
                    ok = hipEventRecord(e0, ctx->stream) == hipSuccess;
                    if (im)
                        launch_region_points<float>(ctx->stream, shapes[s], c->d_node32, c->d_eh32, sizes[b], d_reg,
                                                    d_desc, d_f, TP_MODE_FP32, d_fl);
                    else
                        launch_region_points<double>(ctx->stream, shapes[s], c->d_node64, c->d_eh64, sizes[b], d_reg,
                                                     d_desc, d_f, TP_MODE_FP64, d_fl);
                    ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, ctx->stream) == hipSuccess &&
                         hipEventSynchronize(e1) == hipSuccess;
                    float t = 0;
                    if (ok && rep > 0 && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms += t;
                }
                if (ok && (s == 0 || ms < best)) {
                    best = ms;
                    ctx->region_lpp[im][b] = shapes[s];
                }
            }
        }
        if (d_reg) (void)hipFree(d_reg);
        if (d_desc) (void)hipFree(d_desc);
        if (d_f) (void)hipFree(d_f);
        if (d_fl) (void)hipFree(d_fl);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (!ok) return tp_set_error(TP_EHIP, "tp_launch_regions: launch-shape tuning failed"), -1;
    }
    return ctx->region_lpp[im][region_bucket(np)];
}

extern "C" int tp_set_region_lanes(tp_ctx *ctx, int lanes)
{
    if (!ctx || !(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 ||
                  lanes == 64))
        return tp_set_error(TP_EINVAL,
                            "tp_set_region_lanes: lanes must be 0 (tuned), 1, 2, 4, 8, or 16, 32, 64 (region-cached)");
    ctx->region_lpp_force = lanes;
    return 0;
}

extern "C" int tp_region_lanes_table(tp_ctx *ctx, int mode, int *lanes4)
{
    if (!ctx || !lanes4 || (mode != TP_MODE_FP64 && mode != TP_MODE_FP32))
        return tp_set_error(TP_EINVAL, "tp_region_lanes_table: bad arguments");
    const int im = mode == TP_MODE_FP32 ? 1 : 0;
    for (int b = 0; b < 4; ++b) lanes4[b] = ctx->region_tuned[im] ? ctx->region_lpp[im][b] : 0;
    return 0;
}
