// tp_internal.h -- data layouts shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "threepoint_amd.h"
#include "tp_cubature.h"

#define TP_NNODES 96      // Gauss-Legendre z-nodes of the Limber integral
#define TP_NZ_MAX 1024    // max redshift bins of the tables
#define TP_NPTS_GM3 33    // Genz-Malik points per 3-D region
#define TP_NBASE 8        // Gauss-Legendre z-nodes of the surrogate's baseline Limber sum

// ---------------------------------------------------------------------------
// Per-cosmology constants, FP64, built on the device by tp_cosmo_build.
// Everything that depends on the cosmology only (Eisenstein & Hu constants,
// sigma8 normalisation, Limber prefactor) is folded here once.
// ---------------------------------------------------------------------------
// A tabulated linear P(k) of one cosmology (the reference's Pk_given input,
// classed_version/bispectrum.cpp:81-89, 785-806); nk == 0 selects Eisenstein & Hu.
// idx[c] is the table interval at the start of cell c of a uniform ln k grid (minus half a
// cell), so a lookup starts within a step or two of its interval.
struct TpPkTable {
    const double *k = nullptr;  // [nk] ascending, h/Mpc
    const double *p = nullptr;  // [nk] (Mpc/h)^3
    const int *idx = nullptr;   // [ng]
    int nk = 0, ng = 0;
    double lnk0 = 0, inv_dlnk = 0;
};

struct TpCosmoConst {
    double h, sigma8, omb, omc, ns, w, om, ow, zmax;
    double norm;          // sigma8 / sigma_TopHat(8 Mpc/h) with norm=1
    // Eisenstein & Hu (1998) no-wiggle fit, classed_version/bispectrum.cpp:811-841
    double eh_fb, eh_sqrt_alnu, eh_sh, eh_omh2, eh_theta2, eh_cL;
    double limber_pref;   // 27/8 om^3 (H0/c)^5
    int n_z;
    int pad;
    TpPkTable tab;        // tabulated P(k) or none
};

// ---------------------------------------------------------------------------
// One Limber z-node of one cosmology.  B_kappa(l) = sum_n wz_n * B(l * ks_n, z_n).
// All z-only factors of the BiHalofit fit (Takahashi et al. 2020, app. B) are
// pre-evaluated; the per-point work is what depends on the wavenumbers.
// ---------------------------------------------------------------------------
template <class T>
struct TpNodeT {
    T wz;        // W_i dx * 27/8 om^3 (H0/c)^5 * (g (1+z))^3 / (f_K E(z))
    T ks;        // r_sigma / f_K : q = ell * ks (dimensionless wavenumber)
    T inv_rs;    // 1 / r_sigma   : k = q / r_sigma [h/Mpc]
    T lnA0;      // ln 10 * (-2.167 - 2.944 L - 1.106 L^2 - 2.865 L^3), L = log10(D1 sigma8)
    T c1;        // 10^(0.182 + 0.57 n_eff)
    T m031ln10;  // -0.310 ln 10
    T bn, cn;    // B7 parameters
    T a0ln, a1ln;// alpha_n = 10^(a0 + a1 r2^2) -> exp(a0ln + a1ln r2^2)
    T acap;      // 1 - 2/3 n_s
    T b0ln, b1ln;// beta_n  = exp(b0ln + b1ln r2)
    T fn, gn, hn, mn, nn, mun, nun, pn, en; // B6/B9 parameters
    T D1sq;      // D1^2
    T dn;        // F2 correction 10^(-0.483 + 0.892 L - 0.086 Omega_m(z))
    T lnkoff;    // ln(2997.9 / r_sigma)
    T pk_pref;   // 2 pi^2 h^3 norm^2 / (h / r_sigma)^3 : P = pk_pref * delk' / q^3
    T lnks;      // ln(ks)
};
typedef TpNodeT<double> TpNode64;
typedef TpNodeT<float> TpNode32;

// Per-cosmology EH constants needed inside the per-point kernel.
template <class T>
struct TpEHT {
    T h, sqrt_alnu, one_m_sqrt_alnu, sh043, omh2, theta2, cL, three_ns, e_const;
    TpPkTable tab;  // when tab.nk > 0 the linear P(k) is this table instead of the fit
};
typedef TpEHT<double> TpEH64;
typedef TpEHT<float> TpEH32;

// ---------------------------------------------------------------------------
// Device tables of ncosmo cosmologies (all FP64 except the FP32 node copies).
// ---------------------------------------------------------------------------
struct tp_cosmo {
    int device = 0;
    int ncosmo = 0, n_z = 0;
    double z_max = 0;
    TpCosmoConst *d_const = nullptr;  // [ncosmo]
    double *d_tab = nullptr;          // [ncosmo][6][n_z]: fK, g, D1, rs, neff, ncur
    double *d_nz = nullptr;           // [ncosmo][n_z]
    TpNode64 *d_node64 = nullptr;     // [ncosmo][96]
    TpNode32 *d_node32 = nullptr;     // [ncosmo][96]
    TpEH64 *d_eh64 = nullptr;         // [ncosmo]
    TpEH32 *d_eh32 = nullptr;         // [ncosmo]
    TpNode32 *d_base32 = nullptr;     // [ncosmo][8]: 8-node FP32 Limber tables (surrogate baseline)
    int *d_err = nullptr;             // build error flag
    void *d_pk = nullptr;             // tabulated P(k): k, p [ncosmo][nk] then idx [ncosmo][ng]
    void *d_arena = nullptr;          // one allocation holding every array above except d_pnode
    void *d_pnode = nullptr;          // [ncosmo][96] TpPNode: the P Limber sum's z-nodes (tp_power.hip)
    void *d_pnode32 = nullptr;        // [ncosmo][96] TpPNode32: their FP32 form (same allocation)
    int pk_nk = 0;                    // 0: Eisenstein & Hu
    std::vector<TpCosmoConst> h_const;
};

// ---------------------------------------------------------------------------
// Batched-engine region record and result (device <-> host, FP64).
// ---------------------------------------------------------------------------
struct TpRegion {
    double c[3], hw[3];
    double vol;
    int integral;  // index into the integral table of the launch
    int pad;
};
struct TpRegionOut {
    double val, err;
    double corr;   // surrogate mode: unbiased correction from the exactly evaluated points
    double cvar;   // surrogate mode: variance estimate of corr
    int split;
    int flag;      // nonzero: non-finite integrand value inside the region
};
struct TpIntegralDesc {
    double th[3];  // aperture radii [rad]
    int icosmo;
    int pad;
};

// Growable device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
};
struct HostPinned {
    void *p = nullptr;
    void *dp = nullptr;  // the same memory as seen by kernels (hipHostGetDevicePointer)
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
};

struct TpSurrogate;  // learned B_kappa model (tp_mlp.hip)

// Buffers and events of one group of integrals in the batched engine.  Two groups
// alternate: the device evaluates one group's regions while the host runs the
// other group's heap decisions.
struct TpLane {
    DevBuf d_regions, d_rout, d_fvals, d_flag;
    HostPinned h_regions, h_rout, h_flag;
    hipEvent_t ev0 = nullptr, evm = nullptr, ev1 = nullptr, done = nullptr;
    // the group's own stream (default priority): the groups' rounds run concurrently on the device,
    // so a small round of one group does not wait behind the other group's kernel
    hipStream_t stream = nullptr;
    // the round as one graph (tp_graph_round): copy in, ev0, points kernel, evm, rule kernel, ev1,
    // copy out, done; instantiated once, its node parameters updated every round
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t n_in = nullptr, n_pts = nullptr, n_rule = nullptr, n_out = nullptr;
    int create();
    void release();
};
constexpr int TP_NLANES = 4;  // groups of a batched call (at most; tp_ctx::nlanes are used)

constexpr size_t kTpKeepPoolBytes = size_t(16) << 30;  // default cap on the host region pools kept
// tp_map3_batch: two in-flight rounds of at least this many regions each run one after the other on the
// device (tp_set_serial_regions)
constexpr long long kTpSerialRegions = 65536;

// Guards of the batched cubatures (tp_set_batch_limits): a call whose integrals need more rounds or
// evaluate more regions than these ends with TP_ECAPACITY instead of growing without bound (the
// measured workloads stay below 300 rounds and 1e8 regions); the host-rule batches also evaluate a
// round in slices of at most kTpRoundPoints points, so the per-round point buffers stay bounded.
constexpr long long kTpMaxRounds = 20000;
constexpr long long kTpMaxRegions = 400000000LL;
constexpr long long kTpRoundPoints = 1LL << 23;

struct tp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;  // evm: after the integrand kernel
    hipEvent_t ev_batch = nullptr;  // start of a tp_map3_batch call: the groups' streams wait on it,
                                    // and their rounds are timed against it
    DevBuf d_in, d_out, d_regions, d_rout, d_fvals, d_desc, d_flag, d_misc;
    HostPinned h_in, h_out, h_regions, h_rout;
    tp_batch_stats stats{};
    // Gamma (Ogata) tables
    double gamma_prec_h = 0;
    unsigned gamma_prec_k = 0;
    double *d_ogata = nullptr;  // [6][prec_k]: psi6, prod6, psi2, prod2, ln psi6, ln psi2
    double gamma_q32[2] = {1e-5, 1e6};  // FP32 window of the Ogata terms' wavenumbers
    TpSurrogate *sur = nullptr;
    double sur_check_frac = 0.01, sur_tol = 1e-3;
    uint64_t sur_seed = 0x5eed;
    // surrogate mode: an integral's regions are evaluated exactly (FP32 form) until the integral holds this
    // many regions (tp_surrogate_exact_regions)
    long long sur_exact_regions = 1024;
    // surrogate mode: a triple whose largest radius is at least this many times its smallest is evaluated
    // exactly throughout (tp_surrogate_exact_ratio; 0 = none)
    double sur_exact_ratio = 16.0;
    // launch shape of the region-point kernel: waves per point for [FP64, FP32][size bucket]
    int region_lpp[2][4] = {{1, 1, 1, 1}, {1, 1, 1, 1}};
    bool region_tuned[2] = {false, false};
    int region_lpp_force = 0;   // 0 = tuned; 1, 2, 4, 8 = fixed
    bool graph_rounds = false;  // batched rounds replayed as one graph per lane (tp_set_graph_rounds)
    int nlanes = 2;             // groups in flight in tp_map3_batch (tp_set_batch_groups)
    long long zero_copy_regions = 2048;  // rounds up to this many regions skip the copies (pinned buffers)
    long long serial_regions = kTpSerialRegions;  // tp_set_serial_regions
    int host_threads = 0;             // OpenMP threads of the host-side cubature work (0: default)
    // tp_map3_batch's integrators of the last call, reused by the next (their region pools keep
    // their capacity, so a call of the same size allocates and page-faults nothing); dropped when
    // they hold more than keep_pool_bytes (tp_set_keep_pool_bytes)
    std::vector<tpc::AdaptiveIntegral31> map3_keep;              // also T1's (3-D) regions batches
    std::vector<tpc::AdaptiveIntegralT<2, 1>> map2_keep;        // <Map^2>'s (2-D)
    size_t keep_pool_bytes = kTpKeepPoolBytes;
    long long max_rounds = kTpMaxRounds;     // tp_set_batch_limits
    long long max_regions = kTpMaxRegions;
    long long round_points = kTpRoundPoints;
    bool round_log_on = false;        // tp_set_round_log: per-round trace of tp_map3_batch
    std::vector<double> round_log;    // 7 doubles per round (tp_round_log)
    TpLane lane[TP_NLANES];
};

// OpenMP threads for the host side of the batched cubatures (region records, heaps):
// ctx->host_threads when set, else dflt, at most omp_get_max_threads().  The <Map^3> engine, whose
// host work overlaps the device, defaults to kTpHostThreadsOverlap: that leaves CPU time to the
// calling thread and the HIP runtime (with as many spinning OpenMP threads as the process's CPU
// quota, a whole-process throttle of several ms was seen now and then between two rounds, with the
// device idle).  The synchronous batches (host work on the critical path) default to
// kTpHostThreadsSync.
constexpr int kTpHostThreadsOverlap = 8, kTpHostThreadsSync = 16;
int tp_host_threads(const tp_ctx *ctx, int dflt);

// Error helpers (thread-local last error string).
int tp_set_error(int code, const std::string &msg);
#define TP_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return tp_set_error(TP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// n-point Gauss-Legendre nodes and weights on [-1, 1], ascending (Newton iteration on P_n).
void tp_gauss_legendre(int n, double *x, double *w);

// Builds c->d_pnode from c's tables (tp_power.hip); called at the end of every table build.
int tp_build_pnodes(tp_ctx *ctx, tp_cosmo *c);

// Launchers implemented in the .hip files (all asynchronous on ctx->stream).
int tp_launch_map3_points(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double *th, size_t npts,
                          const double *d_x, double *d_f, int mode, int *d_flag);
int tp_launch_bkappa_points(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *d_l,
                            double *d_out, int mode, int *d_flag);
int tp_launch_regions(tp_ctx *ctx, const tp_cosmo *c, size_t nreg, const TpRegion *d_reg,
                      const TpIntegralDesc *d_desc, double *d_fvals, TpRegionOut *d_out, int mode,
                      int *d_flag, hipEvent_t evm, hipStream_t st);
// Copies `bytes` (a multiple of 4) from device memory to pinned host memory through its device
// pointer, as a kernel on st: the batched <Map^3> rounds read their results back this way, so that
// no DMA copy waiting on a kernel can hold up the other group's copies behind it.
int tp_copy_to_pinned(const void *d_src, void *h_dp, size_t bytes, hipStream_t st);
// One exact-mode round of lane L (nreg regions already in L.h_regions) as a graph replay: the same
// work and events as the copy + tp_launch_regions + copy sequence, queued on ctx->stream.
int tp_graph_round(tp_ctx *ctx, const tp_cosmo *c, TpLane &L, size_t nreg, int mode, hipStream_t st);
int tp_launch_gamma_points(tp_ctx *ctx, const tp_cosmo *c, int icosmo, int which, double x1, double x2,
                           double x3, size_t npts, const double *d_x, double *d_f, int *d_flag);
