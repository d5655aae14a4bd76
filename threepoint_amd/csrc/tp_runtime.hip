// tp_runtime.hip -- C ABI runtime: contexts, staging buffers, the plain per-batch
// integrand path, the drop-in hcubature_v, and the batched many-integral engine.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "tp_cubature.h"
#include "tp_internal.h"

static thread_local std::string g_err;

int tp_set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

extern "C" const char *tp_last_error(void) { return g_err.c_str(); }

int DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return 0;
    size_t want = std::max(bytes, cap * 2);
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return tp_set_error(TP_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    cap = want;
    return 0;
}
void DevBuf::release()
{
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
}
int HostPinned::ensure(size_t bytes)
{
    if (bytes <= cap) return 0;
    size_t want = std::max(bytes, cap * 2);
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) return tp_set_error(TP_EHIP, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    e = hipHostGetDevicePointer(&dp, p, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(p);
        p = dp = nullptr;
        return tp_set_error(TP_EHIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    }
    cap = want;
    return 0;
}
void HostPinned::release()
{
    if (p) hipHostFree(p);
    p = dp = nullptr;
    cap = 0;
}

extern "C" int tp_device_count(int *n)
{
    int k = 0;
    hipError_t e = hipGetDeviceCount(&k);
    if (e != hipSuccess) {
        *n = 0;
        return tp_set_error(TP_ENODEV, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *n = k;
    return 0;
}

// ------------------------------------------------ staging of the entry point that allocates nothing
// tp_map3_eval_noalloc stages its points through `scratch`, a file-scope device array (one copy per
// GPU, held until the process ends): bytes [0, 256) hold the point kernel's flag word, then
// TP_NOALLOC_MAX_POINTS x 3 doubles of points, then TP_NOALLOC_MAX_POINTS doubles of values.
constexpr size_t kNoAllocIn = 256;
constexpr size_t kNoAllocOut = kNoAllocIn + (size_t)TP_NOALLOC_MAX_POINTS * 3 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocOut + (size_t)TP_NOALLOC_MAX_POINTS * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
// The array's device address on each GPU, looked up once per device with that device current: in host
// code `scratch` names a host-side stand-in, never the device memory, so only these addresses are used.
constexpr int kNoAllocMaxDevices = 64;
static void *g_scratch_dev[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_lookup(int device)  // device is current
{
    if (device < 0 || device >= kNoAllocMaxDevices)
        return tp_set_error(TP_EINVAL, "device index beyond the static staging table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return tp_set_error(TP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
    return 0;
}

namespace {
__global__ void k_clear_word(int *w) { *w = 0; }
__global__ void k_copy_words(const unsigned *__restrict__ a, unsigned *__restrict__ b, size_t n)
{
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) b[i] = a[i];
}
}  // namespace

int tp_copy_to_pinned(const void *d_src, void *h_dp, size_t bytes, hipStream_t st)
{
    if (bytes % 4 || bytes > 4096) return tp_set_error(TP_EINVAL, "tp_copy_to_pinned: bad size");
    hipLaunchKernelGGL(k_copy_words, dim3(1), dim3(64), 0, st, (const unsigned *)d_src, (unsigned *)h_dp,
                       bytes / 4);
    TP_HIP(hipGetLastError());
    return 0;
}

extern "C" int tp_ctx_create(int device, tp_ctx **out)
{
    if (!out) return tp_set_error(TP_EINVAL, "tp_ctx_create: NULL out");
    int n = 0;
    if (tp_device_count(&n) || n == 0) return tp_set_error(TP_ENODEV, "tp_ctx_create: no HIP device");
    if (device < 0 || device >= n) return tp_set_error(TP_EINVAL, "tp_ctx_create: bad device index");
    TP_HIP(hipSetDevice(device));
    tp_ctx *c = new tp_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return tp_set_error(TP_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    hipEventCreate(&c->evm);
    hipEventCreate(&c->ev_batch);
    if (c->d_flag.ensure(256)) {
        delete c;
        return TP_EHIP;
    }
    int rc = scratch_lookup(device);  // once per device: the static staging of tp_map3_eval_noalloc
    if (rc) {
        tp_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

int TpLane::create()
{
    if (done) return 0;
    TP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (hipEvent_t *e : {&ev0, &evm, &ev1, &done}) TP_HIP(hipEventCreate(e));
    if (h_flag.ensure(64) || d_flag.ensure(256)) return TP_EHIP;
    return 0;
}
void TpLane::release()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (DevBuf *b : {&d_regions, &d_rout, &d_fvals, &d_flag}) b->release();
    for (HostPinned *b : {&h_regions, &h_rout, &h_flag}) b->release();
    for (hipEvent_t *e : {&ev0, &evm, &ev1, &done})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

void tp_surrogate_free(tp_ctx *ctx);  // tp_mlp.hip

extern "C" int tp_ctx_destroy(tp_ctx *ctx)
{
    if (!ctx) return 0;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    tp_surrogate_free(ctx);
    for (DevBuf *b : {&ctx->d_in, &ctx->d_out, &ctx->d_regions, &ctx->d_rout, &ctx->d_fvals, &ctx->d_desc,
                      &ctx->d_flag, &ctx->d_misc})
        b->release();
    for (HostPinned *b : {&ctx->h_in, &ctx->h_out, &ctx->h_regions, &ctx->h_rout}) b->release();
    for (TpLane &L : ctx->lane) L.release();
    if (ctx->d_ogata) hipFree(ctx->d_ogata);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    hipEventDestroy(ctx->evm);
    hipEventDestroy(ctx->ev_batch);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

namespace {
__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}
}  // namespace

extern "C" int tp_hbm_copy_probe(tp_ctx *ctx, size_t nbytes, double *gbps)
{
    if (!ctx || nbytes < 16 || nbytes % 16) return tp_set_error(TP_EINVAL, "tp_hbm_copy_probe: nbytes must be a positive multiple of 16");
    TP_HIP(hipSetDevice(ctx->device));
    void *a = nullptr, *b = nullptr;
    TP_HIP(hipMalloc(&a, nbytes));
    hipError_t e = hipMalloc(&b, nbytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return tp_set_error(TP_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    const size_t n = nbytes / 16;
    const int blocks = 256 * 16;  // 16 workgroups per CU, grid-stride
    (void)hipMemsetAsync(a, 1, nbytes, ctx->stream);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)a, (uint4 *)b, n);  // warm
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)a, (uint4 *)b, n);
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    e = hipStreamSynchronize(ctx->stream);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return tp_set_error(TP_EHIP, std::string("tp_hbm_copy_probe: ") + hipGetErrorString(e));
    if (gbps) *gbps = ms > 0 ? 2.0 * (double)nbytes / (ms * 1e-3) / 1e9 : 0.0;
    return 0;
}

extern "C" int tp_last_batch_stats(const tp_ctx *ctx, tp_batch_stats *out)
{
    if (!ctx || !out) return tp_set_error(TP_EINVAL, "tp_last_batch_stats: NULL");
    *out = ctx->stats;
    return 0;
}

static int check_cosmo(tp_ctx *ctx, const tp_cosmo *c, int icosmo)
{
    if (!ctx || !c) return tp_set_error(TP_EINVAL, "NULL context or cosmology");
    if (icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "cosmology index out of range");
    if (c->device != ctx->device) return tp_set_error(TP_EINVAL, "cosmology lives on another device");
    return 0;
}

static int read_flag(tp_ctx *ctx, int *flag)
{
    int rc = ctx->h_out.ensure(64);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(ctx->h_out.p, ctx->d_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(flag, ctx->h_out.p, sizeof(int));
    return 0;
}

// Runs one batch of points through a point kernel: host in -> device -> host out.
template <class Launch>
static int run_points(tp_ctx *ctx, size_t npts, size_t in_per, size_t out_per, const double *in, double *out,
                      Launch launch, const char *what)
{
    if (npts == 0) return 0;
    size_t bin = npts * in_per * sizeof(double), bout = npts * out_per * sizeof(double);
    int rc;
    if ((rc = ctx->d_in.ensure(bin)) || (rc = ctx->d_out.ensure(bout)) || (rc = ctx->h_in.ensure(bin)) ||
        (rc = ctx->h_out.ensure(std::max(bout, (size_t)64))))
        return rc;
    std::memcpy(ctx->h_in.p, in, bin);
    TP_HIP(hipMemsetAsync(ctx->d_flag.p, 0, sizeof(int), ctx->stream));
    TP_HIP(hipMemcpyAsync(ctx->d_in.p, ctx->h_in.p, bin, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = launch((const double *)ctx->d_in.p, (double *)ctx->d_out.p, (int *)ctx->d_flag.p))) return rc;
    TP_HIP(hipMemcpyAsync(ctx->h_out.p, ctx->d_out.p, bout, hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out, ctx->h_out.p, bout);
    int flag = 0;
    if ((rc = read_flag(ctx, &flag))) return rc;
    if (flag) return tp_set_error(TP_ENONFINITE, std::string(what) + ": non-finite ell or B_kappa value");
    return 0;
}

extern "C" int tp_bkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ell123,
                              double *out, int mode)
{
    int rc = check_cosmo(ctx, c, icosmo);
    if (rc) return rc;
    if (npts && (!ell123 || !out)) return tp_set_error(TP_EINVAL, "tp_bkappa_eval: NULL buffer");
    if (mode != TP_MODE_FP64 && mode != TP_MODE_FP32 && mode != TP_MODE_ANALYTIC)
        return tp_set_error(TP_EINVAL, "tp_bkappa_eval: mode must be FP64, FP32 or ANALYTIC");
    TP_HIP(hipSetDevice(ctx->device));
    return run_points(
        ctx, npts, 3, 1, ell123, out,
        [&](const double *dx, double *dy, int *fl) { return tp_launch_bkappa_points(ctx, c, icosmo, npts, dx, dy, mode, fl); },
        "tp_bkappa_eval");
}

extern "C" int tp_map3_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], size_t npts,
                            const double *x, double *fval, int mode)
{
    int rc = check_cosmo(ctx, c, icosmo);
    if (rc) return rc;
    if (!theta || (npts && (!x || !fval))) return tp_set_error(TP_EINVAL, "tp_map3_eval: NULL buffer");
    if (mode != TP_MODE_FP64 && mode != TP_MODE_FP32 && mode != TP_MODE_ANALYTIC)
        return tp_set_error(TP_EINVAL, "tp_map3_eval: mode must be FP64, FP32 or ANALYTIC");
    TP_HIP(hipSetDevice(ctx->device));
    return run_points(
        ctx, npts, 3, 1, x, fval,
        [&](const double *dx, double *dy, int *fl) {
            return tp_launch_map3_points(ctx, c, icosmo, theta, npts, dx, dy, mode, fl);
        },
        "tp_map3_eval");
}

extern "C" int tp_map3_eval_noalloc(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], size_t npts,
                                    const double *x, double *fval, int mode)
{
    int rc = check_cosmo(ctx, c, icosmo);
    if (rc) return rc;
    if (!theta || (npts && (!x || !fval))) return tp_set_error(TP_EINVAL, "tp_map3_eval_noalloc: NULL buffer");
    if (mode != TP_MODE_FP64 && mode != TP_MODE_FP32 && mode != TP_MODE_ANALYTIC)
        return tp_set_error(TP_EINVAL, "tp_map3_eval_noalloc: mode must be FP64, FP32 or ANALYTIC");
    if (npts > (size_t)TP_NOALLOC_MAX_POINTS)
        return tp_set_error(TP_ECAPACITY, "tp_map3_eval_noalloc: npts = " + std::to_string(npts) +
                                              " exceeds TP_NOALLOC_MAX_POINTS = " +
                                              std::to_string(TP_NOALLOC_MAX_POINTS) + " (split the batch)");
    if (npts == 0) return 0;
    TP_HIP(hipSetDevice(ctx->device));
    unsigned char *base = ctx->device < kNoAllocMaxDevices ? (unsigned char *)g_scratch_dev[ctx->device] : nullptr;
    if (!base) return tp_set_error(TP_EINVAL, "tp_map3_eval_noalloc: context without static staging");
    int *d_flag = (int *)base;
    double *d_in = (double *)(base + kNoAllocIn), *d_out = (double *)(base + kNoAllocOut);
    // the flag word keeps the last call's contents: cleared by a kernel (no memset node if captured)
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(1), 0, ctx->stream, d_flag);
    TP_HIP(hipGetLastError());
    TP_HIP(hipMemcpyAsync(d_in, x, npts * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if ((rc = tp_launch_map3_points(ctx, c, icosmo, theta, npts, d_in, d_out, mode, d_flag))) return rc;
    TP_HIP(hipMemcpyAsync(fval, d_out, npts * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    TP_HIP(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    if (flag) return tp_set_error(TP_ENONFINITE, "tp_map3_eval_noalloc: non-finite ell or B_kappa value");
    return 0;
}

// ------------------------------------------------------------- hcubature_v
static thread_local long long g_last_neval = 0;
extern "C" long long tp_hcubature_last_neval(void) { return g_last_neval; }

extern "C" int tp_hcubature_v(unsigned fdim, tp_integrand_v f, void *fdata, unsigned dim, const double *xmin,
                              const double *xmax, size_t maxEval, double reqAbsError, double reqRelError, int norm,
                              double *val, double *err)
{
    // classed_version/hcubature.cpp:1023-1068
    g_last_neval = 0;
    if (fdim == 0) return 0;
    if (dim == 0) {
        if (f(0, 1, xmin, fdata, fdim, val)) return 1;
        for (unsigned i = 0; i < fdim; ++i) err[i] = 0;
        return 0;
    }
    tpc::AdaptiveIntegral I;
    if (!I.init(dim, fdim, xmin, xmax, maxEval, reqAbsError, reqRelError, norm)) {
        for (unsigned i = 0; i < fdim; ++i) {
            val[i] = 0;
            err[i] = HUGE_VAL;
        }
        tp_set_error(TP_EINVAL, "tp_hcubature_v: unsupported dim/fdim/norm");
        return 1;
    }
    const unsigned np = tpc::rule_points(dim);
    std::vector<double> pts, vals;
    for (;;) {
        const auto &pend = I.pending();
        size_t nR = pend.size();
        pts.resize(nR * np * dim);
        vals.resize(nR * np * fdim);
        for (size_t r = 0; r < nR; ++r) tpc::rule_points_of(dim, I.regions()[pend[r]], pts.data() + r * np * dim);
        if (f(dim, nR * np, pts.data(), fdata, fdim, vals.data())) {
            tp_set_error(TP_ECALLBACK, "tp_hcubature_v: integrand callback returned nonzero");
            return 1;
        }
        for (size_t r = 0; r < nR; ++r)
            tpc::rule_apply(dim, fdim, I.regions()[pend[r]], vals.data() + r * np * fdim);
        I.absorb();
        if (I.advance()) break;
    }
    I.result(val, err);
    g_last_neval = (long long)I.num_eval();
    return 0;
}

// --------------------------------------------------- one triple, plain path
struct Map3CbData {
    tp_ctx *ctx;
    const tp_cosmo *c;
    int icosmo, mode;
    double th[3];
    int rc;
};

static int map3_cb(unsigned ndim, size_t npt, const double *x, void *fdata, unsigned fdim, double *fval)
{
    // the integrand_v adapter of classed_version/apertureStatistics.cpp:254-279, on the GPU
    Map3CbData *d = (Map3CbData *)fdata;
    if (ndim != 3 || fdim != 1) {
        d->rc = tp_set_error(TP_EINVAL, "map3 integrand: wrong dimensions");
        return 1;
    }
    d->rc = tp_map3_eval(d->ctx, d->c, d->icosmo, d->th, npt, x, fval, d->mode);
    return d->rc ? 1 : 0;
}

extern "C" int tp_map3_single(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], double reltol,
                              int mode, double *val, double *err, long long *neval)
{
    // classed_version/apertureStatistics.cpp:321-353, 3-D branch
    int rc = check_cosmo(ctx, c, icosmo);
    if (rc) return rc;
    if (!theta || !val) return tp_set_error(TP_EINVAL, "tp_map3_single: NULL");
    double tmin = std::min(theta[0], std::min(theta[1], theta[2]));
    if (!(tmin > 0)) return tp_set_error(TP_EINVAL, "tp_map3_single: theta must be positive");
    double lMax = 10. / tmin;
    double vmin[3] = {1., 1., 0.}, vmax[3] = {lMax, lMax, 2 * M_PI};
    Map3CbData d{ctx, c, icosmo, mode, {theta[0], theta[1], theta[2]}, 0};
    double r = 0, e = 0;
    auto t0 = std::chrono::steady_clock::now();
    int st = tp_hcubature_v(1, map3_cb, &d, 3, vmin, vmax, 0, 0, reltol, tpc::L1, &r, &e);
    if (st) return d.rc ? d.rc : tp_set_error(TP_EINVAL, "tp_map3_single: cubature failed");
    *val = r / 8 / M_PI / M_PI / M_PI;
    if (err) *err = e / 8 / M_PI / M_PI / M_PI;
    if (neval) *neval = g_last_neval;
    ctx->stats = tp_batch_stats{};
    ctx->stats.points = g_last_neval;
    ctx->stats.wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

// ----------------------------------------------------------- batched engine
// Guards against integrals that stop converging (every mode): ctx->max_rounds and ctx->max_regions
// (tp_set_batch_limits) end a run that needs more with TP_ECAPACITY.
// adaptive grouping of tp_map3_batch (two groups): split above, merge below these round sizes
static const size_t kSplitRegions = 8192, kMergeRegions = 2048;
// Rounds of at least tp_ctx::serial_regions regions fill the device on their own: such a round's kernels
// wait for the other group's in-flight round's kernels (an event wait on the device), so the two groups'
// rounds run one after the other instead of side by side.  Side by side they end together and the device
// idles while the host absorbs the first of them; one after the other, each round's host work overlaps the
// other's kernels.
static const size_t kSurrogateMaxEval = 4000000;  // points per triple before an exact redo
int tp_region_lanes(tp_ctx *ctx, const tp_cosmo *c, int mode, size_t np);  // tp_map3.hip
// tp_mlp.hip: a surrogate round queued on a group's stream, and its check statistics once it ran
int tp_surrogate_queue(tp_ctx *ctx, const tp_cosmo *c, int lane, size_t nreg, const TpRegion *d_reg,
                       const TpIntegralDesc *d_desc, TpRegionOut *d_out, int *d_flag, hipStream_t st,
                       unsigned long long round_key);
int tp_surrogate_result(tp_ctx *ctx, int lane, size_t nreg, bool *fallback);

extern "C" int tp_map3_batch(tp_ctx *ctx, const tp_cosmo *c, size_t ntriples, const int *icosmo,
                             const double *theta_rad, double reltol, int mode, double *val, double *err,
                             long long *neval)
{
    if (!ctx || !c || !icosmo || !theta_rad || !val) return tp_set_error(TP_EINVAL, "tp_map3_batch: NULL");
    if (mode < TP_MODE_FP64 || mode > TP_MODE_ANALYTIC) return tp_set_error(TP_EINVAL, "tp_map3_batch: bad mode");
    if (mode == TP_MODE_SURROGATE && !ctx->sur)
        return tp_set_error(TP_EINVAL, "tp_map3_batch: surrogate mode needs a trained or loaded surrogate");
    if (mode == TP_MODE_SURROGATE && c->pk_nk)
        return tp_set_error(TP_EINVAL, "tp_map3_batch: the surrogate is trained on the Eisenstein & Hu P(k); "
                                       "tables built from a tabulated P(k) need an exact mode");
    if (c->device != ctx->device) return tp_set_error(TP_EINVAL, "tp_map3_batch: cosmology on another device");
    if (ntriples == 0) return 0;
    TP_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    ctx->stats = tp_batch_stats{};
    std::vector<TpIntegralDesc> desc(ntriples);
    std::vector<char> wide(ntriples, 0);  // surrogate mode: triples evaluated exactly throughout
    std::vector<tpc::AdaptiveIntegral31> I;
    I.swap(ctx->map3_keep);  // the last call's integrators (their pools keep capacity), if kept
    if (I.size() < ntriples) I.resize(ntriples);
    for (size_t t = 0; t < ntriples; ++t) {
        const double *th = theta_rad + 3 * t;
        if (icosmo[t] < 0 || icosmo[t] >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_map3_batch: bad icosmo");
        double tmin = std::min(th[0], std::min(th[1], th[2]));
        if (!(tmin > 0)) return tp_set_error(TP_EINVAL, "tp_map3_batch: theta must be positive");
        desc[t] = TpIntegralDesc{{th[0], th[1], th[2]}, icosmo[t], 0};
        const double tmax = std::max(th[0], std::max(th[1], th[2]));
        wide[t] = ctx->sur_exact_ratio > 0 && tmax >= ctx->sur_exact_ratio * tmin;
        double lMax = 10. / tmin;  // classed_version/apertureStatistics.cpp:325-326
        double vmin[3] = {1., 1., 0.}, vmax[3] = {lMax, lMax, 2 * M_PI};
        // surrogate mode: a per-triple point budget; a triple that exhausts it is redone exactly
        I[t].init(3, 1, vmin, vmax, mode == TP_MODE_SURROGATE ? kSurrogateMaxEval : 0, 0, reltol, tpc::L1);
    }
    int rc;
    if ((rc = ctx->d_desc.ensure(sizeof(TpIntegralDesc) * ntriples))) return rc;
    TP_HIP(hipMemcpyAsync(ctx->d_desc.p, desc.data(), sizeof(TpIntegralDesc) * ntriples, hipMemcpyHostToDevice,
                          ctx->stream));
    // Integrals are split into groups, each with its own stream: while the device evaluates
    // one group's regions, the host absorbs the other group's results and runs its heap
    // decisions (with two groups the split adapts to the round size, see below).  Each
    // integral's decision sequence is unchanged.
    // Surrogate mode: the regions of integrals that hold fewer than ctx->sur_exact_regions regions are
    // evaluated exactly (the FP32 form), so an integral that the reference algorithm settles in a few
    // rounds takes the same decisions as the exact paths; the others go through the surrogate with its
    // checks, whose statistics are read when the round completes (a round that misses the tolerance is
    // then re-evaluated exactly before its results are absorbed).
    const int nl = std::max(1, std::min(ctx->nlanes, TP_NLANES));
    const int exact_mode = mode == TP_MODE_SURROGATE ? TP_MODE_FP32 : mode;
    struct Group {
        std::vector<size_t> active, offs;
        size_t nreg = 0;
        size_t nexact = 0;  // surrogate mode: the round's first nexact regions are evaluated exactly
        long long rounds = 0;
        bool inflight = false;
        double tl0 = 0, tl1 = 0;  // host time of the round's launch [ms since the call began] (round log)
    };
    Group G[TP_NLANES];
    for (size_t t = 0; t < ntriples; ++t) G[nl == 1 ? 0 : (int)(t % nl)].active.push_back(t);
    // launch-shape tuning (first batched call of this precision) before any round is timed
    if (exact_mode == TP_MODE_FP64 || exact_mode == TP_MODE_FP32)
        if (tp_region_lanes(ctx, c, exact_mode, 0) < 0) return TP_EHIP;
    // each group's rounds go to its own stream (one group: the context's stream), ordered after the
    // descriptor upload above
    hipStream_t gst[TP_NLANES];
    TP_HIP(hipEventRecord(ctx->ev_batch, ctx->stream));
    for (int g = 0; g < nl; ++g) {
        if ((rc = ctx->lane[g].create())) return rc;
        gst[g] = nl == 1 ? ctx->stream : ctx->lane[g].stream;
        if (nl > 1) TP_HIP(hipStreamWaitEvent(gst[g], ctx->ev_batch, 0));
        // the flag word is only ever raised, and a raised flag ends the batch: clear it once per call
        TP_HIP(hipMemsetAsync(ctx->lane[g].d_flag.p, 0, sizeof(int), gst[g]));
    }
    const int nthreads = tp_host_threads(ctx, kTpHostThreadsOverlap);
    long long total_regions = 0;
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    ctx->round_log.clear();

    // gathers the group's pending regions and queues their evaluation
    auto launch = [&](int g) -> int {
        Group &Gr = G[g];
        TpLane &L = ctx->lane[g];
        auto th0 = clk::now();
        Gr.tl0 = ms_since(t0);
        if (Gr.rounds >= ctx->max_rounds)
            return tp_set_error(TP_ECAPACITY, "tp_map3_batch: round limit reached (integrals not converging)");
        const size_t na = Gr.active.size();
        size_t na_exact = 0;
        if (mode == TP_MODE_SURROGATE) {  // integrals evaluated exactly (wide, or still small) come first
            auto mid = std::stable_partition(Gr.active.begin(), Gr.active.end(), [&](size_t t) {
                return wide[t] || (long long)I[t].num_regions() < ctx->sur_exact_regions;
            });
            na_exact = (size_t)(mid - Gr.active.begin());
        }
        Gr.offs.assign(na + 1, 0);
        for (size_t a = 0; a < na; ++a) Gr.offs[a + 1] = Gr.offs[a] + I[Gr.active[a]].pending().size();
        const size_t nreg = Gr.offs[na];
        Gr.nreg = nreg;
        // the surrogate part of a round needs enough checked points; a smaller part runs exactly too
        Gr.nexact = mode != TP_MODE_SURROGATE ? nreg : Gr.offs[na_exact];
        if (mode == TP_MODE_SURROGATE && (double)(nreg - Gr.nexact) * TP_NPTS_GM3 * ctx->sur_check_frac < 64.0)
            Gr.nexact = nreg;
        total_regions += (long long)nreg;
        if (total_regions > ctx->max_regions)
            return tp_set_error(TP_ECAPACITY, "tp_map3_batch: region limit reached (integrals not converging)");
        int e;
        if ((e = L.h_regions.ensure(sizeof(TpRegion) * nreg)) || (e = L.d_regions.ensure(sizeof(TpRegion) * nreg)) ||
            (e = L.h_rout.ensure(sizeof(TpRegionOut) * nreg)) || (e = L.d_rout.ensure(sizeof(TpRegionOut) * nreg)) ||
            (e = L.d_fvals.ensure(sizeof(double) * TP_NPTS_GM3 * nreg)))
            return e;
        TpRegion *hr = (TpRegion *)L.h_regions.p;
        const std::vector<size_t> &act = Gr.active;
        const std::vector<size_t> &offs = Gr.offs;
        // a few integrals with many regions (the tail of a batch) are worth threads as well
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if (na > 1 && nreg > 4096)
        for (size_t a = 0; a < na; ++a) {
            tpc::AdaptiveIntegral31 &A = I[act[a]];
            const auto &pend = A.pending();
            for (size_t k = 0; k < pend.size(); ++k) {
                const auto &R = A.regions()[pend[k]];
                TpRegion &o = hr[offs[a] + k];
                for (int d = 0; d < 3; ++d) {
                    o.c[d] = R.c[d];
                    o.hw[d] = R.hw[d];
                }
                o.vol = R.vol;
                o.integral = (int)act[a];
                o.pad = 0;
            }
        }
        ctx->stats.host_ms += ms_since(th0);
        Gr.tl1 = ms_since(t0);
        if (ctx->graph_rounds && mode != TP_MODE_SURROGATE) {
            if ((e = tp_graph_round(ctx, c, L, nreg, mode, gst[g]))) return e;
            Gr.inflight = true;
            Gr.rounds += 1;
            ctx->stats.launches += 1;
            ctx->stats.regions += (long long)nreg;
            ctx->stats.points += (long long)nreg * TP_NPTS_GM3;
            return 0;
        }
        // a small exact round reads its regions from the pinned host buffer directly: no copy in (per-round
        // latency is what the tail rounds cost).  A round below the serial size, which runs side by side
        // with the other group's, writes its results straight to the pinned host buffer: a copy out would
        // be a DMA transfer waiting on the round's kernels, and the other group's copy in, when the runtime
        // puts it on the same DMA engine, would queue behind it and run the two groups' rounds strictly one
        // after the other (C1 FP32: 28.4 against 24.8 ms a call).  A larger round (tens of MB of results,
        // which a kernel writes across PCIe slower than the DMA engine copies them) runs after the other
        // group's anyway and keeps the copy out.
        const bool zc = mode != TP_MODE_SURROGATE && nreg <= (size_t)ctx->zero_copy_regions;
        const bool zout = zc || nreg < (size_t)ctx->serial_regions;
        TpRegionOut *const hout = zout ? (TpRegionOut *)L.h_rout.dp : (TpRegionOut *)L.d_rout.p;
        if (!zc)
            TP_HIP(hipMemcpyAsync(L.d_regions.p, hr, sizeof(TpRegion) * nreg, hipMemcpyHostToDevice, gst[g]));
        // a large round's kernels queue behind the other group's in-flight large round (its copy in still
        // overlaps that round); ev1 marks the end of that round's kernels
        for (int o = 0; o < nl; ++o)
            if (o != g && G[o].inflight && G[o].nreg >= (size_t)ctx->serial_regions &&
                nreg >= (size_t)ctx->serial_regions)
                TP_HIP(hipStreamWaitEvent(gst[g], ctx->lane[o].ev1, 0));
        TP_HIP(hipEventRecord(L.ev0, gst[g]));
        if (mode == TP_MODE_SURROGATE) {
            const TpRegion *dr = (const TpRegion *)L.d_regions.p;
            double *dfv = (double *)L.d_fvals.p;
            if (Gr.nexact > 0 && (e = tp_launch_regions(ctx, c, Gr.nexact, dr, (const TpIntegralDesc *)ctx->d_desc.p,
                                                         dfv, hout, exact_mode, (int *)L.d_flag.p, L.evm, gst[g])))
                return e;
            if (nreg > Gr.nexact &&
                (e = tp_surrogate_queue(ctx, c, g, nreg - Gr.nexact, dr + Gr.nexact,
                                        (const TpIntegralDesc *)ctx->d_desc.p, hout + Gr.nexact, (int *)L.d_flag.p,
                                        gst[g],
                                        (unsigned long long)(Gr.rounds * TP_NLANES + g))))
                return e;
        } else {
            const TpRegion *rin = zc ? (const TpRegion *)L.h_regions.dp : (const TpRegion *)L.d_regions.p;
            e = tp_launch_regions(ctx, c, nreg, rin, (const TpIntegralDesc *)ctx->d_desc.p, (double *)L.d_fvals.p,
                                  hout, mode, (int *)L.d_flag.p, L.evm, gst[g]);
            if (e) return e;
        }
        TP_HIP(hipEventRecord(L.ev1, gst[g]));
        if (!zout)
            TP_HIP(hipMemcpyAsync(L.h_rout.p, L.d_rout.p, sizeof(TpRegionOut) * nreg, hipMemcpyDeviceToHost,
                                  gst[g]));
        if (mode == TP_MODE_SURROGATE && (e = tp_copy_to_pinned(L.d_flag.p, L.h_flag.dp, sizeof(int), gst[g])))
            return e;
        TP_HIP(hipEventRecord(L.done, gst[g]));
        Gr.inflight = true;
        Gr.rounds += 1;
        ctx->stats.launches += 1;
        ctx->stats.regions += (long long)nreg;
        ctx->stats.points += (long long)nreg * TP_NPTS_GM3;
        if (nl == 1) ctx->stats.rounds += 1;
        return 0;
    };

    std::vector<std::pair<float, float>> iv_pts, iv_ker;  // device intervals of the rounds [ms]
    // each integral's final re-sums, taken when it finishes (on the host threads, while the device runs
    // the other integrals' rounds) instead of all after the last round
    std::vector<double> fin_v(ntriples), fin_e(ntriples), fin_cv(ntriples);
    std::vector<char> fin_done(ntriples, 0);
    auto union_ms = [](std::vector<std::pair<float, float>> &iv) {
        std::sort(iv.begin(), iv.end());
        double tot = 0, a = 0, b = -1;
        for (const auto &x : iv) {
            if (x.first > b) {
                if (b > a) tot += b - a;
                a = x.first;
                b = x.second;
            } else if (x.second > b) {
                b = x.second;
            }
        }
        if (b > a) tot += b - a;
        return tot;
    };

    // waits for the group's evaluation, absorbs it and selects the next regions
    auto complete = [&](int g) -> int {
        Group &Gr = G[g];
        TpLane &L = ctx->lane[g];
        auto tw = clk::now();
        TP_HIP(hipEventSynchronize(L.done));
        const double waited = ms_since(tw);
        ctx->stats.wait_ms += waited;
        const double tc0 = ms_since(t0);
        Gr.inflight = false;
        // the round's device intervals against the call's start; the groups' rounds can overlap on
        // the device, so the times reported are the unions of these intervals
        float ta = 0, tm = 0, tb = 0;
        TP_HIP(hipEventElapsedTime(&ta, ctx->ev_batch, L.ev0));
        TP_HIP(hipEventElapsedTime(&tb, ctx->ev_batch, L.ev1));
        iv_ker.push_back({ta, tb});
        if (mode != TP_MODE_SURROGATE || Gr.nexact > 0) {  // the exact points kernel ran (evm follows it)
            TP_HIP(hipEventElapsedTime(&tm, ctx->ev_batch, L.evm));
            iv_pts.push_back({ta, tm});
            ctx->stats.integrand_sum_ms += tm - ta;
        }
        const TpRegionOut *ro = (const TpRegionOut *)L.h_rout.p;
        // the point kernels' flag word: in region 0's flag (exact rounds) or copied back (surrogate rounds)
        const int pflag = mode == TP_MODE_SURROGATE ? *(const int *)L.h_flag.p : (Gr.nreg ? ro[0].flag >> 1 : 0);
        if (pflag) return tp_set_error(TP_ENONFINITE, "tp_map3_batch: non-finite ell or B_kappa value");
        if (mode == TP_MODE_SURROGATE && Gr.nreg > Gr.nexact) {
            bool fb = false;
            const size_t ns = Gr.nreg - Gr.nexact;
            int e = tp_surrogate_result(ctx, g, ns, &fb);
            if (e) return e;
            if (fb) {  // the surrogate missed the tolerance on this round's checks: exact re-evaluation
                ctx->stats.fallback_batches += 1;
                const bool zout = Gr.nreg < (size_t)ctx->serial_regions;  // as the round's launch chose
                TpRegionOut *dout = (zout ? (TpRegionOut *)L.h_rout.dp : (TpRegionOut *)L.d_rout.p) + Gr.nexact;
                if ((e = tp_launch_regions(ctx, c, ns, (const TpRegion *)L.d_regions.p + Gr.nexact,
                                           (const TpIntegralDesc *)ctx->d_desc.p,
                                           (double *)L.d_fvals.p + Gr.nexact * TP_NPTS_GM3, dout, exact_mode,
                                           (int *)L.d_flag.p, L.evm, gst[g])))
                    return e;
                if (!zout)
                    TP_HIP(hipMemcpyAsync((TpRegionOut *)L.h_rout.p + Gr.nexact, dout, sizeof(TpRegionOut) * ns,
                                          hipMemcpyDeviceToHost, gst[g]));
                TP_HIP(hipStreamSynchronize(gst[g]));
                if (ro[Gr.nexact].flag >> 1)
                    return tp_set_error(TP_ENONFINITE, "tp_map3_batch: non-finite ell or B_kappa value");
            }
        }
        auto th0 = clk::now();
        const std::vector<size_t> &act = Gr.active;
        const std::vector<size_t> &offs = Gr.offs;
        const size_t na = act.size();
        std::vector<char> finished(na, 0);
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if (na > 1 && Gr.nreg > 1024) reduction(| : bad)
        for (size_t a = 0; a < na; ++a) {
            tpc::AdaptiveIntegral31 &A = I[act[a]];
            const auto &pend = A.pending();
            for (size_t k = 0; k < pend.size(); ++k) {
                const TpRegionOut &o = ro[offs[a] + k];
                auto &R = A.regions()[pend[k]];
                R.val[0] = o.val;
                R.err[0] = o.err;
                R.corr = o.corr;
                R.cvar = o.cvar;
                R.split = (unsigned)o.split;
                R.errmax = o.err;
                bad |= o.flag & 1;
            }
            A.absorb();
            finished[a] = A.advance() ? 1 : 0;
        }
        if (bad) return tp_set_error(TP_ENONFINITE, "tp_map3_batch: non-finite region estimate");
        std::vector<size_t> next, done_now;
        next.reserve(na);
        for (size_t a = 0; a < na; ++a)
            (finished[a] ? done_now : next).push_back(act[a]);
        // the final re-sums of the integrals that finished in this round (each walks its whole heap)
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if (done_now.size() > 1)
        for (size_t k = 0; k < done_now.size(); ++k) {
            const size_t t = done_now[k];
            I[t].result(&fin_v[t], &fin_e[t]);
            if (mode == TP_MODE_SURROGATE) fin_cv[t] = I[t].correction_var();
            fin_done[t] = 1;
        }
        Gr.active.swap(next);
        ctx->stats.host_ms += ms_since(th0);
        if (ctx->round_log_on)
            ctx->round_log.insert(ctx->round_log.end(),
                                  {(double)g, (double)Gr.nreg, Gr.tl0, Gr.tl1, tc0, ms_since(t0), waited, (double)ta, (double)tb});
        return 0;
    };

    // on an error, let queued work drain before the buffers can be reused
    auto fail = [&](int e) {
        (void)hipStreamSynchronize(ctx->stream);
        for (int g = 0; g < nl; ++g) (void)hipStreamSynchronize(gst[g]);
        return e;
    };
    // With two groups the grouping adapts to the round size: while a round is small (the first rounds
    // and the tail, tens of us of kernel) the two groups' host turnarounds would only queue behind each
    // other, so the integrals run as one group; a round above kSplitRegions splits them into two
    // groups, which merge again below kMergeRegions.  Each integral's decision sequence is unchanged.
    const bool adaptive = nl == 2;
    auto pending_of = [&](int g) {
        size_t n = 0;
        for (size_t a : G[g].active) n += I[a].pending().size();
        return n;
    };
    int ng = nl;
    if (adaptive && pending_of(0) + pending_of(1) < kSplitRegions) {
        G[0].active.insert(G[0].active.end(), G[1].active.begin(), G[1].active.end());
        G[1].active.clear();
        ng = 1;
    }
    for (int g = 0; g < nl; ++g)
        if (!G[g].active.empty() && (rc = launch(g))) return fail(rc);
    for (;;) {
        bool any = false;
        for (int g = 0; g < nl; ++g) {
            if (!G[g].inflight) continue;
            any = true;
            if ((rc = complete(g))) return fail(rc);
            if (adaptive) {
                const int o = 1 - g;
                const size_t next = pending_of(g);
                if (ng == 2 && next + (G[o].inflight ? G[o].nreg : pending_of(o)) < kMergeRegions) {
                    if (G[o].inflight && (rc = complete(o))) return fail(rc);
                    G[g].active.insert(G[g].active.end(), G[o].active.begin(), G[o].active.end());
                    G[o].active.clear();
                    G[g].rounds = std::max(G[g].rounds, G[o].rounds);
                    ng = 1;
                } else if (ng == 1 && next > kSplitRegions && G[g].active.size() > 1) {
                    const size_t h = G[g].active.size() / 2;
                    G[o].active.assign(G[g].active.begin() + h, G[g].active.end());
                    G[g].active.resize(h);
                    G[o].rounds = G[g].rounds;
                    ng = 2;
                    if ((rc = launch(o))) return fail(rc);
                }
            }
            if (!G[g].active.empty() && (rc = launch(g))) return fail(rc);
        }
        if (!any) break;
    }
    if (nl > 1) {  // evaluation rounds of the integral that took the most
        long long r = 0;
        for (size_t t = 0; t < ntriples; ++t) r = std::max(r, I[t].rounds() + 1);
        ctx->stats.rounds = r;
    }
    ctx->stats.kernel_ms = union_ms(iv_ker);
    ctx->stats.integrand_ms = union_ms(iv_pts);
    // results (the final re-sums of any integral that did not finish in a round: none, normally)
    std::vector<char> redo_flag(ntriples, 0);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) if (ntriples > 16)
    for (size_t t = 0; t < ntriples; ++t) {
        if (!fin_done[t]) {
            I[t].result(&fin_v[t], &fin_e[t]);
            if (mode == TP_MODE_SURROGATE) fin_cv[t] = I[t].correction_var();
        }
        const double r = fin_v[t], e = fin_e[t];
        val[t] = r / 8 / M_PI / M_PI / M_PI;
        if (err) err[t] = e / 8 / M_PI / M_PI / M_PI;
        if (neval) neval[t] = (long long)I[t].num_eval();
        // surrogate mode: a triple whose correction has a standard error above tol is redone exactly (FP32 form)
        if (mode == TP_MODE_SURROGATE &&
            (!I[t].converged_flag() || !(std::sqrt(fin_cv[t]) <= ctx->sur_tol * std::fabs(r))))
            redo_flag[t] = 1;
    }
    std::vector<size_t> redo;
    for (size_t t = 0; t < ntriples; ++t)
        if (redo_flag[t]) redo.push_back(t);
    {
        size_t held = 0;
        for (const auto &A : I) held += A.bytes_reserved();
        if (held <= ctx->keep_pool_bytes) ctx->map3_keep.swap(I);
    }
    if (!redo.empty()) {
        tp_batch_stats keep = ctx->stats;
        std::vector<int> ic(redo.size());
        std::vector<double> th(3 * redo.size()), v(redo.size()), e(redo.size());
        std::vector<long long> ne(redo.size());
        for (size_t k = 0; k < redo.size(); ++k) {
            ic[k] = icosmo[redo[k]];
            for (int d = 0; d < 3; ++d) th[3 * k + d] = theta_rad[3 * redo[k] + d];
        }
        rc = tp_map3_batch(ctx, c, redo.size(), ic.data(), th.data(), reltol, exact_mode, v.data(), e.data(),
                           ne.data());
        if (rc) return rc;
        for (size_t k = 0; k < redo.size(); ++k) {
            val[redo[k]] = v[k];
            if (err) err[redo[k]] = e[k];
            if (neval) neval[redo[k]] += ne[k];
        }
        keep.rounds += ctx->stats.rounds;
        keep.regions += ctx->stats.regions;
        keep.points += ctx->stats.points;
        keep.launches += ctx->stats.launches;
        keep.kernel_ms += ctx->stats.kernel_ms;
        keep.integrand_ms += ctx->stats.integrand_ms;
        keep.integrand_sum_ms += ctx->stats.integrand_sum_ms;
        keep.host_ms += ctx->stats.host_ms;
        keep.wait_ms += ctx->stats.wait_ms;
        keep.fallback_integrals += (long long)redo.size();
        ctx->stats = keep;
    }
    ctx->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

extern "C" int tp_surrogate_exact_regions(tp_ctx *ctx, long long nregions)
{
    if (!ctx || nregions < 0) return tp_set_error(TP_EINVAL, "tp_surrogate_exact_regions: nregions must be >= 0");
    ctx->sur_exact_regions = nregions;
    return 0;
}

extern "C" int tp_surrogate_exact_ratio(tp_ctx *ctx, double ratio)
{
    if (!ctx || !(ratio >= 0)) return tp_set_error(TP_EINVAL, "tp_surrogate_exact_ratio: ratio must be >= 0");
    ctx->sur_exact_ratio = ratio;
    return 0;
}

extern "C" int tp_set_batch_limits(tp_ctx *ctx, long long max_rounds, long long max_regions, long long round_points)
{
    if (!ctx || max_rounds < 0 || max_regions < 0 || round_points < 0)
        return tp_set_error(TP_EINVAL, "tp_set_batch_limits: limits must be >= 0 (0 = default)");
    ctx->max_rounds = max_rounds ? max_rounds : kTpMaxRounds;
    ctx->max_regions = max_regions ? max_regions : kTpMaxRegions;
    ctx->round_points = round_points ? round_points : kTpRoundPoints;
    return 0;
}

extern "C" int tp_set_graph_rounds(tp_ctx *ctx, int on)
{
    if (!ctx) return tp_set_error(TP_EINVAL, "tp_set_graph_rounds: NULL context");
    ctx->graph_rounds = on != 0;
    return 0;
}

extern "C" int tp_set_batch_groups(tp_ctx *ctx, int groups)
{
    if (!ctx || groups < 1 || groups > TP_NLANES)
        return tp_set_error(TP_EINVAL, "tp_set_batch_groups: groups must be 1.." + std::to_string(TP_NLANES));
    ctx->nlanes = groups;
    return 0;
}

int tp_host_threads(const tp_ctx *ctx, int dflt)
{
    const int want = ctx && ctx->host_threads > 0 ? ctx->host_threads : dflt;
    return std::max(1, std::min(omp_get_max_threads(), want));
}

extern "C" int tp_set_host_threads(tp_ctx *ctx, int n)
{
    if (!ctx || n < 0 || n > 256) return tp_set_error(TP_EINVAL, "tp_set_host_threads: n must be 0 (default) .. 256");
    ctx->host_threads = n;
    return 0;
}

extern "C" int tp_set_keep_pool_bytes(tp_ctx *ctx, unsigned long long bytes)
{
    if (!ctx) return tp_set_error(TP_EINVAL, "tp_set_keep_pool_bytes: NULL context");
    ctx->keep_pool_bytes = (size_t)bytes;
    if (bytes == 0) {  // release what is kept now
        std::vector<tpc::AdaptiveIntegral31>().swap(ctx->map3_keep);
        std::vector<tpc::AdaptiveIntegralT<2, 1>>().swap(ctx->map2_keep);
    }
    return 0;
}

extern "C" int tp_set_round_log(tp_ctx *ctx, int on)
{
    if (!ctx) return tp_set_error(TP_EINVAL, "tp_set_round_log: NULL context");
    ctx->round_log_on = on != 0;
    if (!on) ctx->round_log.clear();
    return 0;
}

extern "C" int tp_round_log(const tp_ctx *ctx, double *out, size_t cap, size_t *n)
{
    if (!ctx || !n || (cap && !out)) return tp_set_error(TP_EINVAL, "tp_round_log: NULL argument");
    *n = ctx->round_log.size();
    if (cap && !ctx->round_log.empty())
        std::memcpy(out, ctx->round_log.data(), sizeof(double) * std::min(cap, ctx->round_log.size()));
    return 0;
}

extern "C" int tp_set_serial_regions(tp_ctx *ctx, long long min_regions)
{
    if (!ctx || min_regions < 0) return tp_set_error(TP_EINVAL, "tp_set_serial_regions: bad arguments");
    ctx->serial_regions = min_regions ? min_regions : kTpSerialRegions;
    return 0;
}

extern "C" int tp_set_zero_copy_regions(tp_ctx *ctx, long long max_regions)
{
    if (!ctx || max_regions < 0) return tp_set_error(TP_EINVAL, "tp_set_zero_copy_regions: bad arguments");
    ctx->zero_copy_regions = max_regions;
    return 0;
}
