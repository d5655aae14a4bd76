/*
 * threepoint_amd.h -- C ABI of the MI355X-native <Map^3> / Gamma_i modelling path.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success and a
 * nonzero TP_E* code on failure (the reason is in tp_last_error()); no entry
 * point calls exit().  Host pointers are read/written synchronously: the call
 * returns only after the device work it started has finished, and no pointer
 * is retained after return (the contract of the reference's integrand_v
 * callback, classed_version/cubature.h:55-57 and hcubature.cpp:247-265).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef THREEPOINT_AMD_H
#define THREEPOINT_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum {
    TP_OK = 0,
    TP_EINVAL = 1,    /* bad argument (sizes, dims, NULL pointers)            */
    TP_EHIP = 2,      /* HIP runtime failure                                  */
    TP_ENONFINITE = 3,/* non-finite ell or B_kappa value (the reference exits) */
    TP_ECAPACITY = 4, /* a fixed-capacity buffer would overflow               */
    TP_ECALLBACK = 5, /* the integrand callback returned nonzero              */
    TP_ENODEV = 6     /* no usable HIP device                                 */
};
const char *tp_last_error(void);

/* ------------------------------------------------------------ integrand modes */
enum {
    TP_MODE_FP64 = 0,      /* plain FP64 integrand                                  */
    TP_MODE_FP32 = 1,      /* FP32 bispectrum, FP64 z-sum / cubature accumulation    */
    TP_MODE_SURROGATE = 2, /* MLP surrogate for log|B_kappa| + exact-check fallback  */
    TP_MODE_ANALYTIC = 3   /* the reference's test_analytical bispectrum
                              (classed_version/bispectrum.hpp:18, bispectrum.cpp:242-255) */
};

/* ---------------------------------------------------------------- context */
typedef struct tp_ctx tp_ctx;
/* One context per device and host thread; owns a HIP stream and staging buffers. */
int tp_ctx_create(int device, tp_ctx **out);
int tp_ctx_destroy(tp_ctx *ctx);
int tp_device_count(int *n);

/* ------------------------------------------------------------- cosmologies */
/* Replaces cosmology (classed_version/cosmology.hpp:11-29): same fields. */
typedef struct {
    double h, sigma8, omb, ns, w, om, zmax; /* ow = 1-om and omc = om-omb are derived */
} tp_cosmology;

typedef struct tp_cosmo tp_cosmo; /* device-resident tables of ncosmo cosmologies */

/* Replaces BispectrumCalculator(cosmology*, n_of_z, n_z, z_max) + set_cosmology
 * (classed_version/bispectrum.cpp:73-79, 111-164).  Builds, in FP64 on the GPU,
 * f_K, g, D1, r_sigma, n_eff, ncur on z_i = i*z_max/n_z for each cosmology, and
 * the 96-node Limber tables used by the integrand kernels.
 * nz: n_z values per cosmology; nz_stride = 0 shares one n(z) among all. */
int tp_cosmo_build(tp_ctx *ctx, int ncosmo, const tp_cosmology *params, const double *nz,
                   int n_z, size_t nz_stride, double z_max, tp_cosmo **out);
/* Replaces the BispectrumCalculator constructor with a given linear P(k)
 * (classed_version/bispectrum.cpp:81-89, linear_pk's Pk_given branch :785-806), the input
 * the cosmosis module passes (classed_version/interface.cpp:176-194): as tp_cosmo_build, with
 * the linear P(k) of cosmology i the nk pairs (k_h[i*pk_stride + j], p_k[...]) instead of the
 * Eisenstein & Hu fit.  k [h/Mpc] strictly increasing, P [(Mpc/h)^3]; between entries P is
 * linear in k, outside the table it is the end value, and sigma8 does not rescale it (as in
 * the reference).  pk_stride = 0 shares one table.  Surrogate mode rejects such tables. */
int tp_cosmo_build_pk(tp_ctx *ctx, int ncosmo, const tp_cosmology *params, const double *nz, int n_z,
                      size_t nz_stride, double z_max, const double *k_h, const double *p_k, int nk,
                      size_t pk_stride, tp_cosmo **out);
int tp_cosmo_destroy(tp_cosmo *c);
int tp_cosmo_count(const tp_cosmo *c, int *ncosmo, int *n_z);
/* Copies one cosmology's tables to the host (each array n_z doubles). */
int tp_cosmo_tables(const tp_cosmo *c, int icosmo, double *f_K, double *g, double *D1,
                    double *r_sigma, double *n_eff, double *ncur, double *norm);

/* -------------------------------------------------------------- point batches */
/* Replaces BispectrumCalculator::bkappa (classed_version/bispectrum.cpp:238-286):
 * out[i] = B_kappa(ell[3i], ell[3i+1], ell[3i+2]). */
int tp_bkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ell123,
                   double *out, int mode);

/* Replaces BispectrumCalculator::bispec (classed_version/bispectrum.cpp:166-236) with the
 * table interpolation of :506-513: in is npts*4 (k1, k2, k3 [h/Mpc], z); out is B [(Mpc/h)^6];
 * z > 10 gives the tree-level bispectrum (:174-175, 609-612). */
int tp_bispec_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *k123z, double *out);

/* Replaces BispectrumCalculator::bispec(k1, k2, k3, z, idx, didx) (classed_version/bispectrum.hpp:247,
 * bispectrum.cpp:166-236) with the caller's table interval: in is npts*6 (k1, k2, k3 [h/Mpc], z, idx,
 * didx); D1, r_sigma and n_eff are the tables lerped at (idx, didx), z enters Omega_m(z) of F2; z > 10
 * gives the tree-level bispectrum (:174-175, 609-612).  idx must be an integer in [0, n_z - 2]
 * (TP_EINVAL otherwise: the reference would read past its tables). */
int tp_bispec_eval_idx(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *k123z_idx,
                       double *out);

/* Replaces BispectrumCalculator::integrand_bkappa (classed_version/bispectrum.cpp:471-541):
 * in is npts*4 (z, ell1, ell2, ell3); out = g^3 (1+z)^3 B(ell/f_K, z) / (f_K E(z)), 0 at z = 0
 * or ell3 = 0. */
int tp_integrand_bkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *zell,
                             double *out);

/* Replaces the integrand_v adapter ApertureStatistics::integrand
 * (classed_version/apertureStatistics.cpp:254-279): x is npts*3 (l1, l2, phi),
 * fval[i] = l1 l2 B_kappa(l1,l2,l3) uhat(l1 th1) uhat(l2 th2) uhat(l3 th3). */
int tp_map3_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], size_t npts,
                 const double *x, double *fval, int mode);

/* tp_map3_eval for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): no counterpart in the reference.  Same arguments,
 * kernels and results as tp_map3_eval, with at most TP_NOALLOC_MAX_POINTS points per call
 * (TP_ECAPACITY beyond); the points and values are staged through one static device array per GPU
 * (2 MiB, held until the process ends), so the call allocates nothing.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (from several threads or contexts) race. */
#define TP_NOALLOC_MAX_POINTS 65536
int tp_map3_eval_noalloc(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], size_t npts,
                         const double *x, double *fval, int mode);

/* Replaces ApertureStatistics::MapMapMap (classed_version/apertureStatistics.cpp:321-353)
 * exactly as the reference drives it: tp_hcubature_v over [1, 10/theta_min]^2 x [0, 2pi]
 * (ERROR_L1, relative tolerance reltol) whose integrand callback sends each round's
 * points through tp_map3_eval (one host<->device round trip per round, as
 * cuda_version/apertureStatistics.cu:155-229 does).  Result / (2 pi)^3. */
int tp_map3_single(tp_ctx *ctx, const tp_cosmo *c, int icosmo, const double theta[3], double reltol,
                   int mode, double *val, double *err, long long *neval);

/* Replaces ApertureStatistics::MapMapMap (classed_version/apertureStatistics.cpp:321-353)
 * for many triples at once: each integral is the 3-D h-adaptive cubature over
 * [1, 10/theta_min]^2 x [0, 2pi] with relative tolerance reltol (ERROR_L1), result
 * divided by (2 pi)^3.  Integrals advance together; each keeps the decision sequence
 * of the one-at-a-time algorithm.  icosmo[t] selects the cosmology of triple t.
 * neval (optional) receives the integrand points used per triple. */
int tp_map3_batch(tp_ctx *ctx, const tp_cosmo *c, size_t ntriples, const int *icosmo,
                  const double *theta_rad, double reltol, int mode, double *val, double *err,
                  long long *neval);

/* Launch shape of the batched region kernel: a cubature point's 96 z-nodes are summed by 1, 2, 4
 * or 8 waves, or (16, 32, 64) by the region-cached kernel with 1, 2 or 4 groups of 256 threads per
 * workgroup, which evaluates the one-wavenumber terms of BiHalofit once per distinct multipole of
 * a region (47 per 33 points) in LDS.  By default the shape is tuned per launch-size bucket on the
 * first batched call of a context and precision (all shapes give bit-identical results).
 * lanes = 0 restores tuning; 1, 2, 4, 8, 16, 32 or 64 forces that shape.  No counterpart in the
 * reference (its GPU build has one fixed launch shape). */
int tp_set_region_lanes(tp_ctx *ctx, int lanes);
/* The tuned shapes of mode (TP_MODE_FP64 or TP_MODE_FP32) for launches of < 32k, < 160k,
 * < 640k and more points (0 = not tuned yet). */
int tp_region_lanes_table(tp_ctx *ctx, int mode, int *lanes4);
/* on != 0: tp_map3_batch replays each exact-mode cubature round as one HIP graph per group (copy in,
 * points kernel, rule kernel, copy out, with the timing events), instantiated once and updated each
 * round, instead of queuing those operations one by one.  Results are identical either way.  No
 * counterpart in the reference. */
int tp_set_graph_rounds(tp_ctx *ctx, int on);
/* Number of integral groups tp_map3_batch keeps in flight (1-4, default 2): while the device evaluates
 * one group's regions the host runs the others' heap decisions.  Results of the FP64 and FP32 modes do
 * not depend on it; in surrogate mode the exactly checked points are drawn per group and round, so
 * another group count draws another sample (results within the surrogate's tolerance, not bit-equal).
 * No counterpart in the reference. */
int tp_set_batch_groups(tp_ctx *ctx, int groups);
/* Per-round trace of tp_map3_batch (diagnostics, off by default).  While on, every cubature round
 * of a call appends 9 doubles: group, regions, host launch begin and end, completion begin (after
 * the wait) and end [ms since the call began], the time waited for the device [ms], and the
 * round's device begin and end [ms since an event recorded when the call began].  The trace
 * is cleared at the start of each call; tp_round_log copies up to cap doubles and sets *n to the
 * number available.  No counterpart in the reference. */
int tp_set_round_log(tp_ctx *ctx, int on);

/* Host memory kept between batched calls: the integrators of the last tp_map3_batch (and of the
 * last <Map^2> and T1 batch) stay in the context when their region pools hold at most `bytes`
 * (default 16 GiB), so the next call of a similar size allocates and page-faults nothing.
 * bytes = 0 releases what is kept and keeps nothing.  No counterpart in the reference. */
int tp_set_keep_pool_bytes(tp_ctx *ctx, unsigned long long bytes);

/* OpenMP threads of the host side of the batched cubatures (region records and heap decisions
 * between device rounds): n = 0 restores the defaults (8 for tp_map3_batch, whose host work
 * overlaps the device; 16 for the synchronous batches); the count is capped by
 * omp_get_max_threads().  The results do not depend on it.  No counterpart in the reference. */
int tp_set_host_threads(tp_ctx *ctx, int n);
int tp_round_log(const tp_ctx *ctx, double *out, size_t cap, size_t *n);

/* Exact-mode rounds of at most max_regions regions read their regions from pinned host memory directly
 * instead of copying them in (0 = always copy in; default 2048).  Their results, and those of every
 * round below the serial size (tp_set_serial_regions), are written to pinned host memory directly;
 * larger rounds copy them out.  Results do not depend on it.  No counterpart in the reference. */
int tp_set_zero_copy_regions(tp_ctx *ctx, long long max_regions);

/* tp_map3_batch runs the rounds of its two groups of integrals one after the other on the device when
 * both hold at least min_regions regions (each then fills the device; the host absorbs one round while
 * the other's kernels run), side by side otherwise; rounds below min_regions also write their results
 * to pinned host memory directly, larger ones copy them out of device memory.  0 restores the default
 * (65536).  Results do not depend on it.  No counterpart in the reference. */
int tp_set_serial_regions(tp_ctx *ctx, long long min_regions);

/* Guards of every batched cubature on this context (tp_map3_batch, tp_gamma_batch, tp_map2_batch,
 * tp_pkappa_batch and the covariance batches): a call whose integrals need more than max_rounds
 * rounds, or evaluate more than max_regions regions in total, ends with TP_ECAPACITY instead of
 * growing without bound; the batches that apply the rule on the host evaluate a round in slices of at
 * most round_points points.  0 restores a default (20000 rounds, 4e8 regions, 2^23 points).  Results
 * do not depend on round_points.  No counterpart in the reference (its integrals have no bound). */
int tp_set_batch_limits(tp_ctx *ctx, long long max_rounds, long long max_regions, long long round_points);

/* Statistics of the last tp_map3_batch call on this context. */
typedef struct {
    long long rounds, regions, points, launches;
    double kernel_ms;            /* device time of the integrand/rule kernels (tp_map3_batch: the union
                                    of the rounds' intervals, since the groups' rounds can overlap) */
    double integrand_ms;         /* the same for the integrand (point) kernel only */
    double wall_ms;              /* whole call                                     */
    long long surrogate_points;  /* points evaluated by the surrogate              */
    long long exact_checks;      /* points re-evaluated exactly                    */
    long long fallback_batches;  /* rounds that fell back to exact B_kappa         */
    double max_check_relerr;     /* worst relative error seen by the exact checks  */
    long long fallback_integrals;/* triples re-integrated exactly because the standard error
                                    of their surrogate correction exceeded tol          */
    double host_ms;              /* host time in region gathering and heap decisions   */
    double wait_ms;              /* host time blocked waiting for the device           */
    long long graph_rebuilds;    /* round graphs rebuilt because an update was refused */
    double integrand_sum_ms;     /* sum over launches of the integrand kernel's event-timed duration
                                    (launches x a per-launch profiler's average; >= integrand_ms) */
    double mlp_ms;               /* surrogate mode: summed event-timed duration of the network's launches */
    long long mlp_launches;      /* surrogate mode: launches of the network inference kernel           */
    long long mlp_points;        /* surrogate mode: points those launches evaluated (padded to 64)     */
} tp_batch_stats;
int tp_last_batch_stats(const tp_ctx *ctx, tp_batch_stats *out);

/* A plain device-to-device copy of nbytes (a multiple of 16) by a vector copy kernel: it reads
 * nbytes and writes nbytes of HBM, so hardware byte counters can be calibrated against it.
 * gbps receives the copy's (read + write) rate.  No counterpart in the reference: a
 * measurement aid for the kernels' HBM traffic. */
int tp_hbm_copy_probe(tp_ctx *ctx, size_t nbytes, double *gbps);

/* Algorithmic operation census of one <Map^3> integrand point of the FP64/FP32 kernels:
 * out[0] flops per z-node, out[1] exp/log calls per z-node, out[2] flops and out[3]
 * elementary calls per point outside the node loop, out[4] z-nodes per point.  Counted by
 * instantiating the kernel's own device functions with a counting scalar on the host. */
int tp_census_map3_point(double *out);

/* Census of one P_E term of BiHalofit (eq. B6: the linear P(k) and the one-halo-free part of one
 * wavenumber): out[0] flops, out[1] elementary calls.  A z-node evaluation makes three; the
 * region-cached kernel makes 47 per 33 points.  No counterpart in the reference. */
int tp_census_map3_pse(double *out);
/* Census of the Gamma_i integrand (tp_gamma_eval, tp_gamma_batch): out[0] flops and out[1]
 * elementary calls per Ogata term (a cubature point has 3 (prec_k - 1) terms), out[2] flops
 * and out[3] elementary calls of the per-point triangle shape. */
int tp_census_gamma_term(double *out);
/* Census of the <Map^2> integrand (tp_map2_batch): out[0] flops and out[1] elementary calls per point
 * (Halofit, no T17 factor, no shape noise); out[2], out[3] per point of the P_kappa Limber integrand;
 * out[4], out[5] per z-node of the fixed-node sums of tp_pell and the T1 covariance integrand. */
int tp_census_power_point(double *out);

/* ---------------------------------------------------------------- Gamma_i */
/* Replaces GammaCalculator(bk, prec_h, prec_k, "centroid") + initialize_bessel
 * (classed_version/gamma.cpp:6-23, 133-176): Ogata node tables for J6 and J2. */
int tp_gamma_setup(tp_ctx *ctx, double prec_h, double prec_k_arg);
/* Replaces integrand_gamma{0,1}_cubature_and_ogata (classed_version/gamma.cpp:338-359,
 * 522-543): which = 0 (Gamma0 integrand) or 1 (Gamma1 integrand); x is npts*3
 * (z, phi, psi); fval is npts*2 (re, im). */
int tp_gamma_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, int which, double x1, double x2,
                  double x3, size_t npts, const double *x, double *fval);
/* Replaces GammaCalculator::gamma0..gamma3 (classed_version/gamma.cpp:27-81) for many
 * triangles: out_re/out_im[4*t + i] = Gamma_i(x1,x2,x3) of triangle t (rad), centroid
 * convention, hcubature over (z, phi, psi) with relative tolerance reltol. */
int tp_gamma_batch(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                   double reltol, double *out_re, double *out_im, long long *neval);
/* tp_gamma_batch with the integrand's precision: TP_MODE_FP64, TP_MODE_ANALYTIC (the reference's analytic
 * test bispectrum, integrand_bkappa's test_analytical branch, classed_version/bispectrum.cpp:477-499: zero
 * above z = 1), or TP_MODE_FP32 (Ogata terms whose
 * wavenumbers lie in the FP32 window in FP32, the others and any that leave the float range in FP64;
 * FP64 accumulation, rule and heap; the FP32 form needs prec_k <= 10753, TP_EINVAL otherwise).  No
 * counterpart in the reference (its integrand is FP64). */
int tp_gamma_batch_mode(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                        double reltol, int mode, double *out_re, double *out_im, long long *neval);
/* Triangle centers of GammaCalculator (classed_version/gamma.cpp:12-16) */
enum { TP_CENTER_CENTROID = 0, TP_CENTER_ORTHOCENTER = 1 };
/* tp_gamma_batch_mode with the reference's triangle_center choice (GammaCalculator's constructor argument,
 * classed_version/gamma.cpp:6-23): TP_CENTER_CENTROID rotates the natural components from the orthocenter
 * to the centroid (:44-45, 67-68, 624-643); TP_CENTER_ORTHOCENTER returns them unrotated. */
int tp_gamma_batch_center(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t ntri, const double *x123,
                          double reltol, int mode, int center, double *out_re, double *out_im, long long *neval);
/* The FP32 form's window: Ogata terms whose wavenumbers all lie in [qmin, qmax] run in FP32
 * (default [1e-5, 1e6]); no counterpart in the reference. */
int tp_gamma_fp32_window(tp_ctx *ctx, double qmin, double qmax);

/* ------------------------------------- convergence power spectrum and <Map^2> */
/* flags of the power-spectrum entry points */
enum {
    TP_PK_LINEAR = 1, /* D1^2 P_lin instead of revised Halofit (limber_integrand_power_spectrum's
                         nonlinear = false, classed_version/bispectrum.cpp:569-572)               */
    TP_PK_T17 = 2,    /* the cuda_version's T17 correction of the Limber integrand and of <Map^2>
                         (cuda_version/bispectrum.cu:971-979, apertureStatistics.cu:39-46)        */
    TP_PK_SHAPENOISE_ONLY = 4 /* Pell is the shape noise alone: the cuda_version's constant spectrum
                         (constant_powerspectrum, bispectrum.cu:515-516); tp_pell and T1 only     */
};
/* Replaces BispectrumCalculator::P_k_nonlinear(k, z) (classed_version/bispectrum.cpp:288-367) and,
 * with TP_PK_LINEAR, linear_pk_at_z (:369-380): kz is npts*2 (k [h/Mpc], z), out npts values
 * [(Mpc/h)^3].  z outside [0, z_max] reads the nearest table interval (the reference reads past
 * its tables there). */
int tp_pk_at_z_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *kz, int flags,
                    double *out);
/* Replaces BispectrumCalculator::limber_integrand_power_spectrum(ell, z, nonlinear)
 * (classed_version/bispectrum.cpp:543-577): ellz is npts*2 (ell, z). */
int tp_limber_pkappa_eval(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ellz,
                          int flags, double *out);
/* Replaces BispectrumCalculator::Pell (classed_version/bispectrum.cpp:1048-1059, GQ96_of_Pk :1034-1046):
 * out[i] = 96-point Gauss-Legendre Limber integral of P_NL over [0, z_max] at ell[i] plus the shape-noise
 * power 0.5 sigma^2 / n_gal (n_gal in rad^-2; the reference's members sigma and n); ell = 0 gives the
 * shape noise alone. */
int tp_pell(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t n, const double *ell, int flags, double sigma,
            double n_gal, double *out);
/* Replaces BispectrumCalculator::convergence_power_spectrum (classed_version/bispectrum.cpp:382-388)
 * for n multipoles at once: hcubature over z in [0, z_max] (Gauss-Kronrod, the cubature's 1-D rule),
 * relative tolerance reltol (the reference's 1e-6), cosmology icosmo[i] (NULL: all 0). */
int tp_pkappa_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *ell,
                    double reltol, int flags, double *out, double *err, long long *neval);
/* Replaces Map2(theta, ...) of the cuda_version (apertureStatistics.cu:110-136) for one source bin
 * and n aperture radii at once: 2-D hcubature over ell in [1, 10/theta], z in [0, z_max] of
 * ell uHat(ell theta)^2 P-Limber-integrand, plus, when sigma != 0, the shape-noise term of
 * apertureStatistics.cu:43-46 (ngal_arcmin2 in arcmin^-2); result / 2 pi.  reltol is the
 * reference's 1e-5.  flags: TP_PK_T17 applies the T17 factors (the cuda_version's default). */
int tp_map2_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta,
                  double reltol, int flags, double sigma, double ngal_arcmin2, double *out, double *err,
                  long long *neval);

/* ------------------------------------------ Gaussian <Map^3> covariance */
/* Replaces ApertureStatistics::MapMapMap_covariance_Gauss (classed_version/apertureStatistics.cpp:
 * 355-396) for n pairs of radius triples at once: theta6[6i..6i+5] = theta_123, theta_456 [rad];
 * 4-D hcubature over ell1, ell2 in [1, 10 / max(min theta_123, min theta_456)], phi in [0, pi],
 * z in [0, z_max] of uHat products x the single-z triple power spectrum (bispectrum.cpp:580-607),
 * relative tolerance reltol (the reference's 1e-3); result 2 / area / (2 pi)^3. */
int tp_map3_cov_gauss_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta6,
                            double survey_area, double reltol, double *out, double *err, long long *neval);
/* Replaces T1_total (cuda_version/apertureStatisticsCovariance.cu:272-326) for an infinite survey
 * (type 2, T1 :1065-1106, integrand :1958-1985) for n pairs at once: each is a sum of T1 integrals
 * over the distinct permutations of theta_456, each a 3-D hcubature over ell1, ell2 in
 * [lmin, 10 / max(min theta_123, min theta_456)], phi in [0, pi] of Pell(l1) Pell(l2) Pell(l3)
 * times both uHat products, relative tolerance reltol (the reference's 1e-4), times 2 / area /
 * (2 pi)^3 (area in rad^2).  Pell carries the shape noise 0.5 sigma^2 / n_gal (n_gal in rad^-2);
 * flags: TP_PK_T17 (the cuda_version's default), TP_PK_SHAPENOISE_ONLY.  T2 vanishes for an
 * infinite survey (apertureStatisticsCovariance.cu:330-331).  The finite geometries:
 * tp_map3_cov_t1_batch_geometry and tp_map3_cov_t2_batch.
 * neval receives each pair's integrand points. */
int tp_map3_cov_t1_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta6,
                         double area, double lmin, double sigma, double n_gal, int flags, double reltol,
                         double *out, long long *neval);
/* tp_map3_cov_t1_batch with the Pell's precision: TP_MODE_FP64, or TP_MODE_FP32 (FP32 node values of the
 * 96-node Limber sum with native exp/log, added in FP64; FP64 filters, rule and heaps; Eisenstein & Hu
 * P(k) only).  No counterpart in the reference (its integrand is FP64). */
int tp_map3_cov_t1_batch_mode(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta6,
                              double area, double lmin, double sigma, double n_gal, int flags, double reltol,
                              int mode, double *out, long long *neval);

/* Replaces T1_total (cuda_version/apertureStatisticsCovariance.cu:272-326) for every survey geometry
 * (the reference's type codes: 0 circle of radius theta_max, 1 square of side theta_max, 2 infinite,
 * 3 rectangle of sides theta_max, theta_max2; rad).  Geometry 2 is tp_map3_cov_t1_batch (area in
 * rad^2).  The finite ones (T1 :1065-1102) are 6-D hcubatures over [-lMax, 1.02 lMax]^6, lMax = 10 /
 * max(min theta_123, min theta_456), of Pell(l1) Pell(l2) Pell(l3), both uHat products and the
 * window G (integrands :1886-1955, 1984-2012), relative tolerance reltol (the reference's 1e-2),
 * divided by (2 pi)^6; area is not used.  Pell, sigma, n_gal, lmin and flags as
 * tp_map3_cov_t1_batch; neval receives each pair's integrand points. */
int tp_map3_cov_t1_batch_geometry(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta6,
                                  int geometry, double theta_max, double theta_max2, double area, double lmin,
                                  double sigma, double n_gal, int flags, double reltol, double *out,
                                  long long *neval);
/* Replaces T2_total (cuda_version/apertureStatisticsCovariance.cu:328-538) for n pairs at once:
 * geometry 0 circle (radius theta_max), 1 square (side theta_max), 3 rectangle (sides theta_max,
 * theta_max2), all in rad; 2 infinite, where T2 = 0 (:330-331).  Each pair sums T2 over the pairings
 * of theta_123 with theta_456, with the reference's multiplicities; each T2 (:1108-1163) is the
 * product A1 A2 B of the 1-D integrals A of ell Pell uHat uHat / 2 pi (integrand :2014-2033, twice)
 * and B, the window-weighted Pell uHat uHat integral: 1-D over ell for the circle (x 4 J1^2 / x^2,
 * :2035-2058), 2-D over (ell_x, ell_y) for the square and the rectangle (x sinc^2 sinc^2,
 * :2060-2110); every integral at relative tolerance reltol (the reference's 1e-4) with the
 * reference's domains.  Pell, sigma, n_gal (rad^-2), lmin and flags as tp_map3_cov_t1_batch.  Equal
 * integrals within the call are evaluated once; neval receives each pair's integrand points as the
 * reference would evaluate them (every T2 call's three integrals). */
int tp_map3_cov_t2_batch(tp_ctx *ctx, const tp_cosmo *c, size_t n, const int *icosmo, const double *theta6,
                         int geometry, double theta_max, double theta_max2, double lmin, double sigma, double n_gal,
                         int flags, double reltol, double *out, long long *neval);

/* ---------------------------------------------------------- input files */
/* Replaces read_n_of_z: averaging = 0 the classed_version's reader (helper.cpp:31-96, linear
 * resampling to z_i = i z_max / n_bins), 1 the cuda_version's (helpers.cu:9-127, overlap-weighted
 * bin averages when the file is coarser than the internal bins); out receives n_bins values
 * normalised to sum n dz = 1. */
int tp_read_n_of_z(const char *fn, int n_bins, double z_max, int averaging, double *out);
/* Replaces read_triangle_configurations (classed_version/helper.cpp:5-29): the (r, u, v) of each
 * "rind uind vind r u v" row into ruv (3 per row, at most cap rows); *n receives the row count
 * (TP_ECAPACITY when it exceeds cap). */
int tp_read_triangle_configurations(const char *fn, size_t cap, double *ruv, size_t *n);

/* ------------------------------------------------------------- cubature */
/* Replaces hcubature_v (classed_version/cubature.h:90-94, hcubature.cpp:1059-1068):
 * same arguments, same callback layout x[i*ndim+j], fval[i*fdim+k], same return
 * convention (0 success, 1 failure or callback abort).  norm takes the
 * error_norm values of cubature.h:63-70. */
typedef int (*tp_integrand_v)(unsigned ndim, size_t npt, const double *x, void *fdata,
                              unsigned fdim, double *fval);
int tp_hcubature_v(unsigned fdim, tp_integrand_v f, void *fdata, unsigned dim,
                   const double *xmin, const double *xmax, size_t maxEval, double reqAbsError,
                   double reqRelError, int norm, double *val, double *err);
/* integrand points evaluated by the last tp_hcubature_v call on this thread */
long long tp_hcubature_last_neval(void);

/* -------------------------------------------------------------- surrogate */
/* The learned part: an MLP 8 -> 4x256 (SiLU) -> 1 for log|B_kappa(l1,l2,l3; cosmology)|.
 * Inputs: the triangle as its three sides' logs in sorted order (the information of
 * (ln l1, ln l2, cos phi) with B_kappa's permutation symmetry factored out) and
 * Omega_m, sigma8, h, w, n_s.  No counterpart in the reference; it serves TP_MODE_SURROGATE. */
typedef struct {
    long long nsamples;   /* training samples generated by the exact GPU path     */
    int epochs;           /* passes over the samples                              */
    int batch;            /* minibatch size                                       */
    double lr, lr_final;  /* Adam learning rate, cosine-decayed to lr_final       */
    double lnl_min, lnl_max;          /* ln l box                                 */
    double om[2], s8[2], h[2], w[2], ns[2]; /* cosmology box                      */
    double omb;           /* fixed Omega_b                                        */
    uint64_t seed;
    int n_z;              /* redshift bins of the exact tables                    */
    double z_max;
    const double *nz;     /* n(z) used to generate samples (n_z values)           */
} tp_train_config;
typedef struct {
    double final_loss, val_rmse, val_maxerr; /* in log|B| units on held-out samples */
    double gen_ms, train_ms;
    long long steps;
} tp_train_report;
int tp_surrogate_train(tp_ctx *ctx, const tp_train_config *cfg, tp_train_report *rep);
int tp_surrogate_save(tp_ctx *ctx, const char *path);
int tp_surrogate_load(tp_ctx *ctx, const char *path);
/* Surrogate settings used by TP_MODE_SURROGATE: check_frac of each batch's points
 * are re-evaluated exactly; if the surrogate's error on them exceeds tol the batch
 * falls back to exact B_kappa. */
int tp_surrogate_config(tp_ctx *ctx, double check_frac, double tol, uint64_t seed);
/* The surrogate's inference kernel: TP_SUR_INFER_FP16 (default) rounds the hidden activations once to
 * FP16 (one FP16 MFMA per weight fragment); TP_SUR_INFER_HILO_BF16 carries them as BF16 hi + lo pairs
 * (~16 significant bits, two BF16 MFMAs per fragment).  The weights are the trained BF16 values in
 * both.  No counterpart in the reference. */
enum { TP_SUR_INFER_HILO_BF16 = 0, TP_SUR_INFER_FP16 = 1 };
int tp_surrogate_inference(tp_ctx *ctx, int kind);
/* Surrogate mode evaluates an integral's regions exactly (the FP32 form) until the integral holds
 * nregions regions (default 1024; 0: the surrogate from the first round).  An integral that the
 * reference algorithm settles within that many regions -- where its early stopping decides the value --
 * then takes the exact paths' decisions.  No counterpart in the reference. */
int tp_surrogate_exact_regions(tp_ctx *ctx, long long nregions);
/* Surrogate mode evaluates a triple exactly (the FP32 form) throughout when its largest radius is at
 * least ratio times its smallest (default 16; 0: no triple).  Such a triple puts its weight on a ridge
 * near phi = pi of relative width ~theta_min / theta_max, which the reference algorithm finds only after
 * many splits; its decision sequence there, and so its value at the reference's tolerance, moves by
 * more than 1e-3 under 1e-4-level changes of the integrand.  No counterpart in the reference. */
int tp_surrogate_exact_ratio(tp_ctx *ctx, double ratio);
/* y[i] = the network's output at in[8*i..] = (l1, l2, l3, Omega_m, sigma8, h, w, n_s): the learned
 * ln B_kappa - ln B_8, where B_8 is the 8-node FP32 Limber baseline of the same cosmology. */
int tp_surrogate_predict(tp_ctx *ctx, size_t n, const double *in8, double *y);
/* Surrogate B_kappa = B_8 exp(network) at ell123 (npts*3) for cosmology icosmo of c. */
int tp_surrogate_bkappa(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ell123,
                        double *out);
/* Raw MLP access for gradient checks: parameters as one FP32 vector (layout
 * W1,b1,...,W5,b5; Wk row-major [out][in]). */
int tp_mlp_num_params(size_t *n);
int tp_mlp_get_params(tp_ctx *ctx, float *p);
int tp_mlp_set_params(tp_ctx *ctx, const float *p);
/* One forward+backward pass on a batch of normalised inputs x (n*8) and targets t (n):
 * loss = mean (y - t)^2 / 2; grad (n_params) receives d loss / d params. */
int tp_mlp_loss_grad(tp_ctx *ctx, size_t n, const float *x, const float *t, double *loss,
                     float *grad);

#ifdef __cplusplus
}
#endif
#endif
