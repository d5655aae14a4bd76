/*
 * tp_oracle.h -- CPU FP64 oracle for the <Map^3> / Gamma_i modelling path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in threepoint_amd/ links, loads or calls
 * this library.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * It is a plain restatement of the reference's classed_version CPU path
 * (bispectrum.cpp, apertureStatistics.cpp, gamma.cpp, hcubature.cpp,
 * helper.cpp, cosmology.cpp; <Map^2> from the cuda_version, which alone has it), with the GSL pieces that path calls
 * (gsl_integration_qag, gsl_sf_bessel_Jn, gsl_sf_bessel_zero_Jnu) replaced
 * by project-written routines that follow the published algorithms.
 * Each function in tp_oracle.cpp cites the reference file:line it follows.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same callback convention as the reference's cubature.h:55-57. */
typedef int (*orc_integrand_v)(unsigned ndim, size_t npt, const double *x,
                               void *fdata, unsigned fdim, double *fval);

/* error_norm values follow cubature.h:63-70 */
enum { ORC_ERROR_INDIVIDUAL = 0, ORC_ERROR_PAIRED, ORC_ERROR_L2, ORC_ERROR_L1, ORC_ERROR_LINF };

int orc_hcubature_v(unsigned fdim, orc_integrand_v f, void *fdata, unsigned dim,
                    const double *xmin, const double *xmax, size_t maxEval,
                    double reqAbsError, double reqRelError, int norm,
                    double *val, double *err);
/* number of integrand points the last orc_hcubature_v call on this thread evaluated */
long long orc_last_neval(void);
long long orc_last_nregions(void);

/* GSL-qag-equivalent (GK15, key=1) adaptive quadrature of a 1-D function. */
typedef double (*orc_f1)(double x, void *p);
double orc_qag15(orc_f1 f, void *p, double a, double b, double epsabs, double epsrel, int limit);

/* Gauss-Legendre n-point nodes/weights on [-1,1] (full set, ascending). */
void orc_gauss_legendre(int n, double *x, double *w);

/* io (helper.cpp / cosmology.cpp) */
int orc_read_cosmology(const char *fn, double *out7 /* h s8 omb ns w om zmax */);
int orc_read_n_of_z(const char *fn, int n_bins, double zmax, double *nz_out);

/* Bispectrum calculator handle.  params7 = {h, sigma8, Omega_b, n_s, w, Omega_m, z_max}. */
void *orc_create(const double *params7, const double *nz, int n_z, double z_max);
void *orc_create_pk(const double *p, const double *nz, int n_z, double z_max, const double *k, const double *pk,
                    int nk);
void orc_destroy(void *h);
void orc_set_analytic(void *h, int on); /* reference's test_analytical switch (bispectrum.hpp:18) */
void orc_get_tables(void *h, double *fK, double *g, double *D1, double *rs, double *neff,
                    double *ncur, double *norm);
double orc_linear_pk(void *h, double k);
double orc_sigmam(void *h, double r, int j);
double orc_lgr(void *h, double z);
double orc_f_K_at_z(void *h, double z);
double orc_bispec(void *h, double k1, double k2, double k3, double z);
double orc_bispec_idx(void *h, double k1, double k2, double k3, double z, int idx, double didx);
double orc_integrand_bkappa(void *h, double z, double l1, double l2, double l3);
double orc_bkappa(void *h, double l1, double l2, double l3);
void orc_bkappa_batch(void *h, long long n, const double *l123, double *out, int nthreads);
double orc_map3_integrand(void *h, double l1, double l2, double phi, const double *thetas);
void orc_map3_integrand_batch(void *h, long long n, const double *x, const double *thetas,
                              double *out, int nthreads);
int orc_map3(void *h, const double *thetas_rad, double reltol, int nthreads,
             double *val, double *err, long long *neval);

/* Gamma_i (gamma.cpp) */
double orc_bessel_jn(int n, double x);
double orc_bessel_zero_jn(int n, int s);
int orc_gamma_init(void *h, double prec_h, double prec_k_arg);
int orc_gamma_integrand(void *h, int which, double z, double phi, double psi,
                        double x1, double x2, double x3, double *re, double *im);
int orc_gamma_integrand_batch(void *h, int which, double x1, double x2, double x3, size_t npts,
                              const double *x, double *fval, int nthreads);
int orc_gamma(void *h, int which, double x1, double x2, double x3, double reltol,
              int nthreads, double *re, double *im, long long *neval);
void orc_ggg_analytic(double r1, double r2, double r3, double *re, double *im);
int orc_gamma_center(void *h, int which, double x1, double x2, double x3, double reltol, int nthreads, int center,
                     double *re, double *im, long long *neval);

/* convergence power spectrum and <Map^2> (bispectrum.cpp:288-417, 543-577, 1022-1059;
   the cuda_version's Map2 and T17 correction, apertureStatistics.cu:29-136, bispectrum.cu:951-982) */
double orc_pk_nonlinear(void *h, double k, double z);
double orc_linear_pk_at_z(void *h, double k, double z);
double orc_om_m_of_z(void *h, double z);
double orc_om_v_of_z(void *h, double z);
/* flags: 1 = linear D1^2 P_lin instead of Halofit, 2 = T17 correction */
double orc_limber_pkappa(void *h, double ell, double z, int flags);
double orc_gq96_pkappa(void *h, double ell, int flags);
double orc_pell(void *h, double ell, double sigma, double n_gal_rad2, int flags);
double orc_map2_integrand(void *h, double ell, double z, double theta, int flags, double sigma,
                          double ngal_arcmin2);
int orc_convergence_power_spectrum(void *h, double ell, double reltol, int flags, double *val,
                                   double *err, long long *neval);
int orc_map2(void *h, double theta, double reltol, int flags, double sigma, double ngal_arcmin2,
             double *val, double *err, long long *neval);

/* Gaussian <Map^3> covariance: the classed_version's 4-D form and the cuda_version's T1 for an
   infinite survey (flags: 1 linear, 2 T17, 4 shape noise only; th6 = theta_123, theta_456 in rad) */
double orc_limber_triple_pk(void *h, double l1, double l2, double l3, double z);
double orc_cov_gauss_integrand(void *h, double l1, double l2, double phi, double z, const double *th6);
double orc_t1_integrand(void *h, double l1, double l2, double phi, const double *th6, double lmin,
                        double lmax, double sigma, double n_gal, int flags);
int orc_map3_cov_gauss(void *h, const double *th6, double survey_area, double reltol, int nthreads,
                       double *val, double *err, long long *neval);
int orc_t1(void *h, const double *th6, double area, double lmin, double sigma, double n_gal, int flags,
           double reltol, int nthreads, double *val, double *err, long long *neval);
int orc_t1_total(void *h, const double *th123, const double *th456, double area, double lmin, double sigma,
                 double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval);

/* T2 of the Gaussian <Map^3> covariance for the survey geometries of the cuda_version
   (geometry 0 circle, 1 square, 2 infinite (T2 = 0), 3 rectangle; theta_max and theta_max2 the
   side lengths or the circle's radius [rad]; flags as orc_t1) */
/* T1 of the finite geometries (0 circle, 1 square, 3 rectangle): 6-D, no area factor */
int orc_t1_finite(void *h, const double *th6, int geometry, double theta_max, double theta_max2, double lmin,
                  double sigma, double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval);
int orc_t1_total_finite(void *h, const double *th123, const double *th456, int geometry, double theta_max,
                        double theta_max2, double lmin, double sigma, double n_gal, int flags, double reltol,
                        int nthreads, double *val, long long *neval);
int orc_t2(void *h, const double *th6, int geometry, double theta_max, double theta_max2, double lmin,
           double sigma, double n_gal, int flags, double reltol, int nthreads, double *val, long long *neval);
int orc_t2_total(void *h, const double *th123, const double *th456, int geometry, double theta_max,
                 double theta_max2, double lmin, double sigma, double n_gal, int flags, double reltol, int nthreads,
                 double *val, long long *neval);

#ifdef __cplusplus
}
#endif
#endif
