/*
 * threepoint_amd.hpp -- C++ drop-in class API over the C ABI (threepoint_amd.h).
 *
 * Same class names, constructors and member signatures as the reference's
 * classed_version headers, so code written against them recompiles unchanged:
 *   cosmology               classed_version/cosmology.hpp:11-29
 *   read_n_of_z, read_thetas, convert_angle_to_rad, convert_rad_to_angle
 *                           classed_version/helper.hpp (helper.cpp:31-200)
 *   configGamma, read_gamma_config
 *                           cuda_version/helpers.cuh (helpers.cu:242-330)
 *   covarianceParameters    cuda_version/cosmology.cuh:44-57
 *   BispectrumCalculator    classed_version/bispectrum.hpp:113-509
 *   ApertureStatistics      classed_version/apertureStatistics.hpp:18-322 (+ the cuda_version's Map2,
 *                           apertureStatistics.cu:110-136)
 *   GammaCalculator         classed_version/gamma.hpp:17-495
 * All numerics run on the GPU through the C ABI; objects hold non-owning
 * pointers exactly like the reference (the caller keeps cosmology and
 * BispectrumCalculator alive).  Instances are not thread-safe.
 */
#ifndef THREEPOINT_AMD_HPP
#define THREEPOINT_AMD_HPP

#include <array>
#include <complex>
#include <iosfwd>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "threepoint_amd.h"

class cosmology {
  public:
    double h = 0.7, sigma8 = 0.82, omb = 0.047, omc = 0.253, ns = 0.97, w = -1, om = 0.3, ow = 0.7, zmax = 3.0;
    cosmology() {}
    explicit cosmology(const std::string &fn_parameters);  // throws std::runtime_error on bad files
};
std::ostream &operator<<(std::ostream &out, const cosmology &cosmo);

void read_n_of_z(const std::string &fn, const int &n_bins, const double &zMax, std::vector<double> &nz);
void read_thetas(const std::string &fn, std::vector<double> &thetas);
/* the cuda_version's n(z) reader (cuda_version/helpers.cu:9-127): interpolation when the internal bins
 * are coarser than the file's spacing, overlap-weighted bin averages otherwise */
void read_n_of_z_averaging(const std::string &fn, const int &n_bins, const double &zMax, std::vector<double> &nz);

/* TreeCorr triangle bins, classed_version/helper.hpp:9-14 (helper.cpp:5-29): rows "rind uind vind r u v" */
struct treecorr_bin {
    double r, u, v;
};
int read_triangle_configurations(std::string &infile, std::vector<treecorr_bin> &triangle_config);
double convert_angle_to_rad(const double &value, const std::string &unit = "arcmin");
double convert_rad_to_angle(const double &value, const std::string &unit = "arcmin");

struct configGamma {
    int rsteps = 10, usteps = 10, vsteps = 10;
    double umin = 0, umax = 1, vmin = 0, vmax = 1, rmin = 0.1, rmax = 100;
};
void read_gamma_config(const std::string &fn, configGamma &config);
std::ostream &operator<<(std::ostream &out, const configGamma &config);

/* Survey parameters of the covariance executables, cuda_version/cosmology.cuh:44-57 (cosmology.cu:92-179):
 * thetaMax [rad], galaxy_density [rad^-2], area [rad^2] after reading. */
class covarianceParameters {
  public:
    double thetaMax = 0, shapenoise_sigma = 0, galaxy_density = 0, thetaMax_smaller = 0, area = 0;
    bool shapenoiseOnly = false;
    covarianceParameters() {}
    explicit covarianceParameters(const std::string &fn);  // throws std::runtime_error on bad files
};
std::ostream &operator<<(std::ostream &out, const covarianceParameters &covPar);

struct ell_params {
    double ell1, ell2, ell3;
};

/* Shared GPU context of the façade objects (one per process and device). */
tp_ctx *threepoint_default_context(int device = 0);

class BispectrumCalculator {
  public:
    cosmology *cosmo = nullptr;
    double norm = 0;
    double z_max = 0;

    BispectrumCalculator() {}
    BispectrumCalculator(cosmology *cosmo, const std::vector<double> &n_of_z, int n_z, double z_max);
    /* with a given linear P(k), classed_version/bispectrum.hpp:218-219 (bispectrum.cpp:81-89) */
    BispectrumCalculator(cosmology *cosmo, const std::vector<double> &n_of_z, int n_z, double z_max,
                         const std::map<double, double> &P_k);
    ~BispectrumCalculator();
    BispectrumCalculator(const BispectrumCalculator &) = delete;
    BispectrumCalculator &operator=(const BispectrumCalculator &) = delete;

    void set_cosmology(cosmology *cosmo);  // rebuilds the GPU tables
    double bispec(double k1, double k2, double k3, double z);
    /* classed_version/bispectrum.hpp:247: the table interval (idx, didx) given by the caller,
     * idx = int(z / z_max * (n_z - 1)), didx the remainder */
    double bispec(double k1, double k2, double k3, double z, int idx, double didx);
    double bkappa(double ell1, double ell2, double ell3);
    double integrand_bkappa(double z, ell_params p);

    /* batched forms (one GPU launch) */
    void bkappa(const std::vector<double> &ell123, std::vector<double> &out, int mode = TP_MODE_FP64);

    /* second-order statistics, classed_version/bispectrum.hpp:272-299, 505-508 */
    double sigma = 0.3;                                        // shape noise (bispectrum.hpp:120)
    double n = 4096 * 4096 / 10 / 10 * 3.14159265358979323846 * 3.14159265358979323846 / 180 / 180;
                                                               // galaxy density [rad^-2] (:121)
    double P_k_nonlinear(double k, double z);
    double linear_pk_at_z(double k, double z);
    double convergence_power_spectrum(double ell);
    double limber_integrand_power_spectrum(double ell, double z, bool nonlinear);
    double Pell(double ell);
    /* batched forms: many multipoles in one call */
    std::vector<double> convergence_power_spectrum(const std::vector<double> &ell);
    std::vector<double> Pell(const std::vector<double> &ell);

    tp_ctx *ctx() const { return ctx_; }
    const tp_cosmo *tables() const { return tables_; }
    int n_redshift_bins() const { return n_z_; }

  private:
    tp_ctx *ctx_ = nullptr;
    tp_cosmo *tables_ = nullptr;
    std::vector<double> n_of_z_;
    int n_z_ = 0;
    std::vector<double> pk_k_, pk_p_;  // given linear P(k), empty: Eisenstein & Hu
};

class ApertureStatistics;
struct ApertureStatisticsContainer {
    ApertureStatistics *aperturestatistics;
    std::vector<double> thetas;
};

class ApertureStatistics {
  public:
    BispectrumCalculator *Bispectrum_;
    double l1_ = 0, l2_ = 0;
    double phiMin = 0;
    double phiMax = 6.283185307179586;
    double lMin = 1;
    double lMax = 1e4;
    double reltol = 1e-4;     /* hcubature relative tolerance (the GPU CLI's 1e-4) */
    int mode = TP_MODE_FP64;  /* integrand precision / method */

    explicit ApertureStatistics(BispectrumCalculator *Bispectrum);
    double uHat(const double &eta);
    double integrand(const double &l1, const double &l2, const double &phi, std::vector<double> thetas);
    static int integrand(unsigned ndim, size_t npts, const double *vars, void *thisPtr, unsigned fdim,
                         double *value);
    /* one triple, the reference's way: hcubature_v over (l1, l2, phi), GPU integrand per round */
    double MapMapMap(const std::vector<double> &thetas);
    /* many triples in one batched run (same decision sequence per triple) */
    std::vector<double> MapMapMap(const std::vector<std::array<double, 3>> &thetas);

    /* <Map^2>(theta) of one source bin, the cuda_version's Map2 (apertureStatistics.cu:110-136):
     * shape noise sigma_epsilon with ngal [arcmin^-2] when sigma_epsilon != 0; t17 applies the T17
     * factors of that version (its default) */
    double Map2(double theta, double sigma_epsilon = 0, double ngal = 1, bool t17 = false);
    std::vector<double> Map2(const std::vector<double> &thetas, double sigma_epsilon = 0, double ngal = 1,
                             bool t17 = false);
    double reltol_map2 = 1e-5; /* apertureStatistics.cu:133 */

    /* classed_version/apertureStatistics.cpp:355-396 (4-D form, relative tolerance 1e-3) */
    double MapMapMap_covariance_Gauss(const std::vector<double> &thetas_123, const std::vector<double> &thetas_456,
                                      double survey_area);
    double reltol_cov = 1e-3;
};

/* The Gaussian <Map^3> covariance terms of the cuda_version (apertureStatisticsCovariance.cu), for
 * the survey of a covarianceParameters file and a geometry "circle", "square", "infinite" or
 * "rectangular" (:109-124 of its executable), with the survey's thetaMax and thetaMax_smaller.  T1:
 * 3-D for the infinite survey, 6-D with the window for the others; T2: 0 for the infinite one, as
 * the reference returns, the window integrals for the others.  Pell carries the parameters' shape
 * noise; t17 selects the cuda_version's T17 spectrum (its default). */
class ApertureStatisticsCovariance {
  public:
    ApertureStatisticsCovariance(BispectrumCalculator *Bispectrum, const covarianceParameters &covPar,
                                 const std::string &type = "infinite", bool t17 = true);
    double T1_total(const std::vector<double> &thetas_123, const std::vector<double> &thetas_456);
    double T2_total(const std::vector<double> &thetas_123, const std::vector<double> &thetas_456);
    /* many pairs in one batched run: pairs[i] = {theta_123, theta_456} */
    std::vector<double> T1_total(const std::vector<std::array<double, 6>> &pairs);
    std::vector<double> T2_total(const std::vector<std::array<double, 6>> &pairs);
    double lMin = 0;        /* the reference's executable sets lMin = 0 */
    double reltol = 1e-4;   /* apertureStatisticsCovariance.cu:1088 (T1 infinite, T2) */
    double reltol_finite = 1e-2; /* :1096 (T1 of the finite geometries) */
    int mode = TP_MODE_FP64;/* or TP_MODE_FP32: FP32 Pell node values, FP64 sums (not in the reference) */

  private:
    BispectrumCalculator *Bispectrum_;
    covarianceParameters covPar_;
    int flags_;
    int geometry_;  /* the reference's type code: 0 circle, 1 square, 2 infinite, 3 rectangular */
};

class GammaCalculator {
  public:
    GammaCalculator(BispectrumCalculator *Bispectrum, double prec_h, double prec_k, std::string triangle_center);
    std::complex<double> gamma0(double x1, double x2, double x3);
    std::complex<double> gamma1(double x1, double x2, double x3);
    std::complex<double> gamma2(double x1, double x2, double x3);
    std::complex<double> gamma3(double x1, double x2, double x3);
    /* all four natural components of many triangles: out[t] = {Gamma0..Gamma3} */
    std::vector<std::array<std::complex<double>, 4>> gammas(const std::vector<std::array<double, 3>> &x123);
    double reltol = 1e-3;  /* hcubature relative tolerance (classed_version/gamma.cpp:375) */
    int mode = TP_MODE_FP64; /* TP_MODE_FP32: FP32 Ogata terms in their window, FP64 accumulation */

  private:
    BispectrumCalculator *Bispectrum_;
    int center_ = TP_CENTER_CENTROID; /* triangle_center: "centroid" or "orthocenter" */
    std::array<std::complex<double>, 4> all(double x1, double x2, double x3);
};

/* same callback type and entry point as classed_version/cubature.h:55-57, 90-94 */
typedef tp_integrand_v integrand_v;
enum error_norm { ERROR_INDIVIDUAL = 0, ERROR_PAIRED, ERROR_L2, ERROR_L1, ERROR_LINF };
inline int hcubature_v(unsigned fdim, integrand_v f, void *fdata, unsigned dim, const double *xmin,
                       const double *xmax, size_t maxEval, double reqAbsError, double reqRelError, error_norm norm,
                       double *val, double *err)
{
    return tp_hcubature_v(fdim, f, fdata, dim, xmin, xmax, maxEval, reqAbsError, reqRelError, (int)norm, val, err);
}

#endif
