// tp_classes.cpp -- the drop-in class API (threepoint_amd.hpp) over the C ABI.
#include <cmath>
#include <iostream>
#include <mutex>
#include <stdexcept>

#include "threepoint_amd.hpp"

static void check(int rc, const char *what)
{
    if (rc) throw std::runtime_error(std::string(what) + ": " + tp_last_error());
}

tp_ctx *threepoint_default_context(int device)
{
    static std::mutex mu;
    static tp_ctx *ctx[64] = {nullptr};
    std::lock_guard<std::mutex> lk(mu);
    if (device < 0 || device >= 64) throw std::runtime_error("threepoint_default_context: bad device");
    if (!ctx[device]) check(tp_ctx_create(device, &ctx[device]), "tp_ctx_create");
    return ctx[device];
}

BispectrumCalculator::BispectrumCalculator(cosmology *c, const std::vector<double> &n_of_z, int n_z, double zmax)
    : cosmo(c), z_max(zmax), n_of_z_(n_of_z), n_z_(n_z)
{
    // classed_version/bispectrum.cpp:73-79 -> initialize :91-109 -> set_cosmology :111-164
    if ((int)n_of_z.size() < n_z) throw std::runtime_error("BispectrumCalculator: n_of_z shorter than n_z");
    ctx_ = threepoint_default_context(0);
    set_cosmology(c);
}

BispectrumCalculator::BispectrumCalculator(cosmology *c, const std::vector<double> &n_of_z, int n_z, double zmax,
                                           const std::map<double, double> &P_k)
    : cosmo(c), z_max(zmax), n_of_z_(n_of_z), n_z_(n_z)
{
    // classed_version/bispectrum.cpp:81-89: the map's ascending keys are the table
    if ((int)n_of_z.size() < n_z) throw std::runtime_error("BispectrumCalculator: n_of_z shorter than n_z");
    if (P_k.size() < 2) throw std::runtime_error("BispectrumCalculator: P_k needs at least 2 entries");
    for (const auto &kv : P_k) {
        pk_k_.push_back(kv.first);
        pk_p_.push_back(kv.second);
    }
    ctx_ = threepoint_default_context(0);
    set_cosmology(c);
}

BispectrumCalculator::~BispectrumCalculator()
{
    if (tables_) tp_cosmo_destroy(tables_);
}

void BispectrumCalculator::set_cosmology(cosmology *c)
{
    cosmo = c;
    if (tables_) tp_cosmo_destroy(tables_);
    tables_ = nullptr;
    tp_cosmology p{c->h, c->sigma8, c->omb, c->ns, c->w, c->om, z_max};
    if (pk_k_.empty())
        check(tp_cosmo_build(ctx_, 1, &p, n_of_z_.data(), n_z_, 0, z_max, &tables_), "tp_cosmo_build");
    else
        check(tp_cosmo_build_pk(ctx_, 1, &p, n_of_z_.data(), n_z_, 0, z_max, pk_k_.data(), pk_p_.data(),
                                (int)pk_k_.size(), 0, &tables_),
              "tp_cosmo_build_pk");
    check(tp_cosmo_tables(tables_, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &norm), "tp_cosmo_tables");
}

double BispectrumCalculator::bispec(double k1, double k2, double k3, double z)
{
    double in[4] = {k1, k2, k3, z}, out = 0;
    check(tp_bispec_eval(ctx_, tables_, 0, 1, in, &out), "tp_bispec_eval");
    return out;
}

double BispectrumCalculator::bispec(double k1, double k2, double k3, double z, int idx, double didx)
{
    double in[6] = {k1, k2, k3, z, (double)idx, didx}, out = 0;
    check(tp_bispec_eval_idx(ctx_, tables_, 0, 1, in, &out), "tp_bispec_eval_idx");
    return out;
}

double BispectrumCalculator::bkappa(double ell1, double ell2, double ell3)
{
    double in[3] = {ell1, ell2, ell3}, out = 0;
    check(tp_bkappa_eval(ctx_, tables_, 0, 1, in, &out, TP_MODE_FP64), "tp_bkappa_eval");
    return out;
}

void BispectrumCalculator::bkappa(const std::vector<double> &ell123, std::vector<double> &out, int mode)
{
    out.assign(ell123.size() / 3, 0.0);
    check(tp_bkappa_eval(ctx_, tables_, 0, out.size(), ell123.data(), out.data(), mode), "tp_bkappa_eval");
}

double BispectrumCalculator::integrand_bkappa(double z, ell_params p)
{
    double in[4] = {z, p.ell1, p.ell2, p.ell3}, out = 0;
    check(tp_integrand_bkappa_eval(ctx_, tables_, 0, 1, in, &out), "tp_integrand_bkappa_eval");
    return out;
}

double BispectrumCalculator::P_k_nonlinear(double k, double z)
{
    double in[2] = {k, z}, out = 0;
    check(tp_pk_at_z_eval(ctx_, tables_, 0, 1, in, 0, &out), "tp_pk_at_z_eval");
    return out;
}

double BispectrumCalculator::linear_pk_at_z(double k, double z)
{
    double in[2] = {k, z}, out = 0;
    check(tp_pk_at_z_eval(ctx_, tables_, 0, 1, in, TP_PK_LINEAR, &out), "tp_pk_at_z_eval");
    return out;
}

double BispectrumCalculator::limber_integrand_power_spectrum(double ell, double z, bool nonlinear)
{
    double in[2] = {ell, z}, out = 0;
    check(tp_limber_pkappa_eval(ctx_, tables_, 0, 1, in, nonlinear ? 0 : TP_PK_LINEAR, &out),
          "tp_limber_pkappa_eval");
    return out;
}

double BispectrumCalculator::convergence_power_spectrum(double ell)
{
    return convergence_power_spectrum(std::vector<double>{ell})[0];
}

std::vector<double> BispectrumCalculator::convergence_power_spectrum(const std::vector<double> &ell)
{
    // classed_version/bispectrum.cpp:382-388: relative tolerance 1e-6
    std::vector<double> out(ell.size()), err(ell.size());
    check(tp_pkappa_batch(ctx_, tables_, ell.size(), nullptr, ell.data(), 1e-6, 0, out.data(), err.data(), nullptr),
          "tp_pkappa_batch");
    return out;
}

double BispectrumCalculator::Pell(double ell) { return Pell(std::vector<double>{ell})[0]; }

std::vector<double> BispectrumCalculator::Pell(const std::vector<double> &ell)
{
    std::vector<double> out(ell.size());
    check(tp_pell(ctx_, tables_, 0, ell.size(), ell.data(), 0, sigma, n, out.data()), "tp_pell");
    return out;
}

ApertureStatistics::ApertureStatistics(BispectrumCalculator *b) : Bispectrum_(b) {}

double ApertureStatistics::Map2(double theta, double sigma_epsilon, double ngal, bool t17)
{
    return Map2(std::vector<double>{theta}, sigma_epsilon, ngal, t17)[0];
}

std::vector<double> ApertureStatistics::Map2(const std::vector<double> &thetas, double sigma_epsilon, double ngal,
                                             bool t17)
{
    std::vector<double> out(thetas.size()), err(thetas.size());
    check(tp_map2_batch(Bispectrum_->ctx(), Bispectrum_->tables(), thetas.size(), nullptr, thetas.data(), reltol_map2,
                        t17 ? TP_PK_T17 : 0, sigma_epsilon, ngal, out.data(), err.data(), nullptr),
          "tp_map2_batch");
    return out;
}

double ApertureStatistics::uHat(const double &eta)
{
    double t = 0.5 * eta * eta;  // classed_version/apertureStatistics.cpp:10-14
    return t * std::exp(-t);
}

double ApertureStatistics::integrand(const double &l1, const double &l2, const double &phi, std::vector<double> th)
{
    double x[3] = {l1, l2, phi}, f = 0;
    check(tp_map3_eval(Bispectrum_->ctx(), Bispectrum_->tables(), 0, th.data(), 1, x, &f, mode), "tp_map3_eval");
    return f;
}

int ApertureStatistics::integrand(unsigned ndim, size_t npts, const double *vars, void *thisPtr, unsigned fdim,
                                  double *value)
{
    // the integrand_v adapter, classed_version/apertureStatistics.cpp:254-279: one GPU batch
    if (fdim != 1 || ndim != 3) return 1;
    ApertureStatisticsContainer *c = (ApertureStatisticsContainer *)thisPtr;
    ApertureStatistics *a = c->aperturestatistics;
    return tp_map3_eval(a->Bispectrum_->ctx(), a->Bispectrum_->tables(), 0, c->thetas.data(), npts, vars, value,
                        a->mode) ? 1 : 0;
}

double ApertureStatistics::MapMapMap(const std::vector<double> &thetas)
{
    // classed_version/apertureStatistics.cpp:321-353 (3-D branch)
    if (thetas.size() != 3) throw std::runtime_error("MapMapMap: need three aperture radii");
    double val = 0, err = 0;
    long long n = 0;
    check(tp_map3_single(Bispectrum_->ctx(), Bispectrum_->tables(), 0, thetas.data(), reltol, mode, &val, &err, &n),
          "tp_map3_single");
    lMax = 10. / std::min(thetas[0], std::min(thetas[1], thetas[2]));
    return val;
}

std::vector<double> ApertureStatistics::MapMapMap(const std::vector<std::array<double, 3>> &thetas)
{
    std::vector<double> th, val(thetas.size()), err(thetas.size());
    std::vector<int> ic(thetas.size(), 0);
    for (auto &t : thetas) th.insert(th.end(), t.begin(), t.end());
    check(tp_map3_batch(Bispectrum_->ctx(), Bispectrum_->tables(), thetas.size(), ic.data(), th.data(), reltol, mode,
                        val.data(), err.data(), nullptr),
          "tp_map3_batch");
    return val;
}

GammaCalculator::GammaCalculator(BispectrumCalculator *b, double prec_h, double prec_k, std::string center)
    : Bispectrum_(b)
{
    // classed_version/gamma.cpp:6-23 (the reference exits on an unknown center; this throws)
    if (center == "orthocenter")
        center_ = TP_CENTER_ORTHOCENTER;
    else if (center == "centroid")
        center_ = TP_CENTER_CENTROID;
    else
        throw std::runtime_error("GammaCalculator: no valid triangle center given; possible choices: "
                                 "orthocenter, centroid");
    check(tp_gamma_setup(b->ctx(), prec_h, prec_k), "tp_gamma_setup");
}

std::array<std::complex<double>, 4> GammaCalculator::all(double x1, double x2, double x3)
{
    return gammas({{{x1, x2, x3}}})[0];
}

std::vector<std::array<std::complex<double>, 4>> GammaCalculator::gammas(const std::vector<std::array<double, 3>> &x)
{
    std::vector<double> in, re(4 * x.size()), im(4 * x.size());
    for (auto &t : x) in.insert(in.end(), t.begin(), t.end());
    check(tp_gamma_batch_center(Bispectrum_->ctx(), Bispectrum_->tables(), 0, x.size(), in.data(), reltol, mode,
                                center_, re.data(), im.data(), nullptr),
          "tp_gamma_batch");
    std::vector<std::array<std::complex<double>, 4>> out(x.size());
    for (size_t t = 0; t < x.size(); ++t)
        for (int i = 0; i < 4; ++i) out[t][i] = std::complex<double>(re[4 * t + i], im[4 * t + i]);
    return out;
}

std::complex<double> GammaCalculator::gamma0(double x1, double x2, double x3) { return all(x1, x2, x3)[0]; }
std::complex<double> GammaCalculator::gamma1(double x1, double x2, double x3) { return all(x1, x2, x3)[1]; }
std::complex<double> GammaCalculator::gamma2(double x1, double x2, double x3) { return all(x1, x2, x3)[2]; }
std::complex<double> GammaCalculator::gamma3(double x1, double x2, double x3) { return all(x1, x2, x3)[3]; }

double ApertureStatistics::MapMapMap_covariance_Gauss(const std::vector<double> &thetas_123,
                                                      const std::vector<double> &thetas_456, double survey_area)
{
    if (thetas_123.size() != 3 || thetas_456.size() != 3)
        throw std::runtime_error("MapMapMap_covariance_Gauss: need two triples of aperture radii");
    double th6[6] = {thetas_123[0], thetas_123[1], thetas_123[2], thetas_456[0], thetas_456[1], thetas_456[2]};
    double v = 0, e = 0;
    check(tp_map3_cov_gauss_batch(Bispectrum_->ctx(), Bispectrum_->tables(), 1, nullptr, th6, survey_area, reltol_cov,
                                  &v, &e, nullptr),
          "tp_map3_cov_gauss_batch");
    return v;
}

ApertureStatisticsCovariance::ApertureStatisticsCovariance(BispectrumCalculator *b, const covarianceParameters &c,
                                                           const std::string &type, bool t17)
    : Bispectrum_(b), covPar_(c), flags_(t17 ? TP_PK_T17 : 0)
{
    // cuda_version/calculateApertureStatisticsCovariance.cu:109-129
    if (type == "circle") geometry_ = 0;
    else if (type == "square") geometry_ = 1;
    else if (type == "infinite") geometry_ = 2;
    else if (type == "rectangular") geometry_ = 3;
    else throw std::runtime_error("ApertureStatisticsCovariance: Cov type not correctly specified");
    if (covPar_.shapenoiseOnly) flags_ |= TP_PK_SHAPENOISE_ONLY;
}

std::vector<double> ApertureStatisticsCovariance::T1_total(const std::vector<std::array<double, 6>> &pairs)
{
    std::vector<double> th, out(pairs.size());
    for (auto &p : pairs) th.insert(th.end(), p.begin(), p.end());
    if (geometry_ != 2) {  // the finite geometries: 6-D, FP64, the reference's tolerance 1e-2
        if (mode != TP_MODE_FP64)
            throw std::invalid_argument("T1_total: the FP32 form exists for the infinite survey only");
        check(tp_map3_cov_t1_batch_geometry(Bispectrum_->ctx(), Bispectrum_->tables(), pairs.size(), nullptr,
                                            th.data(), geometry_, covPar_.thetaMax, covPar_.thetaMax_smaller,
                                            covPar_.area, lMin, covPar_.shapenoise_sigma, covPar_.galaxy_density,
                                            flags_, reltol_finite, out.data(), nullptr),
              "tp_map3_cov_t1_batch_geometry");
        return out;
    }
    check(tp_map3_cov_t1_batch_mode(Bispectrum_->ctx(), Bispectrum_->tables(), pairs.size(), nullptr, th.data(),
                                    covPar_.area, lMin, covPar_.shapenoise_sigma, covPar_.galaxy_density, flags_,
                                    reltol, mode, out.data(), nullptr),
          "tp_map3_cov_t1_batch_mode");
    return out;
}

double ApertureStatisticsCovariance::T1_total(const std::vector<double> &a, const std::vector<double> &b)
{
    if (a.size() != 3 || b.size() != 3) throw std::invalid_argument("T1_total: Wrong number of aperture radii");
    return T1_total(std::vector<std::array<double, 6>>{{a[0], a[1], a[2], b[0], b[1], b[2]}})[0];
}

std::vector<double> ApertureStatisticsCovariance::T2_total(const std::vector<std::array<double, 6>> &pairs)
{
    if (mode != TP_MODE_FP64) throw std::invalid_argument("T2_total: the FP32 form exists for T1 only");
    std::vector<double> th, out(pairs.size());
    for (auto &p : pairs) th.insert(th.end(), p.begin(), p.end());
    check(tp_map3_cov_t2_batch(Bispectrum_->ctx(), Bispectrum_->tables(), pairs.size(), nullptr, th.data(), geometry_,
                               covPar_.thetaMax, covPar_.thetaMax_smaller, lMin, covPar_.shapenoise_sigma,
                               covPar_.galaxy_density, flags_, reltol, out.data(), nullptr),
          "tp_map3_cov_t2_batch");
    return out;
}

double ApertureStatisticsCovariance::T2_total(const std::vector<double> &a, const std::vector<double> &b)
{
    if (a.size() != 3 || b.size() != 3) throw std::invalid_argument("T2_total: Wrong number of aperture radii");
    return T2_total(std::vector<std::array<double, 6>>{{a[0], a[1], a[2], b[0], b[1], b[2]}})[0];
}
