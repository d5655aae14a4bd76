// threepoint_module.cpp -- cosmosis module computing <Map^3> on the GPU.
//
// Replaces the reference's cosmosis interface (classed_version/interface.cpp:18-260): the same
// "threepoint" options section, the same inputs from the datablock (cosmological_parameters and
// the linear matter power spectrum matter_power_lin), the same output (threepoint/Map3s in the
// order of the aperture-radius triples), the same optional per-cosmology text file.  The
// difference is the engine: one tp_map3_batch call integrates all triples of a sampled cosmology
// together, and no error path calls exit() -- every failure returns a nonzero status.
//
// Written against cosmosis's C datablock API, so the library carries no cosmosis headers; the
// datablock functions are resolved when cosmosis (libcosmosis) loads the module.
//
// Options ("threepoint" section), as the reference: nz_file, zMax, theta_unit, theta_min,
// theta_max, theta_bins, diag_only, output_Map3, output_dir, and zBins.  zBins is read by the
// classed_version interface (interface.cpp:76); the cuda_version interface, which the shipped
// pipeline (cosmosis/pipeline_gpu.ini) loads, has no such option and uses its fixed
// n_redshift_bins = 256 (cuda_version/bispectrum.cuh:35, interface.cu:81), so an absent zBins
// means 256 here.  Additional, optional: reltol (default 1e-4, the reference's tolerance) and
// precision ("fp64" default, or "fp32").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "threepoint_amd.h"
#include "threepoint_amd.hpp"

extern "C" {
// cosmosis/datablock/c_datablock.h (public C API of the cosmosis datablock)
typedef void c_datablock;
typedef int DATABLOCK_STATUS;  // 0 = DBS_SUCCESS
DATABLOCK_STATUS c_datablock_get_int(c_datablock *s, const char *section, const char *name, int *val);
DATABLOCK_STATUS c_datablock_get_bool(c_datablock *s, const char *section, const char *name, bool *val);
DATABLOCK_STATUS c_datablock_get_double(c_datablock *s, const char *section, const char *name, double *val);
DATABLOCK_STATUS c_datablock_get_string(c_datablock *s, const char *section, const char *name, char **val);
DATABLOCK_STATUS c_datablock_get_double_array_1d(c_datablock *s, const char *section, const char *name,
                                                 double **val, int *size);
DATABLOCK_STATUS c_datablock_get_double_array_shape(c_datablock *s, const char *section, const char *name, int ndims,
                                                    int *extents);
DATABLOCK_STATUS c_datablock_get_double_array(c_datablock *s, const char *section, const char *name, double *val,
                                              int ndims, int *extents);
DATABLOCK_STATUS c_datablock_put_double_array_1d(c_datablock *s, const char *section, const char *name,
                                                 const double *val, int size);
}

namespace {

constexpr int kModuleError = 1;  // status returned for failures outside the datablock
constexpr int kDefaultZBins = 256;  // cuda_version/bispectrum.cuh:35 (n_redshift_bins)

struct ModuleConfig {
    int zBins = 0;
    double zMax = 0;
    std::vector<double> nz;
    std::vector<double> thetas;  // N x 3 [rad], the reference's triple order
    bool output = false;
    std::string output_dir;
    double reltol = 1e-4;
    int mode = TP_MODE_FP64;
    tp_ctx *ctx = nullptr;
};

std::string take_string(char *s)
{
    std::string r = s ? s : "";
    std::free(s);
    return r;
}

bool report(DATABLOCK_STATUS st, const char *what)
{
    if (st != 0) std::fprintf(stderr, "threepoint: datablock error %d reading %s\n", st, what);
    return st == 0;
}

}  // namespace

extern "C" {

// classed_version/interface.cpp:69-147
void *setup(c_datablock *options)
{
    const char *sec = "threepoint";
    ModuleConfig *cfg = new ModuleConfig();
    char *s = nullptr;
    bool ok = report(c_datablock_get_string(options, sec, "nz_file", &s), "nz_file");
    const std::string fn_nz = take_string(s);
    // zBins is optional: without it, the cuda_version's fixed bin count
    if (c_datablock_get_int(options, sec, "zBins", &cfg->zBins) != 0) cfg->zBins = kDefaultZBins;
    ok = ok && report(c_datablock_get_double(options, sec, "zMax", &cfg->zMax), "zMax");
    s = nullptr;
    ok = ok && report(c_datablock_get_string(options, sec, "theta_unit", &s), "theta_unit");
    const std::string unit = take_string(s);
    double tmin = 0, tmax = 0;
    int nbins = 0;
    bool diag = false;
    ok = ok && report(c_datablock_get_double(options, sec, "theta_min", &tmin), "theta_min");
    ok = ok && report(c_datablock_get_double(options, sec, "theta_max", &tmax), "theta_max");
    ok = ok && report(c_datablock_get_int(options, sec, "theta_bins", &nbins), "theta_bins");
    ok = ok && report(c_datablock_get_bool(options, sec, "diag_only", &diag), "diag_only");
    ok = ok && report(c_datablock_get_bool(options, sec, "output_Map3", &cfg->output), "output_Map3");
    s = nullptr;
    ok = ok && report(c_datablock_get_string(options, sec, "output_dir", &s), "output_dir");
    cfg->output_dir = take_string(s);
    if (!ok || nbins < 1 || cfg->zBins < 2) {
        std::fprintf(stderr, "threepoint: incomplete or invalid options\n");
        delete cfg;
        return nullptr;
    }
    // optional settings of this implementation
    double rt = 0;
    if (c_datablock_get_double(options, sec, "reltol", &rt) == 0 && rt > 0) cfg->reltol = rt;
    s = nullptr;
    if (c_datablock_get_string(options, sec, "precision", &s) == 0 && take_string(s) == "fp32")
        cfg->mode = TP_MODE_FP32;
    try {
        read_n_of_z(fn_nz, cfg->zBins, cfg->zMax, cfg->nz);
        tmin = convert_angle_to_rad(tmin, unit);
        tmax = convert_angle_to_rad(tmax, unit);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "threepoint: %s\n", e.what());
        delete cfg;
        return nullptr;
    }
    // aperture radii, log-spaced; all i <= j <= k or the diagonal only (interface.cpp:116-141)
    const double dl = nbins > 1 ? (std::log(tmax) - std::log(tmin)) / (nbins - 1) : 0.0;
    for (int i = 0; i < nbins; ++i) {
        const double t1 = tmin * std::exp(i * dl);
        if (diag) {
            cfg->thetas.insert(cfg->thetas.end(), {t1, t1, t1});
            continue;
        }
        for (int j = i; j < nbins; ++j) {
            const double t2 = tmin * std::exp(j * dl);
            for (int k = j; k < nbins; ++k) cfg->thetas.insert(cfg->thetas.end(), {t1, t2, tmin * std::exp(k * dl)});
        }
    }
    if (tp_ctx_create(0, &cfg->ctx) != 0) {
        std::fprintf(stderr, "threepoint: %s\n", tp_last_error());
        delete cfg;
        return nullptr;
    }
    return cfg;
}

// classed_version/interface.cpp:157-247
DATABLOCK_STATUS execute(c_datablock *block, void *config)
{
    ModuleConfig *cfg = (ModuleConfig *)config;
    if (!cfg) return kModuleError;
    const char *cp = "cosmological_parameters";
    double om = 0, s8 = 0, h = 0, omb = 0, ns = 0, w = 0;
    DATABLOCK_STATUS st;
    if ((st = c_datablock_get_double(block, cp, "omega_m", &om)) || (st = c_datablock_get_double(block, cp, "sigma_8", &s8)) ||
        (st = c_datablock_get_double(block, cp, "h0", &h)) || (st = c_datablock_get_double(block, cp, "omega_b", &omb)) ||
        (st = c_datablock_get_double(block, cp, "n_s", &ns)) || (st = c_datablock_get_double(block, cp, "w", &w))) {
        report(st, "cosmological_parameters");
        return st;
    }
    // linear P(k): k_h and the first row (z = 0) of p_k (interface.cpp:176-194)
    double *k = nullptr;
    int nk = 0;
    if ((st = c_datablock_get_double_array_1d(block, "matter_power_lin", "k_h", &k, &nk))) {
        report(st, "matter_power_lin/k_h");
        return st;
    }
    std::vector<double> kh(k, k + nk);
    std::free(k);
    int ext[2] = {0, 0};
    if ((st = c_datablock_get_double_array_shape(block, "matter_power_lin", "p_k", 2, ext))) {
        report(st, "matter_power_lin/p_k shape");
        return st;
    }
    if (ext[1] != nk || ext[0] < 1) {
        std::fprintf(stderr, "threepoint: p_k has shape %d x %d, k_h %d entries\n", ext[0], ext[1], nk);
        return kModuleError;
    }
    std::vector<double> pk((size_t)ext[0] * ext[1]);
    if ((st = c_datablock_get_double_array(block, "matter_power_lin", "p_k", pk.data(), 2, ext))) {
        report(st, "matter_power_lin/p_k");
        return st;
    }
    tp_cosmology p{h, s8, omb, ns, w, om, cfg->zMax};
    tp_cosmo *tab = nullptr;
    if (tp_cosmo_build_pk(cfg->ctx, 1, &p, cfg->nz.data(), cfg->zBins, 0, cfg->zMax, kh.data(), pk.data(), nk, 0,
                          &tab) != 0) {
        std::fprintf(stderr, "threepoint: %s\n", tp_last_error());
        return kModuleError;
    }
    const size_t n = cfg->thetas.size() / 3;
    std::vector<double> map3(n), err(n);
    std::vector<int> ic(n, 0);
    int rc = tp_map3_batch(cfg->ctx, tab, n, ic.data(), cfg->thetas.data(), cfg->reltol, cfg->mode, map3.data(),
                           err.data(), nullptr);
    tp_cosmo_destroy(tab);
    if (rc != 0) {
        std::fprintf(stderr, "threepoint: %s\n", tp_last_error());
        return kModuleError;
    }
    if (cfg->output) {  // interface.cpp:202-236
        const std::string fn = cfg->output_dir + "/om_" + std::to_string(om) + "_sig8_" + std::to_string(s8) + "_h_" +
                               std::to_string(h) + "_w_" + std::to_string(w) + ".dat";
        std::ofstream out(fn);
        if (!out.is_open()) {
            std::fprintf(stderr, "threepoint: could not open %s\n", fn.c_str());
            return kModuleError;
        }
        out << "#theta1[rad] theta2[rad] theta3[rad] Map^3" << std::endl;
        for (size_t i = 0; i < n; ++i)
            out << cfg->thetas[3 * i] << " " << cfg->thetas[3 * i + 1] << " " << cfg->thetas[3 * i + 2] << " "
                << map3[i] << std::endl;
    }
    return c_datablock_put_double_array_1d(block, "threepoint", "Map3s", map3.data(), (int)n);
}

// classed_version/interface.cpp:254-258
int cleanup(void *config)
{
    ModuleConfig *cfg = (ModuleConfig *)config;
    if (!cfg) return 0;
    if (cfg->ctx) tp_ctx_destroy(cfg->ctx);
    delete cfg;
    return 0;
}

}  // extern "C"
