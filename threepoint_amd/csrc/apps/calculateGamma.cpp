// calculateGamma.x -- drop-in for cuda_version/calculateGamma.cu:21-152.
//
//   calculateGamma.x <cosmology file> <gamma config file> <output file> <n(z) file> [device]
//                    [--nz-bins N] [--nz-averaging] [--prec-h H] [--prec-k K] [--reltol X]
//                    [--triangles FILE] [--fp32] [--chunk N] [--resume]
//
// Gamma_0..Gamma_3 (centroid convention) on the (r, u, v) grid of the config file, r2 = r pi/180/60,
// r3 = r2 u, r1 = v r3 + r2 (the TreeCorr binning of cuda_version/calculateGamma.cu:107-121), one
// output line per triangle: Re/Im Gamma0 ... Re/Im Gamma3, r, u, v.  All triangles of one r bin are
// evaluated in one batched GPU pass and written (and flushed) before the next r bin starts (--chunk N: in
// batches of N triangles, each written before the next starts); --resume keeps the complete r bins an
// interrupted run left in the output file and computes the rest.
//
// --triangles FILE: the classed_version's mode (calculateGamma.cpp:64-160): the (r, u, v) bins of a
// TreeCorr triangle list ("rind uind vind r u v" rows, read_triangle_configurations), sides converted
// with convert_angle_to_rad, bins with a zero side left at 0; rows "rind uind vind Re/Im Gamma0..3
// r u v" in %e format.  The config argument is not read then.  --nz-averaging reads n(z) with the
// cuda_version's reader (overlap-weighted bin averages for coarse files).  --fp32 selects the FP32 form
// of the integrand (FP32 Ogata terms inside their wavenumber window, FP64 accumulation).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <iostream>
#include <vector>

#include "app_resume.h"
#include "threepoint_amd.hpp"

int main(int argc, char **argv)
{
    const char *usage = "calculateGamma.x <cosmology> <gamma config> <output> <n(z)> [device] [--nz-bins N]\n"
                        "                 [--nz-averaging] [--prec-h H] [--prec-k K] [--reltol X] [--triangles FILE]\n"
                        "                 [--fp32] [--chunk N] [--resume]\n";
    if (argc < 5) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100, device = 0;
    double prec_h = 1e-3, prec_k = 5.0, reltol = 1e-3;
    bool nz_avg = false, fp32 = false, resume = false;
    size_t chunk = 0;
    std::string trifile;
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--nz-bins" && i + 1 < argc) nbins = std::atoi(argv[++i]);
        else if (a == "--nz-averaging") nz_avg = true;
        else if (a == "--fp32") fp32 = true;
        else if (a == "--resume") resume = true;
        else if (a == "--chunk" && i + 1 < argc) chunk = (size_t)std::atoll(argv[++i]);
        else if (a == "--triangles" && i + 1 < argc) trifile = argv[++i];
        else if (a == "--prec-h" && i + 1 < argc) prec_h = std::atof(argv[++i]);
        else if (a == "--prec-k" && i + 1 < argc) prec_k = std::atof(argv[++i]);
        else if (a == "--reltol" && i + 1 < argc) reltol = std::atof(argv[++i]);
        else if (i == 5 && a.size() && a[0] != '-') device = std::atoi(argv[i]);
        else {
            std::cerr << usage;
            return 1;
        }
    }
    try {
        if (device != 0) threepoint_default_context(device);
        cosmology cosmo(argv[1]);
        std::vector<double> nz;
        if (nz_avg) read_n_of_z_averaging(argv[4], nbins, cosmo.zmax, nz);
        else read_n_of_z(argv[4], nbins, cosmo.zmax, nz);
        if (!trifile.empty()) {
            if (resume) throw std::runtime_error("--resume applies to the (r, u, v) grid, not to --triangles");
            std::ofstream out(argv[3]);
            if (!out.is_open()) {
                std::cerr << "Couldn't open " << argv[3] << std::endl;
                return 1;
            }
            std::vector<treecorr_bin> bins;
            std::vector<std::array<int, 3>> idx;
            {  // the row indices travel with the bins
                std::ifstream fin(trifile);
                if (!fin.is_open()) throw std::runtime_error("Could not open " + trifile);
                std::string line;
                while (std::getline(fin, line)) {
                    if (line.empty() || line[0] == '#') continue;
                    std::istringstream iss(line);
                    std::array<int, 3> ix;
                    treecorr_bin b;
                    if (iss >> ix[0] >> ix[1] >> ix[2] >> b.r >> b.u >> b.v) {
                        bins.push_back(b);
                        idx.push_back(ix);
                    }
                }
            }
            BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
            GammaCalculator gc(&bk, prec_h, prec_k, "centroid");
            gc.reltol = reltol;
            if (fp32) gc.mode = TP_MODE_FP32;
            std::vector<std::array<double, 3>> x;
            std::vector<size_t> which;
            for (size_t t = 0; t < bins.size(); ++t) {
                double r2 = bins[t].r, r3 = r2 * bins[t].u, r1 = bins[t].v * r3 + r2;  // calculateGamma.cpp:101-103
                if (r1 != 0 && r2 != 0 && r3 != 0) {
                    x.push_back({convert_angle_to_rad(r1), convert_angle_to_rad(r2), convert_angle_to_rad(r3)});
                    which.push_back(t);
                }
            }
            std::vector<std::array<std::complex<double>, 4>> g(bins.size());
            auto gs = gc.gammas(x);
            for (size_t q = 0; q < which.size(); ++q) g[which[q]] = gs[q];
            char buf[512];
            for (size_t t = 0; t < bins.size(); ++t) {
                std::snprintf(buf, sizeof(buf), "%d %d %d %e %e %e %e %e %e %e %e %e %e %e \n", idx[t][0], idx[t][1],
                              idx[t][2], g[t][0].real(), g[t][0].imag(), g[t][1].real(), g[t][1].imag(), g[t][2].real(),
                              g[t][2].imag(), g[t][3].real(), g[t][3].imag(), bins[t].r, bins[t].u, bins[t].v);
                out << buf;
            }
            std::cerr << bins.size() << " triangle bins done" << std::endl;
            return 0;
        }
        configGamma config;
        read_gamma_config(argv[2], config);
        std::cerr << "Using cosmology:" << std::endl << cosmo << "Writing to:" << argv[3] << std::endl << config;
        BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
        GammaCalculator gc(&bk, prec_h, prec_k, "centroid");
        gc.reltol = reltol;
        if (fp32) gc.mode = TP_MODE_FP32;
        double lrmin = std::log(config.rmin), lrmax = std::log(config.rmax);
        auto rbin = [&](int i) { return std::exp(lrmin + (lrmax - lrmin) / config.rsteps * (i + 0.5)); };
        // --resume: the r bins the output file holds in full (each line ends with r u v)
        const size_t per_bin = (size_t)config.usteps * config.vsteps;
        int first = 0;
        std::vector<std::string> old;
        if (resume) {
            old = tp_app::complete_lines(argv[3]);
            size_t ok = 0;
            for (; ok < old.size() && ok / per_bin < (size_t)config.rsteps; ++ok) {
                const auto v = tp_app::leading_numbers(old[ok], 9);
                const double r = rbin((int)(ok / per_bin));
                if (v.size() < 9) break;
                if (std::fabs(v[8] - r) > 1e-5 * r)
                    throw std::runtime_error("--resume: line " + std::to_string(ok + 1) + " of " + argv[3] +
                                             " does not belong to this grid");
            }
            first = (int)(ok / per_bin);
            std::cerr << "resuming after " << first << " of " << config.rsteps << " r bins" << std::endl;
        }
        std::ofstream out = tp_app::keep_and_append(argv[3], old, (size_t)first * per_bin);
        if (!out.is_open()) {
            std::cerr << "Couldn't open " << argv[3] << std::endl;
            return 1;
        }
        auto t0 = std::chrono::steady_clock::now();
        for (int i = first; i < config.rsteps; i++) {
            double r = rbin(i);
            std::vector<std::array<double, 3>> x;
            std::vector<std::array<double, 2>> uv;
            for (int j = 0; j < config.usteps; j++) {
                double u = config.umin + (config.umax - config.umin) / config.usteps * (j + 0.5);
                for (int k = 0; k < config.vsteps; k++) {
                    double v = config.vmin + (config.vmax - config.vmin) / config.vsteps * (k + 0.5);
                    double r2 = r * M_PI / 180. / 60.;
                    double r3 = r2 * u;
                    double r1 = v * r3 + r2;
                    x.push_back({r1, r2, r3});
                    uv.push_back({u, v});
                }
            }
            const size_t step = chunk ? chunk : x.size();
            for (size_t s0 = 0; s0 < x.size(); s0 += step) {
                const size_t s1 = std::min(x.size(), s0 + step);
                auto g = gc.gammas(std::vector<std::array<double, 3>>(x.begin() + s0, x.begin() + s1));
                for (size_t t = s0; t < s1; ++t) {
                    const auto &gt = g[t - s0];
                    out << gt[0].real() << " " << gt[0].imag() << " " << gt[1].real() << " " << gt[1].imag() << " "
                        << gt[2].real() << " " << gt[2].imag() << " " << gt[3].real() << " " << gt[3].imag() << " " << r
                        << " " << uv[t][0] << " " << uv[t][1] << std::endl;
                }
                if (chunk && s1 < x.size())
                    std::cerr << "r bin " << i + 1 << ": " << s1 << "/" << x.size() << " triangles" << std::endl;
            }
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cerr << "r bin " << i + 1 << "/" << config.rsteps << " done, " << el << " s" << std::endl;
        }
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
