// calculateApertureStatistics.x -- drop-in for cuda_version/calculateApertureStatistics.cu:20-133.
//
//   calculateApertureStatistics.x <cosmology file> <thetas file [arcmin]> <output file> <n(z) file>
//                                 [--nz-bins N] [--nz-averaging] [--reltol X] [--mode fp64|fp32|surrogate]
//                                 [--surrogate FILE] [--single] [--chunk N] [--resume]
//
// Computes <Map^3>(theta1, theta2, theta3) for all theta1 <= theta2 <= theta3 of the list and writes
// "theta1 theta2 theta3 Map3" lines in arcmin (the reference's output format).  All triples run in one
// batched GPU pass unless --single asks for the reference's one-triple-at-a-time loop; --chunk N runs them
// in batches of N triples, each written (and flushed) before the next starts.  --resume keeps the complete
// lines an interrupted run left in the output file (their radii must be the list's first triples) and
// computes the rest.  The reference's "argc < 4" check (which accepted a missing n(z) argument) is
// "argc < 5" here.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "app_resume.h"
#include "threepoint_amd.hpp"

int main(int argc, char *argv[])
{
    const char *usage =
        "calculateApertureStatistics.x <cosmology> <thetas [arcmin]> <output> <n(z)> [--nz-bins N] [--nz-averaging]\n"
        "                              [--reltol X]\n"
        "                              [--mode fp64|fp32|surrogate] [--surrogate FILE] [--single] [--chunk N]\n"
        "                              [--resume]\n";
    if (argc < 5) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100;
    double reltol = 1e-4;
    int mode = TP_MODE_FP64;
    bool single = false, nz_avg = false, resume = false;
    size_t chunk = 0;
    std::string surrogate;
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--nz-bins" && i + 1 < argc) nbins = std::atoi(argv[++i]);
        else if (a == "--nz-averaging") nz_avg = true;
        else if (a == "--reltol" && i + 1 < argc) reltol = std::atof(argv[++i]);
        else if (a == "--mode" && i + 1 < argc) {
            std::string m = argv[++i];
            mode = m == "fp32" ? TP_MODE_FP32 : (m == "surrogate" ? TP_MODE_SURROGATE : TP_MODE_FP64);
        } else if (a == "--surrogate" && i + 1 < argc) surrogate = argv[++i];
        else if (a == "--single") single = true;
        else if (a == "--resume") resume = true;
        else if (a == "--chunk" && i + 1 < argc) chunk = (size_t)std::atoll(argv[++i]);
        else {
            std::cerr << usage;
            return 1;
        }
    }
    try {
        cosmology cosmo(argv[1]);
        std::vector<double> nz;
        if (nz_avg) read_n_of_z_averaging(argv[4], nbins, cosmo.zmax, nz);  // cuda_version/helpers.cu:9-127
        else read_n_of_z(argv[4], nbins, cosmo.zmax, nz);
        std::vector<double> thetas;
        read_thetas(argv[2], thetas);
        std::cerr << "Using cosmology from " << argv[1] << ":" << std::endl << cosmo;
        std::cerr << "Using thetas in " << argv[2] << std::endl << "Writing to:" << argv[3] << std::endl;
        BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
        ApertureStatistics ap(&bk);
        ap.reltol = reltol;
        ap.mode = mode;
        if (mode == TP_MODE_SURROGATE) {
            if (surrogate.empty() || tp_surrogate_load(bk.ctx(), surrogate.c_str())) {
                std::cerr << "surrogate mode needs --surrogate FILE: " << tp_last_error() << std::endl;
                return 1;
            }
        }
        const size_t N = thetas.size();
        std::vector<std::array<double, 3>> tri;
        std::vector<std::array<double, 3>> tri_arcmin;
        for (size_t i = 0; i < N; i++)
            for (size_t j = i; j < N; j++)
                for (size_t k = j; k < N; k++) {
                    tri.push_back({convert_angle_to_rad(thetas[i]), convert_angle_to_rad(thetas[j]),
                                   convert_angle_to_rad(thetas[k])});
                    tri_arcmin.push_back({thetas[i], thetas[j], thetas[k]});
                }
        // --resume: the complete lines already written, checked against the list's first triples
        size_t done = 0;
        std::vector<std::string> old;
        if (resume) {
            old = tp_app::complete_lines(argv[3]);
            for (; done < old.size() && done < tri.size(); ++done) {
                const auto v = tp_app::leading_numbers(old[done], 4);
                if (v.size() < 4) break;
                for (int d = 0; d < 3; ++d)
                    if (std::fabs(v[d] - tri_arcmin[done][d]) > 1e-5 * std::fabs(tri_arcmin[done][d]))
                        throw std::runtime_error("--resume: line " + std::to_string(done + 1) + " of " + argv[3] +
                                                 " does not hold the triple of this list");
            }
            std::cerr << "resuming after " << done << " of " << tri.size() << " triples" << std::endl;
        }
        std::ofstream out = tp_app::keep_and_append(argv[3], old, done);
        if (!out.is_open()) {
            std::cerr << "Couldn't open " << argv[3] << std::endl;
            return 1;
        }
        auto write = [&](size_t s, double v) {
            out << tri_arcmin[s][0] << " " << tri_arcmin[s][1] << " " << tri_arcmin[s][2] << " " << v << " "
                << std::endl;
        };
        auto t0 = std::chrono::steady_clock::now();
        if (single) {
            for (size_t s = done; s < tri.size(); ++s) {
                std::cerr << s + 1 << "/" << tri.size() << ": Thetas:" << tri_arcmin[s][0] << " " << tri_arcmin[s][1]
                          << " " << tri_arcmin[s][2] << " \r";
                write(s, ap.MapMapMap(std::vector<double>(tri[s].begin(), tri[s].end())));
            }
        } else {
            const size_t step = chunk ? chunk : tri.size();
            for (size_t s0 = done; s0 < tri.size(); s0 += step) {
                const size_t s1 = std::min(tri.size(), s0 + step);
                std::vector<std::array<double, 3>> part(tri.begin() + s0, tri.begin() + s1);
                const std::vector<double> m3 = ap.MapMapMap(part);
                for (size_t s = s0; s < s1; ++s) write(s, m3[s - s0]);
                if (chunk) std::cerr << s1 << "/" << tri.size() << " triples written" << std::endl;
            }
        }
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cerr << std::endl << tri.size() - done << " triples in " << dt << " s" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
