// calculateApertureStatistics.x -- drop-in for cuda_version/calculateApertureStatistics.cu:20-133.
//
//   calculateApertureStatistics.x <cosmology file> <thetas file [arcmin]> <output file> <n(z) file>
//                                 [--nz-bins N] [--nz-averaging] [--reltol X] [--mode fp64|fp32|surrogate]
//                                 [--surrogate FILE] [--single]
//
// Computes <Map^3>(theta1, theta2, theta3) for all theta1 <= theta2 <= theta3 of the list and writes
// "theta1 theta2 theta3 Map3" lines in arcmin (the reference's output format).  All triples run in one
// batched GPU pass unless --single asks for the reference's one-triple-at-a-time loop.  The reference's
// "argc < 4" check (which accepted a missing n(z) argument) is "argc < 5" here.
#include <array>
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "threepoint_amd.hpp"

int main(int argc, char *argv[])
{
    const char *usage =
        "calculateApertureStatistics.x <cosmology> <thetas [arcmin]> <output> <n(z)> [--nz-bins N] [--nz-averaging]\n"
        "                              [--reltol X]\n"
        "                              [--mode fp64|fp32|surrogate] [--surrogate FILE] [--single]\n";
    if (argc < 5) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100;
    double reltol = 1e-4;
    int mode = TP_MODE_FP64;
    bool single = false, nz_avg = false;
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
        std::ofstream out(argv[3]);
        if (!out.is_open()) {
            std::cerr << "Couldn't open " << argv[3] << std::endl;
            return 1;
        }
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
        auto t0 = std::chrono::steady_clock::now();
        std::vector<double> m3;
        if (single) {
            for (size_t s = 0; s < tri.size(); ++s) {
                std::cerr << s + 1 << "/" << tri.size() << ": Thetas:" << tri_arcmin[s][0] << " " << tri_arcmin[s][1]
                          << " " << tri_arcmin[s][2] << " \r";
                m3.push_back(ap.MapMapMap(std::vector<double>(tri[s].begin(), tri[s].end())));
            }
        } else {
            m3 = ap.MapMapMap(tri);
        }
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cerr << std::endl << tri.size() << " triples in " << dt << " s" << std::endl;
        for (size_t s = 0; s < tri.size(); ++s)
            out << tri_arcmin[s][0] << " " << tri_arcmin[s][1] << " " << tri_arcmin[s][2] << " " << m3[s] << " "
                << std::endl;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
