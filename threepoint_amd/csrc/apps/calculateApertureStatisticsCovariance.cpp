// calculateApertureStatisticsCovariance.x -- drop-in for cuda_version/calculateApertureStatisticsCovariance.cu
// (its Gaussian terms).
//
//   calculateApertureStatisticsCovariance.x <cosmology> <thetas [arcmin]> <n(z)> <output folder>
//       <covariance parameters> <T1 0|1> <T2 0|1> <T4 0|1> <T5 0|1> <T6 0|1> <T7 0|1> <T7_2h 0|1> <geometry>
//       [--nz-bins N] [--reltol X] [--no-t17] [--fp32]
//
// The same 13 arguments as the reference.  Provided: T1 and T2 for every geometry (circle, square,
// infinite, rectangular; T2 is 0 for the infinite one, as in the reference; T1 of a finite one is the
// reference's 6-D integral at its tolerance 1e-2, --reltol sets the others); requesting T4-T7 is an
// error (exit status 1).  --fp32 (FP32 Pell node values) exists for T1 of the infinite survey only;
// with a finite geometry or with T2 it is an error (exit status 1), so no run silently writes FP64
// numbers under that flag.  For every pair i <= j of the
// theta_1 <= theta_2 <= theta_3 triples of the list it computes T1_total and T2_total, all pairs of a
// term in one batched GPU pass, and writes each symmetric N x N matrix to
// <folder>cov_<geometry>_term<k>Numerical_sigma_<s>_n_<n>_thetaMax_<t>_gpu.dat as the reference does
// (apertureStatisticsCovariance.cu:46-86; the folder string is prefixed as given).
#include <array>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "threepoint_amd.hpp"

static void write_cov(const std::vector<double> &v, int N, const std::string &fn)
{
    if ((int)v.size() != N * (N + 1) / 2) throw std::out_of_range("writeCov: Values has wrong length");
    std::ofstream out(fn);
    if (!out.is_open()) throw std::runtime_error("writeCov: Could not open " + fn);
    std::vector<double> m((size_t)N * N);
    int ix = 0;
    for (int i = 0; i < N; i++)
        for (int j = i; j < N; j++) m[(size_t)i * N + j] = v[ix++];
    for (int i = 0; i < N; i++)
        for (int j = 0; j < i; j++) m[(size_t)i * N + j] = m[(size_t)j * N + i];
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) out << m[(size_t)i * N + j] << " ";
        out << std::endl;
    }
}

int main(int argc, char *argv[])
{
    const char *usage =
        "calculateApertureStatisticsCovariance.x <cosmology> <thetas [arcmin]> <n(z)> <output folder> <covariance "
        "parameters>\n    <T1> <T2> <T4> <T5> <T6> <T7> <T7_2h> <circle|square|infinite|rectangular> [--nz-bins N] "
        "[--reltol X] [--no-t17] [--fp32]\n"
        "    --fp32: T1 of the infinite geometry only (FP32 Pell node values, FP64 sums)\n";
    if (argc < 14) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100;
    double reltol = 1e-4;
    bool fp32 = false;  // T1 with FP32 Pell node values (FP64 sums), within 1e-3 of FP64
    bool t17 = true;
    for (int i = 14; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--nz-bins" && i + 1 < argc) nbins = std::atoi(argv[++i]);
        else if (a == "--reltol" && i + 1 < argc) reltol = std::atof(argv[++i]);
        else if (a == "--no-t17") t17 = false;
        else if (a == "--fp32") fp32 = true;
        else {
            std::cerr << usage;
            return 1;
        }
    }
    const bool do_t1 = std::stoi(argv[6]) != 0, do_t2 = std::stoi(argv[7]) != 0;
    for (int k = 8; k <= 12; ++k)
        if (std::stoi(argv[k]) != 0) {
            std::cerr << "terms T4-T7 (halo-model spectra) are not provided by this build" << std::endl;
            return 1;
        }
    const std::string type = argv[13];
    if (type != "circle" && type != "square" && type != "infinite" && type != "rectangular") {
        std::cerr << "Cov type not correctly specified" << std::endl;
        return 1;
    }
    if (fp32 && (type != "infinite" || do_t2)) {
        std::cerr << "--fp32 applies to T1 of the infinite geometry only (requested: geometry " << type
                  << (do_t2 ? ", T2" : "") << ")" << std::endl;
        return 1;
    }

    try {
        cosmology cosmo(argv[1]);
        covarianceParameters covPar(argv[5]);
        if (covPar.shapenoiseOnly) std::cerr << "WARNING: Uses constant powerspectrum" << std::endl;
        std::vector<double> nz;
        read_n_of_z(argv[3], nbins, cosmo.zmax, nz);
        std::vector<double> thetas;
        read_thetas(argv[2], thetas);
        BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
        ApertureStatisticsCovariance cov(&bk, covPar, type, t17);
        cov.reltol = reltol;
        if (fp32) cov.mode = TP_MODE_FP32;
        std::vector<std::array<double, 3>> tri;
        const size_t N = thetas.size();
        for (size_t i = 0; i < N; i++)
            for (size_t j = i; j < N; j++)
                for (size_t k = j; k < N; k++)
                    tri.push_back({convert_angle_to_rad(thetas[i]), convert_angle_to_rad(thetas[j]),
                                   convert_angle_to_rad(thetas[k])});
        const int N_ind = (int)tri.size();
        std::vector<std::array<double, 6>> pairs;
        for (int i = 0; i < N_ind; i++)
            for (int j = i; j < N_ind; j++)
                pairs.push_back({tri[i][0], tri[i][1], tri[i][2], tri[j][0], tri[j][1], tri[j][2]});
        std::cerr << "Using survey geometry " << type << ", " << pairs.size() << " pairs" << std::endl;
        const double n_deg = covPar.galaxy_density / convert_rad_to_angle(1, "deg") / convert_rad_to_angle(1, "deg");
        const double thetaMax_deg = convert_rad_to_angle(covPar.thetaMax, "deg");
        const std::string folder = argv[4];
        char fn[512];
        if (do_t1) {
            std::vector<double> t1 = cov.T1_total(pairs);
            std::snprintf(fn, sizeof(fn), "cov_%s_term1Numerical_sigma_%.2f_n_%.2f_thetaMax_%.2f_gpu.dat", type.c_str(),
                          covPar.shapenoise_sigma, n_deg, thetaMax_deg);
            std::cerr << "Writing Term1 to " << folder + fn << std::endl;
            write_cov(t1, N_ind, folder + fn);
        }
        if (do_t2) {
            std::vector<double> t2 = cov.T2_total(pairs);
            std::snprintf(fn, sizeof(fn), "cov_%s_term2Numerical_sigma_%.2f_n_%.2f_thetaMax_%.2f_gpu.dat", type.c_str(),
                          covPar.shapenoise_sigma, n_deg, thetaMax_deg);
            std::cerr << "Writing Term2 to " << folder + fn << std::endl;
            write_cov(t2, N_ind, folder + fn);
        }
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
