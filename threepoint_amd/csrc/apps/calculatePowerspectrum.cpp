// calculatePowerspectrum.x -- drop-in for cuda_version/calculatePowerspectrum.cu:17-101.
//
//   calculatePowerspectrum.x <cosmology file> <output file> <n(z) file> [--nz-bins N] [--no-t17]
//
// Writes "ell P_kappa(ell)" for ell = 1, 1.05, 1.05^2, ... < 5e4 (the reference's hard-coded grid):
// the Limber-integrated revised-Halofit power spectrum, integrated adaptively over z with relative
// tolerance 1e-6 (the host Pell of cuda_version/bispectrum.cu:511-529), without shape noise.  The
// cuda_version applies its T17 correction (bispectrum.cuh:4); --no-t17 gives classed_version's
// convergence_power_spectrum instead.  All multipoles run in one batched GPU pass.
#include <cmath>
#include <fstream>
#include <iostream>
#include <vector>

#include "threepoint_amd.hpp"

int main(int argc, char *argv[])
{
    const char *usage = "calculatePowerspectrum.x <cosmology> <output> <n(z)> [--nz-bins N] [--no-t17]\n";
    if (argc < 4) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100;
    int flags = TP_PK_T17;
    for (int i = 4; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--nz-bins" && i + 1 < argc) nbins = std::atoi(argv[++i]);
        else if (a == "--no-t17") flags = 0;
        else {
            std::cerr << usage;
            return 1;
        }
    }
    try {
        cosmology cosmo(argv[1]);
        std::vector<double> nz;
        read_n_of_z(argv[3], nbins, cosmo.zmax, nz);
        std::ofstream out(argv[2]);
        if (!out.is_open()) {
            std::cerr << "Couldn't open " << argv[2] << std::endl;
            return 1;
        }
        std::cerr << "Using cosmology from " << argv[1] << ":" << std::endl << cosmo;
        std::cerr << "Writing to:" << argv[2] << std::endl;
        BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
        std::vector<double> ell;
        for (double l = 1; l < 5. * std::pow(10, 4); l *= 1.05) ell.push_back(l);  // calculatePowerspectrum.cu:90
        std::vector<double> pk(ell.size()), err(ell.size());
        if (tp_pkappa_batch(bk.ctx(), bk.tables(), ell.size(), nullptr, ell.data(), 1e-6, flags, pk.data(),
                            err.data(), nullptr)) {
            std::cerr << tp_last_error() << std::endl;
            return 1;
        }
        for (size_t i = 0; i < ell.size(); ++i) out << ell[i] << " " << pk[i] << " " << std::endl;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
