// calculateSecondOrderApertureStatistics.x -- drop-in for
// cuda_version/calculateSecondOrderAperturestatistics.cu:19-141.
//
//   calculateSecondOrderApertureStatistics.x <cosmology file> <covariance parameter file>
//                                            <thetas file [arcmin]> <output file> <n(z) file>
//                                            [--nz-bins N] [--reltol X] [--no-t17]
//
// Writes "theta Map2" lines (theta in arcmin) for every aperture radius of the list: <Map^2> of one
// source bin with the shape noise (shapenoise_sigma, galaxy_density) of the covariance parameter file,
// 2-D cubature over (ell, z) with relative tolerance 1e-5 (apertureStatistics.cu:110-136), the T17
// factors of the cuda_version unless --no-t17.  All radii run in one batched GPU pass.  The
// reference's "argc < 5" check (which accepted a missing n(z) argument) is "argc < 6" here.
#include <fstream>
#include <iostream>
#include <vector>

#include "threepoint_amd.hpp"

int main(int argc, char *argv[])
{
    const char *usage =
        "calculateSecondOrderApertureStatistics.x <cosmology> <covariance parameters> <thetas [arcmin]> <output>\n"
        "                                         <n(z)> [--nz-bins N] [--reltol X] [--no-t17]\n";
    if (argc < 6) {
        std::cerr << usage;
        return 1;
    }
    int nbins = 100;
    double reltol = 1e-5;
    bool t17 = true;
    for (int i = 6; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--nz-bins" && i + 1 < argc) nbins = std::atoi(argv[++i]);
        else if (a == "--reltol" && i + 1 < argc) reltol = std::atof(argv[++i]);
        else if (a == "--no-t17") t17 = false;
        else {
            std::cerr << usage;
            return 1;
        }
    }
    try {
        covarianceParameters covPar(argv[2]);
        if (covPar.shapenoiseOnly) std::cerr << "WARNING: Uses shape noise only powerspectrum" << std::endl;
        cosmology cosmo(argv[1]);
        std::vector<double> nz;
        read_n_of_z(argv[5], nbins, cosmo.zmax, nz);
        std::ofstream out(argv[4]);
        if (!out.is_open()) {
            std::cerr << "Couldn't open " << argv[4] << std::endl;
            return 1;
        }
        std::vector<double> thetas;
        read_thetas(argv[3], thetas);
        std::cerr << "Using cosmology from " << argv[1] << ":" << std::endl << cosmo;
        std::cerr << "Covariance from " << argv[2] << ":" << std::endl << covPar;
        const double sigma = covPar.shapenoise_sigma;
        if (sigma != 0) std::cerr << "Calculating Power Spectrum WITH shapenoise, sigma=" << sigma << std::endl;
        else std::cerr << "Calculating Power Spectrum WITHOUT shapenoise" << std::endl;
        BispectrumCalculator bk(&cosmo, nz, nbins, cosmo.zmax);
        ApertureStatistics ap(&bk);
        ap.reltol_map2 = reltol;
        std::vector<double> th;
        for (double t : thetas) th.push_back(convert_angle_to_rad(t));
        // the noise term takes the density per arcmin^2 (apertureStatistics.cu:43-46)
        const double arcmin = convert_angle_to_rad(1.0);
        const double ngal_arcmin2 = covPar.galaxy_density * arcmin * arcmin;
        if (sigma != 0 && !(ngal_arcmin2 > 0)) {
            std::cerr << "galaxy_density must be positive with shape noise" << std::endl;
            return 1;
        }
        std::vector<double> m2 = ap.Map2(th, sigma, ngal_arcmin2, t17);
        // shapenoiseOnly: the reference only warns (its Map2 kernel ignores the switch); here the output
        // is then the shape-noise term alone
        if (covPar.shapenoiseOnly) {
            std::vector<double> sig = ap.Map2(th, 0.0, 1.0, t17);
            for (size_t i = 0; i < m2.size(); ++i) m2[i] -= sig[i];
        }
        for (size_t i = 0; i < thetas.size(); ++i) out << thetas[i] << " " << m2[i] << " " << std::endl;
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
