// tp_io.cpp -- input readers of the drop-in API (file formats of the reference).
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>

#include "threepoint_amd.hpp"
#include "tp_internal.h"

cosmology::cosmology(const std::string &fn_parameters)
{
    // classed_version/cosmology.cpp:10-73: "name value" lines, '#' comments
    std::ifstream input(fn_parameters.c_str());
    if (input.fail()) throw std::runtime_error("cosmology: Could not open " + fn_parameters);
    std::string line;
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string name;
        double value = 0;
        iss >> name >> value;
        if (name == "h") h = value;
        else if (name == "sigma_8") sigma8 = value;
        else if (name == "Omega_b") omb = value;
        else if (name == "n_s") ns = value;
        else if (name == "w") w = value;
        else if (name == "Omega_m") om = value;
        else if (name == "z_max") zmax = value;
        else throw std::runtime_error("cosmology: Parameter file in wrong format");
    }
    ow = 1 - om;
    omc = om - omb;
}

std::ostream &operator<<(std::ostream &out, const cosmology &c)
{
    out << "h:" << c.h << std::endl << "sigma_8:" << c.sigma8 << std::endl << "Omega_b:" << c.omb << std::endl
        << "n_s:" << c.ns << std::endl << "w:" << c.w << std::endl << "Omega_m:" << c.om << std::endl
        << "z_max:" << c.zmax << std::endl;
    return out;
}

void read_n_of_z(const std::string &fn, const int &n_bins, const double &zMax, std::vector<double> &nz)
{
    // classed_version/helper.cpp:31-96: linear resampling to z_i = i zMax / n_bins, sum n dz = 1
    std::ifstream input(fn.c_str());
    if (input.fail()) throw std::runtime_error("read_n_of_z: Could not open " + fn);
    std::vector<double> zs, ns;
    std::string line;
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double z, v;
        iss >> z >> v;
        zs.push_back(z);
        ns.push_back(v);
    }
    if (zs.size() < 2) throw std::runtime_error("read_n_of_z: need at least two rows in " + fn);
    int nf = (int)zs.size();
    double zmin_f = zs.front(), zmax_f = zs.back();
    double dz_f = (zmax_f - zmin_f) / (nf - 1);
    double dz = zMax / n_bins;
    for (int i = 0; i < n_bins; i++) {
        double z = i * dz;
        int ix = int((z - zmin_f) / dz_f);
        double dix = (z - zmin_f) / dz_f - ix;
        double v = 0;
        if (ix >= 0 && ix < nf - 1) v = ns.at(ix + 1) * dix + ns.at(ix) * (1 - dix);
        if (ix == nf - 1) v = ns.at(nf - 1);
        nz.push_back(v);
    }
    double norm = std::accumulate(nz.end() - n_bins, nz.end(), 0.0);
    if (norm == 0) throw std::runtime_error("sum of n(z) is zero! Check " + fn);
    norm *= dz;
    for (int i = 0; i < n_bins; i++) nz.at(nz.size() - n_bins + i) /= norm;
}

static void read_z_n_rows(const std::string &fn, std::vector<double> &zs, std::vector<double> &ns)
{
    std::ifstream input(fn.c_str());
    if (input.fail()) throw std::runtime_error("read_n_of_z: Could not open " + fn);
    std::string line;
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double z, v;
        iss >> z >> v;
        zs.push_back(z);
        ns.push_back(v);
    }
    if (zs.size() < 2) throw std::runtime_error("read_n_of_z: need at least two rows in " + fn);
}

void read_n_of_z_averaging(const std::string &fn, const int &n_bins, const double &zMax, std::vector<double> &nz)
{
    // cuda_version/helpers.cu:9-127.  Internal bins coarser than the file's spacing: linear
    // interpolation at z_i = i zMax / n_bins, 0 from the file's last interval on; otherwise each
    // internal bin [z_low, z_up) takes the file samples that overlap it, weighted by their overlap.
    std::vector<double> zs, ns;
    read_z_n_rows(fn, zs, ns);
    const int nf = (int)zs.size();
    const double zmin_f = zs.front(), zmax_f = zs.back();
    const double dz_f = (zmax_f - zmin_f) / (nf - 1);
    const double dz = zMax / n_bins;
    const size_t first = nz.size();
    if (dz > dz_f) {
        for (int i = 0; i < n_bins; i++) {
            double z = i * dz;
            int ix = int((z - zmin_f) / dz_f);
            double dix = (z - zmin_f) / dz_f - ix;
            double v = 0;
            if (ix >= 0 && ix < nf - 1) v = ns.at(ix + 1) * dix + ns.at(ix) * (1 - dix);
            if (ix >= nf - 1) v = 0.0;
            nz.push_back(v);
        }
    } else {
        for (int i = 0; i < n_bins; i++) {
            double z_low = i * dz, z_up = (i + 1) * dz;
            int k = 0, l = 0;  // last file sample below z_low, below z_up
            for (int j = 0; j < nf; j++) {
                if (zs.at(j) < z_low) k = j;
                if (zs.at(j) < z_up) l = j;
            }
            double v = 0;
            if (l - k > 1)
                std::cerr << "WARNING: Redshift spacing of n(z) file smaller than intern spacing. Increase number "
                             "of redshift bins inside."
                          << std::endl;
            if (z_up < zmin_f || k > nf) {
                v = 0.0;
            } else if (l - k <= 1) {
                if (k == l) v = ns.at(k);
                else if (l > nf) v = ns.at(k) * (zs.at(l) - z_low) / dz;
                else if (z_low < zmin_f) v = ns.at(l) * (z_up - zs.at(l)) / dz;
                else v = ns.at(k) * (zs.at(l) - z_low) / dz + ns.at(l) * (z_up - zs.at(l)) / dz;
            }
            nz.push_back(v);
        }
    }
    double norm = std::accumulate(nz.begin() + first, nz.end(), 0.0);
    if (norm == 0) throw std::runtime_error("sum of n(z) is zero! Check " + fn);
    norm *= dz;
    for (int i = 0; i < n_bins; i++) nz.at(first + i) /= norm;
}

int read_triangle_configurations(std::string &infile, std::vector<treecorr_bin> &triangle_config)
{
    // classed_version/helper.cpp:5-29: "rind uind vind r u v" rows.  The reference's eof loop appends
    // one more bin built from a failed read after the last row; only the file's rows are kept here.
    std::ifstream fin(infile);
    if (!fin.is_open()) throw std::runtime_error("read_triangle_configurations: Could not open " + infile);
    std::string line;
    while (std::getline(fin, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        int ri, ui, vi;
        treecorr_bin b;
        if (iss >> ri >> ui >> vi >> b.r >> b.u >> b.v) triangle_config.push_back(b);
    }
    return 1;
}

extern "C" int tp_read_n_of_z(const char *fn, int n_bins, double z_max, int averaging, double *out)
{
    if (!fn || !out || n_bins < 1 || !(z_max > 0)) return tp_set_error(TP_EINVAL, "tp_read_n_of_z: bad arguments");
    try {
        std::vector<double> nz;
        if (averaging) read_n_of_z_averaging(fn, n_bins, z_max, nz);
        else read_n_of_z(fn, n_bins, z_max, nz);
        std::copy(nz.begin(), nz.end(), out);
    } catch (const std::exception &e) {
        return tp_set_error(TP_EINVAL, e.what());
    }
    return 0;
}

extern "C" int tp_read_triangle_configurations(const char *fn, size_t cap, double *ruv, size_t *n)
{
    if (!fn || !n || (cap && !ruv)) return tp_set_error(TP_EINVAL, "tp_read_triangle_configurations: bad arguments");
    try {
        std::string f(fn);
        std::vector<treecorr_bin> t;
        read_triangle_configurations(f, t);
        *n = t.size();
        if (t.size() > cap) return tp_set_error(TP_ECAPACITY, "tp_read_triangle_configurations: more rows than cap");
        for (size_t i = 0; i < t.size(); ++i) {
            ruv[3 * i] = t[i].r;
            ruv[3 * i + 1] = t[i].u;
            ruv[3 * i + 2] = t[i].v;
        }
    } catch (const std::exception &e) {
        return tp_set_error(TP_EINVAL, e.what());
    }
    return 0;
}

void read_thetas(const std::string &fn, std::vector<double> &thetas)
{
    // classed_version/helper.cpp:178-200: one theta per line, '#' comments
    std::ifstream input(fn.c_str());
    if (input.fail()) throw std::runtime_error("read_thetas: Couldn't open " + fn);
    std::string line;
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double t;
        if (iss >> t) thetas.push_back(t);
    }
}

double convert_angle_to_rad(const double &value, const std::string &unit)
{
    // classed_version/helper.cpp:132-155 (note: 2.9088820866e-4, not pi/10800)
    if (unit == "arcmin") return 2.9088820866e-4 * value;
    if (unit == "deg") return 0.017453 * value;
    if (unit == "rad") return value;
    throw std::runtime_error("Unit not correctly specified. Needs to be arcmin, deg, or rad.");
}

double convert_rad_to_angle(const double &value, const std::string &unit)
{
    // classed_version/helper.cpp:157-176
    if (unit == "arcmin") return 3437.74677 * value;
    if (unit == "deg") return 57.3 * value;
    throw std::runtime_error("Unit not correctly specified. Needs to be arcmin or deg.");
}

void read_gamma_config(const std::string &fn, configGamma &config)
{
    // cuda_version/helpers.cu:242-310: "name value" lines
    std::ifstream input(fn.c_str());
    if (input.fail()) throw std::runtime_error("read_gamma_config: Could not open " + fn);
    std::string line;
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string name;
        double v = 0;
        iss >> name >> v;
        if (name == "rmin") config.rmin = v;
        else if (name == "rmax") config.rmax = v;
        else if (name == "umin") config.umin = v;
        else if (name == "umax") config.umax = v;
        else if (name == "vmin") config.vmin = v;
        else if (name == "vmax") config.vmax = v;
        else if (name == "rsteps") config.rsteps = (int)std::round(v);
        else if (name == "usteps") config.usteps = (int)std::round(v);
        else if (name == "vsteps") config.vsteps = (int)std::round(v);
        else throw std::runtime_error("read_gamma_config: unknown key " + name);
    }
}

std::ostream &operator<<(std::ostream &out, const configGamma &c)
{
    out << "Binning r in " << c.rsteps << " bins from " << c.rmin << " to " << c.rmax << std::endl;
    out << "Binning u in " << c.usteps << " bins from " << c.umin << " to " << c.umax << std::endl;
    out << "Binning v in " << c.vsteps << " bins from " << c.vmin << " to " << c.vmax << std::endl;
    out << "Setting the triangle center as Centroid" << std::endl;
    return out;
}

covarianceParameters::covarianceParameters(const std::string &fn)
{
    // cuda_version/cosmology.cu:92-170: "name value" lines, '#' comments; lengths and densities in `unit`
    std::ifstream input(fn.c_str());
    if (input.fail()) throw std::runtime_error("covarianceParameters: Could not open " + fn);
    std::string line, unit;  // no default: a file without "unit" is rejected, as by the reference
    while (std::getline(input, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string name, value;
        iss >> name >> value;
        if (name == "thetaMax") thetaMax = std::stod(value);
        else if (name == "thetaMax_smaller") thetaMax_smaller = std::stod(value);
        else if (name == "shapenoise_sigma") shapenoise_sigma = std::stod(value);
        else if (name == "galaxy_density") galaxy_density = std::stod(value);
        else if (name == "unit") unit = value;
        else if (name == "shapenoiseOnly") shapenoiseOnly = std::stoi(value) != 0;
        else if (name == "area") area = std::stod(value);
        else throw std::runtime_error("covarianceParameters: parameter file is not in the right format (" + name + ")");
    }
    // to rad and rad^-2 (cosmology.cu:160-169)
    const double r1 = convert_angle_to_rad(1, unit);
    thetaMax = convert_angle_to_rad(thetaMax, unit);
    thetaMax_smaller = convert_angle_to_rad(thetaMax_smaller, unit);
    area = area * r1 * r1;
    galaxy_density = galaxy_density / r1 / r1;
}

std::ostream &operator<<(std::ostream &out, const covarianceParameters &c)
{
    // cuda_version/cosmology.cu:172-179
    out << "# thetaMax and galaxy density in rad and rad^-2" << std::endl;
    out << "thetaMax " << c.thetaMax << std::endl;
    out << "galaxy density " << c.galaxy_density << std::endl;
    out << "shapenoise " << c.shapenoise_sigma << std::endl;
    return out;
}
