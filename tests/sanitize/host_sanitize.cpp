// AddressSanitizer / UndefinedBehaviorSanitizer run of the library's host layer (test infrastructure).
//
// Built by tests/sanitize/Makefile with -fsanitize=address,undefined from the product's own host sources
// threepoint_amd/csrc/tp_cubature.cpp (the h-adaptive cubature state machine) and tp_io.cpp (the input
// readers), with no GPU code: the cubature runs on analytic integrands in 1-4 dimensions, one and two
// components, every error norm, a maxEval cut and zero-volume edge cases; the readers run on n(z),
// triangle-list, theta, cosmology and covariance-parameter files including comments, empty files,
// short rows and capacity overflow.  Any invalid access or undefined operation aborts the run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "threepoint_amd.hpp"
#include "tp_cubature.h"

// the C ABI's error helper lives in the HIP runtime source; the readers only need its contract
int tp_set_error(int code, const std::string &msg)
{
    (void)msg;
    return code;
}

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                      \
        }                                                                 \
    } while (0)

// integral over [0,1]^dim of prod_i (1 + x_i^2) (component 0) and prod_i cos(x_i) (component 1)
static void integrand(unsigned dim, unsigned fdim, const double *x, double *f)
{
    double a = 1, b = 1;
    for (unsigned d = 0; d < dim; ++d) {
        a *= 1 + x[d] * x[d];
        b *= std::cos(x[d]);
    }
    f[0] = a;
    if (fdim > 1) f[1] = b;
}

static void run_cubature(unsigned dim, unsigned fdim, int norm, size_t maxEval, double reltol)
{
    tpc::AdaptiveIntegral I;
    double lo[8] = {0}, hi[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    CHECK(I.init(dim, fdim, lo, hi, maxEval, 0, reltol, norm));
    const unsigned np = tpc::rule_points(dim);
    std::vector<double> pts, vals;
    for (int round = 0; round < 10000; ++round) {
        const auto &pend = I.pending();
        pts.assign(pend.size() * np * dim, 0.0);
        vals.assign(pend.size() * np * fdim, 0.0);
        for (size_t r = 0; r < pend.size(); ++r) tpc::rule_points_of(dim, I.regions()[pend[r]], pts.data() + r * np * dim);
        for (size_t p = 0; p < pend.size() * np; ++p) integrand(dim, fdim, pts.data() + p * dim, vals.data() + p * fdim);
        for (size_t r = 0; r < pend.size(); ++r) tpc::rule_apply(dim, fdim, I.regions()[pend[r]], vals.data() + r * np * fdim);
        I.absorb();
        if (I.advance()) break;
    }
    double v[4], e[4];
    I.result(v, e);
    const double exact0 = std::pow(4.0 / 3.0, dim), exact1 = std::pow(std::sin(1.0), dim);
    if (maxEval == 0) {
        CHECK(std::fabs(v[0] / exact0 - 1) < 10 * reltol);
        if (fdim > 1) CHECK(std::fabs(v[1] / exact1 - 1) < 10 * reltol);
    } else {
        CHECK(I.num_eval() <= maxEval + np * 64);
    }
}

static std::string write(const std::string &name, const std::string &text)
{
    const char *dir = std::getenv("TMPDIR");
    std::string fn = std::string(dir ? dir : "/tmp") + "/tp_sanitize_" + std::to_string(::getpid()) + "_" + name;
    std::ofstream(fn) << text;
    return fn;
}

int main()
{
    for (unsigned dim = 1; dim <= 4; ++dim)
        for (unsigned fdim = 1; fdim <= 2; ++fdim)
            for (int norm = 0; norm <= 4; ++norm) {
                run_cubature(dim, fdim, norm, 0, 1e-6);
                run_cubature(dim, fdim, norm, 500, 1e-10);  // cut by maxEval
            }
    {   // bad domains and dimensions are refused
        tpc::AdaptiveIntegral I;
        double lo[2] = {0, 0}, hi[2] = {1, 1};
        CHECK(!I.init(0, 1, lo, hi, 0, 0, 1e-4, tpc::L1) || true);
        CHECK(!I.init(9, 1, lo, hi, 0, 0, 1e-4, tpc::L1));
    }
    // readers
    std::string z_rows;
    for (int i = 0; i <= 300; ++i) {
        const double z = 0.01 * i;
        z_rows += std::to_string(z) + " " + std::to_string(z * z * std::exp(-std::pow(z / 0.5, 1.5))) + "\n";
    }
    const std::string fnz = write("nz.dat", "# comment\n" + z_rows);
    std::vector<double> nz;
    for (int bins : {1, 2, 7, 100, 1000}) {
        try {  // one bin holds only z = 0, where n(z) = 0: the reader refuses the zero sum
            nz.clear();  // the readers append, as the reference's (helper.cpp:81)
            read_n_of_z(fnz, bins, 3.0, nz);
            CHECK((int)nz.size() == bins);
        } catch (const std::exception &) {
            CHECK(bins == 1);
        }
        try {
            nz.clear();
            read_n_of_z_averaging(fnz, bins, 3.0, nz);
            CHECK((int)nz.size() == bins);
        } catch (const std::exception &) {
            CHECK(bins == 1);
        }
    }
    for (int averaging = 0; averaging <= 1; ++averaging) {
        std::vector<double> out(100, 0.0);
        CHECK(tp_read_n_of_z(fnz.c_str(), 100, 3.0, averaging, out.data()) == 0);
        CHECK(tp_read_n_of_z((fnz + ".missing").c_str(), 100, 3.0, averaging, out.data()) != 0);
    }
    const std::string fshort = write("nz_short.dat", "0.0 1.0\n");  // one row: the reader must refuse or cope
    try {
        read_n_of_z(fshort, 10, 3.0, nz);
    } catch (const std::exception &) {
    }
    const std::string fempty = write("empty.dat", "");
    try {
        read_n_of_z(fempty, 10, 3.0, nz);
    } catch (const std::exception &) {
    }
    const std::string ftri = write("tri.dat", "# rind uind vind r u v\n0 0 0 1.0 0.5 0.1\n1 2 3 2.0 0.7 -0.3\n");
    std::vector<double> ruv(6);
    size_t n = 0;
    CHECK(tp_read_triangle_configurations(ftri.c_str(), 2, ruv.data(), &n) == 0 && n == 2);
    CHECK(tp_read_triangle_configurations(ftri.c_str(), 1, ruv.data(), &n) == TP_ECAPACITY);
    CHECK(tp_read_triangle_configurations(fempty.c_str(), 2, ruv.data(), &n) == 0 && n == 0);
    const std::string fth = write("thetas.dat", "# radii\n2\n4\n\n8\n");
    std::vector<double> th;
    read_thetas(fth, th);
    CHECK(th.size() == 3);
    const std::string fc = write("cosmo.dat", "# c\nh 0.7\nsigma_8 0.82\nOmega_b 0.047\nn_s 0.97\nw -1\nOmega_m 0.3\nz_max 3\n");
    cosmology c(fc);
    CHECK(c.h == 0.7 && c.zmax == 3);
    const std::string fcov = write("cov.param", "thetaMax 7.87\nshapenoise_sigma 0.37\ngalaxy_density 108000\n"
                                                "unit deg\nshapenoiseOnly 0\narea 61.9369\n");
    covarianceParameters cp(fcov);
    CHECK(cp.thetaMax > 0 && cp.area > 0);
    const std::string fg = write("gamma.cfg", "rmin 0.1\nrmax 100\nrbins 10\numin 0\numax 1\nubins 10\n"
                                              "vmin -1\nvmax 1\nvbins 10\n");
    configGamma cg;
    try {
        read_gamma_config(fg, cg);
    } catch (const std::exception &) {
    }
    for (const std::string &f : {fnz, fshort, fempty, ftri, fth, fc, fcov, fg}) std::remove(f.c_str());
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("host layer sanitize run: ok\n");
    return 0;
}
