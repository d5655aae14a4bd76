// AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU FP64 oracle (test infrastructure).
//
// Built by tests/sanitize/Makefile with -fsanitize=address,undefined from oracle/tp_oracle.cpp: the
// reference's table construction (f_K, growth, r_sigma, n_eff, the n(z) lookup at and beyond its last
// sample, where classed_version/bispectrum.cpp:929-932 reads element [n]), B_kappa and the <Map^3> integrand
// at squeezed and zero-side triangles, one <Map^3> cubature, the Gamma integrands on a coarse Ogata grid, the
// power spectra and <Map^2>, the covariance integrands and one small T1 and T2, the tabulated-P(k) form, and
// the readers.  Any invalid access or undefined operation aborts the run.
#include <cmath>
#include <cstdio>
#include <vector>

#include "tp_oracle.h"

static int fails = 0;
#define CHECK(c)                                                                \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                            \
        }                                                                       \
    } while (0)

int main()
{
    const int n_z = 100;
    std::vector<double> nz(n_z);
    double s = 0;
    for (int i = 0; i < n_z; ++i) {  // z^2 exp(-(z/0.5)^1.5) on z_i = 0.03 i, normalised
        const double z = 0.03 * i;
        nz[i] = z * z * std::exp(-std::pow(z / 0.5, 1.5));
        s += nz[i] * 0.03;
    }
    for (double &v : nz) v /= s;
    const double p[7] = {0.7, 0.82, 0.047, 0.97, -1.0, 0.3, 3.0};  // h s8 omb ns w om zmax
    void *h = orc_create(p, nz.data(), n_z, 3.0);
    std::vector<double> t(6 * n_z);
    double norm = 0;
    orc_get_tables(h, t.data(), t.data() + n_z, t.data() + 2 * n_z, t.data() + 3 * n_z, t.data() + 4 * n_z,
                   t.data() + 5 * n_z, &norm);
    CHECK(norm > 0 && t[3 * n_z + 50] > 0);
    for (double z : {0.0, 1e-6, 0.03, 1.5, 2.97, 2.999999, 3.0})
        for (double k : {1e-4, 0.1, 10.0, 1e3}) {
            CHECK(std::isfinite(orc_bispec(h, k, k, k, z)));
            CHECK(std::isfinite(orc_pk_nonlinear(h, k, z)));
        }
    CHECK(std::isfinite(orc_bispec_idx(h, 0.1, 0.2, 0.25, 15.0, 98, 1.0)));  // tree-level branch
    const double ells[][3] = {{100, 100, 100}, {1e4, 1e4, 1.0}, {3, 5e4, 5e4}, {0, 10, 10}, {10, 10, 0}};
    for (const auto &l : ells) CHECK(std::isfinite(orc_bkappa(h, l[0], l[1], l[2])));
    const double th[3] = {4 * 2.9088820866e-4, 8 * 2.9088820866e-4, 16 * 2.9088820866e-4};
    CHECK(orc_map3_integrand(h, 100.0, 100.0, M_PI, th) == 0.0);
    double v = 0, e = 0;
    long long ne = 0;
    CHECK(orc_map3(h, th, 1e-2, 2, &v, &e, &ne) == 0 && v > 0);
    CHECK(orc_gamma_init(h, 0.1, 2.0) == 0);
    double re = 0, im = 0;
    for (int which = 0; which < 2; ++which)
        for (double z : {0.0, 0.5, 2.9}) {
            CHECK(orc_gamma_integrand(h, which, z, 1.0, 0.7, 1e-3, 1.2e-3, 1.5e-3, &re, &im) == 0);
            CHECK(std::isfinite(re) && std::isfinite(im));
        }
    orc_ggg_analytic(1e-3, 1.2e-3, 1.5e-3, &re, &im);
    for (int flags : {0, 1, 2, 4}) {
        CHECK(std::isfinite(orc_limber_pkappa(h, 300.0, 0.5, flags)));
        CHECK(std::isfinite(orc_pell(h, 300.0, 0.3, 1e8, flags)));
    }
    CHECK(orc_map2(h, 4 * 2.9088820866e-4, 1e-3, 2, 0.3, 30.0, &v, &e, &ne) == 0);
    const double th6[6] = {th[0], th[1], th[1], th[0], th[0], th[2]};
    CHECK(std::isfinite(orc_cov_gauss_integrand(h, 300.0, 400.0, 1.0, 0.5, th6)));
    CHECK(orc_t1(h, th6, 1e-3, 0.0, 0.3, 1e8, 4, 1e-2, 2, &v, &e, &ne) == 0);
    long long n2 = 0;
    CHECK(orc_t2(h, th6, 0, 0.02, 0.0, 0.0, 0.3, 1e8, 4, 1e-3, 2, &v, &n2) == 0);
    // a tabulated linear P(k)
    std::vector<double> kk(60), pk(60);
    for (int i = 0; i < 60; ++i) {
        kk[i] = std::pow(10.0, -4 + 6.0 * i / 59);
        pk[i] = orc_linear_pk(h, kk[i]);
    }
    void *hp = orc_create_pk(p, nz.data(), n_z, 3.0, kk.data(), pk.data(), 60);
    CHECK(std::isfinite(orc_bkappa(hp, 100, 200, 250)));
    CHECK(std::isfinite(orc_linear_pk(hp, 1e-6)) && std::isfinite(orc_linear_pk(hp, 1e6)));
    orc_destroy(hp);
    orc_set_analytic(h, 1);
    CHECK(std::isfinite(orc_bkappa(h, 100, 200, 250)));
    orc_destroy(h);
    for (int s_ = 1; s_ <= 50; ++s_) CHECK(orc_bessel_zero_jn(6, s_) > 0);
    for (double x : {0.0, 1.0, 20.0, 100.0}) CHECK(std::isfinite(orc_bessel_jn(7, x)));
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("oracle sanitize run: ok\n");
    return 0;
}
