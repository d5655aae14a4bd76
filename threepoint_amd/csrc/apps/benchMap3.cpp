// benchMap3.x -- the batched <Map^3> engine driven from C++ through the C ABI, for profilers.
//
//   benchMap3.x <cosmologies> <triples> <n(z) file> <fp64|fp32|surrogate> <steps> [--model FILE]
//               [--reltol X] [--nz-bins N] [--warmup N] [--inference fp16|hilo] [--copy-probe MIB]
//
// --copy-probe MIB first copies MIB MiB device to device with the library's plain copy kernel (k_copy16):
// a kernel of known traffic, whose HBM counters calibrate those of the other kernels in the same run.
// <cosmologies>: one cosmology per line, "h sigma8 Omega_b n_s w Omega_m z_max" ('#' comments);
// <triples>: one integral per line, "theta1 theta2 theta3 [arcmin] icosmo"; the n(z) file is read as
// the classed_version reads it (helper.cpp:31-96).  Every step builds all cosmologies' FP64 tables on the
// GPU and integrates every triple in one tp_map3_batch call (as bench.py's steps do); the program prints one
// line per step and a summary.  No counterpart in the reference: it runs the same workload as bench.py
// without a Python interpreter in the process, so a profiler attached to it sees only this library.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "threepoint_amd.h"

static std::vector<std::vector<double>> read_rows(const char *fn, size_t ncol)
{
    std::ifstream in(fn);
    if (!in) {
        std::fprintf(stderr, "benchMap3: cannot open %s\n", fn);
        std::exit(1);
    }
    std::vector<std::vector<double>> rows;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::vector<double> r(ncol);
        for (size_t k = 0; k < ncol; ++k) ss >> r[k];
        if (!ss.fail()) rows.push_back(r);
    }
    return rows;
}

static void check(int rc, const char *what)
{
    if (rc) {
        std::fprintf(stderr, "benchMap3: %s failed (%d): %s\n", what, rc, tp_last_error());
        std::exit(1);
    }
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: benchMap3.x <cosmologies> <triples> <n(z)> <fp64|fp32|surrogate> <steps> "
                             "[--model FILE] [--reltol X] [--nz-bins N] [--warmup N] [--inference fp16|hilo] "
                             "[--copy-probe MIB]\n");
        return 1;
    }
    const std::string mode_s = argv[4];
    const int mode = mode_s == "fp64" ? TP_MODE_FP64 : mode_s == "fp32" ? TP_MODE_FP32
                   : mode_s == "surrogate" ? TP_MODE_SURROGATE : -1;
    if (mode < 0) {
        std::fprintf(stderr, "benchMap3: mode must be fp64, fp32 or surrogate\n");
        return 1;
    }
    const int steps = std::atoi(argv[5]);
    std::string model;
    double reltol = 1e-4;
    int nbins = 100, warmup = 1, inference = TP_SUR_INFER_FP16;
    long long copy_mib = 0;
    for (int i = 6; i + 1 < argc; i += 2) {
        const std::string a = argv[i];
        if (a == "--model") model = argv[i + 1];
        else if (a == "--reltol") reltol = std::atof(argv[i + 1]);
        else if (a == "--nz-bins") nbins = std::atoi(argv[i + 1]);
        else if (a == "--warmup") warmup = std::atoi(argv[i + 1]);
        else if (a == "--copy-probe") copy_mib = std::atoll(argv[i + 1]);
        else if (a == "--inference") {
            const std::string k = argv[i + 1];
            if (k != "fp16" && k != "hilo") {
                std::fprintf(stderr, "benchMap3: --inference must be fp16 or hilo\n");
                return 1;
            }
            inference = k == "fp16" ? TP_SUR_INFER_FP16 : TP_SUR_INFER_HILO_BF16;
        } else {
            std::fprintf(stderr, "benchMap3: unknown option %s\n", a.c_str());
            return 1;
        }
    }
    const auto cos_rows = read_rows(argv[1], 7), tri_rows = read_rows(argv[2], 4);
    if (cos_rows.empty() || tri_rows.empty()) {
        std::fprintf(stderr, "benchMap3: no cosmologies or no triples\n");
        return 1;
    }
    std::vector<tp_cosmology> cp(cos_rows.size());
    for (size_t i = 0; i < cp.size(); ++i) {
        const auto &r = cos_rows[i];
        cp[i] = tp_cosmology{r[0], r[1], r[2], r[3], r[4], r[5], r[6]};
    }
    const double zmax = cp[0].zmax;
    std::vector<double> nz(nbins);
    check(tp_read_n_of_z(argv[3], nbins, zmax, 0, nz.data()), "tp_read_n_of_z");
    const double arcmin = 2.9088820866e-4;  // classed_version/helper.cpp:139
    const size_t n = tri_rows.size();
    std::vector<double> th(3 * n), val(n), err(n);
    std::vector<int> ic(n);
    std::vector<long long> ne(n);
    for (size_t t = 0; t < n; ++t) {
        for (int d = 0; d < 3; ++d) th[3 * t + d] = tri_rows[t][d] * arcmin;
        ic[t] = (int)tri_rows[t][3];
        if (ic[t] < 0 || ic[t] >= (int)cp.size()) {
            std::fprintf(stderr, "benchMap3: triple %zu names cosmology %d of %zu\n", t, ic[t], cp.size());
            return 1;
        }
    }
    tp_ctx *ctx = nullptr;
    check(tp_ctx_create(0, &ctx), "tp_ctx_create");
    if (mode == TP_MODE_SURROGATE) {
        check(tp_surrogate_load(ctx, model.c_str()), "tp_surrogate_load");
        check(tp_surrogate_inference(ctx, inference), "tp_surrogate_inference");
    }
    if (copy_mib > 0) {
        double gbps = 0;
        check(tp_hbm_copy_probe(ctx, (size_t)copy_mib << 20, &gbps), "tp_hbm_copy_probe");
        std::printf("copy probe: %lld MiB device to device at %.0f GB/s\n", copy_mib, gbps);
        std::fflush(stdout);
    }
    using clk = std::chrono::steady_clock;
    double total_ms = 0, tables_ms = 0;
    long long points = 0;
    for (int s = -warmup; s < steps; ++s) {
        auto t0 = clk::now();
        tp_cosmo *tab = nullptr;
        check(tp_cosmo_build(ctx, (int)cp.size(), cp.data(), nz.data(), nbins, 0, zmax, &tab), "tp_cosmo_build");
        auto t1 = clk::now();
        check(tp_map3_batch(ctx, tab, n, ic.data(), th.data(), reltol, mode, val.data(), err.data(), ne.data()),
              "tp_map3_batch");
        check(tp_cosmo_destroy(tab), "tp_cosmo_destroy");
        const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        const double tms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        tp_batch_stats st{};
        check(tp_last_batch_stats(ctx, &st), "tp_last_batch_stats");
        std::printf("step %d: %.2f ms (tables %.2f ms), %lld points, %lld rounds, %lld launches\n", s, ms, tms,
                    st.points, st.rounds, st.launches);
        std::fflush(stdout);
        if (s >= 0) {
            total_ms += ms;
            tables_ms += tms;
            points += st.points;
        }
    }
    if (steps > 0)
        std::printf("benchMap3 %s: %zu integrals, %.2f ms per step (tables %.2f ms), %.1f triples/s, %.3e points per "
                    "step\n", mode_s.c_str(), n, total_ms / steps, tables_ms / steps, n * steps / (total_ms * 1e-3),
                    (double)points / steps);
    tp_ctx_destroy(ctx);
    return 0;
}
