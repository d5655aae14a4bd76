#!/usr/bin/env python
"""<Map^3> throughput on one MI355X: plain FP64 path vs the surrogate form, with the batched exact forms and the
CPU FP64 oracle beside it.

Default workload (one "step"): 512 Latin-hypercube wCDM cosmologies (Omega_m 0.2-0.4, sigma8 0.7-0.95,
h 0.6-0.8, w -1.2 - -0.8, n_s 0.92-1.0, Omega_b 0.047) x the 35 triples of radii {2, 4, 8, 16, 32} arcmin =
17 920 integrals; n(z) ~ z^2 exp(-(z/0.5)^1.5) on [0, 3] (100 bins); each <Map^3> an h-adaptive cubature
over (l1, l2, phi) with relative tolerance 1e-4.  A step builds the 512 cosmologies' FP64 tables on the
GPU (timed) and then integrates.  --workload c1: the 220 triples of 10 log radii 0.5-32 arcmin of the
fiducial flat LCDM cosmology.  Synthetic inputs, no data files.

Paths timed in the same run:
  plain_fp64    one triple after another through tp_map3_single: the reference's hcubature_v driving an
                integrand callback that crosses to the GPU every cubature round
                (cuda_version/apertureStatistics.cu:155-229); on the C3 workload one cosmology's 35 triples.
  surrogate     batched; triples whose largest radius is >= 16x their smallest run the exact FP32 form, the
                others run it until an integral holds 1024 regions and then the learned B_kappa (FP16 MFMA
                network over an 8-node FP32 baseline) with a seeded 1% of points checked exactly.
  batched_fp32  all integrals advance together; FP32 bispectrum, FP64 z-sum and cubature accumulation.
  batched_fp64  the same in FP64: the reference's arithmetic, the parity anchor of the other paths.
The headline (value) is the surrogate form when its max relative error against batched FP64 (itself
within 1e-6 of the CPU FP64 oracle) is <= 1e-3 on every integral, else batched FP64; it runs --steps
timed steps, the comparison paths --compare-steps.  Errors are compared after all paths ran.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "fp16": 2.5e15, "hbm": 8.0e12}
# HBM bytes per integrand point of the region-cached kernels on the C3 step: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over benchMap3.x (profiles/map3_c3_counters.md), each counter scaled by its reading for
# k_copy16 copying 1 GiB (FETCH_SIZE x2.000, WRITE_SIZE x1.000).  Algorithmic: 64/33 B of region in +
# 8 B of integrand value out.
HBM_BYTES_PER_POINT = {"double": 12.92, "float": 11.66}
# The surrogate network 8 -> 4 x 256 -> 1: algorithmic flops per point 2 (8 x 256 + 3 x 256 x 256 + 256); the
# inference kernel k_mlp_h executes the three 256 x 256 layers as one FP16 MFMA per weight fragment and
# layer 1 on the FP32 MFMA, i.e. the algorithmic MFMA work.
MLP_FLOPS_PER_POINT = 2 * (8 * 256 + 3 * 256 * 256 + 256)
MLP_MFMA_FLOPS_EXECUTED = 2 * (8 * 256 + 3 * 256 * 256)
# Counter readings of the C3 step (profiles/map3_c3_counters.md): MFMA flops = SQ_INSTS_VALU_MFMA_MOPS_* x 512,
# VALU flops = SQ_INSTS_VALU_FLOPS_* x 64, MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 XCDs
# x 256 CUs x 4 SIMDs), each rate over the kernel's summed dispatch time in the counter run.
MLP_HBM_BYTES_PER_POINT = 43.5
MLP_COUNTERS = {"kernel": "k_mlp_h", "mfma_tflops": 877.8, "mfma_frac_of_fp16_peak": 0.351, "mfma_busy": 0.46,
                "valu_fp32_tflops": 10.3, "hbm_read_bytes_per_point": 35.5, "hbm_write_bytes_per_point": 8.0}
REGION_COUNTERS = {
    "float": {"kernel": "k_region_cached<float, 1>", "valu_fp32_tflops": 36.2, "frac_of_fp32_peak": 0.230,
              "frac_of_scalar_fp32": 0.460, "valu_fp32_flops_per_point": 22240, "trans_fp32_per_point": 2720,
              "hbm_read_bytes_per_point": 3.59, "hbm_write_bytes_per_point": 8.08},
    "double": {"kernel": "k_region_cached<double, 1>", "valu_fp64_tflops": 39.5, "frac_of_fp64_peak": 0.502,
               "valu_fp64_flops_per_point": 93400, "hbm_read_bytes_per_point": 4.85,
               "hbm_write_bytes_per_point": 8.08},
}
FLOPS_PER_ELEM = 20.0
CENSUS_FILE = os.path.join("tests", "golden", "census.json")


def census_fixture(tp):
    """The committed op census (tests/golden/census.json, written by tests/golden/make_golden.py from the
    library's op-counting instantiation of its device functions); the library's live census must agree."""
    fx = json.load(open(os.path.join(ROOT, CENSUS_FILE)))
    live = {"map3_point": tp.census_map3_point(), "power_point": tp.census_power_point(),
            "gamma_term": tp.census_gamma_term()}
    for k, d in live.items():
        for f, v in d.items():
            if float(v) != fx[k][f]:
                raise RuntimeError("op census of %s.%s (%g) differs from %s (%g)" % (k, f, v, CENSUS_FILE, fx[k][f]))
    if fx["flops_per_elementary_call"] != FLOPS_PER_ELEM:
        raise RuntimeError("%s uses another flops-per-elementary-call convention" % CENSUS_FILE)
    return fx


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_path(tabs, ctx, thetas, path, steps, warmup, reltol, icosmo=None, make_tables=None):
    """Times `steps` steps of one path after `warmup` untimed ones.  With make_tables, every step first
    builds the cosmologies' tables on the GPU (a new set of cosmologies, as a sampler step has) and the
    step's time includes that build."""
    import threepoint_amd as tp
    mode = {"plain_fp64": tp.MODE_FP64, "batched_fp64": tp.MODE_FP64, "batched_fp32": tp.MODE_FP32,
            "batched_fp64_graph": tp.MODE_FP64, "batched_fp32_graph": tp.MODE_FP32,
            "surrogate": tp.MODE_SURROGATE}[path]
    n = thetas.shape[0]
    # *_graph: every cubature round replayed as one HIP graph per group (tp_set_graph_rounds)
    tp._check(tp.lib().tp_set_graph_rounds(ctx.h, 1 if path.endswith("_graph") else 0), "tp_set_graph_rounds")

    def one_step():
        tb = tabs
        t_tab = 0.0
        if make_tables is not None:
            t_tab = time.perf_counter()
            tb = make_tables()
            t_tab = time.perf_counter() - t_tab
        if path == "plain_fp64":
            vals = np.zeros(n)
            pts = 0
            for i in range(n):
                v, e, ne = tb.map3_single(thetas[i], reltol, 0 if icosmo is None else int(icosmo[i]), mode)
                vals[i] = v
                pts += ne
            st = {"points": pts}
        else:
            vals, e, ne = tb.map3_batch(thetas, icosmo, reltol, mode)
            st = ctx.stats()
        st["tables_ms"] = 1e3 * t_tab
        if make_tables is not None:
            tb.close()
        return vals, st

    for _ in range(warmup):
        one_step()
    agg = {"points": 0, "integrand_ms": 0.0, "integrand_sum_ms": 0.0, "kernel_ms": 0.0, "launches": 0, "rounds": 0,
           "host_ms": 0.0, "wait_ms": 0.0, "graph_rebuilds": 0, "tables_ms": 0.0, "mlp_ms": 0.0, "mlp_launches": 0,
           "mlp_points": 0, "surrogate_points": 0, "exact_checks": 0, "fallback_batches": 0,
           "fallback_integrals": 0}
    t0 = time.perf_counter()
    vals = None
    for _ in range(steps):
        vals, st = one_step()
        for k in agg:
            if k in st:
                agg[k] += st[k]
    dt = time.perf_counter() - t0
    agg["steps"] = steps
    agg["warmup"] = warmup
    agg["ms_per_step"] = 1e3 * dt / steps
    agg["triples_per_s"] = n * steps / dt
    if path.startswith("batched_fp"):  # the launch shapes the tuner chose, by launch-size bucket
        import ctypes
        tab = (ctypes.c_int * 4)()
        tp._check(tp.lib().tp_region_lanes_table(ctx.h, mode, tab), "tp_region_lanes_table")
        agg["tuned_shapes"] = list(tab)
    return vals, agg


def cpu_baseline(cos, thetas, icosmo, seconds_target, reltol):
    """The CPU FP64 oracle (OpenMP over each integrand batch) on a bounded sample of the step's triples."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    from threepoint_amd import workloads as W
    cores = os.cpu_count() or 1
    cores = min(cores, 64)
    if icosmo is None:
        # every 11th of the 220 triples (20 spread over all radius combinations)
        order = list(range(0, thetas.shape[0], 11))
        what = "%d of the 220 triples (every 11th, in order)"
    else:
        # the labelled subsample: the first 8 cosmologies x 35 triples, every 7th (40 integrals)
        order = list(range(0, 8 * 35, 7))
        what = "%d of the 280 integrals of the first 8 Latin-hypercube cosmologies x 35 triples (every 7th)"
    oracles = {}
    done, t0 = 0, time.perf_counter()
    for i in order:
        ic = 0 if icosmo is None else int(icosmo[i])
        if ic not in oracles:  # the cosmology's CPU tables are part of the CPU work
            oracles[ic] = Oracle(cos[ic], W.fiducial_nz())
        oracles[ic].map3(thetas[i], reltol, cores)
        done += 1
        if time.perf_counter() - t0 > seconds_target:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "triples/s", "cores": cores, "kind": "port",
            "sample": (what + " through the FP64 CPU oracle incl. its table builds, OpenMP %d threads, %s, %.1f s")
            % (done, cores, cpu_model(), dt)}


def write_inputs(W, workload, out_dir):
    """The C1 or C3 workload as the text inputs of benchMap3.x: cosmologies.txt (h sigma8 Omega_b n_s w Omega_m
    z_max), triples.txt (theta1 theta2 theta3 [arcmin] icosmo) and nz.txt (the n(z) file rows)."""
    os.makedirs(out_dir, exist_ok=True)
    if workload == "c1":
        cos, tri = [W.fiducial()], W.triples(W.map3_c1_radii())
        rows = [(t, 0) for t in tri]
    else:
        cos, tri = W.lhs_c3(512), W.triples(W.map3_c3_radii())
        rows = [(t, i) for i in range(len(cos)) for t in tri]
    with open(os.path.join(out_dir, "cosmologies.txt"), "w") as f:
        for c in cos:
            f.write("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (c.h, c.sigma8, c.omb, c.ns, c.w, c.om, c.zmax))
    with open(os.path.join(out_dir, "triples.txt"), "w") as f:
        for t, i in rows:
            f.write("%.17g %.17g %.17g %d\n" % (t[0], t[1], t[2], i))
    z, n = W.nz_file_rows()
    with open(os.path.join(out_dir, "nz.txt"), "w") as f:
        for a, b in zip(z, n):
            f.write("%.17g %.17g\n" % (a, b))
    print("bench: wrote %d cosmologies and %d triples to %s" % (len(cos), len(rows), out_dir), file=sys.stderr)


def accuracy_study(tabs, ctx, thetas, icosmo, paths, reltol):
    """Each batched path at reltol against batched FP64 at reltol/10 (the converged value)."""
    import threepoint_amd as tp
    modes = {"batched_fp64": tp.MODE_FP64, "batched_fp32": tp.MODE_FP32, "surrogate": tp.MODE_SURROGATE}
    ref, _, nref = tabs.map3_batch(thetas, icosmo, reltol / 10, tp.MODE_FP64)
    nz = ref != 0
    vals = {}
    for p in paths:
        if p not in modes:
            continue
        t0 = time.perf_counter()
        v, _, n = tabs.map3_batch(thetas, icosmo, reltol, modes[p])
        dt = time.perf_counter() - t0
        vals[p] = v
        err = np.abs(v[nz] - ref[nz]) / np.abs(ref[nz])
        w = int(np.nonzero(nz)[0][np.argmax(err)])
        print(json.dumps({"path": p, "reltol": reltol, "vs": "batched_fp64 at reltol %g" % (reltol / 10),
                          "seconds": dt, "points": int(n.sum()), "ref_points": int(nref.sum()),
                          "max": float(err.max()), "p50": float(np.median(err)),
                          "p99": float(np.percentile(err, 99)), "n_over_1e-3": int((err > 1e-3).sum()),
                          "zeros_kept": bool(np.all(v[~nz] == 0)),
                          "worst_theta_arcmin": [float(t) for t in np.degrees(thetas[w]) * 60]}), flush=True)
    if "batched_fp64" in vals:
        for p, v in vals.items():
            if p != "batched_fp64":
                err = np.abs(v[nz] - vals["batched_fp64"][nz]) / np.abs(vals["batched_fp64"][nz])
                print(json.dumps({"path": p, "vs": "batched_fp64 at reltol %g" % reltol, "max": float(err.max()),
                                  "n_over_1e-3": int((err > 1e-3).sum())}), flush=True)


def c2_sample(W, n):
    """n of the 1000 triangles of the 10x10x10 TreeCorr grid: all when n >= 1000, else spread over the
    r bins, each with its own u and v bin."""
    tri = W.gamma_c2_triangles()
    if n >= tri.shape[0]:
        return np.arange(tri.shape[0]), tri
    sel = [100 * (i * 10 // n) + 10 * ((7 * i + 4) % 10) + (3 * i + 5) % 10 for i in range(n)]
    return np.array(sel), tri


def gamma_bench(tp, W, ctx, args):
    """Gamma_0..3 of TreeCorr-grid triangles, Ogata h = 1e-3 with 4999 Bessel-zero nodes, rel tol 1e-3."""
    t0 = time.perf_counter()
    tabs = tp.CosmologyTables([W.fiducial()], W.fiducial_nz(), 100, 3.0, ctx)
    tp.gamma_setup(1e-3, 5.0, ctx)
    setup_ms = 1e3 * (time.perf_counter() - t0)
    prec_k = 5000
    sel, tri = c2_sample(W, args.c2_triangles)
    x = tri[sel, :3]
    fx = census_fixture(tp)
    fe_point = tp.flop_equivalents_per_gamma_point(prec_k, FLOPS_PER_ELEM)
    results, values = {}, {}
    for p in [q for q in args.paths.split(",") if q in ("plain_fp64", "batched_fp64", "batched_fp32")]:
        print("bench: gamma path %s ..." % p, file=sys.stderr, flush=True)
        steps = 1
        agg = {"points": 0, "integrand_ms": 0.0, "launches": 0, "rounds": 0, "host_ms": 0.0}
        tt = time.perf_counter()
        for _ in range(steps):
            if p in ("batched_fp64", "batched_fp32"):
                g, ne = tabs.gamma_batch(x, 1e-3, mode=tp.MODE_FP32 if p == "batched_fp32" else tp.MODE_FP64)
                st = ctx.stats()
                for k in agg:
                    agg[k] += st.get(k, 0)
            else:  # one triangle (its four integrals) per call
                out = []
                for i in range(x.shape[0]):
                    gi, ne = tabs.gamma_batch(x[i:i + 1], 1e-3)
                    out.append(gi[0])
                    st = ctx.stats()
                    for k in agg:
                        agg[k] += st.get(k, 0)
                g = np.array(out)
        dt = time.perf_counter() - tt
        agg["ms_per_step"] = 1e3 * dt / steps
        agg["triangles_per_s"] = x.shape[0] * steps / dt
        agg["points_per_triangle"] = agg["points"] / (steps * x.shape[0])
        if agg["integrand_ms"] > 0:
            ach = fe_point * agg["points"] / (agg["integrand_ms"] * 1e-3)
            pk = PEAK["fp32"] if p == "batched_fp32" else PEAK["fp64"]  # mixed form: priced at the FP32 peak
            agg["integrand_kernel"] = {"achieved_flops": ach, "peak": pk, "frac": ach / pk,
                                       "avg_launch_ms": agg["integrand_ms"] / max(1, agg["launches"])}
        results[p], values[p] = agg, g
    if "plain_fp64" in values and "batched_fp64" in values:
        a, b = values["batched_fp64"], values["plain_fp64"]
        results["plain_fp64"]["max_abs_diff_vs_batched"] = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
    if "batched_fp32" in values and "batched_fp64" in values:
        a, b = values["batched_fp64"], values["batched_fp32"]
        # per triangle and component, |difference| relative to the component's modulus
        rel = np.abs(b - a) / np.maximum(np.abs(a), 1e-300)
        results["batched_fp32"]["max_rel_err_vs_fp64"] = float(rel.max())
        results["batched_fp32"]["rel_err_p50_p99"] = [float(np.percentile(rel, q)) for q in (50, 99)]
    ok = [p for p in results if p != "batched_fp32" or results[p].get("max_rel_err_vs_fp64", 1) <= 1e-3]
    head = max(ok, key=lambda p: results[p]["triangles_per_s"]) if ok else next(iter(results))
    hr = results[head]
    dom = hr["integrand_kernel"]
    line = {
        "metric": "Gamma_i triangles/s on 1 GPU (4 complex natural components per triangle)",
        "value": hr["triangles_per_s"], "unit": "triangles/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": hr["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head == "batched_fp32" else "f64", "data": "synthetic",
        "config": {"workload": "Gamma_0..3 of %d of the 1000 TreeCorr-grid triangles (r 0.1-100 arcmin log x 10, "
                               "u 0-1 x 10, v -1-1 x 10, centroid), fiducial flat LCDM, Ogata h 1e-3 with 4999 "
                               "zeros of J6 and J2, hcubature (z, phi, psi) rel tol 1e-3" % x.shape[0],
                   "triangles": [int(i) for i in sel], "path": head, "setup_ms": setup_ms},
        "paths": results,
        "roofline": {"bound": "valu", "kernel": "k_gamma_points", "achieved": dom["achieved_flops"] / 1e12,
                     "peak": dom["peak"] / 1e12, "unit": "TFLOP/s", "frac": dom["frac"], "traffic": None,
                     "work_per_point_flop_eq": fe_point, "flops_per_elementary_call": FLOPS_PER_ELEM,
                     "census": fx["gamma_term"], "census_file": CENSUS_FILE},
    }
    if not args.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = gamma_cpu_baseline(W, x, hr["points_per_triangle"], args.cpu_seconds)
    print(json.dumps(line))


def gamma_cpu_baseline(W, x, points_per_triangle, seconds_target):
    """The FP64 CPU oracle's Gamma integrands (OpenMP over points) on a bounded sample of cubature
    points; triangles/s follows from the GPU run's integrand points per triangle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    cores = min(os.cpu_count() or 1, 64)
    o = Oracle(W.fiducial(), W.fiducial_nz())
    o.gamma_init(1e-3, 5.0)
    rng = np.random.default_rng(0xC2)
    done, t0, n = 0, time.perf_counter(), 2 * cores
    while time.perf_counter() - t0 < seconds_target:
        pts = np.column_stack([rng.uniform(0, 3.0, n), rng.uniform(0, 2 * np.pi, n), rng.uniform(0, np.pi / 2, n)])
        for which in (0, 1):  # the Gamma_0 and the Gamma_1..3 integrands, in the 1:3 mix of a triangle
            o.gamma_integrand_batch(which, x[0], pts[: n // 3] if which == 0 else pts, cores)
        done += n // 3 + n
    dt = time.perf_counter() - t0
    pps = done / dt
    return {"value": pps / points_per_triangle, "unit": "triangles/s", "cores": cores, "kind": "port",
            "points_per_s": pps,
            "sample": "%d Gamma integrand points (uniform in (z, phi, psi), Gamma_0:Gamma_1 integrands 1:3) through "
                      "the FP64 CPU oracle, OpenMP %d threads, %s, %.1f s; triangles/s = points/s / %.0f points "
                      "per triangle of the GPU run" % (done, cores, cpu_model(), dt, points_per_triangle)}


def map2_bench(tp, W, ctx, args):
    """<Map^2> of one source bin for 512 Latin-hypercube cosmologies x 10 aperture radii (0.5-32 arcmin):
    5120 2-D cubatures over (ell, z), rel tol 1e-5, no shape noise, the classed_version spectrum."""
    t0 = time.perf_counter()
    cos = W.lhs_c3(512)
    tabs = tp.CosmologyTables(cos, W.fiducial_nz(), 100, 3.0, ctx)
    tables_ms = 1e3 * (time.perf_counter() - t0)
    radii = W.to_rad(W.map3_c1_radii())
    theta = np.tile(radii, 512)
    icosmo = np.repeat(np.arange(512, dtype=np.int32), radii.size)
    fx = census_fixture(tp)
    c = fx["power_point"]
    fe_point = fx["flop_eq"]["map2_point"]
    results, values = {}, {}
    for p in [q for q in args.paths.split(",") if q in ("plain_fp64", "batched_fp64")]:
        print("bench: map2 path %s ..." % p, file=sys.stderr, flush=True)
        steps = 1 if p == "plain_fp64" else args.steps
        n = 20 if p == "plain_fp64" else theta.size
        agg = {"points": 0, "integrand_ms": 0.0, "launches": 0, "rounds": 0}
        if p == "batched_fp64":
            for _ in range(args.warmup):
                tabs.map2_batch(theta, icosmo)
        tt = time.perf_counter()
        for _ in range(steps):
            if p == "batched_fp64":
                v = tabs.map2_batch(theta, icosmo)[0]
                st = ctx.stats()
                for k in agg:
                    agg[k] += st.get(k, 0)
            else:  # one integral per call: the reference's one-at-a-time loop
                v = []
                for i in range(n):
                    v.append(tabs.map2_batch(theta[i:i + 1], icosmo[i:i + 1])[0][0])
                    st = ctx.stats()
                    for k in agg:
                        agg[k] += st.get(k, 0)
                v = np.array(v)
        dt = time.perf_counter() - tt
        agg["ms_per_step"] = 1e3 * dt / steps
        agg["integrals_per_s"] = n * steps / dt
        agg["points_per_integral"] = agg["points"] / (steps * n)
        if agg["integrand_ms"] > 0:
            ach = fe_point * agg["points"] / (agg["integrand_ms"] * 1e-3)
            agg["integrand_kernel"] = {"achieved_flops": ach, "peak": PEAK["fp64"], "frac": ach / PEAK["fp64"],
                                       "avg_launch_ms": agg["integrand_ms"] / max(1, agg["launches"])}
        results[p], values[p] = agg, v
    if "plain_fp64" in values and "batched_fp64" in values:
        a, b = values["batched_fp64"][:20], values["plain_fp64"]
        results["plain_fp64"]["max_rel_diff_vs_batched"] = float(np.max(np.abs(a / b - 1)))
    head = "batched_fp64" if "batched_fp64" in results else next(iter(results))
    hr = results[head]
    dom = hr["integrand_kernel"]
    line = {
        "metric": "<Map^2> integrals/s on 1 GPU (one source bin)",
        "value": hr["integrals_per_s"], "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": hr["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "<Map^2> of 512 Latin-hypercube wCDM cosmologies x 10 log radii 0.5-32 arcmin = "
                               "5120 integrals, revised Halofit, 100 z-bins, 2-D hcubature (ell, z) rel tol 1e-5, "
                               "no shape noise", "path": head, "tables_ms": tables_ms},
        "paths": results,
        "roofline": {"bound": "valu", "kernel": "k_power_points", "achieved": dom["achieved_flops"] / 1e12,
                     "peak": dom["peak"] / 1e12, "unit": "TFLOP/s", "frac": dom["frac"], "traffic": None,
                     "algorithmic_bytes_per_point": 16 + 4 + 8,
                     "work_per_point_flop_eq": fe_point, "flops_per_elementary_call": FLOPS_PER_ELEM,
                     "census": c, "census_file": CENSUS_FILE},
    }
    if not args.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = map2_cpu_baseline(W, cos, radii, args.cpu_seconds)
    print(json.dumps(line))


def map2_cpu_baseline(W, cos, radii, seconds_target):
    """The FP64 CPU oracle's <Map^2> (the same cubature and integrand), one integral per thread."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    cores = min(os.cpu_count() or 1, 64)
    nz = W.fiducial_nz()
    orcs = [Oracle(c, nz) for c in cos[:cores]]
    jobs = [(orcs[i % len(orcs)], radii[i % radii.size]) for i in range(100000)]
    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:  # the oracle's ctypes calls release the GIL
        while time.perf_counter() - t0 < seconds_target:
            batch = jobs[done:done + 4 * cores]
            list(ex.map(lambda j: j[0].map2(j[1], 1e-5), batch))
            done += len(batch)
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "integrals/s", "cores": cores, "kind": "port",
            "sample": "%d <Map^2> integrals (the first %d cosmologies x the 10 radii) through the FP64 CPU oracle, "
                      "%d threads, %s, %.1f s" % (done, len(orcs), cores, cpu_model(), dt)}


def cov_bench(tp, W, ctx, args):
    """T1 of the Gaussian <Map^3> covariance (infinite survey) for all 210 pairs of the 20 triples of
    radii {2, 4, 8, 16} arcmin: the survey of the reference's example parameter file (61.94 deg^2,
    sigma_eps 0.37, 108000 galaxies per deg^2), T17 spectrum, rel tol 1e-4."""
    t0 = time.perf_counter()
    tabs = tp.CosmologyTables([W.fiducial()], W.fiducial_nz(), 100, 3.0, ctx)
    tables_ms = 1e3 * (time.perf_counter() - t0)
    deg = 0.017453  # classed_version/helper.cpp:142 (the reference's degree constant)
    area = 61.9369 * deg * deg
    n_gal = 108000 / deg / deg
    tri = W.to_rad(W.triples(W.map3_c0_radii()))
    pairs = np.array([np.concatenate([tri[i], tri[j]]) for i in range(len(tri)) for j in range(i, len(tri))])
    flags = tp.PK_T17
    results = {}
    fx = census_fixture(tp)
    # a T1 region (k_t1_regions): 47 Pell (the 7 distinct l1, 7 distinct l2 and 33 l3 of its 33 points) of
    # 96 z-nodes, each node the k-dependent part of the Limber P (the z-only Halofit coefficients are kept
    # per node), plus 3 flops per node pair; per point l3 (7 flops + cos + sqrt), 6 uHat (4 flops + exp
    # each) and 8 products; the 3-D Genz-Malik rule (~90 flops)
    fe_region = fx["flop_eq"]["t1_region"]
    values = {}
    for path, mode in (("batched_fp32", tp.MODE_FP32), ("batched_fp64", tp.MODE_FP64)):
        print("bench: cov T1 %s ..." % path, file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            tabs.map3_cov_t1(pairs[:8], area, 0.0, 0.37, n_gal, flags, mode=mode)
        tt = time.perf_counter()
        for _ in range(args.steps):
            v, ne = tabs.map3_cov_t1(pairs, area, 0.0, 0.37, n_gal, flags, mode=mode)
        dt = time.perf_counter() - tt
        st = ctx.stats()  # the last step's
        values[path] = v
        agg = {"ms_per_step": 1e3 * dt / args.steps, "pairs_per_s": pairs.shape[0] * args.steps / dt,
               "points": st.get("points", 0), "regions": st.get("regions", 0),
               "integrand_ms": st.get("integrand_ms", 0.0), "integrand_sum_ms": st.get("integrand_sum_ms", 0.0),
               "launches": st.get("launches", 0), "rounds": st.get("rounds", 0)}
        peak = PEAK["fp32"] if mode == tp.MODE_FP32 else PEAK["fp64"]
        if agg["integrand_sum_ms"] > 0:
            ach = fe_region * agg["regions"] / (agg["integrand_sum_ms"] * 1e-3)
            agg["integrand_kernel"] = {"achieved_flops": ach, "peak": peak, "frac": ach / peak,
                                       "avg_launch_ms": agg["integrand_sum_ms"] / max(1, agg["launches"])}
        results[path] = agg
    ref = values["batched_fp64"]
    for path in results:
        results[path]["max_rel_err_vs_fp64"] = float(np.max(np.abs(values[path] / ref - 1)))
    ok = [p for p in results if results[p]["max_rel_err_vs_fp64"] <= 1e-3]
    head = max(ok, key=lambda p: results[p]["pairs_per_s"])
    agg = results[head]
    dom = agg["integrand_kernel"]
    ach = dom["achieved_flops"]
    line = {
        "metric": "Gaussian <Map^3> covariance T1 pairs/s on 1 GPU (infinite survey)",
        "value": agg["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": agg["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head == "batched_fp32" else "f64", "data": "synthetic",
        "config": {"workload": "T1_total of 210 pairs of the 20 triples of radii 2-16 arcmin, infinite survey of "
                               "61.94 deg^2, sigma_eps 0.37, 108000 / deg^2, T17 revised Halofit, 3-D hcubature "
                               "rel tol 1e-4", "path": head, "tables_ms": tables_ms,
                   "t1_integrals": int(sum({3: 6, 2: 3, 1: 1}[len(set(p[3:]))] for p in pairs))},
        "paths": results,
        "roofline": {"bound": "valu", "kernel": "k_t1_regions<%s>" % ("true" if head == "batched_fp32" else "false"),
                     "achieved": ach / 1e12, "peak": dom["peak"] / 1e12, "unit": "TFLOP/s", "frac": dom["frac"],
                     "traffic": None,
                     "work_per_region_flop_eq": fe_region, "flops_per_elementary_call": FLOPS_PER_ELEM,
                     "census_file": CENSUS_FILE,
                     "node_evaluations_per_region": 47 * 96},
    }
    if not args.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import Oracle
        cores = min(os.cpu_count() or 1, 64)
        o = Oracle(W.fiducial(), W.fiducial_nz())
        done, pts_done, t1 = 0, 0, time.perf_counter()
        while time.perf_counter() - t1 < args.cpu_seconds and done < pairs.shape[0]:
            p = pairs[(done * 37) % pairs.shape[0]]
            _, n = o.t1_total(p[:3], p[3:], area, 0.0, 0.37, n_gal, flags, 1e-4, cores)
            done += 1
            pts_done += n
        d1 = time.perf_counter() - t1
        line["cpu_baseline"] = {"value": done / d1, "unit": "pairs/s", "cores": cores, "kind": "port",
                                "sample": "%d of the 210 pairs (every 37th) through the FP64 CPU oracle's T1_total, "
                                          "OpenMP %d threads over each cubature round, %s, %.1f s"
                                          % (done, cores, cpu_model(), d1)}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--compare-steps", type=int, default=2,
                    help="timed steps of the comparison paths (the headline path runs --steps)")
    ap.add_argument("--compare-warmup", type=int, default=1)
    ap.add_argument("--reltol", type=float, default=1e-4)
    ap.add_argument("--paths", default="plain_fp64,surrogate,batched_fp32,batched_fp64",
                    help="paths to time (c2 takes plain_fp64, batched_fp64, batched_fp32)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "map2", "cov"])
    ap.add_argument("--c2-triangles", type=int, default=10,
                    help="Gamma workload: number of the 1000 TreeCorr-grid triangles (one per r bin at 10)")
    ap.add_argument("--surrogate", default=os.path.join(ROOT, "threepoint_amd", "models", "bkappa_mlp.bin"))
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known traffic for "
                         "calibrating HBM byte counters)")
    ap.add_argument("--write-inputs", metavar="DIR",
                    help="write the workload's cosmologies, triples and n(z) as text files for benchMap3.x "
                         "(the same workload driven from C++, for profilers), then exit")
    ap.add_argument("--accuracy-study", action="store_true",
                    help="compare every path at the tolerance against FP64 at a 10x tighter tolerance, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the path shards by cosmology/triple with no exchange "
              "between GPUs, and multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import threepoint_amd as tp
    from threepoint_amd import workloads as W

    if args.write_inputs:
        write_inputs(W, args.workload, args.write_inputs)
        return
    ctx = tp.default_context()
    if args.copy_calibration_mib > 0:
        import ctypes
        gb = ctypes.c_double(0)
        tp._check(tp.lib().tp_hbm_copy_probe(ctx.h, args.copy_calibration_mib << 20, ctypes.byref(gb)),
                  "tp_hbm_copy_probe")
        print("bench: copied %d MiB device-to-device (k_copy16) at %.0f GB/s" % (args.copy_calibration_mib, gb.value),
              file=sys.stderr, flush=True)
    if args.workload == "c2":
        gamma_bench(tp, W, ctx, args)
        return
    if args.workload == "map2":
        map2_bench(tp, W, ctx, args)
        return
    if args.workload == "cov":
        cov_bench(tp, W, ctx, args)
        return
    icosmo = None
    nz = W.fiducial_nz()
    if args.workload == "c1":
        cos = [W.fiducial()]
        thetas = W.to_rad(W.triples(W.map3_c1_radii()))
        wl = ("Map3 of 220 triples (10 log radii 0.5-32 arcmin), fiducial flat LCDM, 100 z-bins, "
              "hcubature rel tol %g; a step builds the cosmology's FP64 tables on the GPU, then integrates"
              % args.reltol)
    else:
        cos = W.lhs_c3(512)
        tri = W.to_rad(W.triples(W.map3_c3_radii()))
        thetas = np.tile(tri, (512, 1))
        icosmo = np.repeat(np.arange(512, dtype=np.int32), tri.shape[0])
        wl = ("Map3 of 512 Latin-hypercube wCDM cosmologies x 35 triples (radii 2-32 arcmin) = 17920 integrals, "
              "100 z-bins, hcubature rel tol %g; a step builds the 512 cosmologies' FP64 tables on the GPU, "
              "then integrates" % args.reltol)

    def make_tables():
        return tp.CosmologyTables(cos, nz, 100, 3.0, ctx)

    def make_tables_first():  # the plain path's step: the first cosmology only
        return tp.CosmologyTables(cos[:1], nz, 100, 3.0, ctx)

    t0 = time.perf_counter()
    tabs = make_tables()
    tables_ms = 1e3 * (time.perf_counter() - t0)

    paths = [p for p in args.paths.split(",") if p]
    have_sur = False
    if "surrogate" in paths:
        have_sur = os.path.exists(args.surrogate) and tp.lib().tp_surrogate_load(ctx.h, args.surrogate.encode()) == 0
        if not have_sur:
            paths.remove("surrogate")

    if args.accuracy_study:
        accuracy_study(tabs, ctx, thetas, icosmo, paths, args.reltol)
        return

    fx = census_fixture(tp)
    census = fx["map3_point"]
    fe_point = fx["flop_eq"]["map3_point"]
    # the headline is the surrogate form (the metric's target form) when a model is present, else batched FP64;
    # it runs the requested steps, the comparison paths --compare-steps (their errors and rates are reported)
    head_want = "surrogate" if have_sur else "batched_fp64"
    results, values = {}, {}
    for p in paths:
        print("bench: path %s ..." % p, file=sys.stderr, flush=True)
        th_p, ic_p, mk = thetas, icosmo, make_tables
        if p == "plain_fp64":
            steps, warm = 1, 0
            if args.workload == "c3":  # one cosmology's 35 triples
                th_p, ic_p, mk = thetas[:35], icosmo[:35], make_tables_first
        elif p == head_want:
            steps, warm = args.steps, args.warmup
        else:
            steps, warm = args.compare_steps, args.compare_warmup
        v, agg = run_path(tabs, ctx, th_p, p, steps, warm, args.reltol, ic_p, mk)
        values[p] = v
        results[p] = agg
    ref = values.get("batched_fp64")
    for p in results:
        if ref is not None and values[p].shape == ref.shape:
            # triples whose first cubature region integrates to exactly 0 end at 0 in every path
            # (the reference algorithm's behaviour); they must stay exactly 0
            nz_ = ref != 0
            err = np.abs(values[p][nz_] - ref[nz_]) / np.abs(ref[nz_])
            zero_ok = bool(np.all(values[p][~nz_] == 0))
            results[p]["max_rel_err_vs_fp64"] = float(err.max()) if zero_ok else float("inf")
            results[p]["zero_triples"] = int((~nz_).sum())
            if p != "batched_fp64":
                idx = np.nonzero(nz_)[0]
                results[p]["rel_err_p50_p99_p999"] = [float(np.percentile(err, q)) for q in (50, 99, 99.9)]
                results[p]["triples_over_1e-3"] = int((err > 1e-3).sum())
                w = int(idx[np.argmax(err)])
                results[p]["worst_triple"] = {"index": w, "theta_rad": [float(t) for t in thetas[w]],
                                              "icosmo": int(icosmo[w]) if icosmo is not None else 0,
                                              "fp64": float(ref[w]), "this_path": float(values[p][w])}
        r = results[p]
        exact_pts = r["points"] - r.get("surrogate_points", 0)
        if r.get("integrand_sum_ms", 0) > 0:
            # per launch: work / event-timed duration (the sum over launches); the two groups' launches can
            # overlap on the device, so the union of their intervals (integrand_ms) is the device-busy time
            ach = fe_point * exact_pts / (r["integrand_sum_ms"] * 1e-3)
            peak = PEAK["fp64"] if p.startswith("batched_fp64") or p == "plain_fp64" else PEAK["fp32"]
            r["integrand_kernel"] = {"achieved_flops": ach, "peak": peak, "frac": ach / peak, "points": exact_pts,
                                     "avg_launch_ms": r["integrand_sum_ms"] / max(1, r["launches"]),
                                     "device_busy_frac": fe_point * exact_pts / (r["integrand_ms"] * 1e-3) / peak}
        if r.get("mlp_ms", 0) > 0:
            ach = MLP_FLOPS_PER_POINT * r["mlp_points"] / (r["mlp_ms"] * 1e-3)
            r["mlp_kernel"] = {"achieved_flops": ach, "peak": PEAK["fp16"], "frac": ach / PEAK["fp16"],
                               "points": r["mlp_points"], "launches": r["mlp_launches"],
                               "avg_launch_ms": r["mlp_ms"] / max(1, r["mlp_launches"])}
    ok = [p for p in results if results[p].get("max_rel_err_vs_fp64", 1.0) <= 1e-3]
    if ref is None:  # no FP64 reference in this run: the headline path, unchecked
        ok = list(results)
    head = head_want if head_want in ok else ("batched_fp64" if "batched_fp64" in results else
                                              max(ok, key=lambda p: results[p]["triples_per_s"]))
    hr = results[head]
    # the dominant kernel of the headline path: the network (MFMA) or the exact region kernel (VALU)
    use_mlp = "mlp_kernel" in hr and ("integrand_kernel" not in hr or hr["mlp_ms"] >= hr["integrand_sum_ms"])
    if use_mlp:
        mk_ = hr["mlp_kernel"]
        roof = {"bound": "mfma", "kernel": "k_mlp_h (FP16 MFMA, FP16 activations, FP32 accumulation)",
                "achieved": mk_["achieved_flops"] / 1e12, "peak": PEAK["fp16"] / 1e12, "unit": "TFLOP/s",
                "frac": mk_["frac"],
                "traffic": MLP_HBM_BYTES_PER_POINT * mk_["points"] / max(1, mk_["launches"]),
                "traffic_unit": "HBM bytes per launch (per-point bytes from the PMC passes x points per launch)",
                "algorithmic_flops_per_point": MLP_FLOPS_PER_POINT,
                "executed_mfma_flops_per_point": MLP_MFMA_FLOPS_EXECUTED,
                "counters": MLP_COUNTERS}
        if "integrand_kernel" in hr:
            roof["secondary"] = {"kernel": "exact FP32 region kernels of the surrogate form",
                                 "achieved": hr["integrand_kernel"]["achieved_flops"] / 1e12,
                                 "peak": hr["integrand_kernel"]["peak"] / 1e12,
                                 "frac": hr["integrand_kernel"]["frac"], "unit": "TFLOP/s",
                                 "ms": hr["integrand_sum_ms"], "mlp_ms": hr["mlp_ms"]}
    else:
        dom = hr.get("integrand_kernel") or next(r["integrand_kernel"] for r in results.values()
                                                 if "integrand_kernel" in r)
        prec = "double" if dom["peak"] == PEAK["fp64"] else "float"
        kname = "k_region_points<%s, 1/2/4/8 waves per point> | k_region_cached<%s> (shape tuned per launch size)" % (
            prec, prec)
        pts_per_launch = dom["points"] / max(1, hr["launches"])
        # the region-cached shape evaluates the P_E terms (3 per node and point) once per distinct multipole of
        # a region: 47 per 33 points instead of 99, so it executes less than the algorithmic work per point
        fe_exec_cached = fx["flop_eq"]["map3_point_region_cached_executed"]
        roof = {"bound": "valu", "kernel": kname,
                "achieved": dom["achieved_flops"] / 1e12, "peak": dom["peak"] / 1e12, "unit": "TFLOP/s",
                "frac": dom["frac"], "traffic": HBM_BYTES_PER_POINT[prec] * pts_per_launch,
                "traffic_unit": "HBM bytes per launch (per-point bytes from the PMC passes x points per launch)",
                "algorithmic_bytes_per_point": 64 / 33 + 8,
                "work_per_point_flop_eq": fe_point, "flops_per_elementary_call": FLOPS_PER_ELEM,
                "executed_work_per_point_flop_eq_region_cached": fe_exec_cached,
                "executed_frac_if_all_region_cached": dom["frac"] * fe_exec_cached / fe_point,
                "census": census, "census_file": CENSUS_FILE, "counters": REGION_COUNTERS.get(prec)}
        if "mlp_kernel" in hr:  # the surrogate form's other kernel: the network
            mk_ = hr["mlp_kernel"]
            roof["secondary"] = {"kernel": "k_mlp_h (FP16 MFMA)", "bound": "mfma",
                                 "achieved": mk_["achieved_flops"] / 1e12, "peak": PEAK["fp16"] / 1e12,
                                 "frac": mk_["frac"], "unit": "TFLOP/s", "mlp_ms": hr["mlp_ms"],
                                 "avg_launch_ms": mk_["avg_launch_ms"], "counters": MLP_COUNTERS}
    line = {
        "metric": "Map3 triples/s on 1 GPU: plain FP64 vs surrogate form at <=1e-3 rel err",
        "value": hr["triples_per_s"], "unit": "triples/s", "n_gpus": 1,
        "steps": hr["steps"], "warmup": hr["warmup"], "ms_per_step": hr["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"batched_fp32": "f32", "surrogate": "fp16 network + f32 exact regions, f64 accumulation"}.get(
            head, "f64"),
        "data": "synthetic",
        "config": {"workload": wl, "path": head, "triples_per_step": int(thetas.shape[0]),
                   "tables_ms_per_step": hr["tables_ms"] / max(1, hr["steps"]), "tables_ms_first_build": tables_ms,
                   "comparison_paths_steps": {"steps": args.compare_steps, "warmup": args.compare_warmup}},
        "paths": results,
        "speedup_vs_plain_fp64": (hr["triples_per_s"] / results["plain_fp64"]["triples_per_s"]
                                  if "plain_fp64" in results else None),
        "roofline": roof,
    }
    if "batched_fp32" in results and head == "surrogate":
        line["surrogate_vs_exact_fp32"] = hr["triples_per_s"] / results["batched_fp32"]["triples_per_s"]
    if not args.no_cpu:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(cos, thetas, icosmo, args.cpu_seconds, args.reltol)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
