"""Host-side checks of the product library (no GPU needed).

* the C-ABI library loads and exports every function include/threepoint_amd.h declares;
* tp_hcubature_v (the drop-in hcubature_v) is bit-identical to the reference's own
  hcubature.cpp (compiled into oracle/_ref) and to the oracle's restatement, for
  fdim 1 and 2, dim 1 (GK15 rule) and dim 2-4 (Genz-Malik), several error norms;
* callback abort propagates as the reference's return value 1;
* workload generators reproduce the reference's binning / resampling formulas.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import _oracle as O
import threepoint_amd as tp
from threepoint_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "threepoint_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tp_[a-z0-9_]+)\s*\(", src)) - {"tp_integrand_v"})


def test_library_exports_every_declared_symbol():
    L = tp.lib()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    assert len(declared_functions()) >= 25


def test_census_is_positive():
    c = tp.census_map3_point()
    assert c["nodes"] == 96 and c["node_flops"] > 100 and c["node_elem"] > 10


def test_census_matches_committed_fixture():
    """bench.py's roofline reads tests/golden/census.json: the library's op census must still give it
    (a change to a device function that changes the work per point must regenerate the fixture)."""
    import json
    import os
    sys_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(sys_path, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fx = json.load(open(os.path.join(sys_path, "census.json")))
    assert mg.census() == fx
    # the split the region-cached kernel relies on: a node is three P_E terms plus the rest
    c = fx["map3_point"]
    assert 3 * c["pse_flops"] < c["node_flops"] and 3 * c["pse_elem"] < c["node_elem"]


def _integrand(fdim, dim):
    def f(ndim, npt, x, fdata, fd, fval):
        for i in range(npt):
            xs = [x[ndim * i + j] for j in range(ndim)]
            r2 = sum((v - 0.3 - 0.1 * j) ** 2 for j, v in enumerate(xs))
            fval[fd * i] = math.exp(-r2 / 0.02) + 0.2 * math.cos(5 * sum(xs))
            if fd == 2:
                fval[fd * i + 1] = math.sin(3 * xs[0]) * xs[-1] + 0.1
        return 0
    return O.INTEGRAND_V(f)


@pytest.mark.parametrize("dim,fdim,norm,tol", [(3, 1, 3, 1e-4), (3, 2, 3, 1e-4), (2, 2, 1, 1e-5), (4, 1, 0, 1e-3),
                                               (1, 2, 2, 1e-8), (1, 1, 3, 1e-10), (3, 2, 4, 1e-4), (2, 2, 0, 1e-6)])
def test_product_cubature_bit_exact(dim, fdim, norm, tol):
    cb = _integrand(fdim, dim)
    lo, hi = np.zeros(dim), np.ones(dim)
    v1, e1 = np.zeros(fdim), np.zeros(fdim)
    r1 = tp.lib().tp_hcubature_v(fdim, ctypes.cast(cb, tp.INTEGRAND_V), None, dim, tp._d(lo), tp._d(hi), 0, 0.0,
                                 tol, norm, tp._d(v1), tp._d(e1))
    assert r1 == 0
    if os.path.exists(O.REF_CUBATURE_SO):
        ref = ctypes.CDLL(O.REF_CUBATURE_SO)
        ref.hcubature_v.restype = ctypes.c_int
        ref.hcubature_v.argtypes = O._SIGS["orc_hcubature_v"][1]
        v2, e2 = np.zeros(fdim), np.zeros(fdim)
        assert ref.hcubature_v(fdim, cb, None, dim, O.d(lo), O.d(hi), 0, 0.0, tol, norm, O.d(v2), O.d(e2)) == 0
        assert np.array_equal(v1, v2) and np.array_equal(e1, e2)
    v3, e3 = np.zeros(fdim), np.zeros(fdim)
    assert O.lib().orc_hcubature_v(fdim, cb, None, dim, O.d(lo), O.d(hi), 0, 0.0, tol, norm, O.d(v3), O.d(e3)) == 0
    assert np.array_equal(v1, v3) and np.array_equal(e1, e3)


def test_product_cubature_maxeval_and_abort():
    cb = _integrand(1, 3)
    lo, hi = np.zeros(3), np.ones(3)
    v, e = np.zeros(1), np.zeros(1)
    assert tp.lib().tp_hcubature_v(1, ctypes.cast(cb, tp.INTEGRAND_V), None, 3, tp._d(lo), tp._d(hi), 500, 0.0,
                                   1e-12, 3, tp._d(v), tp._d(e)) == 0
    assert tp.lib().tp_hcubature_last_neval() <= 500 + 2 * 33

    def bad(ndim, npt, x, fdata, fd, fval):
        return 1
    cbb = tp.INTEGRAND_V(bad)
    assert tp.lib().tp_hcubature_v(1, cbb, None, 3, tp._d(lo), tp._d(hi), 0, 0.0, 1e-4, 3, tp._d(v), tp._d(e)) == 1


def test_python_hcubature_wrapper():
    rc, val, err, n = tp.hcubature_v(1, lambda x: np.exp(-np.sum(x * x, 1)), 2, [0, 0], [1, 1], reqRelError=1e-8)
    exact = (math.sqrt(math.pi) / 2 * math.erf(1)) ** 2
    assert rc == 0 and abs(val[0] / exact - 1) < 1e-7


def test_nz_resampling_matches_oracle_reader(tmp_path):
    z, n = W.nz_file_rows()
    fn = tmp_path / "nz.dat"
    with open(fn, "w") as f:
        for a, b in zip(z, n):
            f.write("%.10g %.10g\n" % (a, b))
    out = np.zeros(100)
    assert O.lib().orc_read_n_of_z(str(fn).encode(), 100, 3.0, O.d(out)) == 0
    zz = np.loadtxt(fn)
    ours = W.resample_n_of_z(zz[:, 0], zz[:, 1])
    assert np.allclose(ours, out, rtol=1e-15, atol=0)
    assert abs(ours.sum() * 0.03 - 1) < 1e-14


def test_triples_and_gamma_grid():
    assert W.triples(W.map3_c0_radii()).shape == (20, 3)
    assert W.triples(W.map3_c1_radii()).shape == (220, 3)
    assert W.triples(W.map3_c3_radii()).shape == (35, 3)
    g = W.gamma_c2_triangles()
    assert g.shape == (1000, 6)
    r, u, v = g[:, 3], g[:, 4], g[:, 5]
    r2 = r * np.pi / 180 / 60
    assert np.allclose(g[:, 1], r2) and np.allclose(g[:, 2], r2 * u) and np.allclose(g[:, 0], v * r2 * u + r2)


def test_executable_inputs_round_trip(tmp_path):
    """The input files written for the executables read back, with the package's readers and the oracle's
    n(z) reader, as the workloads they stand for."""
    p = W.write_executable_inputs(str(tmp_path))
    c = tp.Cosmology.from_file(p["cosmo.param"])
    for k, v in W.FIDUCIAL.items():
        assert getattr(c, k) == v, k
    bad = tmp_path / "bad.param"
    bad.write_text("h 0.7\nOmega_k 0.1\n")
    with pytest.raises(ValueError):
        tp.Cosmology.from_file(str(bad))
    out = np.zeros(100)
    assert O.lib().orc_read_n_of_z(p["nz.dat"].encode(), 100, 3.0, O.d(out)) == 0
    zz = np.loadtxt(p["nz.dat"])
    assert np.allclose(W.resample_n_of_z(zz[:, 0], zz[:, 1]), out, rtol=1e-15, atol=0)
    cfg = dict(ln.split() for ln in open(p["gamma_c2.param"]))
    g = W.gamma_c2_triangles(int(cfg["rsteps"]), int(cfg["usteps"]), int(cfg["vsteps"]), float(cfg["rmin"]),
                             float(cfg["rmax"]), float(cfg["umin"]), float(cfg["umax"]), float(cfg["vmin"]),
                             float(cfg["vmax"]))
    assert np.array_equal(g, W.gamma_c2_triangles())
    th = np.loadtxt(p["thetas_c1.dat"], comments="#")
    assert np.array_equal(th, np.asarray(W.map3_c1_radii(), dtype=np.float64))


def test_lhs_is_latin():
    cs = W.lhs_c3(64)
    om = np.array([c.om for c in cs])
    strata = np.floor((om - 0.2) / 0.2 * 64).astype(int)
    assert sorted(strata) == list(range(64))


CV_CALLER = r"""
// A caller written against the classed_version class API (bispectrum.hpp:207-263,
// apertureStatistics.hpp:223-232, gamma.hpp:33-77, cubature.h:55-94)
#include <complex>
#include <vector>
#include "threepoint_amd.hpp"

static int f(unsigned ndim, size_t npt, const double *x, void *fdata, unsigned fdim, double *fval)
{
    (void)ndim; (void)x; (void)fdata;
    for (size_t i = 0; i < npt * fdim; ++i) fval[i] = 1.0;
    return 0;
}

double caller(cosmology *cosmo, const std::vector<double> &nz)
{
    BispectrumCalculator bk(cosmo, nz, 100, 3.0);
    bk.set_cosmology(cosmo);
    double b = bk.bispec(0.1, 0.2, 0.25, 0.5, 16, 0.5) + bk.bispec(0.1, 0.2, 0.25, 0.5);
    b += bk.bkappa(100., 200., 250.);
    ell_params p{100., 200., 250.};
    b += bk.integrand_bkappa(0.5, p) + bk.z_max;
    ApertureStatistics as(&bk);
    std::vector<double> th{1e-3, 2e-3, 4e-3};
    b += as.MapMapMap(th) + as.integrand(10., 20., 0.5, th);
    double in[3] = {10., 20., 0.5}, out = 0;
    ApertureStatisticsContainer cont{&as, th};
    ApertureStatistics::integrand(3, 1, in, &cont, 1, &out);
    GammaCalculator g(&bk, 1e-3, 5., "orthocenter"), g2(&bk, 1e-3, 5., "centroid");
    std::complex<double> c = g.gamma0(1e-3, 2e-3, 2.5e-3) + g.gamma1(1e-3, 2e-3, 2.5e-3) +
                             g2.gamma2(1e-3, 2e-3, 2.5e-3) + g2.gamma3(1e-3, 2e-3, 2.5e-3);
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1}, v = 0, e = 0;
    hcubature_v(1, f, nullptr, 3, lo, hi, 0, 0, 1e-4, ERROR_L1, &v, &e);
    return b + c.real() + v;
}
"""


def test_cv_style_caller_compiles(tmp_path):
    """A C++ caller using the classed_version class signatures (including the 6-argument bispec and the
    orthocenter triangle center) compiles against include/threepoint_amd.hpp unchanged."""
    import subprocess
    src = tmp_path / "cv_caller.cpp"
    src.write_text(CV_CALLER)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I" + os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_bench_process_loads_no_torch():
    """The bench's process holds no PyTorch: importing bench.py (numpy, ctypes and this package) leaves torch out
    of sys.modules, and the HIP library links no torch library.  (Profiled runs use benchMap3.x, a C++ program
    linked only against the library.)"""
    import subprocess
    import sys
    code = ("import importlib.util, sys; sys.path.insert(0, %r); "
            "spec = importlib.util.spec_from_file_location('bench', %r); m = importlib.util.module_from_spec(spec); "
            "spec.loader.exec_module(m); import threepoint_amd; "
            "print(sorted(k for k in sys.modules if k.split('.')[0] == 'torch'))") % (ROOT, os.path.join(ROOT, "bench.py"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "[]", r.stdout
    lib = os.path.join(ROOT, "threepoint_amd", "_lib", "libthreepoint_amd.so")
    if os.path.exists(lib):
        ldd = subprocess.run(["ldd", lib], capture_output=True, text=True).stdout
        # library names only: the load addresses ldd prints are random and may contain "c10"
        names = [ln.split()[0] for ln in ldd.splitlines() if ln.strip()]
        assert not any("torch" in n or "libc10" in n for n in names), names
