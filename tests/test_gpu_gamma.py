"""GPU parity of the Gamma_i (Ogata + cubature) path against the CPU FP64 oracle."""
import numpy as np
import pytest

import threepoint_amd as tp
from threepoint_amd import workloads as W

pytestmark = pytest.mark.gpu


def _tri():
    rows = W.gamma_c2_triangles()
    return rows[555, :3]  # an interior (r, u, v) bin


@pytest.mark.parametrize("prec", [(0.05, 2.0), (1e-3, 5.0)])
def test_gamma_integrand_matches_oracle(fid_tables, fid_oracle, gpu_ctx, prec):
    tp.gamma_setup(prec[0], prec[1], gpu_ctx)
    fid_oracle.gamma_init(prec[0], prec[1])
    x1, x2, x3 = _tri()
    rng = np.random.default_rng(5)
    npts = 6 if prec[0] < 0.01 else 24
    pts = np.stack([rng.uniform(0.05, 2.9, npts), rng.uniform(0, 2 * np.pi, npts), rng.uniform(0.05, 1.5, npts)], 1)
    for which in (0, 1):
        g = fid_tables.gamma_eval(which, (x1, x2, x3), pts)
        o = np.array([fid_oracle.gamma_integrand(which, *p, x1, x2, x3) for p in pts])
        scale = np.maximum(np.abs(o), 1e-300)
        err = np.abs(g - o) / scale
        print("gamma%d integrand prec=%s max rel err %.3e" % (which, prec, err.max()))
        assert err.max() < 1e-8


def test_gamma_batch_small_matches_oracle(fid_tables, fid_oracle, gpu_ctx):
    """Whole Gamma_0..3 of one triangle with a coarse Ogata grid (the oracle finishes in seconds)."""
    prec = (0.1, 2.0)
    tp.gamma_setup(prec[0], prec[1], gpu_ctx)
    fid_oracle.gamma_init(prec[0], prec[1])
    x = _tri()
    g, n = fid_tables.gamma_batch(x[None, :], reltol=1e-2)
    for i in range(4):
        o, no = fid_oracle.gamma(i, *x, reltol=1e-2)
        assert abs(g[0, i] - o) <= 1e-6 * abs(o), (i, g[0, i], o)


def test_gamma_isosceles_symmetry(fid_tables, gpu_ctx):
    """Im Gamma0 = 0 and Im Gamma1 = -Im Gamma2 for isosceles triangles
    (the sanity relations of classed_version/testGamma_analytical_bispectrum.cpp:200-226)."""
    tp.gamma_setup(0.1, 2.0, gpu_ctx)
    a = 5.0 * np.pi / 180 / 60
    g, _ = fid_tables.gamma_batch(np.array([[a, a, 0.6 * a]]), reltol=1e-3)
    g = g[0]
    print("isosceles Gamma:", g)
    assert abs(g[0].imag) <= 5e-3 * abs(g[0])
    assert abs(g[3].imag) <= 5e-3 * abs(g[3])
    assert abs(g[1].imag + g[2].imag) <= 5e-3 * (abs(g[1]) + abs(g[2]))


def test_gamma_fp32_form_within_1e3(fid_tables, fid_oracle, gpu_ctx):
    """FP32 Ogata terms (inside the wavenumber window) with FP64 accumulation: Gamma_0..3 of two
    triangles within 1e-3 of the FP64 path (itself within 1e-6 of the oracle), at the production Ogata grid."""
    tp.gamma_setup(1e-3, 5.0, gpu_ctx)
    rows = W.gamma_c2_triangles()
    x = rows[[555, 231], :3]
    g64, n64 = fid_tables.gamma_batch(x, reltol=1e-2)
    g32, n32 = fid_tables.gamma_batch(x, reltol=1e-2, mode=tp.MODE_FP32)
    rel = np.abs(g32 - g64) / np.abs(g64)
    print("Gamma FP32 vs FP64 max rel %.3e, points %d / %d" % (rel.max(), n32.sum(), n64.sum()))
    assert rel.max() < 1e-3
    with pytest.raises(tp.TpError):
        fid_tables.gamma_batch(x, reltol=1e-2, mode=tp.MODE_SURROGATE)


def test_gamma_orthocenter_convention(fid_tables, fid_oracle, gpu_ctx):
    """triangle_center "orthocenter" (classed_version/gamma.cpp:12-13) returns the natural components
    unrotated: the same integrals (equal point counts) and moduli as the centroid convention, and the
    phase that the rotation of gamma.cpp:624-643 removes (gamma1..3 conjugate the first angle)."""
    prec = (0.1, 2.0)
    tp.gamma_setup(prec[0], prec[1], gpu_ctx)
    x = _tri()
    gc, nc = fid_tables.gamma_batch(x[None, :], reltol=1e-2, center=tp.CENTER_CENTROID)
    go, no = fid_tables.gamma_batch(x[None, :], reltol=1e-2, center=tp.CENTER_ORTHOCENTER)
    assert np.array_equal(nc, no)
    assert np.allclose(np.abs(go), np.abs(gc), rtol=1e-13, atol=0)

    def angle(a, b, c):  # one_rotation_angle_otc(a, b, c)
        psi3 = np.arccos((a * a + b * b - c * c) / (2 * a * b))
        h3 = 0.5 * np.sqrt(2 * a * a + 2 * b * b - c * c)
        c2 = ((b * b - a * a) ** 2 - 4 * (a * b * np.sin(psi3)) ** 2) / (2 * h3 * c) ** 2
        s2 = (b * b - a * a) * a * b * np.sin(psi3) / (h3 * c) ** 2
        return np.arctan2(s2, c2) / 2

    x1, x2, x3 = x
    perms = [(x1, x2, x3), (x1, x2, x3), (x2, x3, x1), (x3, x1, x2)]
    for i, (a, b, c) in enumerate(perms):
        a1, a2, a3 = angle(b, c, a), angle(c, a, b), angle(a, b, c)
        if i > 0:
            a1 = -a1
        want = go[0, i] * np.exp(-2j * (a1 + a2 + a3))
        assert abs(want - gc[0, i]) <= 1e-12 * abs(gc[0, i]), (i, want, gc[0, i])
    with pytest.raises(tp.TpError):
        fid_tables.gamma_batch(x[None, :], reltol=1e-2, center=7)
    with pytest.raises(ValueError):
        tp.GammaCalculator(None, 0.1, 2.0, "incenter")


def test_gamma0_analytic_closed_form_on_gpu(gpu_ctx):
    """The device's Gamma_0 integrand with the reference's analytic test bispectrum (TP_MODE_ANALYTIC, z_max = 1,
    the grid of classed_version/testGamma_analytical_bispectrum.cpp) in the orthocenter convention: equal to
    the closed form ggg (gamma.cpp:83-94, 567-586) at the agreement the oracle reaches (tests/test_gamma_oracle.py),
    and to the oracle's value with the same cubature decisions."""
    from _oracle import Oracle
    from test_gamma_oracle import grid_triangle
    c = W.fiducial()
    c.zmax = 1.0
    nz = np.full(10, 1.0)
    tabs = tp.CosmologyTables([c], nz, 10, 1.0, gpu_ctx)
    tp.gamma_setup(0.005, 5.0, gpu_ctx)
    o = Oracle(c, nz, 10, 1.0)
    o.set_analytic(True)
    o.gamma_init(0.005, 5.0)
    for ijk, tol in (((8, 10, 2), 1e-6), ((7, 3, 10), 5e-4)):
        x = np.array(grid_triangle(*ijk))
        g, n = tabs.gamma_batch(x[None, :], reltol=1e-5, mode=tp.MODE_ANALYTIC, center=tp.CENTER_ORTHOCENTER)
        ggg = Oracle.ggg_analytic(*x)
        go, no = o.gamma(0, *x, reltol=1e-5, center=1)
        print("bin %s: GPU %s, closed form %s (rel %.2e), oracle rel %.2e, points %d / %d"
              % (ijk, g[0, 0], ggg, abs(g[0, 0] - ggg) / abs(ggg), abs(g[0, 0] - go) / abs(go), n[0, 0], no))
        assert abs(g[0, 0] - ggg) < tol * abs(ggg)
        assert abs(g[0, 0] - go) < 1e-8 * abs(go) and n[0, 0] == no


def test_gamma_production_grid_matches_oracle_fixture(fid_tables, gpu_ctx):
    """Gamma_0..3 of TreeCorr-grid triangles at the production Ogata grid (step 1e-3, 4999 zeros of J6 and J2)
    against the CPU FP64 oracle (tests/golden/gamma_c2_fp64.json, written offline: the oracle needs ~10 CPU
    minutes per triangle there): FP64 within 1e-4 of each component's modulus, the FP32 form within 1e-3.

    1e-4, not 1e-6: at this grid the integrand is ill-conditioned near phi = pi (ell3 -> 0), where one ulp of
    (z, phi, psi) moves the oracle's own integrand by up to 1e-1 relative, and the adaptive cubature's
    decisions follow those digits.  Moving x1 of triangle 555 by one ulp moves the oracle's own Gamma_0 by
    3.9e-5 and its point count from 62007 to 63855; the integrands themselves agree to 1e-8
    (test_gamma_integrand_matches_oracle) and whole Gamma_i at a coarser grid to 1e-6
    (test_gamma_batch_small_matches_oracle)."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gamma_c2_fp64.json")))
    assert fx["prec_h"] == 1e-3 and fx["prec_k_arg"] == 5.0 and fx["rows"]
    tp.gamma_setup(1e-3, 5.0, gpu_ctx)
    for row in fx["rows"]:
        x = np.array(row["x123_rad"])
        want = np.array([complex(a, b) for a, b in row["gamma_re_im"]])
        g64, n64 = fid_tables.gamma_batch(x[None, :], reltol=row["reltol"])
        g32, _ = fid_tables.gamma_batch(x[None, :], reltol=row["reltol"], mode=tp.MODE_FP32)
        e64 = np.abs(g64[0] - want) / np.abs(want)
        e32 = np.abs(g32[0] - want) / np.abs(want)
        print("triangle %d (rel tol %g): fp64 max rel %.2e, fp32 max rel %.2e, points %s vs %s"
              % (row["triangle"], row["reltol"], e64.max(), e32.max(), list(n64[0]), row["neval"]))
        assert e64.max() < 1e-4
        assert e32.max() < 1e-3



def test_gamma_points_beyond_one_launch(fid_tables, gpu_ctx):
    """A round of more points than one launch covers (one 256-thread block per point; a dispatch counts at most
    2^32 work-items, so the host splits a round into launches of 2^22 points): every point's value, on either
    side of a launch boundary, equals the value of the same point evaluated in a call of its own.  Batched
    Gamma rounds of 100 wide triangles exceed 2^24 points; without the split, the points beyond the first
    2^32 work-items of a round were never evaluated and kept the previous round's values."""
    tp.gamma_setup(0.1, 2.0, gpu_ctx)  # 20 Ogata terms per point: the big call stays cheap
    x = _tri()
    rng = np.random.default_rng(11)
    m = 997
    few = np.stack([rng.uniform(0.05, 2.9, m), rng.uniform(0, 2 * np.pi, m), rng.uniform(0.05, 1.5, m)], 1)
    n = (1 << 22) + 2 * m
    pts = np.ascontiguousarray(np.tile(few, (-(-n // m), 1))[:n])
    for which in (0, 1):
        big = fid_tables.gamma_eval(which, tuple(x), pts)
        ref = fid_tables.gamma_eval(which, tuple(x), few)
        assert np.all(np.isfinite(big))
        for lo in (0, (1 << 22) - m, 1 << 22, n - m):
            idx = np.arange(lo, min(lo + m, n))
            assert np.array_equal(big[idx], ref[idx % m]), (which, lo)
