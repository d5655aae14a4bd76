"""GPU parity of the Gaussian <Map^3> covariance (4-D form, T1, T2) against the CPU FP64 oracle.

The adaptive integrals take the oracle's decision sequence (equal point counts); values agree to
rounding (1e-10).  The shape-noise-only T1 is also checked against its closed form.
"""
import numpy as np
import pytest

import threepoint_amd as tp
from test_cov_oracle import t1_closed_form

pytestmark = pytest.mark.gpu
A = 2.9088820866e-4
AREA = (5 * 60 * A) ** 2  # a 5 x 5 deg^2 survey [rad^2]
NGAL = 30 / A ** 2        # 30 galaxies per arcmin^2 [rad^-2]


def test_cov_gauss_matches_oracle(fid_tables, fid_oracle):
    pairs = np.array([[4, 8, 8, 6, 6, 12], [4, 4, 4, 4, 4, 4], [8, 8, 16, 4, 16, 16]], dtype=float) * A
    v, e, n = fid_tables.map3_cov_gauss(pairs, AREA, reltol=1e-3)
    for i, p in enumerate(pairs):
        ov, oe, on = fid_oracle.map3_cov_gauss(p, AREA, 1e-3)
        assert n[i] == on, (i, n[i], on)
        assert abs(v[i] / ov - 1) < 1e-10


@pytest.mark.parametrize("flags,lmin", [(0, 0.0), (tp.PK_T17, 0.0), (tp.PK_SHAPENOISE_ONLY, 0.0), (tp.PK_T17, 20.0)])
def test_t1_total_matches_oracle(fid_tables, fid_oracle, flags, lmin):
    """lmin = 20 cuts regions partly (points with l3 <= lmin are zero): the device's per-region
    Pell cache skips those points' Pell and must still give the oracle's values and point counts."""
    pairs = np.array([[4, 8, 8, 8, 8, 8], [6, 6, 12, 4, 8, 16], [8, 8, 8, 6, 12, 6]], dtype=float) * A
    v, n = fid_tables.map3_cov_t1(pairs, AREA, lmin, 0.3, NGAL, flags, reltol=1e-3)
    for i, p in enumerate(pairs):
        ov, on = fid_oracle.t1_total(p[:3], p[3:], AREA, lmin, 0.3, NGAL, flags, 1e-3)
        assert n[i] == on, (i, n[i], on)
        assert abs(v[i] / ov - 1) < 1e-10


def test_t1_shape_noise_closed_form_on_gpu(fid_tables):
    th = np.array([3.0, 6.0, 12.0]) * A
    pairs = np.repeat(th[:, None], 6, axis=1)
    v, n = fid_tables.map3_cov_t1(pairs, AREA, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY, reltol=1e-7)
    P = 0.5 * 0.09 / NGAL
    for i, t in enumerate(th):
        assert abs(v[i] / (6 * t1_closed_form(t, AREA, P)) - 1) < 1e-6


@pytest.mark.parametrize("flags", [tp.PK_T17, tp.PK_SHAPENOISE_ONLY])
@pytest.mark.parametrize("geometry,tx,ty", [(tp.GEOMETRY_CIRCLE, 300 * A, 0.0), (tp.GEOMETRY_SQUARE, 300 * A, 0.0),
                                            (tp.GEOMETRY_RECTANGLE, 300 * A, 180 * A)])
def test_t2_total_matches_oracle(fid_tables, fid_oracle, geometry, tx, ty, flags):
    """T2 of the finite geometries: the pairings (the third pair takes the reference's theta_1 == theta_3
    branch), the 1-D and 2-D integrals and their domains; equal point counts, values to rounding."""
    pairs = np.array([[4, 8, 8, 8, 8, 8], [6, 6, 12, 4, 8, 16], [8, 4, 8, 6, 12, 6]], dtype=float) * A
    v, n = fid_tables.map3_cov_t2(pairs, geometry, tx, ty, 1.0, 0.3, NGAL, flags, reltol=1e-4)
    for i, p in enumerate(pairs):
        ov, on = fid_oracle.t2_total(p[:3], p[3:], geometry, tx, ty, 1.0, 0.3, NGAL, flags, 1e-4)
        assert n[i] == on, (i, n[i], on)
        assert ov > 0 and abs(v[i] / ov - 1) < 1e-10, (i, v[i], ov)


@pytest.mark.parametrize("geometry,tx,ty", [(tp.GEOMETRY_CIRCLE, 60 * A, 0.0), (tp.GEOMETRY_SQUARE, 60 * A, 0.0),
                                            (tp.GEOMETRY_RECTANGLE, 60 * A, 40 * A)])
def test_t1_finite_geometries_match_oracle(fid_tables, fid_oracle, geometry, tx, ty):
    """T1 of the finite geometries (6-D with the window, the reference's tolerance 1e-2), constant spectrum:
    the cubature domain, the cuts, the window and the permutations.  The device's exp, sin and J1 and the
    C library's differ in the last bits now and then; in a heap of ~6e4 6-D regions that moves a few
    near-tied regions (24 on the circle's first pair, 2 on the square's), so these integrals are held
    to the cubature's own scale: 1e-3 in value (the tolerance is 1e-2), 1% in points."""
    pairs = np.array([[8, 8, 8, 8, 8, 8], [4, 8, 8, 6, 6, 12]], dtype=float) * A
    v, n = fid_tables.map3_cov_t1_geometry(pairs, geometry, tx, ty, 1.0, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY)
    for i, p in enumerate(pairs):
        ov, on = fid_oracle.t1_total_finite(p[:3], p[3:], geometry, tx, ty, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY, 1e-2)
        assert abs(n[i] / on - 1) < 1e-2, (i, n[i], on)
        assert ov > 0 and abs(v[i] / ov - 1) < 1e-3, (i, v[i], ov)


def test_t2_infinite_survey_is_zero_and_bad_arguments(fid_tables):
    pairs = np.full((2, 6), 4 * A)
    v, n = fid_tables.map3_cov_t2(pairs, tp.GEOMETRY_INFINITE, 0.0, 0.0, 1.0, 0.3, NGAL, tp.PK_T17)
    assert np.all(v == 0) and np.all(n == 0)
    with pytest.raises(tp.TpError):
        fid_tables.map3_cov_t2(pairs, 5, 300 * A)
    with pytest.raises(tp.TpError):
        fid_tables.map3_cov_t2(pairs, tp.GEOMETRY_RECTANGLE, 300 * A, 0.0)
    with pytest.raises(tp.TpError):
        fid_tables.map3_cov_t2(np.zeros((1, 6)), tp.GEOMETRY_CIRCLE, 300 * A)


def test_cov_rejects_bad_arguments(fid_tables):
    with pytest.raises(tp.TpError):
        fid_tables.map3_cov_gauss(np.zeros((1, 6)), AREA)
    with pytest.raises(tp.TpError):
        fid_tables.map3_cov_t1(np.full((1, 6), 4 * A), -1.0)


@pytest.mark.parametrize("flags", [0, tp.PK_T17, tp.PK_LINEAR])
def test_t1_fp32_form_within_1e3(fid_tables, fid_oracle, flags):
    """The FP32 form of T1 (FP32 Pell node values, FP64 sums, rule and heaps) against the FP64 path and
    the FP64 oracle, at the 1e-3 bar of the FP32 forms."""
    pairs = np.array([[4, 8, 8, 8, 8, 8], [6, 6, 12, 4, 8, 16], [2, 4, 16, 2, 2, 2]], dtype=float) * A
    v64, n64 = fid_tables.map3_cov_t1(pairs, AREA, 0.0, 0.3, NGAL, flags, reltol=1e-4)
    v32, n32 = fid_tables.map3_cov_t1(pairs, AREA, 0.0, 0.3, NGAL, flags, reltol=1e-4, mode=tp.MODE_FP32)
    r = np.abs(v32 / v64 - 1)
    print("T1 FP32 vs FP64 (flags %d): max rel %.3e, points %s vs %s" % (flags, r.max(), n32, n64))
    assert r.max() < 1e-3
    ov, _ = fid_oracle.t1_total(pairs[0, :3], pairs[0, 3:], AREA, 0.0, 0.3, NGAL, flags, 1e-4)
    assert abs(v32[0] / ov - 1) < 1e-3


def test_t1_fp32_shape_noise_only_is_exact(fid_tables):
    """With shape noise only no node is evaluated: the FP32 form gives the FP64 path's bits."""
    pairs = np.array([[3, 6, 12, 3, 6, 12], [4, 4, 4, 8, 8, 8]], dtype=float) * A
    a = fid_tables.map3_cov_t1(pairs, AREA, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY, reltol=1e-5)
    b = fid_tables.map3_cov_t1(pairs, AREA, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY, reltol=1e-5, mode=tp.MODE_FP32)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_t1_fp32_rejects_tabulated_pk(gpu_ctx, fid_nz):
    from threepoint_amd import workloads as W
    c = W.fiducial()
    kh = np.geomspace(1e-4, 1e2, 50)
    tabs = tp.CosmologyTables([c], fid_nz, 100, 3.0, gpu_ctx, pk=(kh, 1e4 * kh / (1 + (kh / 0.02) ** 3)))
    with pytest.raises(tp.TpError):
        tabs.map3_cov_t1(np.full((1, 6), 4 * A), AREA, mode=tp.MODE_FP32)


def test_batch_limits_end_runaway_integrals(gpu_ctx, fid_tables):
    """The batched cubatures' guards: a tiny region or round cap ends a call with TP_ECAPACITY (code 4)
    instead of letting it grow; slicing the host-rule rounds into small launches changes no bit."""
    pairs = np.array([[4, 8, 8, 6, 6, 12]], dtype=float) * A  # three 6-D integrals (permutations of 6, 6, 12)
    args = (pairs, tp.GEOMETRY_SQUARE, 60 * A, 0.0, 1.0, 0.0, 0.3, NGAL, tp.PK_SHAPENOISE_ONLY)
    v0, n0 = fid_tables.map3_cov_t1_geometry(*args)
    try:
        gpu_ctx.set_batch_limits(round_points=10000)
        v1, n1 = fid_tables.map3_cov_t1_geometry(*args)
        assert np.array_equal(v0, v1) and np.array_equal(n0, n1)
        gpu_ctx.set_batch_limits(max_regions=50)
        with pytest.raises(tp.TpError, match=r"code 4\).*region limit"):
            fid_tables.map3_cov_t1_geometry(*args)
        with pytest.raises(tp.TpError, match=r"code 4\).*region limit"):
            fid_tables.map2_batch(np.array([4 * A]))
        gpu_ctx.set_batch_limits(max_rounds=3)
        with pytest.raises(tp.TpError, match=r"code 4\).*round limit"):
            fid_tables.map3_batch(np.array([[4 * A, 4 * A, 4 * A]]))
        with pytest.raises(tp.TpError, match=r"code 4\).*round limit"):
            fid_tables.pkappa_batch(np.array([100.0]), reltol=1e-12)
    finally:
        gpu_ctx.set_batch_limits()
    v2, n2 = fid_tables.map3_cov_t1_geometry(*args)
    assert np.array_equal(v0, v2) and np.array_equal(n0, n2)
