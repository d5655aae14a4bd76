"""GPU parity of the <Map^3> path against the CPU FP64 oracle.

Tolerances (relative): FP64 integrand / B_kappa 1e-10 (different but exact
rearrangements of the same FP64 arithmetic); plain FP64 <Map^3> 1e-6; FP32
forms 1e-3 -- the bars the task sets for the plain and FP32 paths.
"""
import json
import os

import numpy as np
import pytest

import threepoint_amd as tp
from threepoint_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def test_tables_match_oracle(fid_tables, fid_oracle):
    g = fid_tables.tables(0)
    o = fid_oracle.tables()
    for k in ["f_K", "g", "D1", "r_sigma", "n_eff", "ncur"]:
        a, b = g[k], o[k]
        scale = np.maximum(np.abs(b), 1e-12)
        assert np.max(np.abs(a - b) / scale) < 1e-9, (k, np.max(np.abs(a - b) / scale))
    assert abs(g["norm"] / o["norm"] - 1) < 1e-12


def test_tables_match_fixture(fid_tables):
    fx = json.load(open(os.path.join(GOLD, "tables_fid.json")))
    g = fid_tables.tables(0)
    for k in ["f_K", "g", "D1", "r_sigma", "n_eff"]:
        b = np.array(fx[k])
        assert np.max(np.abs(g[k] - b) / np.maximum(np.abs(b), 1e-12)) < 1e-9, k


def squeeze(ell):
    ell = np.asarray(ell).reshape(-1, 3)
    return ell.max(1) / ell.min(1)


def test_bkappa_fp64_matches_oracle(fid_tables):
    """Same FP64 arithmetic, rearranged: differences are rounding-level (~1e-15), except that
    in squeezed triangles the cancelling 1/k dipole terms of the F2 P P sum multiply them by
    l_max/l_min (the conditioning of B itself), hence the bound 1e-12 + 2e-14 l_max/l_min."""
    fx = np.load(os.path.join(GOLD, "bkappa_c1_2k.npz"))
    g = fid_tables.bkappa(fx["ell"], mode=tp.MODE_FP64)
    r = rel(g, fx["bkappa"])
    sq = squeeze(fx["ell"])
    print("bkappa fp64 max rel err %.3e (ratio<=100: %.3e)" % (r.max(), r[sq <= 100].max()))
    assert np.all(r <= 1e-12 + 2e-14 * sq)


def test_bkappa_fp32_within_1e3(fid_tables):
    """FP32 B_kappa: ~1e-6 typical; in squeezed triangles the 1/k dipole terms of the
    three F2 P P products cancel, which multiplies the FP32 P(k) error by l_max/l_min.
    The bar 1e-3 is checked where l_max/l_min <= 100 (all C0/C1 filter pairs lie there);
    beyond, the error is bounded by 2e-6 * l_max/l_min."""
    fx = np.load(os.path.join(GOLD, "bkappa_c1_2k.npz"))
    ell = fx["ell"]
    g = fid_tables.bkappa(ell, mode=tp.MODE_FP32)
    r = rel(g, fx["bkappa"])
    sq = ell.max(1) / ell.min(1)
    print("bkappa fp32 max rel err %.3e median %.3e; ratio<=100: max %.3e"
          % (r.max(), np.median(r), r[sq <= 100].max()))
    assert r[sq <= 100].max() < 1e-3
    assert np.all(r <= 2e-6 * np.maximum(sq, 100))


def test_bkappa_dump_bit_stable(fid_tables):
    _, ell = W.bkappa_c1_points(200000)
    a = fid_tables.bkappa(ell)
    b = fid_tables.bkappa(ell)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_bkappa_zero_guards(fid_tables):
    ell = np.array([[0.0, 10.0, 10.0], [10.0, 0.0, 10.0], [10.0, 10.0, 0.0], [10.0, 10.0, 10.0]])
    b = fid_tables.bkappa(ell)
    assert np.all(b[:3] == 0.0) and b[3] > 0


def test_map3_integrand_matches_oracle(fid_tables, fid_oracle):
    rng = np.random.default_rng(11)
    n = 512
    x = np.stack([np.exp(rng.uniform(0, 11, n)), np.exp(rng.uniform(0, 11, n)), rng.uniform(0, 2 * np.pi, n)], 1)
    # include the exact l3 = 0 configuration (l1 = l2, phi = pi)
    x[0] = [100.0, 100.0, np.pi]
    th = W.to_rad([2.0, 4.0, 8.0])
    g = fid_tables.map3_integrand(x, th)
    o = fid_oracle.map3_integrand(x, th)
    assert g[0] == 0.0 and o[0] == 0.0
    m = np.abs(o) > 1e-280
    l3 = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + 2 * x[:, 0] * x[:, 1] * np.cos(x[:, 2]))
    sq = np.maximum(x[:, 0], x[:, 1]) / np.minimum(np.minimum(x[:, 0], x[:, 1]), np.maximum(l3, 1e-300))
    assert np.all(rel(g[m], o[m]) <= 1e-12 + 2e-14 * sq[m])


def _c0_fixture():
    fx = json.load(open(os.path.join(GOLD, "map3_c0_fp64.json")))
    th = np.array([r["theta_arcmin"] for r in fx["rows"]])
    return th, np.array([r["map3"] for r in fx["rows"]]), np.array([r["neval"] for r in fx["rows"]])


def test_map3_c0_batch_fp64_within_1e6(fid_tables):
    th, ref, nref = _c0_fixture()
    v, e, n = fid_tables.map3_batch(W.to_rad(th), reltol=1e-4, mode=tp.MODE_FP64)
    r = rel(v, ref)
    print("C0 fp64 batched: max rel err %.3e; points gpu/oracle %s" % (r.max(), (n / nref).round(3)))
    assert r.max() < 1e-6


def test_map3_single_equals_batch(fid_tables):
    th = W.to_rad(np.array([[2.0, 4.0, 8.0], [16.0, 16.0, 16.0]]))
    vb, _, nb = fid_tables.map3_batch(th, reltol=1e-4)
    for i in range(2):
        vs, _, ns = fid_tables.map3_single(th[i], 1e-4)
        assert ns == nb[i]
        assert abs(vs - vb[i]) <= 1e-12 * abs(vb[i])


def test_map3_c0_fp32_within_1e3(fid_tables):
    th, ref, _ = _c0_fixture()
    v, e, n = fid_tables.map3_batch(W.to_rad(th), reltol=1e-4, mode=tp.MODE_FP32)
    r = rel(v, ref)
    print("C0 fp32 batched: max rel err %.3e" % r.max())
    assert r.max() < 1e-3


def test_map3_live_oracle_two_triples(fid_tables, fid_oracle):
    th = W.to_rad(np.array([[8.0, 8.0, 16.0], [4.0, 16.0, 16.0]]))
    v, _, n = fid_tables.map3_batch(th, reltol=1e-3)
    for i in range(2):
        vo, _, no = fid_oracle.map3(th[i], 1e-3)
        assert abs(v[i] - vo) <= 1e-6 * abs(vo)


def test_map3_analytic_matches_oracle(fid_tables, fid_oracle):
    """The reference's analytic test bispectrum through the same cubature (machinery check)."""
    th = W.to_rad(np.array([[1.0, 1.0, 1.0], [1.0, 2.0, 3.0]]))
    v, _, _ = fid_tables.map3_batch(th, reltol=1e-6, mode=tp.MODE_ANALYTIC)
    fid_oracle.set_analytic(True)
    try:
        for i in range(2):
            vo, _, _ = fid_oracle.map3(th[i], 1e-6)
            assert abs(v[i] - vo) <= 1e-9 * abs(vo), (v[i], vo)
    finally:
        fid_oracle.set_analytic(False)


def test_map3_analytic_matches_closed_form(fid_tables):
    """The device's <Map^3> with the analytic test bispectrum against its closed form (exact Gaussian moments,
    tests/test_map3_closed_form.py) at the cubature's tolerance."""
    from test_map3_closed_form import TRIPLES, map3_analytic_closed_form
    th = W.to_rad(np.array(TRIPLES))
    v, _, _ = fid_tables.map3_batch(th, reltol=1e-6, mode=tp.MODE_ANALYTIC)
    for i in range(len(TRIPLES)):
        c = map3_analytic_closed_form(th[i])
        print("theta %s: GPU %.12e closed form %.12e (rel %.2e)" % (TRIPLES[i], v[i], c, abs(v[i] - c) / abs(c)))
        assert abs(v[i] - c) <= 1e-6 * abs(c)


def test_nonfinite_theta_rejected(fid_tables):
    with pytest.raises(tp.TpError):
        fid_tables.map3_batch(np.array([[0.0, 1e-3, 1e-3]]))


def test_launch_shapes_bit_identical(gpu_ctx, fid_tables):
    """Every launch shape of the region kernel (1, 2, 4, 8 waves per point; 16, 32, 64 = the
    region-cached kernel with 1, 2, 4 groups) gives the same bits, so the autotuner's choice never
    changes a result."""
    import ctypes
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    try:
        for mode in (tp.MODE_FP64, tp.MODE_FP32):
            ref = None
            for lanes in (1, 2, 4, 8, 16, 32, 64):
                tp._check(tp.lib().tp_set_region_lanes(gpu_ctx.h, lanes), "tp_set_region_lanes")
                v, e, n = fid_tables.map3_batch(th, None, 1e-4, mode)
                if ref is None:
                    ref = (v, e, n)
                else:
                    assert np.array_equal(v, ref[0]) and np.array_equal(e, ref[1]) and np.array_equal(n, ref[2]), \
                        (mode, lanes)
    finally:
        tp._check(tp.lib().tp_set_region_lanes(gpu_ctx.h, 0), "tp_set_region_lanes")
    fid_tables.map3_batch(th[:2], None, 1e-4, tp.MODE_FP32)  # tuned shapes in use
    tab = (ctypes.c_int * 4)()
    tp._check(tp.lib().tp_region_lanes_table(gpu_ctx.h, tp.MODE_FP32, tab), "tp_region_lanes_table")
    print("tuned waves per point (FP32, by launch size):", list(tab))
    assert all(x in (1, 2, 4, 8, 16, 32, 64) for x in tab)


def test_given_pk_matches_oracle(gpu_ctx, fid_nz):
    """Tables, B_kappa and <Map^3> from a given linear P(k) (the cosmosis input,
    classed_version/bispectrum.cpp:81-89, 785-806) against the oracle's Pk_given branch."""
    from _oracle import Oracle
    eh = Oracle(W.fiducial(), fid_nz)
    k = np.logspace(-5, 4, 150)
    p = np.array([eh.scalar("linear_pk", kk) for kk in k]) * (1 + 0.05 * np.sin(3 * np.log(k)))  # not the fit
    o = Oracle(W.fiducial(), fid_nz, pk=(k, p))
    tabs = tp.CosmologyTables([W.fiducial()], fid_nz, 100, 3.0, gpu_ctx, pk=(k, p))
    ot, gt = o.tables(), tabs.tables(0)
    for name in ("r_sigma", "n_eff", "D1", "f_K", "g"):
        assert np.max(np.abs(gt[name] - ot[name])) <= 1e-9 * np.max(np.abs(ot[name])), name
    rng = np.random.default_rng(5)
    l1 = np.exp(rng.uniform(1, 9, 500)); l2 = np.exp(rng.uniform(1, 9, 500)); phi = rng.uniform(0, np.pi, 500)
    ell = np.column_stack([l1, l2, np.sqrt(l1 ** 2 + l2 ** 2 + 2 * l1 * l2 * np.cos(phi))])
    gb, ob = tabs.bkappa(ell, 0, tp.MODE_FP64), o.bkappa(ell)
    ratio = np.max(ell, axis=1) / np.min(ell, axis=1)
    assert np.all(np.abs(gb / ob - 1) < 1e-12 + 2e-14 * ratio)
    th = W.to_rad(W.triples(W.map3_c0_radii()))[[0, 5, 12, 19]]
    v, _, n = tabs.map3_batch(th, None, 1e-4, tp.MODE_FP64)
    v32, _, _ = tabs.map3_batch(th, None, 1e-4, tp.MODE_FP32)
    for j in range(th.shape[0]):
        vo, _, no = o.map3(th[j], 1e-4)
        assert abs(v[j] / vo - 1) < 1e-6, (j, v[j], vo)
        assert abs(v32[j] / vo - 1) < 1e-3
    with pytest.raises(RuntimeError):
        tabs.map3_batch(th, None, 1e-4, tp.MODE_SURROGATE)


@pytest.mark.parametrize("mode", [tp.MODE_FP64, tp.MODE_FP32])
def test_graph_rounds_bit_identical(fid_tables, gpu_ctx, mode):
    """Rounds replayed as one HIP graph per group (instantiated once, updated every round, with the
    host's decisions between replays) give the eager sequence's values, errors and point counts."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    v0, e0, n0 = fid_tables.map3_batch(th, None, 1e-4, mode)
    try:
        tp._check(tp.lib().tp_set_graph_rounds(gpu_ctx.h, 1), "tp_set_graph_rounds")
        for rep in range(2):
            v1, e1, n1 = fid_tables.map3_batch(th, None, 1e-4, mode)
            assert np.array_equal(v0, v1) and np.array_equal(e0, e1) and np.array_equal(n0, n1), rep
    finally:
        tp.lib().tp_set_graph_rounds(gpu_ctx.h, 0)


def test_bkappa_1e6_dump_bit_stable_and_checked(fid_oracle, tmp_path):
    """The 1e6-point B_kappa dump of the C1 workload: two runs write the same bytes; a seeded 4000-point
    sample of it matches the CPU oracle (FP64 bound of test_bkappa_fp64_matches_oracle) and the FP32
    dump is within 1e-3 where l_max / l_min <= 100."""
    from threepoint_amd import dump_bkappa
    a = dump_bkappa.dump(1000000, tp.MODE_FP64)
    b = dump_bkappa.dump(1000000, tp.MODE_FP64)
    assert a.shape == (1000000, 4) and a.tobytes() == b.tobytes()
    p = tmp_path / "bk.f64"
    a.tofile(p)
    assert np.array_equal(np.fromfile(p, dtype="<f8").reshape(-1, 4), a)
    idx = np.random.default_rng(3).choice(1000000, 4000, replace=False)
    x = a[idx]
    l3 = np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + 2 * x[:, 0] * x[:, 1] * np.cos(x[:, 2]))
    ell = np.column_stack([x[:, 0], x[:, 1], l3])
    o = fid_oracle.bkappa(ell)
    sq = ell.max(1) / ell.min(1)
    assert np.all(rel(x[:, 3], o) <= 1e-12 + 2e-14 * sq)
    f = dump_bkappa.dump(1000000, tp.MODE_FP32)
    r32 = rel(f[:, 3], a[:, 3])
    l3a = np.sqrt(a[:, 0] ** 2 + a[:, 1] ** 2 + 2 * a[:, 0] * a[:, 1] * np.cos(a[:, 2]))
    sqa = np.maximum(np.maximum(a[:, 0], a[:, 1]), l3a) / np.minimum(np.minimum(a[:, 0], a[:, 1]), l3a)
    print("1e6 dump: FP32 max rel %.3e (ratio <= 100: %.3e)" % (r32.max(), r32[sqa <= 100].max()))
    assert r32[sqa <= 100].max() < 1e-3


@pytest.mark.parametrize("mode", [tp.MODE_FP64, tp.MODE_FP32])
def test_zero_copy_rounds_bit_identical(fid_tables, gpu_ctx, mode):
    """Rounds that read regions from and write results to pinned host memory directly give the copied
    rounds' values, errors and point counts (C0, every round zero-copy, then none)."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    try:
        tp._check(tp.lib().tp_set_zero_copy_regions(gpu_ctx.h, 0), "zc")
        v0, e0, n0 = fid_tables.map3_batch(th, None, 1e-4, mode)
        tp._check(tp.lib().tp_set_zero_copy_regions(gpu_ctx.h, 1 << 40), "zc")
        v1, e1, n1 = fid_tables.map3_batch(th, None, 1e-4, mode)
        assert np.array_equal(v0, v1) and np.array_equal(e0, e1) and np.array_equal(n0, n1)
    finally:
        tp.lib().tp_set_zero_copy_regions(gpu_ctx.h, 2048)  # the default


@pytest.mark.parametrize("mode", [tp.MODE_FP64, tp.MODE_FP32])
def test_dma_and_in_place_results_bit_identical(fid_tables, gpu_ctx, mode):
    """A round below the serial size writes its region results to pinned host memory in place; a larger
    round copies them out by DMA.  With the serial size at one region every round above the zero-copy size
    takes the DMA path (and the two groups' rounds run one after the other): values, errors and point
    counts are those of the default (C0)."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    ref = fid_tables.map3_batch(th, None, 1e-4, mode)
    try:
        tp._check(tp.lib().tp_set_serial_regions(gpu_ctx.h, 1), "serial")
        dma = fid_tables.map3_batch(th, None, 1e-4, mode)
    finally:
        tp.lib().tp_set_serial_regions(gpu_ctx.h, 0)  # the default
    for a, b in zip(ref, dma):
        assert np.array_equal(a, b)


def test_batch_groups_bit_identical(fid_tables, gpu_ctx):
    """One to four integral groups in flight (each on its own stream) give the default's values, errors
    and point counts in the FP64 and FP32 modes (C0)."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    try:
        for mode in (tp.MODE_FP64, tp.MODE_FP32):
            ref = fid_tables.map3_batch(th, None, 1e-4, mode)
            for g in (1, 3, 4):
                tp._check(tp.lib().tp_set_batch_groups(gpu_ctx.h, g), "groups")
                got = fid_tables.map3_batch(th, None, 1e-4, mode)
                for a, b in zip(ref, got):
                    assert np.array_equal(a, b), (mode, g)
            tp._check(tp.lib().tp_set_batch_groups(gpu_ctx.h, 2), "groups")
    finally:
        tp.lib().tp_set_batch_groups(gpu_ctx.h, 2)  # the default


def test_host_settings_bit_identical(gpu_ctx, fid_tables):
    """Host threads, kept or fresh region pools, and a reused integrator set change no bit of a result."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    try:
        for mode in (tp.MODE_FP64, tp.MODE_FP32):
            ref = fid_tables.map3_batch(th, None, 1e-4, mode)
            again = fid_tables.map3_batch(th, None, 1e-4, mode)  # the kept integrators reused
            gpu_ctx.set_host_threads(1)
            one = fid_tables.map3_batch(th, None, 1e-4, mode)
            gpu_ctx.set_host_threads(0)
            gpu_ctx.set_keep_pool_bytes(0)  # fresh pools every call
            fresh = fid_tables.map3_batch(th, None, 1e-4, mode)
            small = fid_tables.map3_batch(th[:3], None, 1e-4, mode)  # then a smaller batch
            gpu_ctx.set_keep_pool_bytes(1 << 34)
            for r in (again, one, fresh):
                for a, b in zip(ref, r):
                    assert np.array_equal(a, b), mode
            for a, b in zip(ref, small):
                assert np.array_equal(a[:3], b), mode
    finally:
        gpu_ctx.set_host_threads(0)
        gpu_ctx.set_keep_pool_bytes(1 << 34)


def test_round_log_and_timing_stats(gpu_ctx, fid_tables):
    """The per-round trace accounts for every launch and region, and the device-busy time (union of the
    two groups' intervals) never exceeds the summed launch durations."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    gpu_ctx.set_round_log(True)
    try:
        fid_tables.map3_batch(th, None, 1e-4, tp.MODE_FP32)
        st = gpu_ctx.stats()
        L = gpu_ctx.round_log()
    finally:
        gpu_ctx.set_round_log(False)
    assert L.shape == (st["launches"], 9)
    assert np.all(L[:, 8] >= L[:, 7]) and np.all(L[:, 7] >= 0)
    assert int(L[:, 1].sum()) == st["regions"] and st["points"] == 33 * st["regions"]
    assert set(np.unique(L[:, 0]).astype(int)) <= {0, 1}
    assert np.all(L[:, 3] >= L[:, 2]) and np.all(L[:, 5] >= L[:, 4]) and np.all(L[:, 4] >= L[:, 3])
    assert 0 < st["integrand_ms"] <= st["integrand_sum_ms"] * (1 + 1e-6) + 1e-3
    assert st["kernel_ms"] >= st["integrand_ms"] - 1e-3


def _single_z_ok(r, sq):
    """Single-z bispectra: rounding-level agreement (1e-12) where the triangle is not squeezed; in squeezed
    triangles the three F2 P P terms of eq. B5 cancel to ~1/squeeze of their size, so the rounding
    differences of the per-wavenumber factors (the device's log-space P_lin against pow, its merged
    reciprocals) are amplified by the squeeze: the B_kappa bound 1e-12 + 2e-14 k_max/k_min."""
    print("  max rel err / squeeze = %.3e" % np.max(r / sq))
    return np.all(r[sq <= 100] <= 1e-12) and np.all(r <= 1e-12 + 2e-14 * sq)


def _single_z_points(n, seed):
    """Random (k1, k2, k3) triangles [h/Mpc] over 1e-3..1e2 with random redshifts in (0, z_max]."""
    rng = np.random.default_rng(seed)
    k1 = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), n))
    k2 = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), n))
    c = rng.uniform(-1, 1, n)
    k3 = np.sqrt(k1 * k1 + k2 * k2 + 2 * k1 * k2 * c)
    z = rng.uniform(0, 3.0, n)
    z[:4] = [3.0, 1e-6, 0.03, 2.97]  # the last table interval, the first one, and grid points
    return np.stack([k1, k2, k3, z], 1)


def test_bispec_eval_matches_oracle(fid_tables, fid_oracle):
    """tp_bispec_eval (bispec at z through the reference's table-index rule) against the oracle's bispec at
    the same interval, in the FP64 bound of the B_kappa test (1e-12 + 2e-14 k_max/k_min)."""
    x = _single_z_points(400, 11)
    g = fid_tables.bispec(x)
    o = np.array([fid_oracle.scalar("bispec", *row) for row in x])
    r = rel(g, o)
    sq = squeeze(x[:, :3])
    print("bispec max rel err %.3e (squeeze <= 100: %.3e)" % (r.max(), r[sq <= 100].max()))
    assert _single_z_ok(r, sq)


def test_bispec_eval_idx_matches_oracle(fid_tables, fid_oracle):
    """The reference's 6-argument bispec(k1, k2, k3, z, idx, didx): the caller's table interval (also one
    that does not match z), and z > 10, where the reference returns the tree-level bispectrum."""
    from _oracle import lib
    x = _single_z_points(300, 12)
    rng = np.random.default_rng(13)
    idx = rng.integers(0, 99, x.shape[0]).astype(np.float64)  # n_z = 100: idx in [0, 98]
    didx = rng.uniform(0, 1, x.shape[0])
    didx[:3] = [0.0, 1.0, 0.5]
    x[-20:, 3] = rng.uniform(10.5, 20.0, 20)  # the tree-level branch
    rows = np.column_stack([x, idx, didx])
    g = fid_tables.bispec_idx(rows)
    o = np.array([lib().orc_bispec_idx(fid_oracle.h, *r[:4], int(r[4]), r[5]) for r in rows])
    r = rel(g, o)
    print("bispec(idx) max rel err %.3e (tree branch %.3e)" % (r.max(), r[-20:].max()))
    assert _single_z_ok(r, squeeze(x[:, :3]))
    for bad in (-1.0, 99.0, 2.5):
        with pytest.raises(tp.TpError):
            fid_tables.bispec_idx(np.array([[1.0, 1.0, 1.0, 1.0, bad, 0.5]]))


def test_integrand_bkappa_eval_matches_oracle(fid_tables, fid_oracle):
    """tp_integrand_bkappa_eval against the oracle's integrand_bkappa, with its zero guards (z = 0, l3 = 0)."""
    rng = np.random.default_rng(14)
    n = 400
    l1 = np.exp(rng.uniform(1, 11, n))
    l2 = np.exp(rng.uniform(1, 11, n))
    l3 = np.sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * rng.uniform(-1, 1, n))
    z = rng.uniform(0, 3.0, n)
    z[0], l3[1], z[2] = 0.0, 0.0, 3.0
    x = np.stack([z, l1, l2, l3], 1)
    g = fid_tables.integrand_bkappa(x)
    o = np.array([fid_oracle.scalar("integrand_bkappa", *row) for row in x])
    assert g[0] == 0 and g[1] == 0 and o[0] == 0 and o[1] == 0
    r = rel(g[2:], o[2:])
    print("integrand_bkappa max rel err %.3e" % r.max())
    assert _single_z_ok(r, squeeze(x[2:, 1:]))


def test_map3_c1_all_triples_match_oracle_fixture(fid_tables):
    """Every one of the 220 triples of 10 log radii 0.5-32' (tests/golden/map3_c1_fp64.json: the CPU FP64
    oracle at rel tol 1e-4) through the batched engine: FP64 within 1e-6 with the oracle's point counts,
    the FP32 form within 1e-3, and the exact zeros (first region underflowing) kept."""
    fx = json.load(open(os.path.join(GOLD, "map3_c1_fp64.json")))
    th = W.to_rad(np.array([r["theta_arcmin"] for r in fx["rows"]]))
    ref = np.array([r["map3"] for r in fx["rows"]])
    nref = np.array([r["neval"] for r in fx["rows"]])
    assert th.shape == (220, 3)
    nz = ref != 0
    v64, _, n64 = fid_tables.map3_batch(th, reltol=1e-4, mode=tp.MODE_FP64)
    v32, _, _ = fid_tables.map3_batch(th, reltol=1e-4, mode=tp.MODE_FP32)
    r64, r32 = rel(v64[nz], ref[nz]), rel(v32[nz], ref[nz])
    print("C1 vs oracle: fp64 max rel %.3e (points equal: %d/220), fp32 max rel %.3e, zeros %d"
          % (r64.max(), int((n64 == nref).sum()), r32.max(), int((~nz).sum())))
    assert np.all(v64[~nz] == 0) and np.all(v32[~nz] == 0)
    assert r64.max() < 1e-6 and np.array_equal(n64, nref)
    assert r32.max() < 1e-3
