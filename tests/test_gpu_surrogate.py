"""The learned surrogate on the GPU: gradients vs the CPU oracle, training, deployment."""
import ctypes
import os
import sys

import numpy as np
import pytest

import threepoint_amd as tp
from threepoint_amd import workloads as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402

pytestmark = pytest.mark.gpu
_fp = ctypes.POINTER(ctypes.c_float)


def _gpu_loss_grad(ctx, p, X, T):
    p32 = np.ascontiguousarray(p, dtype=np.float32)
    tp._check(tp.lib().tp_mlp_set_params(ctx.h, p32.ctypes.data_as(_fp)), "set")
    X32 = np.ascontiguousarray(X, dtype=np.float32)
    T32 = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(MO.NPARAM, dtype=np.float32)
    loss = ctypes.c_double()
    tp._check(tp.lib().tp_mlp_loss_grad(ctx.h, X.shape[0], X32.ctypes.data_as(_fp), T32.ctypes.data_as(_fp),
                                        ctypes.byref(loss), g.ctypes.data_as(_fp)), "loss_grad")
    return loss.value, g.astype(np.float64)


@pytest.mark.parametrize("M", [64, 512])
def test_mlp_gradients_match_cpu_oracle(gpu_ctx, M):
    rng = np.random.default_rng(M)
    d = {}
    for name, shp in MO.SHAPES:
        fan_in = shp[1] if len(shp) == 2 else (MO.H if name == "W5" else 1)
        d[name] = rng.uniform(-1, 1, shp) * (np.sqrt(6.0 / fan_in) if name.startswith("W") else 0.1)
    p = MO.pack(d).astype(np.float32).astype(np.float64)
    X = rng.uniform(-1, 1, (M, MO.NIN)).astype(np.float32).astype(np.float64)
    T = rng.normal(0, 1, M).astype(np.float32).astype(np.float64)
    lg, gg = _gpu_loss_grad(gpu_ctx, p, X, T)
    lo, go = MO.loss_grad(p, X, T, bf=True)
    assert abs(lg / lo - 1) < 1e-4
    G, O = MO.unpack(gg), MO.unpack(go)
    for name, _ in MO.SHAPES:
        err = np.linalg.norm(G[name] - O[name]) / max(np.linalg.norm(O[name]), 1e-30)
        print("grad %s rel err %.3e" % (name, err))
        assert err < 2e-3, name


def test_training_learns(gpu_ctx):
    from threepoint_amd.train_surrogate import train
    r = train(samples=1 << 19, epochs=2, batch=4096, lr=2e-3, lr_final=1e-4, out=None, ctx=gpu_ctx)
    print(r)
    assert r["val_rmse_lnB"] < 0.2


@pytest.mark.parametrize("kind", [tp.SUR_INFER_HILO_BF16, tp.SUR_INFER_FP16])
def test_fused_inference_matches_cpu_oracle(gpu_ctx, tmp_path, kind):
    """tp_surrogate_predict (fused MFMA kernels: hi/lo-BF16 or FP16 activations) against the numpy forward
    pass with the same roundings."""
    rng = np.random.default_rng(3)
    d = {}
    for name, shp in MO.SHAPES:
        fan_in = shp[1] if len(shp) == 2 else (MO.H if name == "W5" else 1)
        d[name] = rng.uniform(-1, 1, shp) * (np.sqrt(6.0 / fan_in) if name.startswith("W") else 0.1)
    p = MO.pack(d).astype(np.float32).astype(np.float64)
    lo, hi = [0.2, 0.7, 0.6, -1.2, 0.92], [0.4, 0.95, 0.8, -0.8, 1.0]
    fn = str(tmp_path / "m.bin")
    MO.write_model(fn, (2.0, 11.0), lo, hi, -30.0, 4.0, p)
    tp._check(tp.lib().tp_surrogate_load(gpu_ctx.h, fn.encode()), "load")
    gpu_ctx.set_surrogate_inference(kind)
    n = 1000
    l1, l2 = np.exp(rng.uniform(2, 11, n)), np.exp(rng.uniform(2, 11, n))
    l3 = np.sqrt(l1 ** 2 + l2 ** 2 + 2 * l1 * l2 * np.cos(rng.uniform(0, np.pi, n)))
    cp = np.stack([rng.uniform(a, b, n) for a, b in zip(lo, hi)], 1)
    raw = np.ascontiguousarray(np.concatenate([np.stack([l1, l2, l3], 1), cp], 1))
    y = np.zeros(n)
    try:
        tp._check(tp.lib().tp_surrogate_predict(gpu_ctx.h, n, tp._d(raw), tp._d(y)), "predict")
    finally:
        gpu_ctx.set_surrogate_inference(tp.SUR_INFER_FP16)
    m = MO.read_model(fn)
    fwd = MO.forward_hilo if kind == tp.SUR_INFER_HILO_BF16 else MO.forward_fp16
    ref = fwd(MO.unpack(m["p"]), MO.encode(m, l1, l2, l3, cp)) * m["sd"] + m["mean"]
    err = np.abs(y - ref)
    print("fused inference (kind %d) max abs err (ln B) %.3e" % (kind, err.max()))
    assert err.max() < 2e-3


def test_fp16_and_hilo_inference_agree_on_the_trained_model(gpu_ctx):
    """The shipped network through both inference kernels: the FP16 activations' single rounding moves ln B by
    about 1e-4 at most against the ~16-bit hi/lo activations."""
    _load_model(gpu_ctx)
    rng = np.random.default_rng(11)
    c = W.fiducial()
    l1, l2 = np.exp(rng.uniform(2.5, 10.5, 8192)), np.exp(rng.uniform(2.5, 10.5, 8192))
    l3 = np.sqrt(l1 ** 2 + l2 ** 2 + 2 * l1 * l2 * np.cos(rng.uniform(0, np.pi, 8192)))
    keep = (l3 > np.exp(2.2)) & (l3 < np.exp(10.8))  # inside the training box
    l1, l2, l3 = l1[keep], l2[keep], l3[keep]
    n = l1.shape[0]
    cp = np.tile([c.om, c.sigma8, c.h, c.w, c.ns], (n, 1))
    raw = np.ascontiguousarray(np.concatenate([np.stack([l1, l2, l3], 1), cp], 1))
    y = {}
    try:
        for kind in (tp.SUR_INFER_HILO_BF16, tp.SUR_INFER_FP16):
            gpu_ctx.set_surrogate_inference(kind)
            y[kind] = np.zeros(n)
            tp._check(tp.lib().tp_surrogate_predict(gpu_ctx.h, n, tp._d(raw), tp._d(y[kind])), "predict")
    finally:
        gpu_ctx.set_surrogate_inference(tp.SUR_INFER_FP16)
    d = np.abs(y[tp.SUR_INFER_FP16] - y[tp.SUR_INFER_HILO_BF16])
    print("FP16 vs hi/lo inference: max |d ln B| %.2e, rms %.2e" % (d.max(), np.sqrt((d ** 2).mean())))
    assert d.max() < 1e-3


def test_corrupted_model_rejected(gpu_ctx, tmp_path):
    fn = tmp_path / "bad.bin"
    fn.write_bytes(b"TPMLP002" + b"\0" * 100)
    assert tp.lib().tp_surrogate_load(gpu_ctx.h, str(fn).encode()) != 0


MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "threepoint_amd", "models",
                     "bkappa_mlp.bin")


def _load_model(ctx):
    tp._check(tp.lib().tp_surrogate_load(ctx.h, MODEL.encode()), "tp_surrogate_load")


def _random_triangles(n, seed):
    rng = np.random.default_rng(seed)
    l1 = np.exp(rng.uniform(3.0, 10.0, 4 * n))
    l2 = np.exp(rng.uniform(3.0, 10.0, 4 * n))
    phi = rng.uniform(0.0, np.pi, 4 * n)
    l3 = np.sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * np.cos(phi))
    keep = (l3 > np.exp(2.2)) & (l3 < np.exp(10.8))
    return np.column_stack([l1, l2, l3])[keep][:n]


def test_trained_network_gain_over_baseline(gpu_ctx, fid_tables):
    """What the network adds: the surrogate is B_8 exp(net), B_8 the 8-node FP32 Limber baseline.
    Against the exact FP64 B_kappa the trained network must cut the baseline's log error tenfold."""
    _load_model(gpu_ctx)
    ell = _random_triangles(4096, 0xB0)
    exact = fid_tables.bkappa(ell, 0, tp.MODE_FP64)
    sur = fid_tables.surrogate_bkappa(ell)
    c = W.fiducial()
    in8 = np.column_stack([ell, np.tile([c.om, c.sigma8, c.h, c.w, c.ns], (ell.shape[0], 1))])
    y = np.zeros(ell.shape[0])
    tp._check(tp.lib().tp_surrogate_predict(gpu_ctx.h, ell.shape[0], tp._d(np.ascontiguousarray(in8)), tp._d(y)),
              "tp_surrogate_predict")
    base = sur / np.exp(y)
    e_sur = np.log(sur / exact)
    e_base = np.log(base / exact)
    rms_sur, rms_base = np.sqrt(np.mean(e_sur ** 2)), np.sqrt(np.mean(e_base ** 2))
    print("log-error rms: trained network %.2e, baseline alone %.2e; max %.2e vs %.2e" %
          (rms_sur, rms_base, np.abs(e_sur).max(), np.abs(e_base).max()))
    assert rms_sur < 5e-3
    assert rms_sur * 10 < rms_base


def test_surrogate_mode_gain_in_map3(gpu_ctx, fid_tables):
    """<Map^3> of the 20 C0 triples: with the trained network more than half of the points are served by it
    (each integral's first 1024 regions run exact: about a third of C0's ~3000 regions per integral) and no
    round falls back; with an untrained network of the same shape the checks reject it and the rounds fall
    back to exact B_kappa.  Both stay within 1e-3 of the FP64 path."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    ref, _, _ = fid_tables.map3_batch(th, None, 1e-4, tp.MODE_FP64)
    nz = ref != 0
    _load_model(gpu_ctx)
    v, _, _ = fid_tables.map3_batch(th, None, 1e-4, tp.MODE_SURROGATE)
    st = gpu_ctx.stats()
    frac = st["surrogate_points"] / st["points"]
    err = np.abs(v[nz] / ref[nz] - 1).max()
    print("trained: surrogate share %.3f, fallback rounds %d, max rel err %.2e" % (frac, st["fallback_batches"], err))
    assert frac > 0.5 and st["fallback_batches"] == 0 and err < 1e-3
    rng = np.random.default_rng(7)
    p = (rng.standard_normal(MO.NPARAM) * 0.05).astype(np.float32)
    tp._check(tp.lib().tp_mlp_set_params(gpu_ctx.h, p.ctypes.data_as(_fp)), "set")
    v2, _, _ = fid_tables.map3_batch(th, None, 1e-4, tp.MODE_SURROGATE)
    st2 = gpu_ctx.stats()
    frac2 = st2["surrogate_points"] / st2["points"]
    err2 = np.abs(v2[nz] / ref[nz] - 1).max()
    print("untrained: surrogate share %.3f, fallback rounds %d, max rel err %.2e" % (frac2, st2["fallback_batches"], err2))
    _load_model(gpu_ctx)
    assert frac2 < 0.2 and st2["fallback_batches"] > 0 and err2 < 1e-3


def test_surrogate_dma_and_in_place_results_agree(gpu_ctx, fid_tables):
    """Surrogate rounds below the serial size write their region results to pinned host memory in place
    (the exact part, the network's part and, after a rejected check, the exact redo); larger rounds copy
    them out by DMA.  Every round on the DMA path (serial size one region) gives the default's values and
    point counts, with the trained network and with an untrained one whose rounds fall back (C0)."""
    th = W.to_rad(W.triples(W.map3_c0_radii()))
    rng = np.random.default_rng(7)
    bad = (rng.standard_normal(MO.NPARAM) * 0.05).astype(np.float32)
    try:
        for trained in (True, False):
            _load_model(gpu_ctx)
            if not trained:
                tp._check(tp.lib().tp_mlp_set_params(gpu_ctx.h, bad.ctypes.data_as(_fp)), "set")
            v0, _, n0 = fid_tables.map3_batch(th, None, 1e-4, tp.MODE_SURROGATE)
            fb0 = gpu_ctx.stats()["fallback_batches"]
            tp._check(tp.lib().tp_set_serial_regions(gpu_ctx.h, 1), "serial")
            v1, _, n1 = fid_tables.map3_batch(th, None, 1e-4, tp.MODE_SURROGATE)
            fb1 = gpu_ctx.stats()["fallback_batches"]
            tp._check(tp.lib().tp_set_serial_regions(gpu_ctx.h, 0), "serial")
            assert (fb0 > 0) == (not trained) and fb1 == fb0
            np.testing.assert_allclose(v1, v0, rtol=1e-12, atol=0)
            assert np.array_equal(n0, n1)
    finally:
        tp.lib().tp_set_serial_regions(gpu_ctx.h, 0)
        _load_model(gpu_ctx)


@pytest.mark.parametrize("workload", ["c1", "c3_subset"])
def test_surrogate_form_within_1e3_of_fp64(gpu_ctx, fid_nz, workload):
    """The surrogate form (triples with max/min radius >= 16 exact throughout; the others exact FP32 regions
    until an integral holds 1024 regions, then the network with 1% exact checks) against batched FP64 at the
    same tolerance 1e-4, on the 220 triples of 10 log radii
    0.5-32' and on the labelled Latin-hypercube subset (the first 8 cosmologies and cosmology 14, whose
    (4', 16', 16') integral the reference algorithm stops after 7 regions) x 35 triples of radii 2-32':
    every integral within 1e-3, zeros kept."""
    _load_model(gpu_ctx)
    if workload == "c1":
        tabs = tp.CosmologyTables([W.fiducial()], fid_nz, 100, 3.0, gpu_ctx)
        th = W.to_rad(W.triples(W.map3_c1_radii()))
        ic = None
    else:
        cos = W.lhs_c3(512)
        pick = list(range(8)) + [14]
        tabs = tp.CosmologyTables([cos[i] for i in pick], fid_nz, 100, 3.0, gpu_ctx)
        tri = W.to_rad(W.triples(W.map3_c3_radii()))
        th = np.tile(tri, (len(pick), 1))
        ic = np.repeat(np.arange(len(pick), dtype=np.int32), tri.shape[0])
    ref, _, nref = tabs.map3_batch(th, ic, 1e-4, tp.MODE_FP64)
    v, _, n = tabs.map3_batch(th, ic, 1e-4, tp.MODE_SURROGATE)
    st = gpu_ctx.stats()
    nz = ref != 0
    err = np.abs(v[nz] / ref[nz] - 1)
    print("%s: surrogate share %.3f, fallback rounds %d, redone integrals %d, max rel err %.2e (p99 %.2e), points "
          "%d vs %d" % (workload, st["surrogate_points"] / max(1, st["points"]), st["fallback_batches"],
                        st["fallback_integrals"], err.max(), np.percentile(err, 99), n.sum(), nref.sum()))
    assert np.all(v[~nz] == 0)
    assert err.max() <= 1e-3
    # the network serves what the exact rules leave: ~60% of C1's points, ~47% of the C3 subset's
    assert st["surrogate_points"] > 0.4 * st["points"]
