"""CPU oracle of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

numpy restatement of the network in threepoint_amd/csrc/tp_mlp.hip: 8 -> 4 x 256 SiLU -> 1,
loss = mean (y - t)^2 / 2.  With bf16=True it rounds to BF16 (round to nearest even) at the
same places as the GPU kernels: the three 256x256 weight matrices, the hidden activations
fed to the next layer, and the pre-activation gradients fed to the GEMMs; everything else
is FP64 here (FP32 on the GPU).  With bf16=False it is the plain FP64 network, pinned by
central finite differences in the tests.
"""
import numpy as np

H, NIN = 256, 8
SHAPES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, H)), ("b4", (H,)), ("W5", (H,)), ("b5", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """round to BF16 (nearest even) and back, via the FP32 bit pattern"""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, X, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    z = [X @ P["W1"].T + P["b1"]]
    a = [r(silu(z[0]))]
    for k in (2, 3, 4):
        z.append(a[-1] @ r(P["W%d" % k]).T + P["b%d" % k])
        a.append(r(silu(z[-1])))
    y = a[-1] @ P["W5"] + P["b5"][0]
    return y, z, a


def loss_grad(p, X, T, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    M = X.shape[0]
    y, z, a = forward(P, X, bf)
    loss = 0.5 * np.mean((y - T) ** 2)
    dy = (y - T) / M
    G = {"W5": a[3].T @ dy, "b5": np.array([dy.sum()])}
    dZ = r(np.outer(dy, P["W5"]) * silu_grad(z[3]))
    for k, l in ((4, 3), (3, 2), (2, 1)):
        G["W%d" % k] = dZ.T @ a[l - 1]
        G["b%d" % k] = dZ.sum(0)
        dZ = r((dZ @ r(P["W%d" % k])) * silu_grad(z[l - 1]))
    G["W1"] = dZ.T @ X
    G["b1"] = dZ.sum(0)
    return loss, pack(G)


def hilo(a):
    """a as the sum of two BF16 numbers (hi = bf16(a), lo = bf16(a - hi))"""
    hi = bf16(a)
    return hi + bf16(a - hi)


def forward_hilo(P, X):
    """the fused inference kernel's arithmetic: hi+lo BF16 activations, BF16 weights in layers 2-4"""
    a = hilo(silu(X @ P["W1"].T + P["b1"]))
    for k in (2, 3, 4):
        a = hilo(silu(a @ bf16(P["W%d" % k]).T + P["b%d" % k]))
    return a @ P["W5"] + P["b5"][0]


def fp16(x):
    """round to FP16 (nearest even) and back"""
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64)


NEG_LOG2E = -1.4426950408889634


def forward_fp16(P, X):
    """the FP16 inference kernel's arithmetic: the BF16 weights of layers 2-4 (exact in FP16 within its range),
    pre-activations scaled by -log2(e) (u = -log2(e) z, from W1 and the biases scaled in FP32) and the FP16
    activations a' = u / (1 + 2^u) = -log2(e) SiLU(z), rounded once; the output weights divided by -log2(e)"""
    f32 = lambda v: np.asarray(v, dtype=np.float32).astype(np.float64)  # noqa: E731
    sil = lambda u: u / (1.0 + np.exp2(u))  # noqa: E731
    u = X @ f32(P["W1"] * NEG_LOG2E).T + f32(P["b1"] * NEG_LOG2E)
    for k in (2, 3, 4):
        u = fp16(sil(u)) @ fp16(bf16(P["W%d" % k])).T + f32(P["b%d" % k] * NEG_LOG2E)
    return sil(u) @ f32(P["W5"] / NEG_LOG2E) + P["b5"][0]


MODEL_MAGIC = b"TPMLP002"


def write_model(path, lnl, lo, hi, mean, sd, p):
    with open(path, "wb") as f:
        f.write(MODEL_MAGIC)
        f.write(np.array([lnl[0], lnl[1]] + list(lo) + list(hi), dtype=np.float64).tobytes())
        f.write(np.array([mean, sd], dtype=np.float32).tobytes())
        f.write(np.asarray(p, dtype=np.float32).tobytes())


def read_model(path):
    b = open(path, "rb").read()
    assert b[:8] == MODEL_MAGIC
    box = np.frombuffer(b[8:8 + 96], dtype=np.float64)
    mean, sd = np.frombuffer(b[104:112], dtype=np.float32)
    p = np.frombuffer(b[112:], dtype=np.float32).astype(np.float64)
    return {"lnl": box[:2], "lo": box[2:7], "hi": box[7:12], "mean": float(mean), "sd": float(sd), "p": p}


def encode(model, l1, l2, l3, cosmo5):
    """network inputs: sorted log sides and the cosmology, each mapped to [-1, 1] over the box"""
    s = np.sort(np.stack([l1, l2, l3], 1), axis=1)[:, ::-1]
    mid, half = 0.5 * (model["lnl"][0] + model["lnl"][1]), 0.5 * (model["lnl"][1] - model["lnl"][0])
    x = np.zeros((s.shape[0], NIN))
    x[:, :3] = (np.log(s) - mid) / half
    c = np.asarray(cosmo5, dtype=np.float64).reshape(-1, 5)
    x[:, 3:] = 2 * (c - model["lo"]) / (model["hi"] - model["lo"]) - 1
    return x.astype(np.float32).astype(np.float64)
