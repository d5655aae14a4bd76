"""The MLP oracle itself (CPU): its FP64 backward pass against central finite differences."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402


def test_param_count():
    assert MO.NPARAM == 199937


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.1, MO.NPARAM)
    X = rng.uniform(-1, 1, (16, MO.NIN))
    T = rng.normal(0, 1, 16)
    _, g = MO.loss_grad(p, X, T, bf=False)
    idx = np.concatenate([rng.choice(MO.NPARAM, 40, replace=False), [MO.NPARAM - 1, 0, 2048 + 256 + 5]])
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (MO.loss_grad(pp, X, T, bf=False)[0] - MO.loss_grad(pm, X, T, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], dtype=np.float64)
    r = MO.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
    assert MO.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0          # tie -> even
    assert MO.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6  # tie -> even (up)
