"""tp_map3_eval_noalloc (the entry point that allocates nothing) against tp_map3_eval, through the C ABI.

Both run the same point kernel on the same inputs, so the values must be bit-identical.  Two calls in a
row, the second with fewer points (the static staging keeps the first call's contents), in FP64 and
FP32, and the size limit.
"""
import ctypes

import numpy as np
import pytest

import threepoint_amd as tp

pytestmark = pytest.mark.gpu
_dp = ctypes.POINTER(ctypes.c_double)
MAX_POINTS = 65536  # TP_NOALLOC_MAX_POINTS


def _noalloc(tables, th, x, mode):
    L = tp.lib()
    f = L.tp_map3_eval_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t, _dp, _dp, ctypes.c_int]
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    th = np.ascontiguousarray(th, dtype=np.float64)
    out = np.full(x.shape[0], np.nan)
    rc = f(tables.ctx.h, tables.h, 0, th.ctypes.data_as(_dp), x.shape[0], x.ctypes.data_as(_dp),
           out.ctypes.data_as(_dp), mode)
    return rc, out


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([np.exp(rng.uniform(0, 9, n)), np.exp(rng.uniform(0, 9, n)), rng.uniform(0, 2 * np.pi, n)], 1)


@pytest.mark.parametrize("mode", [tp.MODE_FP64, tp.MODE_FP32])
def test_noalloc_matches_map3_eval(fid_tables, mode):
    th = np.array([4.0, 8.0, 16.0]) * tp.ARCMIN
    for n, seed in ((MAX_POINTS, 1), (1000, 2)):  # the second call is smaller than the first
        x = _points(n, seed)
        x[0] = [10.0, 10.0, np.pi]  # l3 = 0: the integrand's zero
        want = fid_tables.map3_integrand(x, th, 0, mode)
        rc, got = _noalloc(fid_tables, th, x, mode)
        assert rc == 0, tp.lib().tp_last_error()
        assert np.array_equal(got, want), (mode, n, np.max(np.abs(got - want)))


def test_noalloc_rejects_oversize_batch(fid_tables):
    th = np.array([4.0, 8.0, 16.0]) * tp.ARCMIN
    rc, _ = _noalloc(fid_tables, th, _points(MAX_POINTS + 1, 3), tp.MODE_FP64)
    assert rc == 4  # TP_ECAPACITY
    assert b"TP_NOALLOC_MAX_POINTS" in tp.lib().tp_last_error()
