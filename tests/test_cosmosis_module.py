"""The cosmosis module (threepoint_amd/_lib/threepoint_cosmosis.so) against an in-memory datablock.

The module replaces classed_version/interface.cpp: setup reads the "threepoint" options, execute
reads cosmological_parameters and matter_power_lin from the block and puts threepoint/Map3s.
cosmosis itself is not installed here, so tests/cosmosis_mock provides the C datablock functions
the module calls.
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

import threepoint_amd as tp
from threepoint_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "cosmosis_mock", "_build", "libmock_datablock.so")
MODULE = os.path.join(ROOT, "threepoint_amd", "_lib", "threepoint_cosmosis.so")
_dp = ctypes.POINTER(ctypes.c_double)


def _load():
    if not os.path.exists(MOCK):
        pytest.fail("datablock stand-in not built (run the build)")
    mock = ctypes.CDLL(MOCK, mode=ctypes.RTLD_GLOBAL)
    mock.mock_block_new.restype = ctypes.c_void_p
    for f in ("mock_put_int", "mock_put_bool"):
        getattr(mock, f).argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    mock.mock_put_double.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double]
    mock.mock_put_string.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
    mock.mock_put_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
    mock.mock_get_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, _dp, ctypes.c_int]
    mock.mock_block_free.argtypes = [ctypes.c_void_p]
    mod = ctypes.CDLL(MODULE)
    mod.setup.restype = ctypes.c_void_p
    mod.setup.argtypes = [ctypes.c_void_p]
    mod.execute.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    mod.cleanup.argtypes = [ctypes.c_void_p]
    return mock, mod


def _options(mock, tmp_path, nbins=3, diag=False, output=True, zbins=100):
    z, n = W.nz_file_rows()
    nf = tmp_path / "nz.dat"
    nf.write_text("".join("%.17g %.17g\n" % (a, b) for a, b in zip(z, n)))
    o = mock.mock_block_new()
    s = b"threepoint"
    mock.mock_put_string(o, s, b"nz_file", str(nf).encode())
    if zbins is not None:
        mock.mock_put_int(o, s, b"zBins", zbins)
    mock.mock_put_double(o, s, b"zMax", 3.0)
    mock.mock_put_string(o, s, b"theta_unit", b"arcmin")
    mock.mock_put_double(o, s, b"theta_min", 4.0)
    mock.mock_put_double(o, s, b"theta_max", 16.0)
    mock.mock_put_int(o, s, b"theta_bins", nbins)
    mock.mock_put_bool(o, s, b"diag_only", 1 if diag else 0)
    mock.mock_put_bool(o, s, b"output_Map3", 1 if output else 0)
    mock.mock_put_string(o, s, b"output_dir", str(tmp_path).encode())
    return o


def _put_array(mock, b, sec, name, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    ext = (ctypes.c_int * a.ndim)(*a.shape)
    mock.mock_put_array(b, sec, name, a.ctypes.data_as(_dp), a.ndim, ext)


def test_module_exports_and_rejects_incomplete_options(tmp_path):
    mock, mod = _load()
    o = mock.mock_block_new()
    mock.mock_put_string(o, b"threepoint", b"nz_file", b"/nonexistent")
    assert mod.setup(o) is None  # missing options: no config, no exit()
    mock.mock_block_free(o)
    assert mod.cleanup(None) == 0


@pytest.mark.gpu
def test_module_execute_matches_batched_engine(gpu_ctx, tmp_path):
    """A sampled cosmology with its CAMB-style linear P(k) (two z rows): Map3s for the 10 triples of
    radii {4, 8, 16}' agree with tp_map3_batch on the same tabulated P(k) (1e-12), in the reference's order."""
    mock, mod = _load()
    opts = _options(mock, tmp_path)
    cfg = mod.setup(opts)
    assert cfg
    c = tp.Cosmology(h=0.68, sigma8=0.8, omb=0.048, ns=0.96, w=-0.95, om=0.31)
    blk = mock.mock_block_new()
    cp = b"cosmological_parameters"
    for k, v in ((b"omega_m", c.om), (b"sigma_8", c.sigma8), (b"h0", c.h), (b"omega_b", c.omb), (b"n_s", c.ns),
                 (b"w", c.w)):
        mock.mock_put_double(blk, cp, k, v)
    # a linear P(k) table: the fit of this cosmology, normalised (rows: z = 0 and a second z)
    nz = W.fiducial_nz()
    ref_tabs = tp.CosmologyTables([c], nz, 100, 3.0, gpu_ctx)
    kh = np.logspace(-4.5, 3.5, 200)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    eh = Oracle(c, nz)
    p0 = np.array([eh.scalar("linear_pk", k) for k in kh])
    _put_array(mock, blk, b"matter_power_lin", b"k_h", kh)
    _put_array(mock, blk, b"matter_power_lin", b"p_k", np.stack([p0, 0.5 * p0]))
    assert mod.execute(blk, cfg) == 0
    out = np.zeros(64)
    n = mock.mock_get_array(blk, b"threepoint", b"Map3s", out.ctypes.data_as(_dp), 64)
    assert n == 10
    # the module's radii: log-spaced from theta_min to theta_max (interface.cpp:116-141)
    tmin, tmax = tp.ARCMIN * 4.0, tp.ARCMIN * 16.0
    dl = (math.log(tmax) - math.log(tmin)) / 2
    r = [tmin * math.exp(i * dl) for i in range(3)]
    th = np.array([[r[i], r[j], r[k]] for i in range(3) for j in range(i, 3) for k in range(j, 3)])
    assert np.allclose(th, W.to_rad(W.triples([4.0, 8.0, 16.0])), rtol=1e-14, atol=0)
    tabs = tp.CosmologyTables([c], nz, 100, 3.0, gpu_ctx, pk=(kh, p0))
    v, _, _ = tabs.map3_batch(th, None, 1e-4, tp.MODE_FP64)
    assert np.allclose(out[:n], v, rtol=1e-12, atol=0)
    # the fit itself gives the same <Map^3> to the table's interpolation error
    v_eh, _, _ = ref_tabs.map3_batch(th, None, 1e-4, tp.MODE_FP64)
    assert np.max(np.abs(v / v_eh - 1)) < 5e-3
    files = list(tmp_path.glob("om_*.dat"))
    assert len(files) == 1 and len(files[0].read_text().splitlines()) == 11
    assert mod.cleanup(ctypes.c_void_p(cfg)) == 0
    mock.mock_block_free(blk)
    mock.mock_block_free(opts)


@pytest.mark.gpu
def test_module_setup_without_zbins_uses_256_bins(gpu_ctx, tmp_path):
    """The shipped pipeline (cosmosis/pipeline_gpu.ini, the cuda_version interface) sets no zBins: setup
    still succeeds and uses the cuda_version's 256 redshift bins (cuda_version/bispectrum.cuh:35)."""
    mock, mod = _load()
    opts = _options(mock, tmp_path, nbins=1, diag=True, output=False, zbins=None)
    cfg = mod.setup(opts)
    assert cfg
    c = W.fiducial()
    blk = mock.mock_block_new()
    cp = b"cosmological_parameters"
    for k, v in ((b"omega_m", c.om), (b"sigma_8", c.sigma8), (b"h0", c.h), (b"omega_b", c.omb), (b"n_s", c.ns),
                 (b"w", c.w)):
        mock.mock_put_double(blk, cp, k, v)
    kh = np.logspace(-4.5, 3.5, 200)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle
    p0 = np.array([Oracle(c, W.fiducial_nz()).scalar("linear_pk", k) for k in kh])
    _put_array(mock, blk, b"matter_power_lin", b"k_h", kh)
    _put_array(mock, blk, b"matter_power_lin", b"p_k", p0[None, :])
    assert mod.execute(blk, cfg) == 0
    out = np.zeros(4)
    assert mock.mock_get_array(blk, b"threepoint", b"Map3s", out.ctypes.data_as(_dp), 4) == 1
    z, n = W.nz_file_rows()
    nz256 = W.resample_n_of_z(z, n, n_bins=256, z_max=3.0)
    th = np.full((1, 3), tp.ARCMIN * 4.0)
    v256, _, _ = tp.CosmologyTables([c], nz256, 256, 3.0, gpu_ctx, pk=(kh, p0)).map3_batch(th, None, 1e-4,
                                                                                              tp.MODE_FP64)
    assert abs(out[0] / v256[0] - 1) < 1e-12
    assert mod.cleanup(ctypes.c_void_p(cfg)) == 0
    mock.mock_block_free(blk)
    mock.mock_block_free(opts)
