"""Input formats of the reference: n(z) files (both readers), TreeCorr triangle lists, theta lists.

The cuda_version's n(z) reader (helpers.cu:9-127) is checked against an independent formulation:
for a file coarser than the internal bins, each internal bin is the exact average over the bin of
the step function s(z) = n_j on [z_j, z_j+1) (0 below the file, the last value above it), computed
here from the step function's cumulative integral; for a finer file it is the classed_version's
linear resampling (oracle reader) with zeros from the file's last interval on.
"""
import ctypes
import os

import numpy as np
import pytest

import threepoint_amd as tp
from _oracle import lib as olib, d as od
from threepoint_amd import workloads as W


def _write_nz(path, z, n):
    path.write_text("# z n(z)\n" + "".join("%.17g %.17g\n" % (a, b) for a, b in zip(z, n)))
    return str(path).encode()


def _read(fn, nb, zmax, avg):
    out = np.zeros(nb)
    tp._check(tp.lib().tp_read_n_of_z(fn, nb, zmax, avg, tp._d(out)), "tp_read_n_of_z")
    return out


def _step_average(z, n, nb, zmax):
    dz = zmax / nb

    def cum(x):  # integral of the step function from -inf to x
        x = np.asarray(x, dtype=float)
        tot = np.zeros_like(x)
        for j in range(len(z)):
            hi = z[j + 1] if j + 1 < len(z) else np.inf
            tot += n[j] * np.clip(np.minimum(x, hi) - z[j], 0, None)
        return tot

    edges = np.arange(nb + 1) * dz
    v = (cum(edges[1:]) - cum(edges[:-1])) / dz
    return v / (v.sum() * dz)


def test_classed_reader_matches_oracle(tmp_path):
    z, n = W.nz_file_rows()
    fn = _write_nz(tmp_path / "nz.dat", z, n)
    ref = np.zeros(100)
    assert olib().orc_read_n_of_z(fn, 100, 3.0, od(ref)) == 0
    assert np.array_equal(_read(fn, 100, 3.0, 0), ref)
    assert np.array_equal(ref, W.fiducial_nz())
    # a file finer than the internal bins: the cuda_version's reader interpolates the same way
    assert np.array_equal(_read(fn, 100, 3.0, 1), ref)


def test_averaging_reader_coarse_file(tmp_path):
    z = np.arange(31) * 0.1
    n = 1.0 + z + 0.3 * np.sin(7 * z)
    fn = _write_nz(tmp_path / "coarse.dat", z, n)
    got = _read(fn, 100, 3.0, 1)
    want = _step_average(z, n, 100, 3.0)
    assert np.max(np.abs(got / want - 1)) < 1e-12
    # a file starting above 0: empty bins below its first sample.  The bin across the first sample
    # takes that sample's full value (the reference's "last sample below z_low" index defaults to 0,
    # so its partial-overlap branch is not reached there); every other bin is the step average
    z2 = 0.35 + np.arange(20) * 0.1
    n2 = 1.0 + 0.1 * np.arange(20)
    fn2 = _write_nz(tmp_path / "offset.dat", z2, n2)
    got2 = _read(fn2, 100, 3.0, 1)
    want2 = _step_average(z2, n2, 100, 3.0)
    assert np.all(got2[:11] == 0)
    assert abs(got2[11] / got2[12] - 1) < 1e-14  # full weight of n_0 in the straddling bin
    assert np.max(np.abs(got2[12:] / got2[12] - want2[12:] / want2[12])) < 1e-12


def test_interpolating_readers_at_the_file_end(tmp_path):
    z = np.arange(201) * 0.01  # ends at 2.0 < z_max
    n = np.exp(-z)
    fn = _write_nz(tmp_path / "short.dat", z, n)
    cv = _read(fn, 100, 3.0, 0)
    gv = _read(fn, 100, 3.0, 1)
    zi = np.arange(100) * 0.03
    inside = zi < 2.0 - 1e-12
    # same interior shape (the normalisations differ by the end bins)
    assert np.allclose(cv[inside] / cv[0], gv[inside] / gv[0], rtol=1e-13)
    assert np.all(gv[~inside] == 0) and np.all(cv[zi > 2.0 + 0.01] == 0)


def test_triangle_list_reader(tmp_path):
    rows = [(0, 0, 11, 0.74276, 0.08302, 0.04412), (0, 0, 12, 0.75011, 0.083105, 0.13627),
            (9, 9, 19, 88.1, 0.95, 0.95)]
    fn = tmp_path / "triangles.dat"
    fn.write_text("".join("%d\t%d\t%d\t%g\t%g\t%g\n" % r for r in rows))
    ruv = np.zeros(3 * 8)
    n = ctypes.c_size_t()
    tp._check(tp.lib().tp_read_triangle_configurations(str(fn).encode(), 8, tp._d(ruv), ctypes.byref(n)), "read")
    assert n.value == 3
    assert np.array_equal(ruv[:9].reshape(3, 3), np.array([r[3:] for r in rows]))
    assert tp.lib().tp_read_triangle_configurations(str(fn).encode(), 2, tp._d(ruv), ctypes.byref(n)) == 4
    assert n.value == 3
    r, u, v = W.read_triangle_configurations(str(fn))
    assert np.array_equal(np.stack([r, u, v], 1), np.array([x[3:] for x in rows]))
    x = W.treecorr_sides(r, u, v)
    assert np.allclose(x[:, 1], r) and np.allclose(x[:, 2], r * u) and np.allclose(x[:, 0], v * r * u + r)
