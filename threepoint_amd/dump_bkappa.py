"""Per-point B_kappa dump: 1e6 points (ln l1, ln l2 ~ U[1, 12], phi ~ U[0, pi], counter-based seed
0xC1B0D5) of the fiducial cosmology, written as little-endian f64 rows (l1, l2, phi, B_kappa).

    python -m threepoint_amd.dump_bkappa --out bkappa_c1.f64 [--n 1000000] [--mode fp64|fp32]

Two runs give the same bytes (the z-sum is a fixed tree); the SHA-256 of the file is printed.
"""
import argparse
import hashlib
import sys

import numpy as np

from . import MODE_FP32, MODE_FP64, CosmologyTables, default_context
from . import workloads as W


def dump(n=1000000, mode=MODE_FP64, ctx=None):
    ctx = ctx or default_context()
    tabs = CosmologyTables([W.fiducial()], W.fiducial_nz(), 100, 3.0, ctx)
    x, ell = W.bkappa_c1_points(n)
    b = tabs.bkappa(ell, mode=mode)
    rows = np.ascontiguousarray(np.column_stack([x, b]), dtype="<f8")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--mode", choices=["fp64", "fp32"], default="fp64")
    a = ap.parse_args(argv)
    rows = dump(a.n, MODE_FP32 if a.mode == "fp32" else MODE_FP64)
    rows.tofile(a.out)
    print("%s: %d rows, sha256 %s" % (a.out, rows.shape[0], hashlib.sha256(rows.tobytes()).hexdigest()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
