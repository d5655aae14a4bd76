"""Writes the oracle-side fixtures under tests/golden/ (run in the build container).

These are outputs of the CPU FP64 oracle (oracle/tp_oracle.cpp), stored so the
GPU parity tests do not have to rerun minutes of CPU cubature on the GPU host:
  map3_c0_fp64.json   <Map^3> of the 20 C0 triples (rel tol 1e-4) + points used
  tables_fid.json     f_K, g, D1, r_sigma, n_eff, ncur, norm of the fiducial cosmology
  bkappa_c1_2k.npz    B_kappa at the first 2000 C1 dump points
Usage: python tests/golden/make_golden.py [map3|tables|bkappa|all]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from threepoint_amd import workloads as W  # noqa: E402
from _oracle import Oracle  # noqa: E402


def main(what):
    o = Oracle(W.fiducial(), W.fiducial_nz())
    if what in ("tables", "all"):
        t = o.tables()
        json.dump({k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in t.items()},
                  open(os.path.join(HERE, "tables_fid.json"), "w"), indent=1)
    if what in ("bkappa", "all"):
        x, ell = W.bkappa_c1_points()
        ell = ell[:2000]
        b = o.bkappa(ell)
        np.savez(os.path.join(HERE, "bkappa_c1_2k.npz"), ell=ell, bkappa=b)
    if what in ("map3", "all"):
        tr = W.triples(W.map3_c0_radii())
        rows = []
        t0 = time.time()
        for th in tr:
            v, e, n = o.map3(W.to_rad(th), 1e-4)
            rows.append({"theta_arcmin": th.tolist(), "map3": v, "err": e, "neval": n})
            print(th, v, e, n, "%.1fs" % (time.time() - t0), flush=True)
        json.dump({"reltol": 1e-4, "rows": rows}, open(os.path.join(HERE, "map3_c0_fp64.json"), "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "all")
