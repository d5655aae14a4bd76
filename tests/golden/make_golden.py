"""Writes the oracle-side fixtures under tests/golden/ (run in the build container).

These are outputs of the CPU FP64 oracle (oracle/tp_oracle.cpp), stored so the
GPU parity tests do not have to rerun minutes of CPU cubature on the GPU host:
  map3_c0_fp64.json   <Map^3> of the 20 C0 triples (rel tol 1e-4) + points used
  map3_c1_fp64.json   <Map^3> of the 220 triples of 10 log radii 0.5-32 arcmin (rel tol 1e-4) + points used
  gamma_c2_fp64.json  Gamma_0..3 of TreeCorr-grid triangles at the production Ogata grid (h 1e-3, 4999
                      Bessel zeros), centroid convention, at the relative tolerance given (argument 2),
                      + points used per component
  tables_fid.json     f_K, g, D1, r_sigma, n_eff, ncur, norm of the fiducial cosmology
  bkappa_c1_2k.npz    B_kappa at the first 2000 C1 dump points
  census.json         the op census of the integrand points (flops and elementary calls per
                      z-node, per point, per P_E term, per Ogata term, per <Map^2> / Limber /
                      node-record point), counted by instantiating the kernels' device functions
                      with an op-counting scalar on the host, and the flop-equivalents per unit at
                      20 flops per elementary call; bench.py's roofline reads it
Usage: python tests/golden/make_golden.py [map3|map3_c1|tables|bkappa|census|all]
       python tests/golden/make_golden.py gamma_c2 RELTOL TRIANGLE_INDEX...
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
    if what in ("census", "all"):
        json.dump(census(), open(os.path.join(HERE, "census.json"), "w"), indent=1)
    if what in ("map3", "all"):
        tr = W.triples(W.map3_c0_radii())
        rows = []
        t0 = time.time()
        for th in tr:
            v, e, n = o.map3(W.to_rad(th), 1e-4)
            rows.append({"theta_arcmin": th.tolist(), "map3": v, "err": e, "neval": n})
            print(th, v, e, n, "%.1fs" % (time.time() - t0), flush=True)
        json.dump({"reltol": 1e-4, "rows": rows}, open(os.path.join(HERE, "map3_c0_fp64.json"), "w"), indent=1)
    if what == "map3_c1":
        tr = W.triples(W.map3_c1_radii())
        rows = []
        t0 = time.time()
        for th in tr:
            v, e, n = o.map3(W.to_rad(th), 1e-4)
            rows.append({"theta_arcmin": th.tolist(), "map3": v, "err": e, "neval": n})
            print(th, v, e, n, "%.1fs" % (time.time() - t0), flush=True)
        json.dump({"reltol": 1e-4, "rows": rows}, open(os.path.join(HERE, "map3_c1_fp64.json"), "w"), indent=1)


def gamma_c2(reltol, indices):
    """Gamma_0..3 of the given TreeCorr-grid triangles (rows of workloads.gamma_c2_triangles()), appended to
    gamma_c2_fp64.json (one entry per triangle and tolerance)."""
    o = Oracle(W.fiducial(), W.fiducial_nz())
    o.gamma_init(1e-3, 5.0)
    fn = os.path.join(HERE, "gamma_c2_fp64.json")
    out = json.load(open(fn)) if os.path.exists(fn) else {"prec_h": 1e-3, "prec_k_arg": 5.0, "rows": []}
    tri = W.gamma_c2_triangles()
    for i in indices:
        x = tri[i, :3]
        g, n = [], []
        t0 = time.time()
        for c in range(4):
            v, ne = o.gamma(c, *x, reltol=reltol)
            g.append([v.real, v.imag])
            n.append(ne)
            print(i, c, v, ne, "%.1fs" % (time.time() - t0), flush=True)
        out["rows"].append({"triangle": int(i), "x123_rad": [float(a) for a in x], "reltol": reltol,
                            "gamma_re_im": g, "neval": n, "cpu_seconds": time.time() - t0})
        json.dump(out, open(fn, "w"), indent=1)


def census(flops_per_elem=20.0):
    """The op census of the library's integrand device functions (tp_census_*) and the work per unit."""
    import threepoint_amd as tp
    m3 = {k: float(v) for k, v in tp.census_map3_point().items()}
    pw = {k: float(v) for k, v in tp.census_power_point().items()}
    gt = {k: float(v) for k, v in tp.census_gamma_term().items()}
    fe_node = m3["node_flops"] + flops_per_elem * m3["node_elem"]
    fe_pse = m3["pse_flops"] + flops_per_elem * m3["pse_elem"]
    fe_point = m3["nodes"] * fe_node + m3["point_flops"] + flops_per_elem * m3["point_elem"]
    fe_pnode = pw["node_flops"] + flops_per_elem * pw["node_elem"]
    return {
        "flops_per_elementary_call": flops_per_elem,
        "map3_point": m3, "power_point": pw, "gamma_term": gt,
        "flop_eq": {
            "map3_point": fe_point,
            "map3_point_region_cached_executed": fe_point - m3["nodes"] * fe_pse * (99 - 47) / 33,
            "map2_point": pw["map2_flops"] + flops_per_elem * pw["map2_elem"],
            "t1_region": 47 * (96 * fe_pnode + 48 * 3 + 2) + 33 * ((7 + 2 * flops_per_elem) +
                                                                   6 * (4 + flops_per_elem) + 8) + 90,
            "gamma_term": gt["term_flops"] + flops_per_elem * gt["term_elem"],
        },
    }


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "gamma_c2":
        gamma_c2(float(sys.argv[2]), [int(a) for a in sys.argv[3:]])
    else:
        main(sys.argv[1] if len(sys.argv) > 1 else "all")
