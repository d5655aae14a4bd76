"""The oracle's model functions pinned to the reference's own source text (no GPU).

tests/golden/fit_constants.json holds the decimal literals of the reference's model functions, read
mechanically from its source (tests/golden/fit_constants.py): BiHalofit (classed_version/bispectrum.cpp:
166-236), F2 and F2_tree (:614-633), the Eisenstein & Hu transfer function (:811-841), revised Halofit
(:288-358) and the T17 factors (cuda_version/bispectrum.cu:951-982, apertureStatistics.cu:39-42).

1. The same extraction over the oracle's restatement (oracle/tp_oracle.cpp) gives the same sequences, up
   to the listed differences of structure (index code and a flatness check that moved elsewhere).
2. A numpy restatement written from the published equations (Takahashi et al. 2020 eqs. B4-B9,
   Eisenstein & Hu 1998 eqs. 26-31, Takahashi et al. 2012 appendix, the T17 ratio), taking every fit
   constant from the fixture by position, reproduces the oracle's bispec, linear_pk, P_k_nonlinear and the
   T17 factor: a wrong sign, power or operand in the oracle fails here even where its constants agree.
"""
import json
import math
import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import fit_constants as FC  # noqa: E402
from _oracle import Oracle, lib  # noqa: E402
from threepoint_amd import workloads as W  # noqa: E402

FX = json.load(open(os.path.join(HERE, "golden", "fit_constants.json")))
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "tp_oracle.cpp")

# The oracle's counterparts of the reference functions (first lines in oracle/tp_oracle.cpp)
ORACLE_FUNCTIONS = {
    "bispec": r"double bispec\(double k1, double k2, double k3, double z, int idx, double didx\) const",
    "F2": r"double F2\(double k1, double k2, double k3, double z, double D1v, double r_sigma\) const",
    "F2_tree": r"static double F2_tree\(double k1, double k2, double k3\)",
    "linear_pk_eh": r"double linear_pk_eh\(double k\) const",
    "P_k_nonlinear": r"double P_k_nonlinear\(double k, double z\) const",
    "t17": r"static double t17_correction\(double k\)",
}


def _oracle_literals(name):
    src = FC.strip_comments(open(ORACLE_SRC).read())
    return [float(x) for x in FC.decimal_literals(FC.function_body(src, ORACLE_FUNCTIONS[name])[2])]


def _ref(name):
    return [float(x) for x in FX[name]["literals"]]


@pytest.mark.parametrize("name", ["bispec", "F2", "F2_tree", "linear_pk_eh"])
def test_oracle_literals_equal_reference(name):
    """Literal for literal, in order: the BiHalofit, F2 and E&H constants of the oracle are the reference's."""
    assert _oracle_literals(name) == _ref(name)


def test_oracle_halofit_and_t17_literals_equal_reference():
    """Revised Halofit and the T17 ratio: equal after the reference's index code (didx = 1. at the last
    table interval, :291-298), its flatness check (|om + ow - 1| > 1e-4, :316-320; ow = 1 - om here) and the
    Limber z guard (z < 1e-5, bispectrum.cu:953-954) that the oracle keeps in z_index and
    limber_integrand_power_spectrum."""
    ref = _ref("P_k_nonlinear")
    assert ref[:2] == [1.0, 1e-4]
    assert _oracle_literals("P_k_nonlinear") == ref[2:]
    t17 = _ref("t17")
    assert t17[0] == 1e-5 and _oracle_literals("t17") == t17[1:]
    src = FC.strip_comments(open(ORACLE_SRC).read())
    assert "1e-5" in FC.function_body(src, r"double limber_integrand_power_spectrum\(double ell, double z")[2]
    assert "1." in FC.decimal_literals(FC.function_body(src, r"void z_index\(double z")[2])
    m2 = _ref("map2_t17")
    assert m2 == [1.6, 8192.0]
    assert "v /= 1 + std::pow(ell / (1.6 * 8192.), 2)" in src


# --------------------------------------------------------- constants by position
def _consts(name, layout):
    """Names the fixture's literals by position.  '=v' entries are structural numbers (1, 2, 10, ...) that
    must have the value v; a name used twice must carry the same value both times."""
    vals = _ref(name)
    assert len(vals) == len(layout), (name, len(vals), len(layout))
    K = {}
    for nm, v in zip(layout, vals):
        if nm.startswith("="):
            assert v == float(nm[1:]), (name, nm, v)
        elif nm in K:
            assert K[nm] == v, (name, nm, K[nm], v)
        else:
            K[nm] = v
    return SimpleNamespace(**K)


_10, _1, _2, _3 = "=10", "=1", "=2", "=3"
BISPEC = (["ztree", _10, "a0", "a1", "a2", "a3", "a4", _10, "a5", "a6", _10, "b0", "b1", "b2", "b3", _10, "c0", "c1",
           _10, "al0", "al1", "al2", _10, "al3", "al4", _1, _2, _3, _1, _2, _3, _10, "be0", "be1", "be2", "be3", "be4",
           _1, _1, _1, _1, _10, "f0", "f1", "f2", "f3", _10, "g0", "g1", "g2", _10, "h0", "h1", _10, "m0", "m1", "m2",
           _10, "n0", "n1", "n2", _10, "mu0", "mu1", "mu2", "mu3", _10, "nu0", "nu1", "nu2", _10, "p0", "p1", _10,
           "e0", "e1", _1, _1, _1, _1, _2, _1, _1])
F2 = [_1, _1, _3, _10, "d0", "d1", "d2"]
F2_TREE = ["=0.5", "=5", "=7", "=0.5", _2, "=7"]
EH = (["t0", "t1", "pc0", "=5", _1, "pc1", "zeq0", "b1a", "b1b", _1, "b1c", "b1d", "b2a", "b2b", "zda", "zdb", _1,
       "zdc", "zdd", _1, _1, _1, "sha", "shb", _1, "shc", "shd", "=5", _2, "=5", _1, "an1", "an2", _1, _1, "=0.5", _1,
       _1, "=7", _3, "=4", _1, _1, _1, "x0", "e_", "L1", _1, "L2", "C0", "C1", _1, "C2", "C3", "cH", _3, _2])
HALOFIT = (["=1", "=1e-4", "f1a", "f2a", "f3a", "f1b", "f2b", "f3b", _1, "a0", "a1", "a2", "a3", "a4", "a5", "a6", _1,
            _10, _10, "b0", "b1", "b2", "b3", "b4", _1, _10, "c0", "c1", "c2", "c3", "g0", "g1", "g2", "al0", "al1",
            "al2", "al3", "be0", "be1", "be2", "be3", "be4", "be5", _10, "x0", "x1", "=4", "=8"])
T17 = ["=1e-5", "c1", "c2", "a1", "a2", "a3", _1, _1]


# ----------------------------------------------------- the numpy restatement
def eh_pk(K, c, k, norm):
    """Eisenstein & Hu (1998) no-wiggle P_lin(k) [(Mpc/h)^3], k [h/Mpc] (eqs. 26-31 with the reference's
    sigma8 normalisation)."""
    k = k * c.h
    fc, fb = c.omc / c.om, c.omb / c.om
    theta = K.t0 / K.t1                                   # T_CMB / 2.7 K
    pc = K.pc0 * (5 - math.sqrt(1 + K.pc1 * fc))
    omh2, ombh2 = c.om * c.h * c.h, c.omb * c.h * c.h
    zeq = K.zeq0 * omh2 / theta ** 4
    b1 = K.b1a * omh2 ** (-K.b1b) * (1 + K.b1c * omh2 ** K.b1d)
    b2 = K.b2a * omh2 ** K.b2b
    zd = K.zda * omh2 ** K.zdb / (1 + K.zdc * omh2 ** K.zdd) * (1 + b1 * ombh2 ** b2)
    yd = (1 + zeq) / (1 + zd)
    sh = K.sha * math.log(K.shb / omh2) / math.sqrt(1 + K.shc * ombh2 ** K.shd)   # sound horizon, eq. 26
    alnu = (fc * (5 - 2 * pc) / 5 * (1 - K.an1 * fb + K.an2 * fb ** 3) * (1 + yd) ** (-pc) *
            (1 + 0.5 * pc * (1 + 1 / (7 * (3 - 4 * pc))) / (1 + yd)))
    geff = omh2 * (math.sqrt(alnu) + (1 - math.sqrt(alnu)) / (1 + (K.x0 * k * sh) ** 4))   # eq. 31
    qeff = k / geff * theta * theta
    L = np.log(K.e_ + K.L1 * math.sqrt(alnu) * qeff / (1 - K.L2 * fb))
    C = K.C0 + K.C1 / (1 + K.C2 * qeff ** K.C3)
    delk = norm ** 2 * (k * K.cH / c.h) ** (3 + c.ns) * (L / (L + C * qeff * qeff)) ** 2
    return c.h ** 3 * 2 * math.pi ** 2 / k ** 3 * delk


def _lerp(tab, idx, d):
    return tab[idx] * (1 - d) + tab[idx + 1] * d


def bihalofit(KB, KF, KE, c, tabs, k, z, idx, d):
    """Takahashi et al. (2020) eqs. B4-B9 at wavenumbers k[3] [h/Mpc] with the tables lerped at (idx, d)."""
    D1, rs, ne = _lerp(tabs["D1"], idx, d), _lerp(tabs["r_sigma"], idx, d), _lerp(tabs["n_eff"], idx, d)
    k = np.asarray(k, dtype=np.float64)
    P = np.array([eh_pk(KE, c, kk, tabs["norm"]) for kk in k])

    def f2_tree(a, b, cc):
        c12 = 0.5 * (cc * cc - a * a - b * b) / (a * b)
        return 5 / 7 + 0.5 * c12 * (a / b + b / a) + 2 / 7 * c12 * c12

    if z > KB.ztree:
        return D1 ** 4 * 2 * (f2_tree(*k) * P[0] * P[1] + f2_tree(k[1], k[2], k[0]) * P[1] * P[2] +
                              f2_tree(k[2], k[0], k[1]) * P[2] * P[0])
    q = k * rs
    qs = np.sort(q)[::-1]
    r1, r2 = qs[2] / qs[0], (qs[1] + qs[2] - qs[0]) / qs[0]          # eq. B8
    L = math.log10(D1 * c.sigma8)
    an = 10 ** (-KB.a0 - KB.a1 * L - KB.a2 * L ** 2 - KB.a3 * L ** 3 - KB.a4 * r1 ** (10 ** (KB.a5 + KB.a6 * ne)))
    bn = 10 ** (-KB.b0 - KB.b1 * L + KB.b2 * L ** 2 - KB.b3 * L ** 3)
    cn = 10 ** (KB.c0 - KB.c1 * ne)
    alphan = min(10 ** (-KB.al0 - KB.al1 * ne - KB.al2 * ne ** 2 + 10 ** (-KB.al3 + KB.al4 * ne) * r2 ** 2),
                 1 - 2 / 3 * c.ns)
    betan = 10 ** (-KB.be0 - KB.be1 * ne - KB.be2 * ne ** 2 - KB.be3 * ne ** 3 + KB.be4 * r2)
    bs1h = np.prod(1 / (an * q ** alphan + bn * q ** betan) / (1 + 1 / (cn * q)))    # eqs. B4, B7
    fn = 10 ** (-KB.f0 - KB.f1 * ne - KB.f2 * ne ** 2 - KB.f3 * ne ** 3)
    gn = 10 ** (KB.g0 + KB.g1 * ne + KB.g2 * ne ** 2)
    hn = 10 ** (-KB.h0 - KB.h1 * ne)
    mn = 10 ** (-KB.m0 - KB.m1 * L + KB.m2 * L ** 2)
    nn = 10 ** (-KB.n0 - KB.n1 * L + KB.n2 * L ** 2)
    mun = 10 ** (KB.mu0 + KB.mu1 * ne + KB.mu2 * ne ** 2 + KB.mu3 * ne ** 3)
    nun = 10 ** (KB.nu0 + KB.nu1 * ne + KB.nu2 * ne ** 2)
    pn = 10 ** (KB.p0 - KB.p1 * ne)
    en = 10 ** (-KB.e0 + KB.e1 * ne)
    pse = ((1 + fn * q ** 2) / (1 + gn * q + hn * q ** 2) * D1 ** 2 * P +
           1 / (mn * q ** mun + nn * q ** nun) / (1 + (pn * q) ** -3))                 # eq. B6
    omz = c.om / (c.om + c.ow * (1 / (1 + z)) ** (-3 * c.w))
    dn = 10 ** (-KF.d0 + KF.d1 * L - KF.d2 * omz)

    def f2(i, j, l):
        return f2_tree(k[i], k[j], k[l]) + dn * q[l]

    bs3h = 2 * (f2(0, 1, 2) * pse[0] * pse[1] + f2(1, 2, 0) * pse[1] * pse[2] + f2(2, 0, 1) * pse[2] * pse[0])
    return bs1h + bs3h * np.prod(1 / (1 + en * q))                                    # eq. B5


def halofit(KH, KE, c, tabs, k, z, idx, d):
    """Revised Halofit (Takahashi et al. 2012, with the reference's w term) at k [h/Mpc], z."""
    D1, rs, ne = _lerp(tabs["D1"], idx, d), _lerp(tabs["r_sigma"], idx, d), _lerp(tabs["n_eff"], idx, d)
    nc = _lerp(tabs["ncur"], idx, d)
    a_ = 1 / (1 + z)
    den = c.om + a_ * (a_ * a_ * c.ow + (1 - c.om - c.ow))
    om_m, om_v = c.om / den, c.ow * a_ ** 3 / den
    frac = om_v / (1 - om_m)
    f1 = frac * om_m ** -KH.f1b + (1 - frac) * om_m ** -KH.f1a
    f2 = frac * om_m ** -KH.f2b + (1 - frac) * om_m ** -KH.f2a
    f3 = frac * om_m ** KH.f3b + (1 - frac) * om_m ** KH.f3a
    n2 = ne * ne
    a = 10 ** (KH.a0 + KH.a1 * ne + KH.a2 * n2 + KH.a3 * ne * n2 + KH.a4 * n2 * n2 - KH.a5 * nc + KH.a6 * om_v * (1 + c.w))
    b = 10 ** (-KH.b0 + KH.b1 * ne + KH.b2 * n2 - KH.b3 * nc + KH.b4 * om_v * (1 + c.w))
    cc = 10 ** (KH.c0 + KH.c1 * ne + KH.c2 * n2 + KH.c3 * nc)
    gam = KH.g0 - KH.g1 * ne + KH.g2 * nc
    alpha = abs(KH.al0 + KH.al1 * ne - KH.al2 * n2 - KH.al3 * nc)
    beta = KH.be0 - KH.be1 * ne + KH.be2 * n2 + KH.be3 * ne * n2 + KH.be4 * n2 * n2 - KH.be5 * nc
    xnu = 10 ** (KH.x0 + KH.x1 * ne)
    plin = eh_pk(KE, c, k, tabs["norm"]) * D1 * D1 * k ** 3 / (2 * math.pi ** 2)
    y = k * rs
    ph = a * y ** (f1 * 3) / (1 + b * y ** f2 + (f3 * cc * y) ** (3 - gam)) / (1 + xnu / (y * y))
    pq = plin * (1 + plin) ** beta / (1 + plin * alpha) * math.exp(-y / 4 - y * y / 8)
    return 2 * math.pi ** 2 * (pq + ph) / k ** 3


@pytest.fixture(scope="module")
def model():
    c = W.fiducial()
    o = Oracle(c, W.fiducial_nz())
    return SimpleNamespace(c=c, o=o, t=o.tables(), KB=_consts("bispec", BISPEC), KF=_consts("F2", F2),
                           KE=_consts("linear_pk_eh", EH), KH=_consts("P_k_nonlinear", HALOFIT),
                           KT=_consts("t17", T17))


def test_layouts_are_consistent():
    _consts("F2_tree", F2_TREE)  # only structural numbers: 1/2, 5/7, 2/7


def test_eisenstein_hu_restatement_matches_oracle(model):
    for k in np.geomspace(1e-4, 1e3, 57):
        want = model.o.scalar("linear_pk", k)
        assert abs(eh_pk(model.KE, model.c, k, model.t["norm"]) / want - 1) < 1e-13, k


def test_bihalofit_restatement_matches_oracle(model):
    """50 triangles of squeeze below 30 (no cancellation) at random table intervals, and 5 beyond z = 10."""
    rng = np.random.default_rng(21)
    n = 0
    while n < 55:
        k1, k2 = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), 2))
        k3 = math.sqrt(k1 * k1 + k2 * k2 + 2 * k1 * k2 * rng.uniform(-0.95, 0.95))
        ks = [k1, k2, k3]
        if max(ks) / min(ks) > 30:
            continue
        idx, d = int(rng.integers(0, 99)), float(rng.uniform())
        z = float(rng.uniform(0.0, 3.0)) if n < 50 else float(rng.uniform(10.5, 20))
        want = lib().orc_bispec_idx(model.o.h, k1, k2, k3, z, idx, d)
        got = bihalofit(model.KB, model.KF, model.KE, model.c, model.t, ks, z, idx, d)
        assert abs(got / want - 1) < 1e-11, (ks, z, idx, d, got, want)
        n += 1


def test_halofit_restatement_matches_oracle(model):
    rng = np.random.default_rng(22)
    for _ in range(50):
        k = float(np.exp(rng.uniform(np.log(1e-3), np.log(1e2))))
        z = float(rng.uniform(0.0, 3.0))
        dd = z / 3.0 * 99
        idx = int(dd)
        d = dd - idx
        if idx == 99:
            idx, d = 98, 1.0
        want = model.o.scalar("pk_nonlinear", k, z)
        assert abs(halofit(model.KH, model.KE, model.c, model.t, k, z, idx, d) / want - 1) < 1e-11, (k, z)


def test_t17_factor_restatement_matches_oracle(model):
    """The ratio of the oracle's Limber integrand with and without the T17 factor is the T17 ratio at
    k = ell / f_K (cuda_version/bispectrum.cu:971-979)."""
    K = model.KT
    for ell, z in ((10.0, 0.5), (300.0, 1.2), (3e3, 0.3), (4e4, 2.2), (2.0, 2.9)):
        dd = z / 3.0 * 99
        idx = int(dd)
        k = ell / _lerp(model.t["f_K"], idx, dd - idx)
        want = (1 + K.c1 * k ** (-K.a1)) ** K.a1 / (1 + K.c2 * k ** (-K.a2)) ** K.a3
        r = model.o.scalar("limber_pkappa", ell, z, 2) / model.o.scalar("limber_pkappa", ell, z, 0)
        assert abs(r / want - 1) < 1e-13, (ell, z, r, want)
