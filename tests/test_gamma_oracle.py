"""The oracle's Gamma machinery pinned by the closed form of the analytic test bispectrum (no GPU).

classed_version/testGamma_analytical_bispectrum.cpp:79-149 integrates Gamma_0 of the reference's analytic test
bispectrum (bispectrum.cpp:477-499, 989-996: a sum of Gaussians in ell, zero above z = 1) on a TreeCorr-style
grid -- r in 15 log bins over [0.1, 300] arcmin, u and v in 15 bins over (0, 1), z_max = 1 -- and compares
it with the closed form ggg (gamma.cpp:83-94, 567-586).  The closed form is the natural component in the
orthocenter frame (the centroid convention multiplies it by the phase of gamma.cpp:624-643), so the oracle's
Gamma_0 with triangle_center "orthocenter" must equal it: this checks the Ogata nodes and weights, the
geometry (A, alpha, betabar, the interior angles), the J6 term sums, the (z, phi, psi) cubature and the
prefactors at once.  The numerical value converges to the closed form as the Ogata step shrinks; the
tolerances are the agreement reached at step 5e-3 with 1000 nodes and cubature tolerance 1e-5.
"""
import math

import numpy as np
import pytest

from _oracle import Oracle
from threepoint_amd import workloads as W

ARCMIN = 2.9088820866e-4


def grid_triangle(i, j, k, steps=15):
    """(r1, r2, r3) [rad] of bin (i, j, k) of the reference test's grid (testGamma_analytical_bispectrum.cpp:79-130)."""
    lr = math.log(0.1) + (math.log(300.0) - math.log(0.1)) / steps * (i + 0.5)
    u = (j + 0.5) / steps
    v = (k + 0.5) / steps
    r2 = math.exp(lr) * ARCMIN
    r3 = r2 * u
    return v * r3 + r2, r2, r3


@pytest.fixture(scope="module")
def analytic_oracle():
    c = W.fiducial()
    c.zmax = 1.0
    o = Oracle(c, np.full(10, 1.0), 10, 1.0)  # the test's BispectrumCalculator(&cosmo, 10, 1., ...)
    o.set_analytic(True)
    o.gamma_init(0.005, 5.0)
    return o


@pytest.mark.parametrize("ijk,tol", [((8, 10, 2), 1e-6), ((7, 3, 10), 5e-4), ((6, 12, 12), 5e-3)])
def test_gamma0_orthocenter_equals_closed_form(analytic_oracle, ijk, tol):
    x = grid_triangle(*ijk)
    g, n = analytic_oracle.gamma(0, *x, reltol=1e-5, center=1)
    ggg = Oracle.ggg_analytic(*x)
    err = abs(g - ggg) / abs(ggg)
    print("bin %s: Gamma0 %s, closed form %s, rel %.2e, %d points" % (ijk, g, ggg, err, n))
    assert err < tol


def test_centroid_rotation_is_a_phase(analytic_oracle):
    """The centroid convention changes the phase only (|Gamma_0| equal), by the rotation angles of the triangle."""
    x = grid_triangle(8, 10, 2)
    go, _ = analytic_oracle.gamma(0, *x, reltol=1e-5, center=1)
    gc, _ = analytic_oracle.gamma(0, *x, reltol=1e-5, center=0)
    assert abs(abs(gc) / abs(go) - 1) < 1e-13
    assert abs(gc - go) > 1e-3 * abs(go)  # the rotation is not the identity on this triangle


def test_integrand_is_ill_conditioned_near_phi_pi():
    """Why whole Gamma_i at the production grid agree only to the reference algorithm's own reproducibility:
    at phi -> pi with psi = pi/4 the Ogata terms' ell3 = |ell1 + ell2| -> 0 by cancellation, and one ulp of
    phi moves the oracle's own integrand by orders of magnitude more than at a generic point."""
    o = Oracle(W.fiducial(), W.fiducial_nz())
    o.gamma_init(1e-3, 5.0)
    x = W.gamma_c2_triangles()[555, :3]

    def spread(z, phi, psi):
        a = o.gamma_integrand(0, z, phi, psi, *x)
        b = o.gamma_integrand(0, z, phi * (1 + 2.2e-16), psi, *x)
        return abs(b - a) / abs(a)

    generic = spread(0.7, 1.3, 0.8)
    near = spread(0.7, math.pi - 1e-6, math.pi / 4 + 1e-7)
    print("one-ulp spread: generic %.2e, near phi = pi %.2e" % (generic, near))
    assert generic < 1e-12
    assert near > 1e-4
