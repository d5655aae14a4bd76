"""ctypes binding of the CPU FP64 oracle (oracle/_build/libtp_oracle.so) for the tests.

The oracle is test infrastructure: it is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libtp_oracle.so")
REF_CUBATURE_SO = os.path.join(ROOT, "oracle", "_ref", "libhcubature_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)
INTEGRAND_V = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, _dp, ctypes.c_void_p, ctypes.c_uint, _dp)
F1 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_void_p)

_SIGS = {
    "orc_hcubature_v": (ctypes.c_int, [ctypes.c_uint, INTEGRAND_V, ctypes.c_void_p, ctypes.c_uint, _dp, _dp,
                                       ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "orc_last_neval": (ctypes.c_longlong, []),
    "orc_last_nregions": (ctypes.c_longlong, []),
    "orc_qag15": (ctypes.c_double, [F1, ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int]),
    "orc_gauss_legendre": (None, [ctypes.c_int, _dp, _dp]),
    "orc_create": (ctypes.c_void_p, [_dp, _dp, ctypes.c_int, ctypes.c_double]),
    "orc_create_pk": (ctypes.c_void_p, [_dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_int]),
    "orc_destroy": (None, [ctypes.c_void_p]),
    "orc_set_analytic": (None, [ctypes.c_void_p, ctypes.c_int]),
    "orc_get_tables": (None, [ctypes.c_void_p] + [_dp] * 7),
    "orc_linear_pk": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double]),
    "orc_sigmam": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    "orc_lgr": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double]),
    "orc_f_K_at_z": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double]),
    "orc_bispec": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 4),
    "orc_integrand_bkappa": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 4),
    "orc_bispec_idx": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 4 + [ctypes.c_int,
                                                                                      ctypes.c_double]),
    "orc_bkappa": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 3),
    "orc_bkappa_batch": (None, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, ctypes.c_int]),
    "orc_map3_integrand": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 3 + [_dp]),
    "orc_map3_integrand_batch": (None, [ctypes.c_void_p, ctypes.c_longlong, _dp, _dp, _dp, ctypes.c_int]),
    "orc_map3": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _llp]),
    "orc_bessel_jn": (ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    "orc_bessel_zero_jn": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
    "orc_gamma_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "orc_gamma_integrand": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 6 + [_dp, _dp]),
    "orc_gamma_integrand_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 3 +
                                  [ctypes.c_size_t, _dp, _dp, ctypes.c_int]),
    "orc_gamma": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int, _dp, _dp,
                                                                                             _llp]),
    "orc_ggg_analytic": (None, [ctypes.c_double] * 3 + [_dp, _dp]),
    "orc_gamma_center": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 4 +
                         [ctypes.c_int, ctypes.c_int, _dp, _dp, _llp]),
    "orc_pk_nonlinear": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "orc_linear_pk_at_z": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "orc_om_m_of_z": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double]),
    "orc_om_v_of_z": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double]),
    "orc_limber_pkappa": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "orc_gq96_pkappa": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    "orc_pell": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int]),
    "orc_map2_integrand": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "orc_convergence_power_spectrum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_int, _dp, _dp, _llp]),
    "orc_map2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, _dp, _dp, _llp]),
    "orc_limber_triple_pk": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 4),
    "orc_cov_gauss_integrand": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 4 + [_dp]),
    "orc_t1_integrand": (ctypes.c_double, [ctypes.c_void_p] + [ctypes.c_double] * 3 + [_dp] +
                         [ctypes.c_double] * 4 + [ctypes.c_int]),
    "orc_map3_cov_gauss": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                          _dp, _llp]),
    "orc_t1": (ctypes.c_int, [ctypes.c_void_p, _dp] + [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double,
                                                                                ctypes.c_int, _dp, _dp, _llp]),
    "orc_t1_total": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp] + [ctypes.c_double] * 4 +
                     [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _llp]),
    "orc_t1_total_finite": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int] + [ctypes.c_double] * 5 +
                            [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _llp]),
    "orc_t2": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int] + [ctypes.c_double] * 5 +
               [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _llp]),
    "orc_t2_total": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, ctypes.c_int] + [ctypes.c_double] * 5 +
                     [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _llp]),
    "orc_read_cosmology": (ctypes.c_int, [ctypes.c_char_p, _dp]),
    "orc_read_n_of_z": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _dp]),
}

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(ORACLE_SO)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def d(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    def __init__(self, cosmo, nz, n_z=100, z_max=3.0, pk=None):
        p = np.array([cosmo.h, cosmo.sigma8, cosmo.omb, cosmo.ns, cosmo.w, cosmo.om, cosmo.zmax], dtype=np.float64)
        nz = np.ascontiguousarray(nz, dtype=np.float64)
        if pk is None:
            self.h = lib().orc_create(d(p), d(nz), n_z, z_max)
        else:
            kk = np.ascontiguousarray(pk[0], dtype=np.float64)
            pp = np.ascontiguousarray(pk[1], dtype=np.float64)
            self.h = lib().orc_create_pk(d(p), d(nz), n_z, z_max, d(kk), d(pp), kk.size)
        self.n_z = n_z

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def tables(self):
        out = [np.zeros(self.n_z) for _ in range(6)]
        norm = ctypes.c_double()
        lib().orc_get_tables(self.h, *[d(a) for a in out], ctypes.byref(norm))
        return dict(zip(["f_K", "g", "D1", "r_sigma", "n_eff", "ncur"], out), norm=norm.value)

    def set_analytic(self, on=True):
        lib().orc_set_analytic(self.h, 1 if on else 0)

    def bkappa(self, ell, nthreads=0):
        ell = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(ell.shape[0])
        lib().orc_bkappa_batch(self.h, ell.shape[0], d(ell), d(out), nthreads)
        return out

    def map3_integrand(self, x, th, nthreads=0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        th = np.ascontiguousarray(th, dtype=np.float64)
        out = np.zeros(x.shape[0])
        lib().orc_map3_integrand_batch(self.h, x.shape[0], d(x), d(th), d(out), nthreads)
        return out

    def map3(self, th, reltol=1e-4, nthreads=0):
        th = np.ascontiguousarray(th, dtype=np.float64)
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        rc = lib().orc_map3(self.h, d(th), reltol, nthreads, ctypes.byref(v), ctypes.byref(e), ctypes.byref(n))
        assert rc == 0
        return v.value, e.value, n.value

    def gamma_init(self, prec_h, prec_k_arg):
        lib().orc_gamma_init(self.h, prec_h, prec_k_arg)

    def gamma_integrand(self, which, z, phi, psi, x1, x2, x3):
        re, im = ctypes.c_double(), ctypes.c_double()
        lib().orc_gamma_integrand(self.h, which, z, phi, psi, x1, x2, x3, ctypes.byref(re), ctypes.byref(im))
        return re.value + 1j * im.value

    def gamma_integrand_batch(self, which, x123, pts, nthreads=0):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((pts.shape[0], 2))
        lib().orc_gamma_integrand_batch(self.h, which, x123[0], x123[1], x123[2], pts.shape[0], d(pts), d(out),
                                        nthreads)
        return out[:, 0] + 1j * out[:, 1]

    def gamma(self, which, x1, x2, x3, reltol=1e-3, nthreads=0, center=0):
        """Gamma_which of the triangle; center 0 = centroid (the reference's rotation), 1 = orthocenter."""
        re, im, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib().orc_gamma_center(self.h, which, x1, x2, x3, reltol, nthreads, center, ctypes.byref(re),
                               ctypes.byref(im), ctypes.byref(n))
        return re.value + 1j * im.value, n.value

    @staticmethod
    def ggg_analytic(r1, r2, r3):
        """The closed-form Gamma_0 of the analytic test bispectrum (CV/gamma.cpp:83-94, 567-586)."""
        re, im = ctypes.c_double(), ctypes.c_double()
        lib().orc_ggg_analytic(r1, r2, r3, ctypes.byref(re), ctypes.byref(im))
        return re.value + 1j * im.value

    def convergence_power_spectrum(self, ell, reltol=1e-6, flags=0):
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        rc = lib().orc_convergence_power_spectrum(self.h, ell, reltol, flags, ctypes.byref(v), ctypes.byref(e),
                                                  ctypes.byref(n))
        assert rc == 0
        return v.value, e.value, n.value

    def map2(self, theta, reltol=1e-5, flags=0, sigma=0.0, ngal=1.0):
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        rc = lib().orc_map2(self.h, theta, reltol, flags, sigma, ngal, ctypes.byref(v), ctypes.byref(e),
                            ctypes.byref(n))
        assert rc == 0
        return v.value, e.value, n.value

    def map3_cov_gauss(self, th6, area, reltol=1e-3, nthreads=0):
        th = np.ascontiguousarray(th6, dtype=np.float64)
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_map3_cov_gauss(self.h, d(th), area, reltol, nthreads, ctypes.byref(v), ctypes.byref(e),
                                        ctypes.byref(n)) == 0
        return v.value, e.value, n.value

    def t1(self, th6, area, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0, reltol=1e-4, nthreads=0):
        th = np.ascontiguousarray(th6, dtype=np.float64)
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_t1(self.h, d(th), area, lmin, sigma, n_gal, flags, reltol, nthreads, ctypes.byref(v),
                            ctypes.byref(e), ctypes.byref(n)) == 0
        return v.value, e.value, n.value

    def t1_total(self, th123, th456, area, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0, reltol=1e-4, nthreads=0):
        a = np.ascontiguousarray(th123, dtype=np.float64)
        b = np.ascontiguousarray(th456, dtype=np.float64)
        v, n = ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_t1_total(self.h, d(a), d(b), area, lmin, sigma, n_gal, flags, reltol, nthreads,
                                  ctypes.byref(v), ctypes.byref(n)) == 0
        return v.value, n.value

    def t1_total_finite(self, th123, th456, geometry, theta_max, theta_max2=0.0, lmin=0.0, sigma=0.3, n_gal=1.0,
                        flags=0, reltol=1e-2, nthreads=0):
        a = np.ascontiguousarray(th123, dtype=np.float64)
        b = np.ascontiguousarray(th456, dtype=np.float64)
        v, n = ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_t1_total_finite(self.h, d(a), d(b), geometry, theta_max, theta_max2, lmin, sigma, n_gal,
                                         flags, reltol, nthreads, ctypes.byref(v), ctypes.byref(n)) == 0
        return v.value, n.value

    def t2(self, th6, geometry, theta_max, theta_max2=0.0, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0, reltol=1e-4,
           nthreads=0):
        th = np.ascontiguousarray(th6, dtype=np.float64)
        v, n = ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_t2(self.h, d(th), geometry, theta_max, theta_max2, lmin, sigma, n_gal, flags, reltol,
                            nthreads, ctypes.byref(v), ctypes.byref(n)) == 0
        return v.value, n.value

    def t2_total(self, th123, th456, geometry, theta_max, theta_max2=0.0, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0,
                 reltol=1e-4, nthreads=0):
        a = np.ascontiguousarray(th123, dtype=np.float64)
        b = np.ascontiguousarray(th456, dtype=np.float64)
        v, n = ctypes.c_double(), ctypes.c_longlong()
        assert lib().orc_t2_total(self.h, d(a), d(b), geometry, theta_max, theta_max2, lmin, sigma, n_gal, flags,
                                  reltol, nthreads, ctypes.byref(v), ctypes.byref(n)) == 0
        return v.value, n.value

    def scalar(self, name, *args):
        return getattr(lib(), "orc_" + name)(self.h, *args)
