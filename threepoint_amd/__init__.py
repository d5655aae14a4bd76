"""threepoint_amd -- MI355X-native <Map^3> / Gamma_i modelling path.

Python mirror of the reference's class API (classed_version/bispectrum.hpp,
apertureStatistics.hpp, gamma.hpp) over the C ABI in include/threepoint_amd.h.
All numerics run in the HIP library threepoint_amd/_lib/libthreepoint_amd.so;
there is no CPU fallback: importing works without a GPU, but every compute call
needs the library and a HIP device and raises otherwise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libthreepoint_amd.so")

MODE_FP64, MODE_FP32, MODE_SURROGATE, MODE_ANALYTIC = 0, 1, 2, 3
SUR_INFER_HILO_BF16, SUR_INFER_FP16 = 0, 1  # surrogate inference kernels (tp_surrogate_inference)
CENTER_CENTROID, CENTER_ORTHOCENTER = 0, 1  # triangle centers of the Gamma_i (gamma.cpp:12-16)
# survey geometries of the covariance (the cuda_version's type codes)
GEOMETRY_CIRCLE, GEOMETRY_SQUARE, GEOMETRY_INFINITE, GEOMETRY_RECTANGLE = 0, 1, 2, 3
PK_LINEAR, PK_T17, PK_SHAPENOISE_ONLY = 1, 2, 4  # flags of the power-spectrum calls
ERROR_INDIVIDUAL, ERROR_PAIRED, ERROR_L2, ERROR_L1, ERROR_LINF = 0, 1, 2, 3, 4
ARCMIN = 2.9088820866e-4  # classed_version/helper.cpp:139

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)

INTEGRAND_V = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, _dp, ctypes.c_void_p,
                               ctypes.c_uint, _dp)


class TpError(RuntimeError):
    pass


class tp_cosmology(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("h", "sigma8", "omb", "ns", "w", "om", "zmax")]


class tp_batch_stats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_longlong), ("regions", ctypes.c_longlong), ("points", ctypes.c_longlong),
                ("launches", ctypes.c_longlong), ("kernel_ms", ctypes.c_double), ("integrand_ms", ctypes.c_double),
                ("wall_ms", ctypes.c_double),
                ("surrogate_points", ctypes.c_longlong), ("exact_checks", ctypes.c_longlong),
                ("fallback_batches", ctypes.c_longlong), ("max_check_relerr", ctypes.c_double),
                ("fallback_integrals", ctypes.c_longlong), ("host_ms", ctypes.c_double),
                ("wait_ms", ctypes.c_double), ("graph_rebuilds", ctypes.c_longlong),
                ("integrand_sum_ms", ctypes.c_double), ("mlp_ms", ctypes.c_double),
                ("mlp_launches", ctypes.c_longlong), ("mlp_points", ctypes.c_longlong)]


class tp_train_config(ctypes.Structure):
    _fields_ = [("nsamples", ctypes.c_longlong), ("epochs", ctypes.c_int), ("batch", ctypes.c_int),
                ("lr", ctypes.c_double), ("lr_final", ctypes.c_double),
                ("lnl_min", ctypes.c_double), ("lnl_max", ctypes.c_double),
                ("om", ctypes.c_double * 2), ("s8", ctypes.c_double * 2), ("h", ctypes.c_double * 2),
                ("w", ctypes.c_double * 2), ("ns", ctypes.c_double * 2), ("omb", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("n_z", ctypes.c_int), ("z_max", ctypes.c_double),
                ("nz", _dp)]


class tp_train_report(ctypes.Structure):
    _fields_ = [("final_loss", ctypes.c_double), ("val_rmse", ctypes.c_double), ("val_maxerr", ctypes.c_double),
                ("gen_ms", ctypes.c_double), ("train_ms", ctypes.c_double), ("steps", ctypes.c_longlong)]


_SIGS = {
    "tp_last_error": (ctypes.c_char_p, []),
    "tp_device_count": (ctypes.c_int, [_ip]),
    "tp_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "tp_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tp_cosmo_build": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(tp_cosmology), _dp, ctypes.c_int,
                                      ctypes.c_size_t, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    "tp_cosmo_build_pk": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(tp_cosmology), _dp,
                                         ctypes.c_int, ctypes.c_size_t, ctypes.c_double, _dp, _dp, ctypes.c_int,
                                         ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "tp_cosmo_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tp_cosmo_count": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    "tp_cosmo_tables": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "tp_bkappa_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp, _dp,
                                      ctypes.c_int]),
    "tp_map3_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t, _dp, _dp,
                                    ctypes.c_int]),
    "tp_map3_single": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double,
                                      ctypes.c_int, _dp, _dp, _llp]),
    "tp_map3_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp, ctypes.c_double,
                                     ctypes.c_int, _dp, _dp, _llp]),
    "tp_last_batch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(tp_batch_stats)]),
    "tp_gamma_setup": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "tp_gamma_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_size_t, _dp, _dp]),
    "tp_pk_at_z_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                       ctypes.c_int, _dp]),
    "tp_limber_pkappa_eval": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                             ctypes.c_int, _dp]),
    "tp_pell": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp, ctypes.c_int,
                               ctypes.c_double, ctypes.c_double, _dp]),
    "tp_pkappa_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp, ctypes.c_double,
                                       ctypes.c_int, _dp, _dp, _llp]),
    "tp_map2_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _llp]),
    "tp_map3_cov_gauss_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp,
                                               ctypes.c_double, ctypes.c_double, _dp, _dp, _llp]),
    "tp_map3_cov_t1_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_double, _dp, _llp]),
    "tp_map3_cov_t1_batch_geometry": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp,
                                                     ctypes.c_int] + [ctypes.c_double] * 6 +
                                      [ctypes.c_int, ctypes.c_double, _dp, _llp]),
    "tp_map3_cov_t2_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp, ctypes.c_int] +
                             [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_double, _dp, _llp]),
    "tp_map3_cov_t1_batch_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, _ip, _dp,
                                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _llp]),
    "tp_read_n_of_z": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
    "tp_read_triangle_configurations": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, _dp,
                                                       ctypes.POINTER(ctypes.c_size_t)]),
    "tp_gamma_batch_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                           ctypes.c_double, ctypes.c_int, _dp, _dp, _llp]),
    "tp_gamma_fp32_window": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "tp_gamma_batch_center": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _llp]),
    "tp_bispec_eval_idx": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                          _dp]),
    "tp_set_graph_rounds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_set_batch_groups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_set_zero_copy_regions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "tp_set_serial_regions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "tp_set_round_log": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_set_host_threads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_set_keep_pool_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_ulonglong]),
    "tp_set_batch_limits": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                           ctypes.c_longlong]),
    "tp_round_log": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "tp_gamma_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                      ctypes.c_double, _dp, _dp, _llp]),
    "tp_hcubature_v": (ctypes.c_int, [ctypes.c_uint, INTEGRAND_V, ctypes.c_void_p, ctypes.c_uint, _dp, _dp,
                                      ctypes.c_size_t, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp]),
    "tp_hcubature_last_neval": (ctypes.c_longlong, []),
    "tp_census_map3_point": (ctypes.c_int, [_dp]),
    "tp_census_map3_pse": (ctypes.c_int, [_dp]),
    "tp_census_gamma_term": (ctypes.c_int, [_dp]),
    "tp_census_power_point": (ctypes.c_int, [_dp]),
    "tp_hbm_copy_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _dp]),
    "tp_set_region_lanes": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_region_lanes_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "tp_surrogate_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(tp_train_config),
                                          ctypes.POINTER(tp_train_report)]),
    "tp_surrogate_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "tp_surrogate_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "tp_surrogate_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
    "tp_surrogate_exact_regions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    "tp_surrogate_exact_ratio": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "tp_surrogate_inference": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "tp_surrogate_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _dp, _dp]),
    "tp_surrogate_bkappa": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, _dp,
                                           _dp]),
    "tp_mlp_num_params": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "tp_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "tp_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "tp_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, _fp, _fp, _dp, _fp]),
}

_lib = None


def lib():
    """The loaded HIP library (raises TpError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TpError("libthreepoint_amd.so is not built; run __graft_entry__.build() "
                          "or python -m threepoint_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().tp_last_error().decode(errors="replace")
        raise TpError("%s failed (code %d): %s" % (what, rc, msg))


def _d(a):
    return a.ctypes.data_as(_dp)


def device_count():
    n = ctypes.c_int(0)
    rc = lib().tp_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


class Context:
    """One device + stream (tp_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().tp_ctx_create(device, ctypes.byref(h)), "tp_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().tp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = tp_batch_stats()
        _check(lib().tp_last_batch_stats(self.h, ctypes.byref(s)), "tp_last_batch_stats")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def set_host_threads(self, n=0):
        """OpenMP threads of the host side of the batched cubatures (0: the default)."""
        _check(lib().tp_set_host_threads(self.h, int(n)), "tp_set_host_threads")

    def set_keep_pool_bytes(self, nbytes):
        """Cap on the host region pools kept between batched calls (0: release and keep nothing)."""
        _check(lib().tp_set_keep_pool_bytes(self.h, int(nbytes)), "tp_set_keep_pool_bytes")

    def set_batch_limits(self, max_rounds=0, max_regions=0, round_points=0):
        """Guards of the batched cubatures (TP_ECAPACITY beyond them) and the host-rule batches' slice size;
        0 restores a default."""
        _check(lib().tp_set_batch_limits(self.h, int(max_rounds), int(max_regions), int(round_points)),
               "tp_set_batch_limits")

    def set_surrogate_exact_regions(self, n=1024):
        """Surrogate mode: integrals are evaluated exactly (FP32 form) until they hold n regions."""
        _check(lib().tp_surrogate_exact_regions(self.h, int(n)), "tp_surrogate_exact_regions")

    def set_surrogate_inference(self, kind=1):
        """Surrogate inference kernel: SUR_INFER_FP16 (default) or SUR_INFER_HILO_BF16."""
        _check(lib().tp_surrogate_inference(self.h, int(kind)), "tp_surrogate_inference")

    def set_surrogate_exact_ratio(self, ratio=16.0):
        """Surrogate mode: triples with max(theta) >= ratio * min(theta) are evaluated exactly (0: none)."""
        _check(lib().tp_surrogate_exact_ratio(self.h, float(ratio)), "tp_surrogate_exact_ratio")

    def set_round_log(self, on=True):
        _check(lib().tp_set_round_log(self.h, 1 if on else 0), "tp_set_round_log")

    def round_log(self):
        """Per-round trace of the last batched <Map^3> call: rows (group, regions, launch begin, launch end,
        completion begin, completion end [ms since the call began], wait [ms], device begin, device end [ms
        since an event recorded when the call began])."""
        n = ctypes.c_size_t()
        _check(lib().tp_round_log(self.h, None, 0, ctypes.byref(n)), "tp_round_log")
        out = np.zeros(n.value)
        _check(lib().tp_round_log(self.h, _d(out), n.value, ctypes.byref(n)), "tp_round_log")
        return out.reshape(-1, 9)


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Cosmology:
    """classed_version/cosmology.hpp:11-29 (ow and omc derived as in cosmology.cpp:71-72)."""

    KEYS = {"h": "h", "sigma_8": "sigma8", "Omega_b": "omb", "n_s": "ns", "w": "w", "Omega_m": "om", "z_max": "zmax"}

    def __init__(self, h=0.7, sigma8=0.82, omb=0.047, ns=0.97, w=-1.0, om=0.3, zmax=3.0):
        self.h, self.sigma8, self.omb, self.ns, self.w, self.om, self.zmax = h, sigma8, omb, ns, w, om, zmax

    @property
    def ow(self):
        return 1 - self.om

    @property
    def omc(self):
        return self.om - self.omb

    @classmethod
    def from_file(cls, fn):
        """classed_version/cosmology.cpp:10-73: "name value" lines, '#' comments and empty lines skipped,
        an unknown name is an error; parameters the file does not name keep their defaults."""
        kw = {}
        with open(fn) as f:
            for line in f:
                line = line.rstrip("\n")
                if not line or line[0] == "#":
                    continue
                parts = line.split()
                if not parts or parts[0] not in cls.KEYS:
                    raise ValueError("cosmology: Parameter file in wrong format: %r" % line)
                kw[cls.KEYS[parts[0]]] = float(parts[1]) if len(parts) > 1 else 0.0
        return cls(**kw)

    def as_struct(self):
        return tp_cosmology(self.h, self.sigma8, self.omb, self.ns, self.w, self.om, self.zmax)


class CosmologyTables:
    """Device-resident tables of one or more cosmologies (tp_cosmo)."""

    def __init__(self, cosmos, nz, n_z=None, z_max=3.0, ctx=None, pk=None):
        """pk = (k_h, p_k): a tabulated linear P(k) (shared, or one row per cosmology) instead of the
        Eisenstein & Hu fit, the reference's Pk_given input (classed_version/bispectrum.cpp:81-89)."""
        self.ctx = ctx or default_context()
        if isinstance(cosmos, Cosmology):
            cosmos = [cosmos]
        self.cosmos = list(cosmos)
        nz = np.ascontiguousarray(nz, dtype=np.float64)
        if nz.ndim == 1:
            n_z = nz.size if n_z is None else n_z
            stride = 0
        else:
            n_z = nz.shape[1]
            stride = n_z
        arr = (tp_cosmology * len(self.cosmos))(*[c.as_struct() for c in self.cosmos])
        h = ctypes.c_void_p()
        if pk is None:
            _check(lib().tp_cosmo_build(self.ctx.h, len(self.cosmos), arr, _d(nz), n_z, stride, z_max,
                                        ctypes.byref(h)), "tp_cosmo_build")
        else:
            kk = np.ascontiguousarray(pk[0], dtype=np.float64)
            pp = np.ascontiguousarray(pk[1], dtype=np.float64)
            if kk.shape != pp.shape:
                raise ValueError("k and P(k) tables differ in shape")
            nk = kk.shape[-1]
            pstride = nk if kk.ndim == 2 else 0
            _check(lib().tp_cosmo_build_pk(self.ctx.h, len(self.cosmos), arr, _d(nz), n_z, stride, z_max, _d(kk),
                                           _d(pp), nk, pstride, ctypes.byref(h)), "tp_cosmo_build_pk")
        self.h = h
        self.n_z = n_z
        self.z_max = z_max

    def close(self):
        if getattr(self, "h", None):
            lib().tp_cosmo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cosmos)

    def tables(self, i=0):
        out = [np.zeros(self.n_z) for _ in range(6)]
        norm = ctypes.c_double()
        _check(lib().tp_cosmo_tables(self.h, i, *[_d(a) for a in out], ctypes.byref(norm)), "tp_cosmo_tables")
        return dict(zip(["f_K", "g", "D1", "r_sigma", "n_eff", "ncur"], out), norm=norm.value)

    def bkappa(self, ell123, icosmo=0, mode=MODE_FP64):
        ell = np.ascontiguousarray(ell123, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(ell.shape[0])
        _check(lib().tp_bkappa_eval(self.ctx.h, self.h, icosmo, ell.shape[0], _d(ell), _d(out), mode),
               "tp_bkappa_eval")
        return out

    def surrogate_bkappa(self, ell123, icosmo=0):
        ell = np.ascontiguousarray(ell123, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(ell.shape[0])
        _check(lib().tp_surrogate_bkappa(self.ctx.h, self.h, icosmo, ell.shape[0], _d(ell), _d(out)),
               "tp_surrogate_bkappa")
        return out

    def map3_integrand(self, x, thetas, icosmo=0, mode=MODE_FP64):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        th = np.ascontiguousarray(thetas, dtype=np.float64)
        out = np.zeros(x.shape[0])
        _check(lib().tp_map3_eval(self.ctx.h, self.h, icosmo, _d(th), x.shape[0], _d(x), _d(out), mode),
               "tp_map3_eval")
        return out

    def map3_single(self, thetas, reltol=1e-4, icosmo=0, mode=MODE_FP64):
        th = np.ascontiguousarray(thetas, dtype=np.float64)
        v, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().tp_map3_single(self.ctx.h, self.h, icosmo, _d(th), reltol, mode, ctypes.byref(v),
                                    ctypes.byref(e), ctypes.byref(n)), "tp_map3_single")
        return v.value, e.value, n.value

    def map3_batch(self, thetas, icosmo=None, reltol=1e-4, mode=MODE_FP64):
        th = np.ascontiguousarray(thetas, dtype=np.float64).reshape(-1, 3)
        n = th.shape[0]
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, err = np.zeros(n), np.zeros(n)
        ne = np.zeros(n, dtype=np.int64)
        _check(lib().tp_map3_batch(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(th), reltol, mode, _d(val),
                                   _d(err), ne.ctypes.data_as(_llp)), "tp_map3_batch")
        return val, err, ne

    def gamma_eval(self, which, x123, pts, icosmo=0):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((pts.shape[0], 2))
        _check(lib().tp_gamma_eval(self.ctx.h, self.h, icosmo, which, x123[0], x123[1], x123[2], pts.shape[0],
                                   _d(pts), _d(out)), "tp_gamma_eval")
        return out[:, 0] + 1j * out[:, 1]

    def gamma_batch(self, x123, reltol=1e-3, icosmo=0, mode=MODE_FP64, center=CENTER_CENTROID):
        x = np.ascontiguousarray(x123, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        re, im = np.zeros(4 * n), np.zeros(4 * n)
        ne = np.zeros(4 * n, dtype=np.int64)
        _check(lib().tp_gamma_batch_center(self.ctx.h, self.h, icosmo, n, _d(x), reltol, mode, center, _d(re),
                                           _d(im), ne.ctypes.data_as(_llp)), "tp_gamma_batch")
        return (re + 1j * im).reshape(n, 4), ne.reshape(n, 4)

    def bispec(self, k123z, icosmo=0):
        """bispec(k1, k2, k3, z) at rows (k1, k2, k3 [h/Mpc], z): the table interval of z."""
        x = np.ascontiguousarray(k123z, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(x.shape[0])
        _check(lib().tp_bispec_eval(self.ctx.h, self.h, icosmo, x.shape[0], _d(x), _d(out)), "tp_bispec_eval")
        return out

    def bispec_idx(self, k123z_idx, icosmo=0):
        """bispec(k1, k2, k3, z, idx, didx) at rows of six (the reference's 6-argument form)."""
        x = np.ascontiguousarray(k123z_idx, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(x.shape[0])
        _check(lib().tp_bispec_eval_idx(self.ctx.h, self.h, icosmo, x.shape[0], _d(x), _d(out)),
               "tp_bispec_eval_idx")
        return out

    def integrand_bkappa(self, zell, icosmo=0):
        """integrand_bkappa at rows (z, ell1, ell2, ell3)."""
        x = np.ascontiguousarray(zell, dtype=np.float64).reshape(-1, 4)
        out = np.zeros(x.shape[0])
        _check(lib().tp_integrand_bkappa_eval(self.ctx.h, self.h, icosmo, x.shape[0], _d(x), _d(out)),
               "tp_integrand_bkappa_eval")
        return out


    # ---------------- convergence power spectrum and <Map^2>
    def pk_at_z(self, kz, icosmo=0, flags=0):
        """P_NL(k, z) (or D1^2 P_lin with PK_LINEAR) at points kz[n, 2] = (k [h/Mpc], z)."""
        x = np.ascontiguousarray(kz, dtype=np.float64).reshape(-1, 2)
        out = np.zeros(x.shape[0])
        _check(lib().tp_pk_at_z_eval(self.ctx.h, self.h, icosmo, x.shape[0], _d(x), flags, _d(out)),
               "tp_pk_at_z_eval")
        return out

    def limber_pkappa(self, ellz, icosmo=0, flags=0):
        """limber_integrand_power_spectrum at points ellz[n, 2] = (ell, z)."""
        x = np.ascontiguousarray(ellz, dtype=np.float64).reshape(-1, 2)
        out = np.zeros(x.shape[0])
        _check(lib().tp_limber_pkappa_eval(self.ctx.h, self.h, icosmo, x.shape[0], _d(x), flags, _d(out)),
               "tp_limber_pkappa_eval")
        return out

    def pell(self, ell, icosmo=0, flags=0, sigma=0.0, n_gal=1.0):
        """Pell: 96-node Limber sum of P_NL plus 0.5 sigma^2 / n_gal (n_gal in rad^-2)."""
        e = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1)
        out = np.zeros(e.size)
        _check(lib().tp_pell(self.ctx.h, self.h, icosmo, e.size, _d(e), flags, sigma, n_gal, _d(out)), "tp_pell")
        return out

    def pkappa_batch(self, ell, icosmo=None, reltol=1e-6, flags=0):
        """convergence_power_spectrum for many multipoles: (values, error estimates, points)."""
        e = np.ascontiguousarray(ell, dtype=np.float64).reshape(-1)
        n = e.size
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, err, ne = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_pkappa_batch(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(e), reltol, flags, _d(val),
                                     _d(err), ne.ctypes.data_as(_llp)), "tp_pkappa_batch")
        return val, err, ne

    def map2_batch(self, theta, icosmo=None, reltol=1e-5, flags=0, sigma=0.0, ngal=1.0):
        """<Map^2>(theta [rad]) for one source bin; sigma != 0 adds shape noise (ngal in arcmin^-2)."""
        t = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        n = t.size
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, err, ne = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_map2_batch(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(t), reltol, flags, sigma, ngal,
                                   _d(val), _d(err), ne.ctypes.data_as(_llp)), "tp_map2_batch")
        return val, err, ne


    # ---------------- Gaussian <Map^3> covariance
    def map3_cov_gauss(self, theta6, area, icosmo=None, reltol=1e-3):
        """The classed_version's MapMapMap_covariance_Gauss for pairs theta6[n, 6] [rad]."""
        t = np.ascontiguousarray(theta6, dtype=np.float64).reshape(-1, 6)
        n = t.shape[0]
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, err, ne = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_map3_cov_gauss_batch(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(t), area, reltol,
                                             _d(val), _d(err), ne.ctypes.data_as(_llp)), "tp_map3_cov_gauss_batch")
        return val, err, ne

    def map3_cov_t1(self, theta6, area, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0, icosmo=None, reltol=1e-4,
                    mode=MODE_FP64):
        """T1_total of the cuda_version's covariance for an infinite survey, pairs theta6[n, 6] [rad];
        mode MODE_FP64 or MODE_FP32 (FP32 Pell node values, FP64 sums)."""
        t = np.ascontiguousarray(theta6, dtype=np.float64).reshape(-1, 6)
        n = t.shape[0]
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, ne = np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_map3_cov_t1_batch_mode(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(t), area, lmin,
                                               sigma, n_gal, flags, reltol, mode, _d(val), ne.ctypes.data_as(_llp)),
               "tp_map3_cov_t1_batch_mode")
        return val, ne

    def map3_cov_t1_geometry(self, theta6, geometry, theta_max, theta_max2=0.0, area=1.0, lmin=0.0, sigma=0.3,
                             n_gal=1.0, flags=0, icosmo=None, reltol=1e-2):
        """T1_total of the cuda_version's covariance for any survey geometry (GEOMETRY_*; the finite ones
        are 6-D integrals at the reference's tolerance 1e-2, GEOMETRY_INFINITE is map3_cov_t1)."""
        t = np.ascontiguousarray(theta6, dtype=np.float64).reshape(-1, 6)
        n = t.shape[0]
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, ne = np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_map3_cov_t1_batch_geometry(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(t),
                                                   int(geometry), theta_max, theta_max2, area, lmin, sigma, n_gal,
                                                   flags, reltol, _d(val), ne.ctypes.data_as(_llp)),
               "tp_map3_cov_t1_batch_geometry")
        return val, ne

    def map3_cov_t2(self, theta6, geometry, theta_max, theta_max2=0.0, lmin=0.0, sigma=0.3, n_gal=1.0, flags=0,
                    icosmo=None, reltol=1e-4):
        """T2_total of the cuda_version's covariance, pairs theta6[n, 6] [rad]; geometry GEOMETRY_CIRCLE
        (theta_max the radius), GEOMETRY_SQUARE (the side), GEOMETRY_RECTANGLE (sides theta_max,
        theta_max2) or GEOMETRY_INFINITE (T2 = 0)."""
        t = np.ascontiguousarray(theta6, dtype=np.float64).reshape(-1, 6)
        n = t.shape[0]
        ic = np.zeros(n, dtype=np.int32) if icosmo is None else np.ascontiguousarray(icosmo, dtype=np.int32)
        val, ne = np.zeros(n), np.zeros(n, dtype=np.int64)
        _check(lib().tp_map3_cov_t2_batch(self.ctx.h, self.h, n, ic.ctypes.data_as(_ip), _d(t), int(geometry),
                                          theta_max, theta_max2, lmin, sigma, n_gal, flags, reltol, _d(val),
                                          ne.ctypes.data_as(_llp)), "tp_map3_cov_t2_batch")
        return val, ne


def census_map3_point():
    """Algorithmic op census of one <Map^3> integrand point (see tp_census_map3_point)."""
    out = np.zeros(5)
    _check(lib().tp_census_map3_point(_d(out)), "tp_census_map3_point")
    pse = np.zeros(2)
    _check(lib().tp_census_map3_pse(_d(pse)), "tp_census_map3_pse")
    return dict(node_flops=out[0], node_elem=out[1], point_flops=out[2], point_elem=out[3], nodes=out[4],
                pse_flops=pse[0], pse_elem=pse[1])


def flop_equivalents_per_point(flops_per_elem=20.0):
    c = census_map3_point()
    return c["nodes"] * (c["node_flops"] + flops_per_elem * c["node_elem"]) + c["point_flops"] + \
        flops_per_elem * c["point_elem"]


def census_power_point():
    """Algorithmic op census of one <Map^2> and one P_kappa integrand point (see tp_census_power_point)."""
    out = np.zeros(6)
    _check(lib().tp_census_power_point(_d(out)), "tp_census_power_point")
    return dict(map2_flops=out[0], map2_elem=out[1], pk_flops=out[2], pk_elem=out[3], node_flops=out[4],
                node_elem=out[5])


def census_gamma_term():
    """Algorithmic op census of one Ogata term of the Gamma_i integrand (see tp_census_gamma_term)."""
    out = np.zeros(4)
    _check(lib().tp_census_gamma_term(_d(out)), "tp_census_gamma_term")
    return dict(term_flops=out[0], term_elem=out[1], shape_flops=out[2], shape_elem=out[3])


def flop_equivalents_per_gamma_point(prec_k, flops_per_elem=20.0):
    """Work of one Gamma cubature point: 3 (prec_k - 1) Ogata terms plus the per-point shape."""
    c = census_gamma_term()
    return 3 * (prec_k - 1) * (c["term_flops"] + flops_per_elem * c["term_elem"]) + c["shape_flops"] + \
        flops_per_elem * c["shape_elem"]


def gamma_setup(prec_h=1e-3, prec_k_arg=5.0, ctx=None):
    ctx = ctx or default_context()
    _check(lib().tp_gamma_setup(ctx.h, prec_h, prec_k_arg), "tp_gamma_setup")


def hcubature_v(fdim, f, dim, xmin, xmax, maxEval=0, reqAbsError=0.0, reqRelError=1e-4, norm=ERROR_L1):
    """classed_version/cubature.h:90-94 with a Python callable f(x[npt, dim]) -> fval[npt, fdim]."""
    def cb(ndim, npt, x, fdata, fd, fval):
        try:
            xs = np.ctypeslib.as_array(x, shape=(npt * ndim,)).reshape(npt, ndim)
            v = np.asarray(f(xs), dtype=np.float64).reshape(npt * fd)
            np.ctypeslib.as_array(fval, shape=(npt * fd,))[:] = v
            return 0
        except Exception:
            return 1
    cfun = INTEGRAND_V(cb)
    xmin = np.ascontiguousarray(xmin, dtype=np.float64)
    xmax = np.ascontiguousarray(xmax, dtype=np.float64)
    val, err = np.zeros(fdim), np.zeros(fdim)
    rc = lib().tp_hcubature_v(fdim, cfun, None, dim, _d(xmin), _d(xmax), maxEval, reqAbsError, reqRelError, norm,
                              _d(val), _d(err))
    return rc, val, err, lib().tp_hcubature_last_neval()


class BispectrumCalculator:
    """classed_version/bispectrum.hpp:113-509 (n(z) constructor, :207)."""

    def __init__(self, cosmo, n_of_z, n_z, z_max, ctx=None):
        self.cosmo = cosmo
        self.z_max = z_max
        self.tables_ = CosmologyTables([cosmo], np.asarray(n_of_z, dtype=np.float64)[:n_z], n_z, z_max, ctx)

    def bkappa(self, ell1, ell2, ell3, mode=MODE_FP64):
        return float(self.tables_.bkappa(np.array([[ell1, ell2, ell3]]), 0, mode)[0])

    def bispec(self, k1, k2, k3, z, idx=None, didx=None):
        """bispec(k1, k2, k3, z, idx, didx) (bispectrum.hpp:247); without idx/didx the interval of z."""
        if idx is None:
            return float(self.tables_.bispec(np.array([[k1, k2, k3, z]]))[0])
        return float(self.tables_.bispec_idx(np.array([[k1, k2, k3, z, idx, didx]]))[0])

    def integrand_bkappa(self, z, ell1, ell2, ell3):
        return float(self.tables_.integrand_bkappa(np.array([[z, ell1, ell2, ell3]]))[0])


class ApertureStatistics:
    """classed_version/apertureStatistics.hpp:18-322: MapMapMap over (l1, l2, phi) with hcubature."""

    def __init__(self, bispectrum, reltol=1e-4, mode=MODE_FP64):
        self.bk = bispectrum
        self.reltol = reltol
        self.mode = mode

    def MapMapMap(self, thetas):
        return self.bk.tables_.map3_single(thetas, self.reltol, 0, self.mode)[0]


class GammaCalculator:
    """classed_version/gamma.hpp: Gamma_i via cubature + Ogata; triangle_center "centroid" or "orthocenter"
    (gamma.cpp:6-23)."""

    def __init__(self, bispectrum, prec_h, prec_k, triangle_center="centroid", reltol=1e-3):
        centers = {"centroid": CENTER_CENTROID, "orthocenter": CENTER_ORTHOCENTER}
        if triangle_center not in centers:
            raise ValueError("no valid triangle center given; possible choices: orthocenter, centroid")
        self.center = centers[triangle_center]
        self.bk = bispectrum
        self.reltol = reltol
        gamma_setup(prec_h, prec_k, self.bk.tables_.ctx)

    def gammas(self, x1, x2, x3):
        g, _ = self.bk.tables_.gamma_batch(np.array([[x1, x2, x3]]), self.reltol, center=self.center)
        return g[0]

    def gamma0(self, x1, x2, x3):
        return self.gammas(x1, x2, x3)[0]

    def gamma1(self, x1, x2, x3):
        return self.gammas(x1, x2, x3)[1]

    def gamma2(self, x1, x2, x3):
        return self.gammas(x1, x2, x3)[2]

    def gamma3(self, x1, x2, x3):
        return self.gammas(x1, x2, x3)[3]
