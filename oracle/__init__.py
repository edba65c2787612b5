"""ctypes binding of the FP64 CPU oracle (liby3o.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, as the checker.  The product package
(y3_cluster_cpp_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liby3o.so")

COUNT3D, COUNT5D, DSIGMA6D = 0, 1, 2


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "omega_m", "h", "hmf_amp", "hmf_A", "hmf_a", "hmf_c", "mor_lnl0",
        "mor_slope", "mor_zslope", "mor_sigma_ln", "f_prj", "prj_scale",
        "f_mis", "mis_rlambda", "z_source", "sigz0", "hmf_s", "hmf_q")]


class TableSpec(C.Structure):
    _fields_ = [("lnm_lo", C.c_double), ("lnm_hi", C.c_double),
                ("n_lnm", C.c_int), ("z_lo", C.c_double),
                ("z_hi", C.c_double), ("n_z", C.c_int),
                ("da_zmax", C.c_double), ("n_da", C.c_int),
                ("sci_zmax", C.c_double), ("n_sci", C.c_int)]


class Tables(C.Structure):
    _fields_ = [("spec", TableSpec)] + [(n, C.POINTER(C.c_double)) for n in (
        "lnm", "z", "hmf", "da_z", "da", "sci_z", "sci", "hmfc")] + [("smooth_hmf", C.c_int)]


class Integral(C.Structure):
    _fields_ = [("kind", C.c_int), ("ndim", C.c_int),
                ("lo_lo", C.c_double), ("lo_hi", C.c_double),
                ("zo_lo", C.c_double), ("zo_hi", C.c_double),
                ("radius", C.c_double), ("sigma_mis", C.c_double),
                ("rmis_norm", C.c_double),
                ("lt_lo", C.c_double), ("lt_hi", C.c_double),
                ("lows", C.c_double * 8), ("highs", C.c_double * 8)]


class Result(C.Structure):
    _fields_ = [("value", C.c_double), ("error", C.c_double),
                ("prob", C.c_double), ("neval", C.c_longlong),
                ("nregions", C.c_int), ("status", C.c_int)]


class CubOpts(C.Structure):
    _fields_ = [("eps_rel", C.c_double), ("eps_abs", C.c_double),
                ("max_eval", C.c_longlong), ("max_regions", C.c_int),
                ("max_split_per_step", C.c_int), ("init_split", C.c_int),
                ("err_c", C.c_double), ("err_smin", C.c_double),
                ("tl_c", C.c_double), ("tl_smin", C.c_double), ("hmf_smooth", C.c_int),
                ("init_split_lnm", C.c_int), ("err_scale", C.c_double),
                ("ext_regions", C.c_int), ("ext_ratio", C.c_double)]


class VegasOpts(C.Structure):
    _fields_ = [("eps_rel", C.c_double), ("eps_abs", C.c_double),
                ("max_eval", C.c_longlong), ("calls_per_iter", C.c_longlong),
                ("itmax", C.c_int), ("itskip", C.c_int), ("nbins", C.c_int),
                ("alpha", C.c_double), ("seed", C.c_uint64), ("hmf_smooth", C.c_int)]


_D = C.c_double
_PD = C.POINTER(C.c_double)
_FN1 = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
_FNN = C.CFUNCTYPE(C.c_double, _PD, C.c_void_p)


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(_LIB_PATH)
    sig = {
        "y3o_gaussian": (_D, [_D, _D, _D]),
        "y3o_roffset_gamma": (_D, [_D, _D]),
        "y3o_lo_lc": (_D, [_D] * 7),
        "y3o_lo_lc_bin": (_D, [_D] * 8),
        "y3o_sigma_photoz_des": (_D, [_D]),
        "y3o_zo_zt_des": (_D, [_D, _D]),
        "y3o_int_zo_zt_des": (_D, [_D, _D, _D]),
        "y3o_ez": (_D, [_D] * 4),
        "y3o_interp1d": (C.c_int, [_PD, _PD, C.c_int, _D, _PD]),
        "y3o_interp1d_clamp": (_D, [_PD, _PD, C.c_int, _D]),
        "y3o_interp2d": (C.c_int, [_PD, C.c_int, _PD, C.c_int, _PD, _D, _D, _PD]),
        "y3o_interp2d_clamp": (_D, [_PD, C.c_int, _PD, C.c_int, _PD, _D, _D]),
        "y3o_dv_do_dz": (_D, [_D, _PD, _PD, C.c_int, _D, _D, _D, _D]),
        "y3o_lc_lt_eq15": (_D, [_D] * 7),
        "y3o_hmf_formula": (_D, [C.POINTER(Params), _D, _D]),
        "y3o_mor_lognormal": (_D, [C.POINTER(Params), _D, _D, _D]),
        "y3o_lo_lt_pdf": (_D, [C.POINTER(Params), _D, _D]),
        "y3o_lo_lt_bin": (_D, [C.POINTER(Params), _D, _D, _D]),
        "y3o_emg_pdf": (_D, [_D] * 4),
        "y3o_emg_cdf": (_D, [_D] * 4),
        "y3o_zo_pdf": (_D, [C.POINTER(Params), _D, _D]),
        "y3o_zo_bin": (_D, [C.POINTER(Params), _D, _D, _D]),
        "y3o_comoving_distance": (_D, [_D, _D, _D]),
        "y3o_nfw_conc": (_D, [_D]),
        "y3o_nfw_g": (_D, [_D]),
        "y3o_nfw_sigma_unit": (_D, [_D]),
        "y3o_nfw_dsigma": (_D, [_D, _D, _D, _D]),
        "y3o_dsigma_offset": (_D, [_D] * 6),
        "y3o_dsigma_mis_ring": (_D, [_D] * 6),
        "y3o_dsigma_mis_gamma": (_D, [_D] * 6),
        "y3o_dsigma_mis_rayleigh": (_D, [_D] * 7),
        "y3o_lc_lt": (_D, [_D, _D, _D]),
        "y3o_sigma_photoz_table": (_D, [_D]),
        "y3o_mor_shifted_poisson": (_D, [_PD, _D, _D, _D]),
        "y3o_pop_area_overlap": (_D, [_D, _D, _D]),
        "y3o_pop_theta_excl": (_D, [_D, _D, _D]),
        "y3o_pop_theta_term": (_D, [C.c_int, _D, _D, _D, _D, _D, _D, _PD, C.c_int, _PD,
                                    C.c_int, _PD]),
        "y3o_integrand_pop": (_D, [C.POINTER(Params), C.POINTER(Tables), _PD,
                                   C.POINTER(Integral), _PD]),
        "y3o_lc_lt_params": (None, [_D, _D, _PD]),
        "y3o_nfw_ring_levels": (C.c_int, [_D, _D]),
        "y3o_nfw_ring": (_D, [_D, _D]),
        "y3o_nfw_mis_unit": (_D, [_D, _D]),
        "y3o_nfw_mis_unit_abs": (_D, [_D, _D]),
        "y3o_nfw_mis_axes": (None, [_PD, _PD]),
        "y3o_nfw_mis_entry": (_D, [_D, _D]),
        "y3o_nfw_dsigma_mis_lookup": (_D, [_PD, C.c_int, _PD, C.c_int, _PD] + [_D] * 5),
        "y3o_default_params": (None, [C.POINTER(Params)]),
        "y3o_spec_config0": (None, [C.POINTER(TableSpec)]),
        "y3o_spec_wide": (None, [C.POINTER(TableSpec)]),
        "y3o_tables_build": (C.c_int, [C.POINTER(Params), C.POINTER(TableSpec), C.POINTER(Tables)]),
        "y3o_tables_free": (None, [C.POINTER(Tables)]),
        "y3o_hmf": (_D, [C.POINTER(Params), C.POINTER(Tables), _D, _D]),
        "y3o_dvdz": (_D, [C.POINTER(Params), C.POINTER(Tables), _D]),
        "y3o_sci": (_D, [C.POINTER(Tables), _D]),
        "y3o_integrand": (_D, [C.POINTER(Params), C.POINTER(Tables), C.POINTER(Integral), _PD]),
        "y3o_integrand_des": (_D, [C.POINTER(Params), C.POINTER(Tables), _PD,
                                   C.POINTER(Integral), _PD]),
        "y3o_integrand_mapped": (_D, [C.POINTER(Params), C.POINTER(Tables), C.POINTER(Integral), _PD]),
        "y3o_integrand_mapped_batch": (None, [C.POINTER(Params), C.POINTER(Tables),
                                              C.POINTER(Integral), _PD, C.c_long, _PD]),
        "y3o_make_integral": (None, [C.POINTER(Params), C.POINTER(TableSpec), C.c_int,
                                     _D, _D, _D, _D, _D, C.POINTER(Integral)]),
        "y3o_gk21_adapt": (_D, [_FN1, C.c_void_p, _D, _D, _D, _D, _PD]),
        "y3o_gl_rule": (None, [C.c_int, _PD, _PD]),
        "y3o_omega_z_des": (_D, [_D]),
        "y3o_int_lc_lt_des": (_D, [_PD, _D, _D, _D]),
        "y3o_mor_des_sigma_skew": (None, [_D, _D, _PD, _PD]),
        "y3o_mor_des": (_D, [_PD, _D, _D, _D]),
        "y3o_mor_des_log": (_D, [_PD, _D, _D, _D]),
        "y3o_emg_des_params": (None, [_PD, _D, _D, _PD]),
        "y3o_emg_des_cdf": (_D, [_PD, _D, _D, _D]),
        "y3o_emg_des_pdf": (_D, [_PD, _D, _D, _D]),
        "y3o_sig_max": (_D, [_PD, _D, _D, _D]),
        "y3o_sel_function": (_D, [_PD, C.c_int, _PD, C.c_int, _PD, _D, _D]),
        "y3o_gl_nodes_ab": (None, [C.c_int, _D, _D, _PD, _PD]),
        "y3o_selgl_counts": (_D, [C.POINTER(Params), C.POINTER(Tables), _PD, C.c_int, _PD,
                                  C.c_int, _PD, C.c_int, C.c_int, _D, _D, _D, _D]),
        "y3o_hmf_smooth": (_D, [C.POINTER(Params), C.POINTER(Tables), _D, _D]),
        "y3o_hmf_coefs": (None, [_PD, C.c_int, C.c_int, _PD]),
        "y3o_spline_coefs": (None, [_PD, C.c_int, C.c_long, _PD, C.c_long]),
        "y3o_b5_weights": (None, [_D, _PD]),
        "y3o_smooth_bound": (_D, [C.c_int, C.c_int]),
        "y3o_gk21_nodes": (_PD, []),
        "y3o_gk21_wk": (_PD, []),
        "y3o_gk21_wg": (_PD, []),
        "y3o_golden": (C.c_int, [C.POINTER(Params), C.POINTER(Tables), C.c_int, _PD, C.c_int,
                                 _PD, C.c_int, _PD, C.c_int, _PD, _PD]),
        "y3o_golden_set_orders": (None, [C.c_int, C.c_int]),
        "y3o_cubature": (None, [C.POINTER(Params), C.POINTER(Tables), C.POINTER(Integral),
                                C.POINTER(CubOpts), C.POINTER(Result)]),
        "y3o_cubature_des": (None, [C.POINTER(Params), C.POINTER(Tables), _PD,
                                    C.POINTER(Integral), C.POINTER(CubOpts), C.POINTER(Result)]),
        "y3o_cubature_fn": (None, [_FNN, C.c_void_p, C.c_int, _PD, _PD,
                                   C.POINTER(CubOpts), C.POINTER(Result)]),
        "y3o_gm_rule": (_D, [_FNN, C.c_void_p, C.c_int, _PD, _PD, _PD, _PD]),
        "y3o_gm_num_nodes": (C.c_int, [C.c_int]),
        "y3o_vegas": (None, [C.POINTER(Params), C.POINTER(Tables), C.POINTER(Integral),
                             C.POINTER(VegasOpts), C.POINTER(Result)]),
        "y3o_vegas_fn": (None, [_FNN, C.c_void_p, C.c_int, _PD, _PD,
                                C.POINTER(VegasOpts), C.POINTER(Result)]),
        "y3o_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "y3o_test_fn": (_D, [C.c_int, _PD, C.c_int]),
        "y3o_batch_integrate": (C.c_int, [C.POINTER(Params), C.POINTER(TableSpec), C.c_int,
                                          C.POINTER(Integral), C.POINTER(C.c_int), _PD,
                                          C.POINTER(C.c_longlong), C.c_int,
                                          C.c_int, C.POINTER(CubOpts), C.POINTER(VegasOpts),
                                          C.c_int, C.POINTER(Result)]),
        "y3o_batch_golden": (C.c_int, [C.POINTER(Params), C.POINTER(TableSpec), C.c_int, C.c_int,
                                       _PD, C.c_int, _PD, C.c_int, _PD, C.c_int, C.c_int, _PD]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def ptr(a):
    """Pointer to a contiguous float64 numpy array."""
    return a.ctypes.data_as(_PD)


def default_params(**kw):
    p = Params()
    lib.y3o_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def spec_config0():
    s = TableSpec()
    lib.y3o_spec_config0(C.byref(s))
    return s


def spec_wide():
    s = TableSpec()
    lib.y3o_spec_wide(C.byref(s))
    return s


class OracleTables:
    """Owns one y3o_tables; arrays exposed as numpy views."""

    def __init__(self, params, spec):
        self.t = Tables()
        self.params = params
        if lib.y3o_tables_build(C.byref(params), C.byref(spec), C.byref(self.t)):
            raise MemoryError("y3o_tables_build")
        s = self.t.spec
        self.lnm = np.ctypeslib.as_array(self.t.lnm, (s.n_lnm,)).copy()
        self.z = np.ctypeslib.as_array(self.t.z, (s.n_z,)).copy()
        self.hmf = np.ctypeslib.as_array(self.t.hmf, (s.n_z * s.n_lnm,)).copy()
        self.da_z = np.ctypeslib.as_array(self.t.da_z, (s.n_da,)).copy()
        self.da = np.ctypeslib.as_array(self.t.da, (s.n_da,)).copy()
        self.sci_z = np.ctypeslib.as_array(self.t.sci_z, (s.n_sci,)).copy()
        self.sci = np.ctypeslib.as_array(self.t.sci, (s.n_sci,)).copy()

    def __del__(self):
        try:
            lib.y3o_tables_free(C.byref(self.t))
        except Exception:
            pass


def make_integral(params, spec, kind, lo_lo=20.0, lo_hi=30.0, zo_lo=0.2,
                  zo_hi=0.35, radius=1.0):
    ig = Integral()
    lib.y3o_make_integral(C.byref(params), C.byref(spec), kind, lo_lo, lo_hi,
                          zo_lo, zo_hi, radius, C.byref(ig))
    return ig


def integrand(params, tables, ig, x):
    xa = np.ascontiguousarray(x, dtype=np.float64)
    return lib.y3o_integrand(C.byref(params), C.byref(tables.t), C.byref(ig), ptr(xa))


def cub_opts(eps_rel=1e-4, eps_abs=1e-300, max_eval=50_000_000, max_regions=65536,
             max_split_per_step=1024, init_split=0, err_c=0.0, err_smin=1.0, tl_c=0.0,
             tl_smin=1.0 / 16.0, hmf_smooth=1, init_split_lnm=4, err_scale=1.5,
             ext_regions=0, ext_ratio=0.0):
    return CubOpts(eps_rel, eps_abs, max_eval, max_regions, max_split_per_step, init_split,
                   err_c, err_smin, tl_c, tl_smin, hmf_smooth, init_split_lnm, err_scale,
                   ext_regions, ext_ratio)


def vegas_opts(eps_rel=1e-4, eps_abs=1e-300, max_eval=50_000_000,
               calls_per_iter=200_000, itmax=5, itskip=1, nbins=64, alpha=1.5,
               seed=0x5933_4331, hmf_smooth=1):
    return VegasOpts(eps_rel, eps_abs, max_eval, calls_per_iter, itmax, itskip,
                     nbins, alpha, seed, hmf_smooth)


def cubature(params, tables, ig, opts):
    r = Result()
    lib.y3o_cubature(C.byref(params), C.byref(tables.t), C.byref(ig),
                     C.byref(opts), C.byref(r))
    return r


def vegas(params, tables, ig, opts):
    r = Result()
    lib.y3o_vegas(C.byref(params), C.byref(tables.t), C.byref(ig),
                  C.byref(opts), C.byref(r))
    return r


def golden(params, tables, count_kind, lo_edges, zo_edges=(), radii=()):
    lo = np.ascontiguousarray(lo_edges, dtype=np.float64)
    zo = np.ascontiguousarray(zo_edges, dtype=np.float64) if len(zo_edges) else np.zeros(1)
    rr = np.ascontiguousarray(radii, dtype=np.float64) if len(radii) else np.zeros(1)
    n_lo = len(lo) - 1
    n_zo = max(len(zo_edges) - 1, 0)
    nzb = max(n_zo, 1)
    n_r = len(radii)
    counts = np.zeros(n_lo * nzb)
    dsig = np.zeros(max(n_lo * nzb * n_r, 1))
    rc = lib.y3o_golden(C.byref(params), C.byref(tables.t), count_kind, ptr(lo), n_lo,
                        ptr(zo), n_zo, ptr(rr), n_r, ptr(counts), ptr(dsig))
    if rc:
        raise RuntimeError("y3o_golden failed")
    return counts.reshape(n_lo, nzb), (dsig.reshape(n_lo, nzb, n_r) if n_r else None)
