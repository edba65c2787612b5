"""MI355X (gfx950) cluster-observable integration: Python binding of liby3g.

The product is the HIP library ``liby3g.so`` (C ABI in ``include/y3g.h``)
and the CosmoSIS modules built on it.  This module is a ctypes binding used
by tests, the benchmark and Python callers; it loads the in-tree library and
raises if it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liby3g.so")

(COUNT3D, COUNT5D, DSIGMA6D, DSIGMA6D_DES, DSIGMA3D_DES, COUNT3D_DES,
 COUNT5D_DES, COUNT4D_CARD, POP_P1, POP_I1, POP_I2, POP_J) = range(12)
PAGANI, MCUBES = 0, 1
FP32SPLIT, FP64 = 0, 1

(F_ROFFSET_GAMMA, F_LO_LC, F_ZO_ZT_DES, F_INT_ZO_ZT_DES, F_SIGMA_PHOTOZ,
 F_DV_DO_DZ, F_HMF, F_MOR, F_LO_LT_PDF, F_LO_LT_BIN, F_NFW_DSIGMA,
 F_DSIGMA_OFFSET, F_SCI, F_INTEGRAND, F_INTEGRAND_PAIRED,
 F_INTEGRAND_NODE, F_INT_LC_LT_DES, F_MOR_DES, F_EMG_DES_PDF, F_EMG_DES_CDF,
 F_SIG_MAX, F_MOR_DES_LOG, F_NFW_DSIGMA_MIS, F_LC_LT, F_SCI_AVERAGE,
 F_SCI_RZ) = range(26)


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


class Integral(C.Structure):
    _fields_ = [("kind", C.c_int), ("ndim", C.c_int), ("point", C.c_int),
                ("algorithm", C.c_int),
                ("lo_lo", C.c_double), ("lo_hi", C.c_double),
                ("zo_lo", C.c_double), ("zo_hi", C.c_double),
                ("radius", C.c_double), ("sigma_mis", C.c_double),
                ("rmis_norm", C.c_double), ("lnlt_lo", C.c_double),
                ("lnlt_hi", C.c_double),
                ("lows", C.c_double * 8), ("highs", C.c_double * 8),
                ("eps_rel", C.c_double), ("eps_abs", C.c_double),
                ("max_eval", C.c_longlong), ("seed", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("value", C.c_double), ("error", C.c_double),
                ("prob", C.c_double), ("neval", C.c_longlong),
                ("nregions", C.c_int), ("status", C.c_int)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_regions", C.c_int),
                ("max_split_per_step", C.c_int), ("init_split", C.c_int),
                ("block_threads", C.c_int), ("slots", C.c_int),
                ("vegas_calls", C.c_longlong), ("vegas_itmax", C.c_int),
                ("vegas_itskip", C.c_int), ("vegas_nbins", C.c_int),
                ("vegas_alpha", C.c_double), ("single", C.c_int),
                ("err_c", C.c_double), ("err_smin", C.c_double),
                ("n_bands", C.c_int), ("band_eps", C.c_double * 8),
                ("band_regions", C.c_int * 8),
                ("tl_c", C.c_double), ("tl_smin", C.c_double), ("hmf_smooth", C.c_int),
                ("init_split_lnm", C.c_int), ("err_scale", C.c_double),
                ("ext_regions", C.c_int), ("ext_ratio", C.c_double)]


class SelGLSpec(C.Structure):
    """y3g_selgl_spec: fixed Gauss-Legendre node counts and windows."""
    _fields_ = [("n_lnm", C.c_int), ("n_z", C.c_int), ("lnm_lo", C.c_double),
                ("lnm_hi", C.c_double), ("zt_lo", C.c_double), ("zt_hi", C.c_double)]


def selgl_spec(lnm_lo, lnm_hi, zt_lo, zt_hi, n_lnm=96, n_z=64):
    return SelGLSpec(n_lnm, n_z, lnm_lo, lnm_hi, zt_lo, zt_hi)


class ServerOptions(C.Structure):
    _fields_ = [("ring_integrals", C.c_int), ("table_slots", C.c_int),
                ("hbm_budget_bytes", C.c_size_t), ("cache_entries", C.c_longlong),
                ("idle_ms", C.c_double), ("table_wait_ms", C.c_double),
                ("options", Options), ("spec", TableSpec)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in (
        "requests_submitted", "requests_done", "integrals_computed", "cache_hits",
        "evaluations", "queued", "launches", "table_builds", "table_reuses",
        "slab_bytes", "table_bytes")] + [("table_slots", C.c_int), ("workgroups", C.c_int)]


STATUS_NO_TABLES = 3  # a served integral whose tables were not ready in time


RESULT_DTYPE = np.dtype([("value", "f8"), ("error", "f8"), ("prob", "f8"),
                         ("neval", "i8"), ("nregions", "i4"), ("status", "i4")])
assert RESULT_DTYPE.itemsize == C.sizeof(Result)

_D = C.c_double
_PD = C.POINTER(C.c_double)

# exported symbols and signatures of include/y3g.h
SIGNATURES = {
    "y3g_last_error": (C.c_char_p, []),
    "y3g_version": (C.c_char_p, []),
    "y3g_runtime_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "y3g_runtime_destroy": (C.c_int, [C.c_void_p]),
    "y3g_device_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), _PD,
                                  C.POINTER(C.c_size_t), C.c_char_p, C.c_int]),
    "y3g_options_default": (None, [C.POINTER(Options)]),
    "y3g_default_params": (None, [C.POINTER(Params)]),
    "y3g_spec_config0": (None, [C.POINTER(TableSpec)]),
    "y3g_spec_wide": (None, [C.POINTER(TableSpec)]),
    "y3g_tables_build": (C.c_int, [C.c_void_p, C.POINTER(Params),
                                   C.POINTER(TableSpec), C.c_int]),
    "y3g_tables_upload": (C.c_int, [C.c_void_p, C.c_int, _PD, _PD, _PD]),
    "y3g_tables_download": (C.c_int, [C.c_void_p, C.c_int, _PD, _PD, _PD]),
    "y3g_make_integral": (None, [C.POINTER(Params), C.POINTER(TableSpec), C.c_int,
                                 _D, _D, _D, _D, _D, C.POINTER(Integral)]),
    "y3g_make_integral_volume": (C.c_int, [C.POINTER(Params), C.POINTER(TableSpec), C.c_int,
                                           _PD, _PD, _PD, C.c_int, C.POINTER(Integral)]),
    "y3g_integrate": (C.c_int, [C.c_void_p, C.POINTER(Integral), C.c_size_t,
                                C.POINTER(Options), C.c_void_p]),
    "y3g_integrate_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Integral), C.c_size_t,
                                        C.POINTER(Options), C.c_void_p]),
    "y3g_runtime_footprint": (C.c_int, [C.c_void_p, C.POINTER(Options), C.POINTER(C.c_size_t)]),
    "y3g_server_last_error": (C.c_char_p, []),
    "y3g_server_options_default": (None, [C.POINTER(ServerOptions)]),
    "y3g_server_create": (C.c_int, [C.c_int, C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]),
    "y3g_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Integral),
                                    C.c_size_t, C.c_void_p, C.POINTER(C.c_ulonglong)]),
    "y3g_server_wait": (C.c_int, [C.c_void_p, C.c_ulonglong, C.c_double, _PD]),
    "y3g_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "y3g_server_destroy": (C.c_int, [C.c_void_p]),
    "y3g_last_timing": (C.c_int, [C.c_void_p, _PD, C.POINTER(C.c_longlong)]),
    "y3g_graph_create": (C.c_int, [C.c_void_p, C.POINTER(TableSpec), C.c_int, C.c_size_t,
                                   C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "y3g_graph_params": (C.c_void_p, [C.c_void_p]),
    "y3g_graph_integrals": (C.c_void_p, [C.c_void_p]),
    "y3g_graph_results": (C.c_void_p, [C.c_void_p]),
    "y3g_graph_run": (C.c_int, [C.c_void_p]),
    "y3g_graph_destroy": (C.c_int, [C.c_void_p]),
    "y3g_runtime_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_size_t)]),
    "y3g_selfn_upload": (C.c_int, [C.c_void_p, C.c_int, _PD, C.c_int, _PD, C.c_int, _PD,
                                   C.c_int]),
    "y3g_selgl_counts": (C.c_int, [C.c_void_p, C.POINTER(SelGLSpec), C.c_int, C.c_int, _PD]),
    "y3g_des_upload": (C.c_int, [C.c_void_p, C.c_int, _PD, C.c_int]),
    "y3g_nfw_mis_table_build": (C.c_int, [C.c_void_p]),
    "y3g_nfw_mis_table_upload": (C.c_int, [C.c_void_p, _PD, C.c_int, _PD, C.c_int, _PD]),
    "y3g_nfw_mis_table_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "y3g_nfw_mis_table_download": (C.c_int, [C.c_void_p, _PD, _PD, _PD]),
    "y3g_eval_functor": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _PD, C.c_int,
                                   C.c_size_t, _PD, C.c_int,
                                   C.POINTER(Integral), _PD]),
}


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(
            f"liby3g.so not found at {path}: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def last_error():
    return lib().y3g_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


def default_params(**kw):
    p = Params()
    lib().y3g_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def spec_config0():
    s = TableSpec()
    lib().y3g_spec_config0(C.byref(s))
    return s


def spec_wide():
    s = TableSpec()
    lib().y3g_spec_wide(C.byref(s))
    return s


def options(bands=None, **kw):
    """y3g_options with defaults; bands = [(eps_edge, regions), ...]
    ascending in eps_edge sets the tolerance-band budgets."""
    o = Options()
    lib().y3g_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    if bands:
        if len(bands) > 8:
            raise ValueError("at most 8 tolerance bands")
        o.n_bands = len(bands)
        for i, (e, r) in enumerate(bands):
            o.band_eps[i] = e
            o.band_regions[i] = int(r)
    return o


def make_integral(params, spec, kind, lo_lo=20.0, lo_hi=30.0, zo_lo=0.2,
                  zo_hi=0.35, radius=1.0, point=0, eps_rel=1e-4,
                  algorithm=PAGANI, max_eval=1 << 40, seed=0x59334331):
    ig = Integral()
    lib().y3g_make_integral(C.byref(params), C.byref(spec), kind, lo_lo, lo_hi,
                            zo_lo, zo_hi, radius, C.byref(ig))
    ig.point = point
    ig.eps_rel = eps_rel
    ig.algorithm = algorithm
    ig.max_eval = max_eval
    ig.seed = seed
    return ig


def make_integral_volume(params, spec, kind, lows, highs, grid=(), eps_rel=1e-4,
                         algorithm=PAGANI, max_eval=1 << 40, seed=0x59334331):
    """Integral descriptor from a physical volume and grid point (the module
    path): see y3g_make_integral_volume in include/y3g.h for the variables."""
    lo = np.ascontiguousarray(lows, dtype=np.float64)
    hi = np.ascontiguousarray(highs, dtype=np.float64)
    g = np.ascontiguousarray(grid if len(grid) else [0.0], dtype=np.float64)
    ig = Integral()
    _check(lib().y3g_make_integral_volume(C.byref(params), C.byref(spec), kind,
                                          lo.ctypes.data_as(_PD), hi.ctypes.data_as(_PD),
                                          g.ctypes.data_as(_PD), len(grid), C.byref(ig)),
           "y3g_make_integral_volume")
    ig.eps_rel = eps_rel
    ig.algorithm = algorithm
    ig.max_eval = max_eval
    ig.seed = seed
    return ig


class Runtime:
    """One device binding (y3g_runtime)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        self._graphs = []  # weak references: graphs are destroyed before the runtime
        _check(lib().y3g_runtime_create(device, C.byref(self._h)), "y3g_runtime_create")

    def close(self):
        for ref in self._graphs:
            g = ref()
            if g is not None:
                g.close()
        self._graphs = []
        if self._h:
            lib().y3g_runtime_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        n = C.c_int()
        clk = C.c_double()
        hbm = C.c_size_t()
        name = C.create_string_buffer(64)
        _check(lib().y3g_device_info(self._h, C.byref(n), C.byref(clk), C.byref(hbm),
                                     name, 64), "y3g_device_info")
        return dict(n_cu=n.value, clock_ghz=clk.value, hbm_bytes=hbm.value,
                    name=name.value.decode())

    def tables_build(self, params, spec):
        if isinstance(params, Params):
            params = [params]
        arr = (Params * len(params))(*params)
        _check(lib().y3g_tables_build(self._h, arr, C.byref(spec), len(params)),
               "y3g_tables_build")

    def tables_upload(self, point, hmf=None, da=None, sci=None):
        def p(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            return a.ctypes.data_as(_PD)
        _check(lib().y3g_tables_upload(self._h, point, p(hmf), p(da), p(sci)),
               "y3g_tables_upload")

    def tables_download(self, point, spec):
        hmf = np.zeros(spec.n_lnm * spec.n_z)
        da = np.zeros(spec.n_da)
        sci = np.zeros(spec.n_sci)
        _check(lib().y3g_tables_download(self._h, point, hmf.ctypes.data_as(_PD),
                                         da.ctypes.data_as(_PD), sci.ctypes.data_as(_PD)),
               "y3g_tables_download")
        return hmf, da, sci

    def integrate(self, integrals, opts=None):
        n = len(integrals)
        arr = integrals if isinstance(integrals, C.Array) else (Integral * n)(*integrals)
        out = np.zeros(n, dtype=RESULT_DTYPE)
        o = opts if opts is not None else options()
        _check(lib().y3g_integrate(self._h, arr, n, C.byref(o),
                                   out.ctypes.data_as(C.c_void_p)), "y3g_integrate")
        return out

    def integrate_noalloc(self, integrals, opts=None):
        """y3g_integrate_noalloc: as integrate(), with the call's device memory
        in the library's fixed scratch array (no device allocation)."""
        n = len(integrals)
        arr = integrals if isinstance(integrals, C.Array) else (Integral * n)(*integrals)
        out = np.zeros(n, dtype=RESULT_DTYPE)
        o = opts if opts is not None else options()
        _check(lib().y3g_integrate_noalloc(self._h, arr, n, C.byref(o),
                                           out.ctypes.data_as(C.c_void_p)),
               "y3g_integrate_noalloc")
        return out

    def des_upload(self, point, des):
        """DES data of a point for DSIGMA6D_DES integrals (y3g_des_upload):
        {roffset_tau, LO_LC alpha a b R_lambda, MOR_DES A B alpha sigma eps
        z_pivot, SIG_MAX n_r n_m n_z r lnM z sigma_1 sigma_2 bias}."""
        d = np.ascontiguousarray(des, dtype=np.float64).reshape(-1)
        _check(lib().y3g_des_upload(self._h, point, d.ctypes.data_as(_PD), d.size),
               "y3g_des_upload")

    def nfw_mis_table_build(self):
        """Build NFW_DSIGMA_MIS's table on the device (y3g_nfw_mis_table_build)."""
        _check(lib().y3g_nfw_mis_table_build(self._h), "y3g_nfw_mis_table_build")

    def nfw_mis_table_upload(self, logx, logxmis, log_dsigma):
        """The reference's three nfw_off_center files' contents
        (y3g_nfw_mis_table_upload); log_dsigma [len(logxmis), len(logx)]."""
        a = np.ascontiguousarray(logx, dtype=np.float64)
        b = np.ascontiguousarray(logxmis, dtype=np.float64)
        t = np.ascontiguousarray(log_dsigma, dtype=np.float64).reshape(-1)
        if t.size != a.size * b.size:
            raise ValueError("log_dsigma must hold len(logxmis) x len(logx) values")
        _check(lib().y3g_nfw_mis_table_upload(self._h, a.ctypes.data_as(_PD), a.size,
                                              b.ctypes.data_as(_PD), b.size,
                                              t.ctypes.data_as(_PD)),
               "y3g_nfw_mis_table_upload")

    def nfw_mis_table_download(self):
        """(ln x, ln xmis, T [n_xmis, n_x]) of the runtime's table."""
        nx, nm = C.c_int(), C.c_int()
        _check(lib().y3g_nfw_mis_table_shape(self._h, C.byref(nx), C.byref(nm)),
               "y3g_nfw_mis_table_shape")
        a, b = np.zeros(nx.value), np.zeros(nm.value)
        t = np.zeros((nm.value, nx.value))
        _check(lib().y3g_nfw_mis_table_download(self._h, a.ctypes.data_as(_PD),
                                                b.ctypes.data_as(_PD),
                                                t.ctypes.data_as(_PD)),
               "y3g_nfw_mis_table_download")
        return a, b, t

    def selfn_upload(self, point, lnm, z, s_stack):
        """Selection function of a point: axes lnM [n_lnm], z [n_z] and
        S_stack [n_bins, n_z, n_lnm] (y3g_selfn_upload)."""
        lnm = np.ascontiguousarray(lnm, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        st = np.ascontiguousarray(s_stack, dtype=np.float64)
        if st.ndim != 3 or st.shape[1] != z.size or st.shape[2] != lnm.size:
            raise ValueError("S_stack must have shape (n_bins, n_z, n_lnm)")
        _check(lib().y3g_selfn_upload(self._h, point, lnm.ctypes.data_as(_PD), lnm.size,
                                      z.ctypes.data_as(_PD), z.size, st.ctypes.data_as(_PD),
                                      st.shape[0]), "y3g_selfn_upload")

    def selgl_counts(self, spec, n_points, max_bins):
        """Fixed-GL number counts N[point, bin] (y3g_selgl_counts)."""
        out = np.zeros((n_points, max_bins))
        _check(lib().y3g_selgl_counts(self._h, C.byref(spec), n_points, max_bins,
                                      out.ctypes.data_as(_PD)), "y3g_selgl_counts")
        return out

    def memory(self):
        """(free, total, runtime) device bytes (y3g_runtime_memory)."""
        f, t, r = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(lib().y3g_runtime_memory(self._h, C.byref(f), C.byref(t), C.byref(r)),
               "y3g_runtime_memory")
        return f.value, t.value, r.value

    def last_timing(self):
        ms = C.c_double()
        nev = C.c_longlong()
        _check(lib().y3g_last_timing(self._h, C.byref(ms), C.byref(nev)), "y3g_last_timing")
        return ms.value, nev.value

    def eval_functor(self, which, x, par=(), precision=FP64, aux=None):
        x = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        n, nin = x.shape
        par = np.ascontiguousarray(np.asarray(par, dtype=np.float64).reshape(-1))
        if par.size == 0:
            par = np.zeros(1)
        out = np.zeros(n)
        _check(lib().y3g_eval_functor(self._h, which, precision, x.ctypes.data_as(_PD), nin, n,
                                      par.ctypes.data_as(_PD), par.size,
                                      C.byref(aux) if aux is not None else None,
                                      out.ctypes.data_as(_PD)), "y3g_eval_functor")
        return out


class Graph:
    """One batched step captured as a HIP graph (y3g_graph_*): write
    .params[k] and .integrals[i] (point = 0 .. n_points - 1), run(), read
    results() (a copy).  The runtime must stay open while the graph lives."""

    def __init__(self, runtime, spec, n_points, n, opts=None):
        self._h = C.c_void_p()
        self._rt = runtime
        o = opts if opts is not None else options()
        _check(lib().y3g_graph_create(runtime._h, C.byref(spec), n_points, n, C.byref(o),
                                      C.byref(self._h)), "y3g_graph_create")
        import weakref
        runtime._graphs.append(weakref.ref(self))
        self.n_points, self.n = n_points, n
        self.params = (Params * n_points).from_address(lib().y3g_graph_params(self._h))
        self.integrals = (Integral * n).from_address(lib().y3g_graph_integrals(self._h))

    def run(self):
        _check(lib().y3g_graph_run(self._h), "y3g_graph_run")

    def results(self):
        addr = lib().y3g_graph_results(self._h)
        buf = (C.c_char * (self.n * RESULT_DTYPE.itemsize)).from_address(addr)
        return np.frombuffer(bytes(buf), dtype=RESULT_DTYPE)

    def close(self):
        if self._h:
            lib().y3g_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_options(**kw):
    o = ServerOptions()
    lib().y3g_server_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Server:
    """Request server (y3g_server_*): submit(params, integrals) -> ticket,
    wait(ticket) -> (results, latency seconds).  One persistent streaming
    launch serves every request; requests join it as they arrive."""

    def __init__(self, device=0, opts=None):
        self._h = C.c_void_p()
        self._pending = {}
        o = opts if opts is not None else server_options()
        rc = lib().y3g_server_create(device, C.byref(o), C.byref(self._h))
        if rc != 0:
            raise RuntimeError(f"y3g_server_create failed ({rc}): "
                               f"{lib().y3g_server_last_error().decode()}")

    def submit(self, params, integrals):
        n = len(integrals)
        arr = integrals if isinstance(integrals, C.Array) else (Integral * n)(*integrals)
        out = np.zeros(n, dtype=RESULT_DTYPE)
        t = C.c_ulonglong()
        rc = lib().y3g_server_submit(self._h, C.byref(params), arr, n,
                                     out.ctypes.data_as(C.c_void_p), C.byref(t))
        if rc != 0:
            raise RuntimeError(f"y3g_server_submit failed ({rc}): "
                               f"{lib().y3g_server_last_error().decode()}")
        self._pending[t.value] = (out, arr)   # keep both alive until wait
        return t.value

    def wait(self, ticket, timeout=-1.0):
        lat = C.c_double()
        rc = lib().y3g_server_wait(self._h, ticket, timeout, C.byref(lat))
        if rc != 0:
            raise RuntimeError(f"y3g_server_wait failed ({rc}): "
                               f"{lib().y3g_server_last_error().decode()}")
        out, _ = self._pending.pop(ticket)
        return out, lat.value

    def stats(self):
        st = ServerStats()
        rc = lib().y3g_server_stats_get(self._h, C.byref(st))
        if rc != 0:
            raise RuntimeError("y3g_server_stats_get failed")
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib().y3g_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
