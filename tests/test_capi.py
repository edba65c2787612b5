"""The C ABI (include/y3g.h) without a GPU: the library loads, exports every
declared function, and its host-only helpers agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

import oracle as O
import y3_cluster_cpp_amd as y3
from conftest import ROOT


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(y3g_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    names = declared("y3g.h")
    assert "y3g_integrate" in names and len(names) >= 15
    lib = C.CDLL(y3.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert set(names) <= set(y3.SIGNATURES), set(names) - set(y3.SIGNATURES)


def test_make_integral_matches_oracle():
    for kind, spec_y, spec_o, pf in [(y3.COUNT3D, y3.spec_config0(), O.spec_config0(), 0.0),
                                     (y3.COUNT5D, y3.spec_wide(), O.spec_wide(), 0.2),
                                     (y3.DSIGMA6D, y3.spec_wide(), O.spec_wide(), 0.2)]:
        for lo_lo, lo_hi, zo_lo, zo_hi, R in [(20, 30, 0.2, 0.35, 1.0), (60, 300, 0.5, 0.65, 12.0)]:
            g = y3.make_integral(y3.default_params(f_prj=pf), spec_y, kind, lo_lo, lo_hi, zo_lo, zo_hi, R)
            o = O.make_integral(O.default_params(f_prj=pf), spec_o, kind, lo_lo, lo_hi, zo_lo, zo_hi, R)
            assert g.ndim == o.ndim
            for k in range(g.ndim):
                assert g.lows[k] == o.lows[k] and g.highs[k] == o.highs[k]
            for f in ("sigma_mis", "rmis_norm", "radius"):
                assert getattr(g, f) == getattr(o, f), f
            assert abs(g.lnlt_lo - np.log(o.lt_lo)) < 1e-15 and abs(g.lnlt_hi - np.log(o.lt_hi)) < 1e-15


def test_params_and_specs_match_oracle():
    a, b = y3.default_params(), O.default_params()
    for n, _ in y3.Params._fields_:
        assert getattr(a, n) == getattr(b, n), n
    for sy, so in [(y3.spec_config0(), O.spec_config0()), (y3.spec_wide(), O.spec_wide())]:
        for n, _ in y3.TableSpec._fields_:
            assert getattr(sy, n) == getattr(so, n), n


def test_runtime_create_without_device_fails_cleanly():
    h = C.c_void_p()
    rc = y3.lib().y3g_runtime_create(0, C.byref(h))
    if rc == 0:  # a GPU is present: nothing to check here
        y3.lib().y3g_runtime_destroy(h)
        return
    assert rc == 4 and "device" in y3.last_error()
    # null runtime is an argument error, not a crash
    assert y3.lib().y3g_integrate(None, None, 0, None, None) == 1


def test_server_options_layout():
    """The ctypes mirror of y3g_server_options reads the C defaults field by
    field (a layout slip shifts every field after it)."""
    o = y3.server_options()
    assert (o.ring_integrals, o.table_slots) == (65536, 256)
    assert o.hbm_budget_bytes == 64 << 30 and o.cache_entries == 1 << 20
    assert (o.idle_ms, o.table_wait_ms) == (50.0, 2000.0)
    d = y3.options()
    for n, _ in y3.Options._fields_:
        a, b = getattr(o.options, n), getattr(d, n)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), n
        else:
            assert a == b, n
    w = y3.spec_wide()
    for n, _ in y3.TableSpec._fields_:
        assert getattr(o.spec, n) == getattr(w, n), n
