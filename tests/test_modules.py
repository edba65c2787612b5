"""The CosmoSIS module boundary without a GPU.

- the C++ host test (tests/cpp/test_module_host.cc): DataBlock subset,
  get_vector_double coercions (datablock_reader.cc:3-28), volumes and grid
  points (make_cuda_integration_volumes.cuh, make_grid_points.hh), config
  errors (gpu_integrator.cuh:54-65, CosmoSISSICUDAModule.hh:191-198), the
  integral list order (CosmoSISSICUDAModule.hh:282-335), the output layout
  (:350-467), and setup/execute/cleanup through dlopen;
- the C API of the DataBlock library from Python;
- every module library exports exactly the three CosmoSIS entry points.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from y3_cluster_cpp_amd import cosmosis as cs

MOD = os.path.join(ROOT, "y3_cluster_cpp_amd", "modules")


@pytest.fixture(scope="module")
def host_test_binary(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "test_module_host")
    lib = os.path.join(ROOT, "y3_cluster_cpp_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I" + MOD, "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_module_host.cc"), "-o", out,
           "-L" + MOD, "-ldatablock", "-L" + lib, "-ly3g", "-ldl",
           "-Wl,-rpath," + MOD, "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_module_host_cpp(host_test_binary):
    r = subprocess.run([host_test_binary, os.path.join(MOD, "SigmaMiscent6DIntegrand.so")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert " 0 failures" in r.stdout


def test_module_libraries_export_entry_points():
    for name in cs.MODULES:
        path = os.path.join(MOD, name + ".so")
        assert os.path.exists(path), path
        syms = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                              text=True, check=True).stdout.split()
        for s in ("setup", "execute", "cleanup"):
            assert s in syms, (name, s)


def test_c_datablock_roundtrip():
    b = cs.DataBlock()
    b.put("Cosmological_Parameters", "omega_m", 0.3)
    b.put("cosmological_parameters", "n", 5)
    b.put("sec", "flag", True)
    b.put("sec", "name", "pagani")
    b.put("sec", "v", [1.0, 2.5, 4.0])
    b.put("sec", "iv", np.array([3, 4], dtype=np.int64))
    b.put("sec", "m", np.arange(6.0).reshape(2, 3))
    assert b.has("COSMOLOGICAL_PARAMETERS") and b.has("sec", "V") and not b.has("sec", "w")
    assert b.get_double("cosmological_parameters", "OMEGA_M") == 0.3
    assert b.get_int("cosmological_parameters", "n") == 5
    np.testing.assert_array_equal(b.get_array("sec", "v"), [1.0, 2.5, 4.0])
    iv = b.get_array("sec", "iv")
    assert iv.dtype == np.int32 and list(iv) == [3, 4]
    np.testing.assert_array_equal(b.get_array("sec", "m"), np.arange(6.0).reshape(2, 3))
    with pytest.raises(cs.DataBlockError):
        b.put("sec", "v", [0.0])            # DBS_NAME_ALREADY_EXISTS
    with pytest.raises(cs.DataBlockError):
        b.get_double("cosmological_parameters", "n")  # int, no coercion
    with pytest.raises(cs.DataBlockError):
        b.get_double("nosuch", "x")
    lib = cs.datablock_lib()
    assert lib.c_datablock_get_array_length(b.ptr, b"sec", b"v") == 3
    assert lib.c_datablock_get_array_length(b.ptr, b"sec", b"m") == -1
    buf = (C.c_double * 2)()
    n = C.c_int()
    rc = lib.c_datablock_get_double_array_1d_preallocated(b.ptr, b"sec", b"v", buf,
                                                          C.byref(n), 2)
    assert rc != 0 and n.value == 3           # DBS_SIZE_INSUFFICIENT, size reported
    b.close()


def test_module_setup_without_device_returns_null():
    """No GPU here: setup binds no device and returns NULL rather than
    throwing across the C boundary; execute on NULL is an error status."""
    import y3_cluster_cpp_amd as y3
    try:
        rt = y3.Runtime(0)
        rt.close()
        pytest.skip("a GPU is present")
    except RuntimeError:
        pass
    m = cs.Module("NumberCounts5DIntegrand")
    cfg = cs.DataBlock()
    L = "NumberCounts5DIntegrand"
    for k, v in [("algorithm", "pagani"), ("eps_rel", 1e-4), ("eps_abs", 0.0),
                 ("max_eval", 1e8), ("use_cartesian_product", False)]:
        cfg.put(L, k, v)
    for var, lo, hi in [("lo", 20, 30), ("zo", 0.2, 0.35), ("lc", 1, 300), ("zt", 0.1, 0.45),
                        ("lnm", 29.9, 35.6)]:
        cfg.put(L, var + "_low", [float(lo)])
        cfg.put(L, var + "_high", [float(hi)])
    assert not m.setup(cfg)
    assert m.execute(cs.DataBlock()) != 0
    ev = cs.Module("NumCountsSel")
    cfg = cs.DataBlock()
    for k, v in (("zt_low", 0.1), ("zt_high", 0.8), ("lnm_low", 30.0), ("lnm_high", 35.0)):
        cfg.put("NumCountsSel", k, v)
    cfg.put("NumCountsSel", "bin_index", [0.0, 1.0])
    assert not ev.setup(cfg)
    assert ev.execute(cs.DataBlock()) != 0
    # a real-data module: same behaviour, and fp32split is refused before
    # any device is touched (these integrands have no FP32 path)
    L = "SigmaCentY1MortCUDAScalarIntegrand"
    rd = cs.Module(L)
    cfg = cs.DataBlock()
    for k, v in [("algorithm", "pagani"), ("eps_rel", 1e-3), ("eps_abs", 0.0),
                 ("max_eval", 1e8), ("use_cartesian_product", False)]:
        cfg.put(L, k, v)
    for var, lo, hi in [("lo", 20, 30), ("zt", 0.1, 0.45), ("lnm", 29.9, 35.6)]:
        cfg.put(L, var + "_low", [float(lo)])
        cfg.put(L, var + "_high", [float(hi)])
    for var, v in [("zo_low", 0.2), ("zo_high", 0.35), ("radii", 1.0)]:
        cfg.put(L, var, [v])
    assert not rd.setup(cfg)
    cfg.put(L, "precision", "fp32split")
    assert not rd.setup(cfg)


def test_datablock_library_exports_c_api():
    import re
    src = open(os.path.join(MOD, "cosmosis", "datablock", "c_datablock.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b((?:make_|destroy_)?c_datablock[a-z0-9_]*)\s*\(", src)))
    assert "make_c_datablock" in names and len(names) >= 20
    lib = C.CDLL(cs.DATABLOCK_LIB)
    assert not [n for n in names if not hasattr(lib, n)]
    assert set(names) <= set(cs._API), set(names) - set(cs._API)


def test_make_integral_volume_standard_box():
    """The module path's descriptor for the standard volumes equals the
    library's standard descriptor, except the offset-radius normalisation
    (1 over the full Rayleigh range instead of 1/(1-e^-18) for the 6-sigma cut)."""
    import y3_cluster_cpp_amd as y3
    p, s = y3.default_params(), y3.spec_wide()
    lnm = (s.lnm_lo, s.lnm_hi)
    cases = [
        (y3.COUNT5D, [20, 0.2, 1, 0.1, lnm[0]], [30, 0.35, 300, 0.45, lnm[1]], (),
         (20, 30, 0.2, 0.35, 0.0)),
        (y3.DSIGMA6D, [60, 1, 0.4, lnm[0], 0, 0], [300, 300, 0.75, lnm[1], 1e300, 2 * np.pi],
         (0.5, 0.65, 3.0), (60, 300, 0.5, 0.65, 3.0)),
    ]
    for kind, lo, hi, grid, std in cases:
        a = y3.make_integral_volume(p, s, kind, lo, hi, grid)
        b = y3.make_integral(p, s, kind, *std)
        assert a.ndim == b.ndim and a.kind == b.kind
        for k in range(a.ndim):
            assert a.lows[k] == pytest.approx(b.lows[k], rel=1e-15, abs=1e-15)
            assert a.highs[k] == pytest.approx(b.highs[k], rel=1e-15, abs=1e-15)
        for f in ("lo_lo", "lo_hi", "zo_lo", "zo_hi", "radius", "sigma_mis", "lnlt_lo"):
            assert getattr(a, f) == pytest.approx(getattr(b, f), rel=1e-15, abs=1e-15), f
        if kind == y3.DSIGMA6D:
            assert a.rmis_norm == 1.0 and b.rmis_norm == pytest.approx(1 / (1 - np.exp(-18.0)))
    # COUNT3D: lambda_obs bin from the grid point
    a = y3.make_integral_volume(p, s, y3.COUNT3D, [1, 0.2, lnm[0]], [300, 0.35, lnm[1]], (20, 30))
    assert (a.lo_lo, a.lo_hi, a.ndim) == (20, 30, 3)
    # errors: wrong grid arity, empty axis, angle beyond 2 pi, zt beyond the tables
    with pytest.raises(RuntimeError, match="grid values"):
        y3.make_integral_volume(p, s, y3.COUNT5D, [20, 0.2, 1, 0.1, 30], [30, 0.35, 300, 0.45, 35],
                                (1.0,))
    with pytest.raises(RuntimeError, match="low > high"):
        y3.make_integral_volume(p, s, y3.COUNT5D, [30, 0.2, 1, 0.1, 30], [20, 0.35, 300, 0.45, 35])
    with pytest.raises(RuntimeError, match="theta"):
        y3.make_integral_volume(p, s, y3.DSIGMA6D, [20, 1, 0.1, 30, 0, 0],
                                [30, 300, 0.45, 35, 1, 7], (0.2, 0.35, 1.0))
    with pytest.raises(RuntimeError, match="zt range"):
        y3.make_integral_volume(p, s, y3.DSIGMA6D, [20, 1, 0.1, 30, 0, 0],
                                [30, 300, 1.2, 35, 1, 1], (0.2, 0.35, 1.0))
