"""CosmoSIS module boundary from Python: DataBlock and module loaders.

``DataBlock`` wraps the C API of ``modules/libdatablock.so`` (the in-tree
subset of CosmoSIS's ``c_datablock.h``); ``Module`` dlopens one of the module
libraries (``NumberCounts3DIntegrand.so``, ``NumberCounts5DIntegrand.so``,
``SigmaMiscent6DIntegrand.so``) and calls its ``setup`` / ``execute`` /
``cleanup`` entry points the way CosmoSIS does (reference:
src/utils/cuda_module_macros.cuh:20-45, exercised like
src/modules/ExampleScalar/test_ExampleScalarIntegrationModule.cc:14-47).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MODULE_DIR = os.path.join(_HERE, "modules")
DATABLOCK_LIB = os.path.join(MODULE_DIR, "libdatablock.so")
MODULES = ("NumberCounts3DIntegrand", "NumberCounts5DIntegrand", "SigmaMiscent6DIntegrand",
           "NumCountsSel", "SigmaMiscentY1MortCUDAScalarIntegrand",
           "SigmaCentY1MortCUDAScalarIntegrand", "NCCentY1MortCUDAScalarIntegrand",
           "NCMiscentY1MortCUDAScalarIntegrand", "numberCountsCard", "bSelMargP1GPU",
           "bSelMargI1GPU", "bSelMargI2GPU", "bSelMargJGPU")

DBS_SUCCESS = 0

_VP = C.c_void_p
_S = C.c_char_p
_API = {
    "make_c_datablock": (_VP, []),
    "destroy_c_datablock": (C.c_int, [_VP]),
    "c_datablock_has_section": (C.c_bool, [_VP, _S]),
    "c_datablock_has_value": (C.c_bool, [_VP, _S, _S]),
    "c_datablock_put_int": (C.c_int, [_VP, _S, _S, C.c_int]),
    "c_datablock_put_double": (C.c_int, [_VP, _S, _S, C.c_double]),
    "c_datablock_put_bool": (C.c_int, [_VP, _S, _S, C.c_bool]),
    "c_datablock_put_string": (C.c_int, [_VP, _S, _S, _S]),
    "c_datablock_put_int_array_1d": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_int), C.c_int]),
    "c_datablock_put_double_array_1d": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_double),
                                                  C.c_int]),
    "c_datablock_get_int": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_int)]),
    "c_datablock_get_double": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_double)]),
    "c_datablock_get_bool": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_bool)]),
    "c_datablock_get_array_length": (C.c_int, [_VP, _S, _S]),
    "c_datablock_get_array_ndim": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_int)]),
    "c_datablock_get_double_array_shape": (C.c_int, [_VP, _S, _S, C.c_int,
                                                     C.POINTER(C.c_int)]),
    "c_datablock_get_int_array_shape": (C.c_int, [_VP, _S, _S, C.c_int, C.POINTER(C.c_int)]),
    "c_datablock_get_double_array": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_double), C.c_int,
                                               C.POINTER(C.c_int)]),
    "c_datablock_get_int_array": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_int), C.c_int,
                                            C.POINTER(C.c_int)]),
    "c_datablock_put_double_array": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_double), C.c_int,
                                               C.POINTER(C.c_int)]),
    "c_datablock_get_int_array_1d_preallocated": (C.c_int, [_VP, _S, _S, C.POINTER(C.c_int),
                                                            C.POINTER(C.c_int), C.c_int]),
    "c_datablock_get_double_array_1d_preallocated": (C.c_int, [_VP, _S, _S,
                                                               C.POINTER(C.c_double),
                                                               C.POINTER(C.c_int), C.c_int]),
}

_db = None


def datablock_lib():
    global _db
    if _db is None:
        if not os.path.exists(DATABLOCK_LIB):
            raise RuntimeError(f"{DATABLOCK_LIB} not built: run the package build")
        lib = C.CDLL(DATABLOCK_LIB)
        for n, (res, args) in _API.items():
            f = getattr(lib, n)
            f.restype = res
            f.argtypes = args
        _db = lib
    return _db


def _b(s):
    return s.encode() if isinstance(s, str) else s


class DataBlockError(RuntimeError):
    pass


class DataBlock:
    """A cosmosis::DataBlock owned from Python (section/name case-insensitive)."""

    def __init__(self):
        self._lib = datablock_lib()
        self.ptr = self._lib.make_c_datablock()

    def close(self):
        if self.ptr:
            self._lib.destroy_c_datablock(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != DBS_SUCCESS:
            raise DataBlockError(f"{what}: DATABLOCK_STATUS {rc}")

    def put(self, section, name, value):
        """Store value with the C API type its Python type implies: bool, int,
        float, str, or a 1-d (int or float) sequence / numpy array."""
        L, s, n = self._lib, _b(section), _b(name)
        if isinstance(value, (bool, np.bool_)):
            rc = L.c_datablock_put_bool(self.ptr, s, n, bool(value))
        elif isinstance(value, (int, np.integer)):
            rc = L.c_datablock_put_int(self.ptr, s, n, int(value))
        elif isinstance(value, (float, np.floating)):
            rc = L.c_datablock_put_double(self.ptr, s, n, float(value))
        elif isinstance(value, str):
            rc = L.c_datablock_put_string(self.ptr, s, n, _b(value))
        else:
            a = np.asarray(value)
            if a.ndim == 0:
                return self.put(section, name, a.item())
            if np.issubdtype(a.dtype, np.integer) and a.ndim == 1:
                a = np.ascontiguousarray(a, dtype=np.int32)
                rc = L.c_datablock_put_int_array_1d(
                    self.ptr, s, n, a.ctypes.data_as(C.POINTER(C.c_int)), a.size)
            elif a.ndim == 1:
                a = np.ascontiguousarray(a, dtype=np.float64)
                rc = L.c_datablock_put_double_array_1d(
                    self.ptr, s, n, a.ctypes.data_as(C.POINTER(C.c_double)), a.size)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                ext = (C.c_int * a.ndim)(*a.shape)
                rc = L.c_datablock_put_double_array(
                    self.ptr, s, n, a.ctypes.data_as(C.POINTER(C.c_double)), a.ndim, ext)
        self._chk(rc, f"put {section}/{name}")

    def has(self, section, name=None):
        if name is None:
            return bool(self._lib.c_datablock_has_section(self.ptr, _b(section)))
        return bool(self._lib.c_datablock_has_value(self.ptr, _b(section), _b(name)))

    def get_double(self, section, name):
        x = C.c_double()
        self._chk(self._lib.c_datablock_get_double(self.ptr, _b(section), _b(name), C.byref(x)),
                  f"get {section}/{name}")
        return x.value

    def get_int(self, section, name):
        x = C.c_int()
        self._chk(self._lib.c_datablock_get_int(self.ptr, _b(section), _b(name), C.byref(x)),
                  f"get {section}/{name}")
        return x.value

    def get_array(self, section, name):
        """An int or double array of any rank as a numpy array."""
        L, s, n = self._lib, _b(section), _b(name)
        nd = C.c_int()
        self._chk(L.c_datablock_get_array_ndim(self.ptr, s, n, C.byref(nd)),
                  f"ndim {section}/{name}")
        ext = (C.c_int * nd.value)()
        if L.c_datablock_get_double_array_shape(self.ptr, s, n, nd.value, ext) == DBS_SUCCESS:
            out = np.zeros(tuple(ext), dtype=np.float64)
            self._chk(L.c_datablock_get_double_array(
                self.ptr, s, n, out.ctypes.data_as(C.POINTER(C.c_double)), nd.value, ext),
                f"get {section}/{name}")
            return out
        self._chk(L.c_datablock_get_int_array_shape(self.ptr, s, n, nd.value, ext),
                  f"shape {section}/{name}")
        out = np.zeros(tuple(ext), dtype=np.int32)
        self._chk(L.c_datablock_get_int_array(
            self.ptr, s, n, out.ctypes.data_as(C.POINTER(C.c_int)), nd.value, ext),
            f"get {section}/{name}")
        return out


class Module:
    """One CosmoSIS module library: setup(cfg) / execute(sample) / cleanup()."""

    def __init__(self, name_or_path):
        path = name_or_path
        if os.path.sep not in path:
            path = os.path.join(MODULE_DIR, name_or_path + ".so")
        if not os.path.exists(path):
            raise RuntimeError(f"module library {path} not built")
        datablock_lib()  # the module resolves the DataBlock symbols from it
        self.lib = C.CDLL(path)
        self.lib.setup.restype = _VP
        self.lib.setup.argtypes = [_VP]
        self.lib.execute.restype = C.c_int
        self.lib.execute.argtypes = [_VP, _VP]
        self.lib.cleanup.restype = C.c_int
        self.lib.cleanup.argtypes = [_VP]
        self.handle = None

    def setup(self, cfg):
        self.handle = self.lib.setup(cfg.ptr)
        return self.handle is not None

    def execute(self, sample):
        return self.lib.execute(sample.ptr, self.handle)

    def cleanup(self):
        rc = self.lib.cleanup(self.handle)
        self.handle = None
        return rc
