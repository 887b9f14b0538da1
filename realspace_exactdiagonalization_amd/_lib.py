"""ctypes binding of ``libsed.so`` (the C ABI declared in ``include/sed.h``).

The library is loaded from this package directory only; there is no fallback.  A
missing library raises ``SedLibraryMissing`` so a GPU run can never silently use
another code path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsed.so")

SED_OK = 0
SED_EINVAL = -1
SED_ENOMEM = -2
SED_EDEVICE = -3
SED_ENOCONV = -4
SED_EUNSUPPORTED = -5
SED_ESTATE = -6

SED_BOSONIC = 0
SED_FERMIONIC = 1

PREC_FP64 = 0
PREC_FP32 = 1
PREC_FP32_REFINE = 2

_ERRNAMES = {SED_EINVAL: "EINVAL", SED_ENOMEM: "ENOMEM", SED_EDEVICE: "EDEVICE", SED_ENOCONV: "ENOCONV",
             SED_EUNSUPPORTED: "EUNSUPPORTED", SED_ESTATE: "ESTATE"}


class SedLibraryMissing(RuntimeError):
    pass


class SedError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sed error {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class SolveStats(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iters_fp32", C.c_int32), ("iters_fp64", C.c_int32),
                ("restarts", C.c_int32), ("matvecs", C.c_int64), ("residual_max", C.c_double),
                ("seconds", C.c_double), ("full_orth_steps", C.c_int32), ("repeated", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol of include/sed.h
SIGNATURES = {
    "sed_last_error": (C.c_char_p, []),
    "sed_version": (C.c_int, []),
    "sed_create": (C.c_int, [C.c_int, _pp]),
    "sed_destroy": (C.c_int, [_vp]),
    "sed_catalog_build": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _f64p, _pp]),
    "sed_catalog_info": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "sed_catalog_copy": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "sed_catalog_stabilizers": (C.c_int, [_vp, _vp, _vp]),
    "sed_catalog_destroy": (C.c_int, [_vp]),
    "sed_catalog_update_phases": (C.c_int, [_vp, _f64p]),
    "sed_sector_build": (C.c_int, [_vp, _f64p, C.c_double, _pp, C.POINTER(C.c_int64)]),
    "sed_sector_basis": (C.c_int, [_vp, _vp, _vp]),
    "sed_sector_set_terms": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _f64p, C.c_int, _i32p, _i32p, _f64p]),
    "sed_sector_hop_counts": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "sed_matvec": (C.c_int, [_vp, _f64p, _f64p, C.c_int]),
    "sed_matvec_device": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "sed_matvec_device_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "sed_matvec_benchmark": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "sed_sector_set_shape": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "sed_sector_shape": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _vp, C.POINTER(C.c_int)]),
    "sed_solve": (C.c_int, [_vp, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int, _f64p, _vp,
                            C.POINTER(SolveStats)]),
    "sed_sector_density": (C.c_int, [_vp, C.c_int, _f64p]),
    "sed_sector_density_correlation": (C.c_int, [_vp, C.c_int, _f64p, _f64p]),
    "sed_sector_one_body": (C.c_int, [_vp, C.c_int, _f64p]),
    "sed_sector_matrix_csc": (C.c_int, [_vp, C.POINTER(C.c_int64), _vp, _vp, _vp]),
    "sed_sector_fock_expansion": (C.c_int, [_vp, _vp, C.c_double, C.c_int64, C.POINTER(C.c_int64), _vp, _vp]),
    "sed_sector_position_operator": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sed_sector_destroy": (C.c_int, [_vp]),
    "sed_server_create": (C.c_int, [_vp, C.c_int64, C.c_int, _pp]),
    "sed_server_submit": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                    C.POINTER(C.c_int64)]),
    "sed_server_run": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64)]),
    "sed_server_result": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_int), _f64p, C.POINTER(SolveStats),
                                    C.POINTER(C.c_double)]),
    "sed_server_memory": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]),
    "sed_server_destroy": (C.c_int, [_vp]),
}

_lib = None


def lib():
    """The loaded library (raises SedLibraryMissing when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SedLibraryMissing(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                                    "or python -m realspace_exactdiagonalization_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != SED_OK:
        msg = lib().sed_last_error()
        raise SedError(rc, msg.decode() if msg else "")
    return rc


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None
