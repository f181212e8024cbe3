"""ctypes binding of libfeelmri_hip.so (the C ABI declared in include/feelmri_hip.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the soname libamdhip64.so.7).  A missing
library is an error, never a fallback: build it with
``python -m feelmri_amd.build``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (load torch's HIP runtime before the library)

# FEELMRI_HIP_LIB selects another build of the same C ABI (e.g. to compare
# kernel revisions); the default is the in-tree build.
LIB_PATH = Path(os.environ.get("FEELMRI_HIP_LIB") or
                Path(__file__).resolve().parent / "_lib" / "libfeelmri_hip.so")

FM_OK, FM_ERR_ARG, FM_ERR_HIP, FM_ERR_UNSUPPORTED, FM_ERR_WORKSPACE = range(5)
PREC_F64, PREC_MIXED, PREC_F32 = 0, 1, 2
DTYPE_F32, DTYPE_F64 = 0, 1
METHOD_AUTO, METHOD_DIRECT, METHOD_CARTESIAN = 0, 1, 2
EVENTS_MAX_CANDIDATES, EVENTS_MAX_PARTS = 8192, 16

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_p = ctypes.c_void_p
_sz = ctypes.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPE)
_SIGNATURES = {
    "fm_last_error": [],
    "fm_version": [],
    "fm_signal_plan_create": [ctypes.POINTER(_p), _i64, _i64, _i64, _p, _p, _p, _p, _int],
    "fm_signal_plan_destroy": [_p],
    "fm_signal_plan_analyse": [_i64, _i64, _i64, _p, _p, _p, _p, _int, ctypes.POINTER(_int),
                               ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                               ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)],
    "fm_signal_plan_info": [_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_int), _p],
    "fm_signal_plan_runs": [_p, ctypes.POINTER(_i64), ctypes.POINTER(_int), ctypes.POINTER(_int)],
    "fm_signal_plan_lines": [_p, ctypes.POINTER(_i64), ctypes.POINTER(_int)],
    "fm_signal_workspace_size": [_p, _int, _i64, _int, _int, ctypes.POINTER(_sz)],
    "fm_signal_nodes_pack": [_i64, _int, _int, _p, _p, _p, _p, _p, _p, _p],
    "fm_signal_execute": [_p, _int, _i64, _int, _p, _p, _int, _p, _p, _sz, _p, _p],
    "fm_signal_execute_noalloc": [_p, _int, _i64, _int, _p, _p, _int, _p, _p, _p],
    "fm_signal_plan_set_timing": [_p, _int],
    "fm_signal_plan_kernel_ms": [_p, ctypes.POINTER(_dbl), ctypes.POINTER(_int)],
    "fm_csr_spmv_c128": [_i64, _int, _p, _p, _p, _p, _p, _p],
    "fm_bloch_nodes_pack": [_i64, _int, _p, _p, _p, _p, _p, _p],
    "fm_bloch_execute": [_int, _i64, _i64, _p, _dbl, _dbl, _p, _p, _p, _p, _p, _p, _int, _p, _p,
                         _p, _int, _p],
    "fm_bloch_store_spoil": [_i64, _p, _p, _p, _p, _i64, _int, _p],
    "fm_bloch_chain": [_int, _i64, _p, _dbl, _dbl, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _p, _p,
                       _p],
    "fm_bloch_chain_marked": [_int, _i64, _p, _dbl, _dbl, _p, _i64, _p, _p, _p, _p, _p, _int, _p, _p, _p, _p, _p,
                              _i64, _p, _p, _p],
    "fm_events_bound": [_i64, _p, _p, _i64, _p, ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "fm_events_build": [_i64, _p, _p, _p, _p, _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p],
    "fm_mass_p1_color": [_i64, _i64, _p, _p, _p, _int, ctypes.POINTER(_int)],
    "fm_mass_p1_lumped": [_i64, _i64, _p, _int, _p, _p, _p, _int, _p, _p],
    "fm_mass_p1_incidence": [_i64, _i64, _p, _p, _p, _p, _p, _i64, ctypes.POINTER(_i64)],
    "fm_mass_p1_lumped_gather": [_i64, _i64, _p, _p, _int, _p, _p, _p, _p, _i64, _p, _p, _p],
    "fm_mass_p1_element_matrices": [_i64, _p, _int, _p, _p, _p],
    "fm_pod_tensordot": [_int, _i64, _i64, _i64, _p, _p, _p, _p],
    "fm_pod_positions": [_i64, _p, _p, _int, _p, _p, _p],
    "fm_modes_pack": [_i64, _int, _p, _p, _p],
}
_RESTYPE = {"fm_last_error": ctypes.c_char_p}

_LIB: ctypes.CDLL | None = None


class FeelmriHipError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed (status {code}): {msg}")
        self.code = code


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library with `python -m feelmri_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int)
    _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int, name: str) -> None:
    """Raise for a non-zero status returned by entry point ``name``."""
    if rc != FM_OK:
        raise FeelmriHipError(name, rc, load().fm_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)


def last_error() -> str:
    return load().fm_last_error().decode(errors="replace")
