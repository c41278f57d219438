"""ctypes bindings for the two native libraries of the package.

libfastpart_host.so  host helpers (include/fastpart_host.h)
libfastpart.so       the HIP engine (include/fastpart.h)

Both are built in-tree by __graft_entry__.build(). Loading the engine never
falls back to anything: if the library is missing, the call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_HERE, "libfastpart_host.so")
ENGINE_LIB = os.environ.get("FP_ENGINE_LIB") or os.path.join(_HERE, "libfastpart.so")

FP_OK = 0
STATUS_NAMES = {0: "FP_OK", 1: "FP_ERR_ARG", 2: "FP_ERR_HIP", 3: "FP_ERR_OOM", 4: "FP_ERR_STATE",
                5: "FP_ERR_UNSUPPORTED", 6: "FP_ERR_BUDGET", 7: "FP_ERR_IO", 8: "FP_ERR_NODEV"}


class FastPartError(RuntimeError):
    def __init__(self, code: int, what: str, detail: str = ""):
        self.code = code
        msg = f"{what} failed: {STATUS_NAMES.get(code, code)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


def check(rc: int, what: str, detail: str = "") -> None:
    if rc != FP_OK:
        raise FastPartError(rc, what, detail)


class GenParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("degree_mode", ctypes.c_int), ("deg_lo", ctypes.c_int),
                ("deg_hi", ctypes.c_int), ("zipf_alpha", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


_host = None
_engine = None

_U32P = ctypes.POINTER(ctypes.c_uint32)
_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)


def host():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise FastPartError(7, "load libfastpart_host.so", f"{HOST_LIB} not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(HOST_LIB)
        lib.fp_ref_capacity.restype = ctypes.c_double
        lib.fp_ref_capacity.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_double]
        for f in (lib.fp_ref_max_pw, lib.fp_ref_min_pw):
            f.restype = ctypes.c_longlong
            f.argtypes = [ctypes.c_longlong, ctypes.c_int]
        lib.fp_balance_cap.restype = ctypes.c_longlong
        lib.fp_balance_cap.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_double]
        lib.fp_gen_weights.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, _I32P]
        _host = lib
    return _host


def engine():
    """The HIP engine library. Raises if it is not built: there is no CPU path."""
    global _engine
    if _engine is None:
        if not os.path.exists(ENGINE_LIB):
            raise FastPartError(7, "load libfastpart.so", f"{ENGINE_LIB} not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(ENGINE_LIB)
        from . import engine as _e
        _e.declare(lib)
        _engine = lib
    return _engine
