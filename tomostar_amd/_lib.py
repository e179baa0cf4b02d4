"""ctypes binding of libtomostar_hip.so (the C ABI declared in include/tomostar_hip.h).

The library is built in-tree by ``tomostar_amd.build`` (``__graft_entry__.build()``). Loading fails
loudly when it is missing: there is no fallback, every solve of this package runs on the HIP device.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TOMOSTAR_LIB") or os.path.join(HERE, "libtomostar_hip.so")  # TOMOSTAR_LIB: another build

TS_OK, TS_DECLINED, TS_EINVAL, TS_CANCELLED, TS_EDEVICE = 0, 1, -1, -2, -3
TS_CREATE_SKIP_SELFTEST = 1
TS_EIK_PLAIN_FP64, TS_EIK_FAST_FP32, TS_EIK_FAST_FP64 = 0, 1, 2
TS_LSQR_PLAIN_FP64, TS_LSQR_FAST_FP32, TS_LSQR_FAST_MIXED, TS_LSQR_FAST_FP64 = 0, 1, 2, 3


class TsGrid(C.Structure):
    """ts_grid: node counts and the GridDefinition bounds (Geo/SphericalGrid.cs:13-31)."""

    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("min_lon", C.c_double), ("max_lon", C.c_double), ("min_lat", C.c_double),
                ("max_lat", C.c_double), ("min_depth", C.c_double), ("max_depth", C.c_double)]


_lib = None

_d = C.c_double
_i = C.c_int
_vp = C.c_void_p
_cp = C.c_char_p


def _sig(L, name, restype, *args):
    f = getattr(L, name)
    f.restype = restype
    f.argtypes = list(args)


def load() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m tomostar_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    G = C.POINTER(TsGrid)
    _sig(L, "ts_hip_create", _i, _i, _i, C.POINTER(_vp), _cp, _i)
    _sig(L, "ts_hip_destroy", None, _vp)
    _sig(L, "ts_hip_device_info", _i, _vp, _cp, _i)
    _sig(L, "ts_hip_selftest", _i, _vp, C.POINTER(_d), C.POINTER(_d))
    _sig(L, "ts_grid_metric", _i, G, _vp, _vp, C.POINTER(_d))
    _sig(L, "ts_point_source", _i, G, _vp, _d, _d, _d, _d, _d, _vp, _vp, _i)
    _sig(L, "ts_eik_suggest_batch", _i, _vp, _i)
    _sig(L, "ts_eik_solve", _i, _vp, G, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _cp, _i)
    _sig(L, "ts_tables_create", _i, _vp, G, _i, _vp, _i, _vp, _vp, _vp, _vp, C.POINTER(_vp), _cp, _i)
    _sig(L, "ts_tables_solve", _i, _vp, _i, _vp, _vp, _cp, _i)
    _sig(L, "ts_tables_read", _i, _vp, _i, _i, _vp)
    _sig(L, "ts_tables_read_f32", _i, _vp, _i, _i, _vp)
    _sig(L, "ts_tables_rounds", _i, _vp, _vp)
    _sig(L, "ts_tables_stats", _i, _vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), C.POINTER(_d))
    _sig(L, "ts_tables_sweep_stats", _i, _vp, _vp, _vp, _vp)
    _sig(L, "ts_tables_round_log", _i, _vp, _i, _vp, _vp)
    _sig(L, "ts_tables_release", None, _vp)
    for name, res, args in _EXTRA:
        if hasattr(L, name):
            _sig(L, name, res, *args)
    _lib = L
    return L


# Entry points of later layers, declared when the library exports them.
_EXTRA: list = []


def ptr(a: np.ndarray | None):
    """Address of a C-contiguous array (None -> NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class HipError(RuntimeError):
    def __init__(self, status: int, reason: str):
        super().__init__(f"status {status}: {reason}")
        self.status = status
        self.reason = reason


def reason_buffer():
    return C.create_string_buffer(512)


def check(status: int, reason=None) -> int:
    if status < 0:
        msg = reason.value.decode(errors="replace") if reason is not None else ""
        raise HipError(status, msg)
    return status
