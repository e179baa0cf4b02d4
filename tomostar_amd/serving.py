"""The serving loop over libtomostar_hip (ts_server): table and LSQR requests against resident
velocity models and matrices, admitted by a device-memory budget and run concurrently on worker
streams. No counterpart in the reference, whose device types serialise their solves
(Forward/EikonalOpenCl.cs:64,165; Numerics/LsqrOpenCl.cs:184,257)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import TsGrid, check, ptr, reason_buffer
from .eikonal import MODES as EIK_MODES
from .eikonal import Device
from .grid import GridDefinition
from .lsqr import MODES as LSQR_MODES
from .lsqr import Matrix


def _declare():
    L = _lib.load()
    if getattr(L, "_server_declared", False):
        return L
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.ts_server_create.restype = i
    L.ts_server_create.argtypes = [vp, i, C.c_uint64, C.POINTER(vp), C.c_char_p, i]
    L.ts_server_add_model.restype = i
    L.ts_server_add_model.argtypes = [vp, C.POINTER(TsGrid), vp, vp, C.c_char_p, i]
    L.ts_server_submit_tables.restype = i
    L.ts_server_submit_tables.argtypes = [vp, C.c_int32, C.c_int32, i, vp, i, vp, C.c_char_p, i]
    L.ts_server_submit_lsqr.restype = i
    L.ts_server_submit_lsqr.argtypes = [vp, vp, i, vp, vp, i, d, d, d, i, vp, C.c_char_p, i]
    L.ts_server_wait.restype = i
    L.ts_server_wait.argtypes = [vp, C.c_int64, d, vp, vp, C.c_char_p, i]
    L.ts_server_result_tables.restype = i
    L.ts_server_result_tables.argtypes = [vp, C.c_int64, i, i, vp]
    L.ts_server_result_lsqr.restype = i
    L.ts_server_result_lsqr.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
    L.ts_server_release_request.restype = i
    L.ts_server_release_request.argtypes = [vp, C.c_int64]
    L.ts_server_stats.restype = i
    L.ts_server_stats.argtypes = [vp, vp, vp, vp]
    L.ts_server_destroy.restype = None
    L.ts_server_destroy.argtypes = [vp]
    L._server_declared = True
    return L


class Server:
    def __init__(self, device: Device, workers: int = 4, memory_budget_gb: float = 96.0):
        L = _declare()
        h = C.c_void_p()
        r = reason_buffer()
        rc = L.ts_server_create(device.handle, workers, int(memory_budget_gb * 1e9), C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self._h = h
        self.device = device
        self.models = {}

    def add_model(self, grid: GridDefinition, slowness: np.ndarray) -> int:
        s = np.ascontiguousarray(slowness, np.float64).ravel()
        if s.size != grid.count:
            raise ValueError("slowness size differs from the grid's node count")
        g = grid.to_ts()
        mid = C.c_int32()
        r = reason_buffer()
        rc = _declare().ts_server_add_model(self._h, C.byref(g), ptr(s), C.byref(mid), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self.models[mid.value] = grid
        return mid.value

    def submit_tables(self, model_p: int, model_s: int, sources: np.ndarray, mode: str = "fast_fp32") -> int:
        src = np.ascontiguousarray(sources, np.float64).reshape(-1, 3)
        rid = C.c_int64()
        r = reason_buffer()
        rc = _declare().ts_server_submit_tables(self._h, model_p, model_s, src.shape[0], ptr(src), EIK_MODES[mode],
                                                C.byref(rid), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        return rid.value

    def submit_lsqr(self, a: Matrix, b: np.ndarray, damp, iterations: int, atol=0.0, btol=0.0, conlim=0.0,
                    mode: str = "fast_mixed") -> int:
        b = np.ascontiguousarray(np.atleast_2d(b), np.float64)
        k = b.shape[0]
        dmp = np.ascontiguousarray(np.broadcast_to(np.asarray(damp, np.float64), (k,)))
        rid = C.c_int64()
        r = reason_buffer()
        rc = _declare().ts_server_submit_lsqr(self._h, a.handle, k, ptr(b), ptr(dmp), iterations, atol, btol, conlim,
                                              LSQR_MODES[mode], C.byref(rid), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        return rid.value

    def wait(self, rid: int, timeout_s: float = 600.0):
        """(status, reason, (submitted, started, finished) in seconds since the server started)."""
        st = C.c_int32()
        t = np.zeros(3)
        r = reason_buffer()
        rc = _declare().ts_server_wait(self._h, rid, timeout_s, C.byref(st), ptr(t), r, len(r))
        if rc != 0:
            raise TimeoutError(f"request {rid} not finished within {timeout_s} s")
        return st.value, r.value.decode(errors="replace"), tuple(t)

    def tables(self, rid: int, first: int, count: int, nodes: int) -> np.ndarray:
        out = np.zeros((count, nodes), np.float32)
        check(_declare().ts_server_result_tables(self._h, rid, first, count, ptr(out)))
        return out

    def lsqr_result(self, rid: int, k: int, n: int):
        x = np.zeros((k, n))
        its, stop = np.zeros(k, np.int32), np.zeros(k, np.int32)
        stats = np.zeros((k, 3))
        check(_declare().ts_server_result_lsqr(self._h, rid, ptr(x), ptr(its), ptr(stop), ptr(stats)))
        return x, its, stop

    def release_request(self, rid: int):
        check(_declare().ts_server_release_request(self._h, rid))

    def stats(self):
        pk, pi, bud = C.c_uint64(), C.c_int32(), C.c_uint64()
        check(_declare().ts_server_stats(self._h, C.byref(pk), C.byref(pi), C.byref(bud)))
        return {"peak_bytes": pk.value, "peak_inflight": pi.value, "budget_bytes": bud.value}

    def close(self):
        if getattr(self, "_h", None) is not None:
            _declare().ts_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
