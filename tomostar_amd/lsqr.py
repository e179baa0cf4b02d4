"""Host mirror of the reference's device LSQR (Numerics/LsqrOpenCl.cs) over libtomostar_hip.

``Matrix`` is a CSR matrix resident in HBM with its stored transpose (CsrMatrix, SparseMatrix.cs:24-97).
``LsqrHip`` keeps LsqrOpenCl's surface: ``try_create`` (self-test gate, None on decline), ``try_solve``
(None plus a reason on decline; cancellation raises), ``device`` and ``SelfTestTolerance``; and adds
``solve_many`` for k right-hand sides with their own damping sharing every pass over A.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from . import _lib
from ._lib import TS_LSQR_FAST_FP32, TS_LSQR_FAST_FP64, TS_LSQR_FAST_MIXED, TS_LSQR_PLAIN_FP64, check, ptr, reason_buffer
from .eikonal import Device

MODES = {"plain_fp64": TS_LSQR_PLAIN_FP64, "fast_fp32": TS_LSQR_FAST_FP32, "fast_mixed": TS_LSQR_FAST_MIXED,
         "fast_fp64": TS_LSQR_FAST_FP64}

# Lsqr.cs:141-147 and :48-62, in the ABI's stop-code order
STOP_REASONS = {
    0: "iteration budget exhausted", 1: "residual reduced to tolerance (btol)",
    2: "normal equations satisfied (atol)", 3: "condition-number limit reached (conlim)",
    4: "residual at machine precision", 5: "normal equations at machine precision",
    6: "condition number at machine precision", 7: "bidiagonalisation terminated", 8: "empty system",
    9: "zero right-hand side", 10: "A^T b = 0: the data do not see the model",
}


@dataclass
class LsqrResult:
    """LsqrResult (Numerics/Lsqr.cs:7-13)."""

    solution: np.ndarray
    iterations: int
    initial_residual_norm: float
    residual_norm: float
    estimated_condition_number: float
    stop_reason: str
    stop_code: int
    residual_history: np.ndarray | None = None


def _declare():
    L = _lib.load()
    if getattr(L, "_lsqr_declared", False):
        return L
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.ts_matrix_create.restype = i
    L.ts_matrix_create.argtypes = [vp, i, i, vp, vp, vp, i, C.POINTER(vp), C.c_char_p, i]
    L.ts_matrix_stats.restype = i
    L.ts_matrix_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(d)]
    L.ts_matrix_release.restype = None
    L.ts_matrix_release.argtypes = [vp]
    L.ts_matrix_multiply.restype = i
    L.ts_matrix_multiply.argtypes = [vp, i, vp, vp, i, C.c_char_p, i]
    L.ts_matrix_multiply_noalloc.restype = i
    L.ts_matrix_multiply_noalloc.argtypes = [vp, i, vp, vp, i, C.c_char_p, i]
    L.ts_lsqr_solve.restype = i
    L.ts_lsqr_solve.argtypes = [vp, i, vp, vp, i, d, d, d, i, vp, vp, vp, vp, vp, vp, C.c_char_p, i]
    L.ts_hip_selftest_lsqr_fast.restype = i
    L.ts_hip_selftest_lsqr_fast.argtypes = [vp, C.POINTER(d)]
    L._lsqr_declared = True
    return L


class Matrix:
    """A CSR matrix in HBM with its transpose (built once, as CsrMatrix.Transpose caches it)."""

    def __init__(self, device: Device, rows: int, cols: int, row_start, column, values, keep_fp64: bool = True):
        L = _declare()
        self.device = device
        self.rows, self.cols = int(rows), int(cols)
        self._rs = np.ascontiguousarray(row_start, np.int32)
        self._ci = np.ascontiguousarray(column, np.int32)
        self._va = np.ascontiguousarray(values, np.float64)
        if self._rs.size != self.rows + 1 or self._ci.size != self._va.size or self._rs[-1] != self._va.size:
            raise ValueError("inconsistent CSR arrays")
        h = C.c_void_p()
        r = reason_buffer()
        rc = L.ts_matrix_create(device.handle, self.rows, self.cols, ptr(self._rs), ptr(self._ci), ptr(self._va),
                                1 if keep_fp64 else 0, C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self._h = h
        self.nnz = int(self._va.size)
        self.has_fp64 = bool(keep_fp64)

    @staticmethod
    def from_scipy(device: Device, a, keep_fp64: bool = True) -> "Matrix":
        a = a.tocsr()
        return Matrix(device, a.shape[0], a.shape[1], a.indptr, a.indices, a.data, keep_fp64)

    @property
    def handle(self):
        return self._h

    def multiply(self, x: np.ndarray, transpose: bool = False) -> np.ndarray:
        """A x (or A^T x) on the device for x[cols] or x[k, cols] (CsrMatrix.MultiplyInto)."""
        return multiply(self, x, transpose)

    def last_seconds(self) -> float:
        n, s = C.c_int64(), C.c_double()
        check(_declare().ts_matrix_stats(self._h, C.byref(n), C.byref(s)))
        return s.value

    def release(self):
        if self._h is not None:
            _declare().ts_matrix_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def multiply(a: Matrix, x: np.ndarray, transpose: bool = False) -> np.ndarray:
    """y = A x (transpose: A^T x) on the device in FP64 sums, x[n] or x[k, n] in the given orders
    (CsrMatrix.MultiplyInto, Numerics/SparseMatrix.cs:46-73)."""
    L = _declare()
    x = np.asarray(x, np.float64)
    single = x.ndim == 1
    x2 = np.ascontiguousarray(np.atleast_2d(x))
    nin, nout = (a.rows, a.cols) if transpose else (a.cols, a.rows)
    if x2.shape[1] != nin:
        raise ValueError(f"vector length {x2.shape[1]} differs from {nin}")
    y = np.zeros((x2.shape[0], nout))
    r = reason_buffer()
    rc = L.ts_matrix_multiply(a.handle, x2.shape[0], ptr(x2), ptr(y), 1 if transpose else 0, r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    return y[0] if single else y


def default_mode(a: Matrix) -> str:
    """The fast form that keeps the FP64 solver's arithmetic: FP64 matrix values when the matrix keeps them
    (TS_LSQR_FAST_FP64), else FP32 values with FP64 vectors (TS_LSQR_FAST_MIXED)."""
    return "fast_fp64" if getattr(a, "has_fp64", False) else "fast_mixed"


def solve_many(a: Matrix, b: np.ndarray, damp: Sequence[float], max_iterations: int = 500, atol: float = 1e-8,
               btol: float = 1e-8, conlim: float = 1e8, mode: str | None = None, cancel: C.c_int32 | None = None,
               history: bool = False):
    """k right-hand sides b[k, m] with damping damp[k]; returns (status, reason, [LsqrResult] * k).
    mode None: default_mode(a)."""
    L = _declare()
    mode = default_mode(a) if mode is None else mode
    b = np.ascontiguousarray(np.atleast_2d(b), np.float64)
    k = b.shape[0]
    if b.shape[1] != a.rows:
        raise ValueError("right-hand side length differs from the row count")
    dmp = np.ascontiguousarray(np.broadcast_to(np.asarray(damp, np.float64), (k,)))
    x = np.zeros((k, a.cols))
    its, stop = np.zeros(k, np.int32), np.zeros(k, np.int32)
    stats = np.zeros((k, 3))
    hist = np.zeros((k, max(1, max_iterations))) if history else None
    r = reason_buffer()
    rc = L.ts_lsqr_solve(a.handle, k, ptr(b), ptr(dmp), max_iterations, atol, btol, conlim, MODES[mode],
                         C.byref(cancel) if cancel is not None else None, ptr(x), ptr(its), ptr(stop), ptr(stats),
                         ptr(hist), r, len(r))
    if rc != 0:
        return rc, r.value.decode(errors="replace"), None
    out = []
    for j in range(k):
        out.append(LsqrResult(x[j], int(its[j]), float(stats[j, 0]), float(stats[j, 1]), float(stats[j, 2]),
                              STOP_REASONS.get(int(stop[j]), "?"), int(stop[j]),
                              hist[j, :its[j]].copy() if history else None))
    return 0, "", out


class LsqrHip:
    """LsqrOpenCl's surface (Numerics/LsqrOpenCl.cs:30-584) on libtomostar_hip."""

    SelfTestTolerance = 1e-8  # LsqrOpenCl.cs:33
    MinNonZerosForDevice = 2_000_000  # LsqrOpenCl.cs:36

    def __init__(self, device: Device, mode: str | None = None):
        """mode None: per matrix, the form that keeps the FP64 solver's arithmetic (default_mode), as
        LsqrOpenCl computes in FP64."""
        self._dev = device
        self.mode = mode

    @staticmethod
    def try_create(log: Callable[[str], None] | None = None, device: int = 0,
                   mode: str | None = None) -> "LsqrHip | None":
        try:
            dev = Device.open(device, self_test=True)
        except _lib.HipError as e:
            if log:
                log(f"HIP LSQR declined ({e.reason}); using CPU.")
            return None
        if log:
            log(f"HIP LSQR ready on {dev.info}.")
        return LsqrHip(dev, mode)

    @property
    def device(self) -> str:
        return self._dev.info

    def self_test(self):
        """(plain FP64 error, fast FP32 error) of the gate, relative to max |x|."""
        L = _declare()
        e = C.c_double()
        L.ts_hip_selftest_lsqr_fast(self._dev.handle, C.byref(e))
        return self._dev.self_test_errors()[1], e.value

    def try_solve(self, a: Matrix, b: np.ndarray, damp: float = 0.0, max_iterations: int = 500, atol: float = 1e-8,
                  btol: float = 1e-8, conlim: float = 1e8, cancel: C.c_int32 | None = None,
                  on_iteration: Callable[[int, float], None] | None = None):
        """LsqrOpenCl.TrySolve (:244-270): (LsqrResult or None, reason)."""
        rc, why, res = solve_many(a, b, [damp], max_iterations, atol, btol, conlim, self.mode, cancel,
                                  history=on_iteration is not None)
        if rc == _lib.TS_CANCELLED:
            raise KeyboardInterrupt("cancelled")
        if rc != 0:
            return None, why
        r = res[0]
        if on_iteration is not None and r.residual_history is not None:
            for it, rn in enumerate(r.residual_history):
                on_iteration(it + 1, float(rn))
        return r, ""
