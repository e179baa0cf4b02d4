"""Host mirror of the reference's device eikonal (Forward/EikonalOpenCl.cs) over libtomostar_hip.

``EikonalHip`` keeps EikonalOpenCl's surface: ``try_create`` (self-test gate, None on decline),
``suggest_batch``, ``try_solve`` (False plus a reason on decline), ``self_test``, ``device`` and
``last_stats``. ``TableSet`` is the resident form used by the ray tracer, the benchmark and the
request loop: many (station, phase) fields of one grid, kept in HBM.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Sequence

import numpy as np

from . import _lib
from ._lib import TS_EIK_FAST_FP32, TS_EIK_FAST_FP64, TS_EIK_PLAIN_FP64, check, ptr, reason_buffer
from .grid import GridDefinition

MODES = {"plain_fp64": TS_EIK_PLAIN_FP64, "fast_fp32": TS_EIK_FAST_FP32, "fast_fp64": TS_EIK_FAST_FP64}


class Device:
    """An opened, self-tested HIP device (ts_context)."""

    def __init__(self, handle, info: str):
        self._h = handle
        self.info = info

    @staticmethod
    def open(device: int = 0, self_test: bool = True) -> "Device":
        L = _lib.load()
        h = C.c_void_p()
        r = reason_buffer()
        rc = L.ts_hip_create(device, 0 if self_test else _lib.TS_CREATE_SKIP_SELFTEST, C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        buf = C.create_string_buffer(512)
        L.ts_hip_device_info(h, buf, len(buf))
        return Device(h, buf.value.decode())

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("device closed")
        return self._h

    def self_test_errors(self):
        e, l = C.c_double(), C.c_double()
        _lib.load().ts_hip_selftest(self.handle, C.byref(e), C.byref(l))
        return e.value, l.value

    def close(self):
        if self._h is not None:
            _lib.load().ts_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def metric(grid: GridDefinition):
    """GridMetric (Forward/SphericalEikonal.cs:15-35): (hlon[ny*nz], hlat[nz], hr)."""
    hlon, hlat, hr = np.zeros(grid.ny * grid.nz), np.zeros(grid.nz), C.c_double()
    g = grid.to_ts()
    check(_lib.load().ts_grid_metric(C.byref(g), ptr(hlon), ptr(hlat), C.byref(hr)))
    return hlon, hlat, hr.value


def point_source(grid: GridDefinition, slowness: np.ndarray, lon: float, lat: float, depth: float,
                 origin_time: float = 0.0, radius_cells: float = 3.0):
    """SphericalEikonal.PointSource (Forward/SphericalEikonal.cs:68-100): (index int32[], time f64[])."""
    s = np.ascontiguousarray(slowness, dtype=np.float64).ravel()
    if s.size != grid.count:
        raise ValueError("slowness and grid differ in size")
    g = grid.to_ts()
    cap = 4096
    while True:
        idx, t = np.zeros(cap, np.int32), np.zeros(cap)
        n = check(_lib.load().ts_point_source(C.byref(g), ptr(s), lon, lat, depth, origin_time, radius_cells,
                                              ptr(idx), ptr(t), cap))
        if n <= cap:
            return idx[:n].copy(), t[:n].copy()
        cap = n


def _pack_seeds(seed_sets):
    offs = np.zeros(len(seed_sets) + 1, np.int64)
    for f, (idx, _) in enumerate(seed_sets):
        offs[f + 1] = offs[f] + len(idx)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.int32) for s in seed_sets]))
    tim = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.float64) for s in seed_sets]))
    return offs, idx, tim


class TableSet:
    """Resident travel-time fields (the device half of TravelTimeTableSet, TravelTimeTable.cs:251-430).

    ``slowness`` is a list of models (e.g. [P, S]) on ``grid``; field f is the point-source problem
    ``seed_sets[f]`` in model ``model_of_field[f]``.
    """

    def __init__(self, device: Device, grid: GridDefinition, slowness: Sequence[np.ndarray],
                 seed_sets: Sequence, model_of_field: Sequence[int] | None = None):
        L = _lib.load()
        self.device = device
        self.grid = grid
        self.nfields = len(seed_sets)
        self._models = [np.ascontiguousarray(s, dtype=np.float64).ravel() for s in slowness]
        for s in self._models:
            if s.size != grid.count:
                raise ValueError("slowness and grid differ in size")
        mptrs = (C.c_void_p * len(self._models))(*[s.ctypes.data for s in self._models])
        mof = np.ascontiguousarray(np.zeros(self.nfields, np.int32) if model_of_field is None
                                   else np.asarray(model_of_field, np.int32))
        offs, idx, tim = _pack_seeds(seed_sets)
        g = grid.to_ts()
        h = C.c_void_p()
        r = reason_buffer()
        rc = L.ts_tables_create(device.handle, C.byref(g), len(self._models), mptrs, self.nfields, ptr(mof),
                                ptr(offs), ptr(idx), ptr(tim), C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self._h = h
        self.rounds = (0, 0)

    @property
    def handle(self):
        return self._h

    def solve(self, mode: str | int = "fast_fp32", cancel: C.c_int32 | None = None):
        """Runs the solve on the device; returns (status, reason). Status > 0 is a decline."""
        m = MODES.get(mode, mode) if isinstance(mode, str) else mode
        r = reason_buffer()
        rounds = (C.c_int32 * 2)()
        rc = _lib.load().ts_tables_solve(self._h, m, C.byref(cancel) if cancel is not None else None, rounds,
                                         r, len(r))
        self.rounds = (rounds[0], rounds[1])
        return rc, r.value.decode(errors="replace")

    def read(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.nfields - first if count is None else count
        out = np.empty((count, self.grid.count), np.float64)
        check(_lib.load().ts_tables_read(self._h, first, count, ptr(out)))
        return out

    def read_f32(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.nfields - first if count is None else count
        out = np.empty((count, self.grid.count), np.float32)
        check(_lib.load().ts_tables_read_f32(self._h, first, count, ptr(out)))
        return out

    def field_rounds(self) -> np.ndarray:
        out = np.zeros(2 * self.nfields, np.int32)
        check(_lib.load().ts_tables_rounds(self._h, ptr(out)))
        return out.reshape(self.nfields, 2)

    def stats(self):
        """dict: node updates on FP32 and FP64 times, last solve wall seconds and sweep-graph device seconds."""
        a, b, s, k = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        L = _lib.load()
        check(L.ts_tables_stats(self._h, C.byref(a), C.byref(b), C.byref(s), C.byref(k)))
        rounds, per_round, secs = np.zeros(2, np.int32), C.c_int32(), np.zeros(2)
        check(L.ts_tables_sweep_stats(self._h, _lib.ptr(rounds), C.byref(per_round), _lib.ptr(secs)))
        return {"updates_fp32": a.value, "updates_fp64": b.value, "seconds": s.value, "sweep_seconds": k.value,
                "rounds_fp32": int(rounds[0]), "rounds_fp64": int(rounds[1]),
                "plane_launches_per_round": per_round.value,
                "sweep_seconds_fp32": float(secs[0]), "sweep_seconds_fp64": float(secs[1])}

    def round_log(self):
        """[(width, order, seconds)] per round of the last fast solve: width 32 or 64 (the times swept),
        order 0 (flag with every load), 1 (flag first), 2 (a few flags per wave), 3 (streaming tiles),
        4 (more flags per wave)."""
        L = _lib.load()
        n = L.ts_tables_round_log(self._h, 0, None, None)
        check(min(n, 0))
        kind, secs = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1))
        L.ts_tables_round_log(self._h, n, _lib.ptr(kind), _lib.ptr(secs))
        return [(64 if k & 16 else 32, int(k & 15), float(x)) for k, x in zip(kind[:n], secs[:n])]

    def release(self):
        if self._h is not None:
            _lib.load().ts_tables_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class EikonalHip:
    """EikonalOpenCl's surface (Forward/EikonalOpenCl.cs:43-385) on libtomostar_hip."""

    SelfTestTolerance = 1e-6  # EikonalOpenCl.cs:45

    def __init__(self, device: Device, mode: str = "fast_fp32"):
        self._dev = device
        self.mode = mode
        self.last_stats = ""

    @staticmethod
    def try_create(log: Callable[[str], None] | None = None, device: int = 0,
                   mode: str = "fast_fp32") -> "EikonalHip | None":
        """EikonalOpenCl.TryCreate (:88-118): None when there is no device or the self-test fails."""
        try:
            dev = Device.open(device, self_test=True)
        except _lib.HipError as e:
            if log:
                log(f"HIP eikonal declined ({e.reason}); using CPU.")
            return None
        if log:
            log(f"HIP eikonal ready on {dev.info}.")
        return EikonalHip(dev, mode)

    @property
    def device(self) -> str:
        return self._dev.info

    def suggest_batch(self, nodes: int) -> int:
        """EikonalOpenCl.SuggestBatch (:121-132)."""
        return check(_lib.load().ts_eik_suggest_batch(self._dev.handle, nodes))

    def try_solve(self, grid: GridDefinition, slowness: np.ndarray, seed_sets: Sequence, output: np.ndarray,
                  cancel: C.c_int32 | None = None):
        """EikonalOpenCl.TrySolve (:138-166): (ok, reason); output is field-major batch*N FP64."""
        batch = len(seed_sets)
        if output.size < batch * grid.count:
            return False, "output too small"
        if output.dtype != np.float64 or not output.flags["C_CONTIGUOUS"]:
            raise ValueError("output must be a C-contiguous float64 array")
        s = np.ascontiguousarray(slowness, dtype=np.float64).ravel()
        offs, idx, tim = _pack_seeds(seed_sets)
        g = grid.to_ts()
        r = reason_buffer()
        rounds = (C.c_int32 * 2)()
        rc = _lib.load().ts_eik_solve(self._dev.handle, C.byref(g), ptr(s), batch, ptr(offs), ptr(idx), ptr(tim),
                                      ptr(output), MODES[self.mode], C.byref(cancel) if cancel is not None else None,
                                      rounds, r, len(r))
        if rc == _lib.TS_CANCELLED:
            raise KeyboardInterrupt("cancelled")
        if rc < 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        if rc > 0:
            return False, r.value.decode(errors="replace")
        self.last_stats = f"{rounds[0]}+{rounds[1]} rounds of 8 sweeps"
        return True, ""

    def self_test(self):
        """The gate's error measured at creation (EikonalOpenCl.SelfTest, :346-374)."""
        return self._dev.self_test_errors()[0]
