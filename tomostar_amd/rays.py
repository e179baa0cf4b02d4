"""Host mirror of the reference's ray forward step (Tomography/RayKernels.cs) over libtomostar_hip.

``trace_rows`` is the device half of RayKernels.Compute: for every arrival, the back-trace on its
station's table (held in a solved ``TableSet``), TCalc and the hypocentre gradient at the event, and
the ray's kernel on the inversion grid; ``RayRow`` matches the reference's record.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr, reason_buffer
from .eikonal import TableSet
from .grid import GridDefinition


@dataclass
class RayRow:
    """RayRow (Tomography/RayKernels.cs:15): one traced arrival."""

    arrival: int
    nodes: np.ndarray    # int32, ascending
    length: np.ndarray   # float32, km
    tcalc: float
    ge: float
    gn: float
    gz: float
    outside_time: float = 0.0  # RayRow.OutsideTime: time of the pieces outside the inversion grid


class Background(C.Structure):
    """ts_background: the 1-D model of an extended forward domain (Model/VelocityModel1D.cs)."""

    _fields_ = [("count", C.c_int), ("depth_km", C.c_void_p), ("vp", C.c_void_p), ("vs", C.c_void_p)]


def _declare():
    L = _lib.load()
    if getattr(L, "_rays_declared", False):
        return L
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.ts_rays_trace.restype = i
    L.ts_rays_trace.argtypes = [vp, C.POINTER(_lib.TsGrid), vp, i, vp, vp, d, C.POINTER(vp), C.c_char_p, i]
    L.ts_rays_trace_extended.restype = i
    L.ts_rays_trace_extended.argtypes = [vp, C.POINTER(_lib.TsGrid), vp, i, vp, vp, vp, d, C.POINTER(Background),
                                         C.POINTER(vp), C.c_char_p, i]
    L.ts_rays_outside_time.restype = i
    L.ts_rays_outside_time.argtypes = [vp, vp]
    L.ts_rays_rows.restype = i
    L.ts_rays_rows.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.ts_rays_stats.restype = i
    L.ts_rays_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(d)]
    L.ts_rays_release.restype = None
    L.ts_rays_release.argtypes = [vp]
    L._rays_declared = True
    return L


def ray_step(forward: GridDefinition, inversion: GridDefinition) -> float:
    """RayKernels.cs:62: half the smaller of the two grids' smallest spacings."""
    return 0.5 * min(forward.min_spacing_km(), inversion.min_spacing_km())


class RayBatch:
    """Rows of many arrivals traced on the device; arrays in arrival order.

    ``background`` = (depth_km, vp, vs) of the 1-D model around an extended forward domain
    (RayKernels.cs:62-64): the pieces of a ray outside the inversion grid then go to the row's
    ``outside_time`` (dl / v of the arrival's phase, ``phase_of_ray`` 0 P, 1 S) instead of the grid's
    edge nodes (RayTracing.Kernel, RayTracing.cs:126-147)."""

    def __init__(self, tables: TableSet, inversion: GridDefinition, field_station: np.ndarray,
                 field_of_ray: np.ndarray, events: np.ndarray, step_km: float | None = None, host_rows: bool = True,
                 background=None, phase_of_ray: np.ndarray | None = None):
        L = _declare()
        fs = np.ascontiguousarray(field_station, np.float64).reshape(-1, 3)
        if fs.shape[0] != tables.nfields:
            raise ValueError("one station position per field is needed")
        fr = np.ascontiguousarray(field_of_ray, np.int32)
        ev = np.ascontiguousarray(events, np.float64).reshape(-1, 3)
        if ev.shape[0] != fr.size:
            raise ValueError("one event per arrival")
        step = ray_step(tables.grid, inversion) if step_km is None else step_km
        g = inversion.to_ts()
        h = C.c_void_p()
        r = reason_buffer()
        if background is None:
            rc = L.ts_rays_trace(tables.handle, C.byref(g), ptr(fs), int(fr.size), ptr(fr), ptr(ev), step,
                                 C.byref(h), r, len(r))
        else:
            bz, bp, bs = (np.ascontiguousarray(a, np.float64).ravel() for a in background)
            if not (bz.size == bp.size == bs.size) or bz.size == 0:
                raise ValueError("the background needs equal, non-empty depth, Vp and Vs arrays")
            if phase_of_ray is None:
                raise ValueError("a background needs the phase of every arrival")
            ph = np.ascontiguousarray(phase_of_ray, np.int32)
            if ph.size != fr.size:
                raise ValueError("one phase per arrival")
            bg = Background(int(bz.size), bz.ctypes.data, bp.ctypes.data, bs.ctypes.data)
            rc = L.ts_rays_trace_extended(tables.handle, C.byref(g), ptr(fs), int(fr.size), ptr(fr), ptr(ph),
                                          ptr(ev), step, C.byref(bg), C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self._h = h
        self.nrays = int(fr.size)
        n = self.nrays
        self.counts = np.zeros(n, np.int32)
        self.row_start = np.zeros(n + 1, np.int64)
        check(L.ts_rays_rows(h, ptr(self.counts), ptr(self.row_start), None, None, None, None))
        nnz = int(self.row_start[-1])
        self.nnz = nnz
        # the rows stay in HBM for the system assembly; host copies only when asked for
        self.nodes = np.zeros(max(1, nnz) if host_rows else 1, np.int32)
        self.lengths = np.zeros(max(1, nnz) if host_rows else 1, np.float32)
        self.tcalc_grad = np.zeros((n, 4))
        self.path_points = np.zeros(n, np.int32)
        check(L.ts_rays_rows(h, None, None, ptr(self.nodes) if host_rows else None,
                             ptr(self.lengths) if host_rows else None, ptr(self.tcalc_grad), ptr(self.path_points)))
        self.nodes, self.lengths = self.nodes[:nnz], self.lengths[:nnz]
        self.outside_time = np.zeros(n)
        if n > 0:
            check(L.ts_rays_outside_time(h, ptr(self.outside_time)))
        e, s = C.c_int64(), C.c_double()
        check(L.ts_rays_stats(h, C.byref(e), C.byref(s)))
        self.seconds = s.value

    @property
    def handle(self):
        return self._h

    def row(self, r: int) -> RayRow | None:
        if self.counts[r] < 0:
            return None
        a, b = self.row_start[r], self.row_start[r + 1]
        t = self.tcalc_grad[r]
        return RayRow(r, self.nodes[a:b], self.lengths[a:b], *t, float(self.outside_time[r]))

    def rows(self):
        """RayKernels.Compute's output: the rows of the arrivals that have one, in arrival order."""
        return [rw for rw in (self.row(r) for r in range(self.nrays)) if rw is not None]

    def release(self):
        if self._h is not None:
            _declare().ts_rays_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
