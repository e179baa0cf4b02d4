"""The tomographic linear system assembled on the device (TravelTimeTomography.BuildSystem,
Tomography/TravelTimeTomography.cs:860-1034, plain node grid) over libtomostar_hip.

``build_system`` takes the traced rows of a ``RayBatch`` (still resident in HBM), the arrivals' data and
the models, and returns a ``System``: the LSQR ``Matrix`` (with its stored transpose), the right-hand
side and the ``MatrixLayout`` offsets (:1031).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import TsGrid, check, ptr, reason_buffer
from .eikonal import Device
from .grid import GridDefinition
from .lsqr import Matrix, _declare as _declare_lsqr
from .rays import RayBatch


class SystemOpts(C.Structure):
    _fields_ = [("vpvs", C.c_int32), ("invert_p", C.c_int32), ("invert_s", C.c_int32), ("any_p", C.c_int32),
                ("any_s", C.c_int32), ("joint_hypocentres", C.c_int32), ("station_corrections", C.c_int32),
                ("gradient_smoothing", C.c_int32), ("smoothing_in_nodes", C.c_int32), ("nevents", C.c_int32),
                ("nstations", C.c_int32), ("damping", C.c_double), ("smoothing", C.c_double),
                ("vertical_weight", C.c_double), ("damping_hypocentre", C.c_double),
                ("damping_station", C.c_double), ("s_factor", C.c_double), ("absolute_weight", C.c_double)]


@dataclass
class SystemSettings:
    """The TomographySettings fields BuildSystem reads (TravelTimeTomography.cs:30-110), reference defaults."""

    vpvs: bool = False                 # Parameterization == VpVpVs
    invert_p: bool = True
    invert_s: bool = True
    joint_hypocentres: bool = True
    station_corrections: bool = True
    smoothing_method: str = "laplacian"  # or "gradient" (first differences)
    smoothing_scale: str = "km"        # or "nodes"
    vertical_weight: float = 0.1       # VerticalSmoothingWeight
    damping_hypocentre: float = 1.0    # DampingHypocentre
    damping_station: float = 5.0       # DampingStation
    s_factor: float = 1.0              # SRegularisationFactor


@dataclass
class System:
    matrix: Matrix
    rhs: np.ndarray
    layout: dict = field(default_factory=dict)
    build_seconds: float = 0.0

    def read(self):
        """(row_start int32, column int32, values float64) of the assembled matrix."""
        m = self.matrix
        rs = np.zeros(m.rows + 1, np.int32)
        ci = np.zeros(max(1, m.nnz), np.int32)
        va = np.zeros(max(1, m.nnz))
        check(_declare().ts_matrix_read(m.handle, ptr(rs), ptr(ci), ptr(va), None))
        return rs, ci[:m.nnz], va[:m.nnz]

    def stored_order(self):
        """(row_perm, col_perm): stored row r is given row row_perm[r], stored column c is given column
        col_perm[c] (ts_matrix_stored_order)."""
        m = self.matrix
        rp = np.zeros(max(1, m.rows), np.int32)
        cp = np.zeros(max(1, m.cols), np.int32)
        L = _declare()
        L.ts_matrix_stored_order.restype = C.c_int
        L.ts_matrix_stored_order.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        check(L.ts_matrix_stored_order(m.handle, ptr(rp), ptr(cp)))
        return rp[:m.rows], cp[:m.cols]

    def read_transpose(self):
        m = self.matrix
        rs = np.zeros(m.cols + 1, np.int32)
        ci = np.zeros(max(1, m.nnz), np.int32)
        va = np.zeros(max(1, m.nnz), np.float32)
        check(_declare().ts_matrix_read_transpose(m.handle, ptr(rs), ptr(ci), ptr(va)))
        return rs, ci[:m.nnz], va[:m.nnz]


def _declare():
    L = _declare_lsqr()
    if getattr(L, "_system_declared", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.ts_system_build.restype = i
    L.ts_system_build.argtypes = [vp, C.POINTER(TsGrid), C.POINTER(SystemOpts), vp, i] + [vp] * 13 + \
        [i, C.POINTER(vp), vp, C.c_char_p, i]
    L.ts_matrix_shape.restype = i
    L.ts_matrix_shape.argtypes = [vp, vp, vp, vp, vp]
    L.ts_matrix_read.restype = i
    L.ts_matrix_read.argtypes = [vp, vp, vp, vp, vp]
    L.ts_matrix_read_transpose.restype = i
    L.ts_matrix_read_transpose.argtypes = [vp, vp, vp, vp]
    L._system_declared = True
    return L


class _AdoptedMatrix(Matrix):
    """A Matrix whose device arrays were assembled by the library (no host copy)."""

    def __init__(self, device: Device, handle: C.c_void_p):
        L = _declare()
        self.device = device
        self._h = handle
        rows, cols, nnz, secs = C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        check(L.ts_matrix_shape(handle, C.byref(rows), C.byref(cols), C.byref(nnz), C.byref(secs)))
        self.rows, self.cols, self.nnz = rows.value, cols.value, nnz.value
        self.build_seconds = secs.value
        self.has_fp64 = False  # set by the builder that knows


def build_system(device: Device, grid: GridDefinition, settings: SystemSettings, rays: RayBatch, ray_of_row,
                 phase, event, station, residual, sigma, event_fixed, event_outside, nstations: int, sP, sS, sRefP,
                 sRefS, kRef, damping: float, smoothing: float, any_p: bool | None = None, any_s: bool | None = None,
                 absolute_weight: float = 1.0, keep_fp64: bool = False) -> System:
    """BuildSystem for the used rows: row r is arrival ray_of_row[r] of ``rays`` with its phase (0 P, 1 S),
    event, station, residual and sigma; event_fixed / event_outside per event, nstations the station
    count (data.Stations.Count); any_p / any_s default to the phases present among the rows."""
    L = _declare()
    phase = np.ascontiguousarray(phase, np.int32)
    nrows = phase.size
    nev, nst = len(event_fixed), int(nstations)
    o = SystemOpts(vpvs=int(settings.vpvs), invert_p=int(settings.invert_p), invert_s=int(settings.invert_s),
                   any_p=int(np.any(phase == 0) if any_p is None else any_p),
                   any_s=int(np.any(phase == 1) if any_s is None else any_s),
                   joint_hypocentres=int(settings.joint_hypocentres),
                   station_corrections=int(settings.station_corrections),
                   gradient_smoothing=int(settings.smoothing_method == "gradient"),
                   smoothing_in_nodes=int(settings.smoothing_scale == "nodes"), nevents=nev, nstations=nst,
                   damping=damping, smoothing=smoothing, vertical_weight=settings.vertical_weight,
                   damping_hypocentre=settings.damping_hypocentre, damping_station=settings.damping_station,
                   s_factor=settings.s_factor, absolute_weight=absolute_weight)
    arrs = [np.ascontiguousarray(ray_of_row, np.int32), phase, np.ascontiguousarray(event, np.int32),
            np.ascontiguousarray(station, np.int32), np.ascontiguousarray(residual, np.float64),
            np.ascontiguousarray(sigma, np.float64), np.ascontiguousarray(event_fixed, np.uint8),
            np.ascontiguousarray(event_outside, np.uint8)]
    models = [np.ascontiguousarray(a, np.float64).ravel() for a in (sP, sS, sRefP, sRefS, kRef)]
    for a in models:
        if a.size != grid.count:
            raise ValueError("model size differs from the grid's node count")
    layout = np.zeros(8, np.int64)
    h = C.c_void_p()
    r = reason_buffer()
    g = grid.to_ts()
    rc = L.ts_system_build(device.handle, C.byref(g), C.byref(o), rays.handle, nrows, *[ptr(a) for a in arrs],
                           *[ptr(a) for a in models], 1 if keep_fp64 else 0, C.byref(h), ptr(layout), r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    mat = _AdoptedMatrix(device, h)
    rhs = np.zeros(max(1, mat.rows))
    check(L.ts_matrix_read(h, None, None, None, ptr(rhs)))
    names = ["offP", "offS", "offH", "offCP", "offCS", "columns", "rows", "data_rows"]
    mat.has_fp64 = bool(keep_fp64)
    return System(mat, rhs[:mat.rows], dict(zip(names, (int(x) for x in layout))), mat.build_seconds)
