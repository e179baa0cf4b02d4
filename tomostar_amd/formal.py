"""The exact formal resolution on the device (FormalResolution, Tomography/FormalResolution.cs:62-352)
over libtomostar_hip.

``compute`` is FormalResolution.Compute: the resolution diagonal, the posterior standard deviations
and, with places, the resolution lengths of the kept columns of a weighted system, the hypocentre
groups separated by their Schur complements. The dense factorisation runs in HBM on FP64 MFMA, so the
reference's limit of 20,000 unknowns (16 n^2 bytes) is the caller's choice here, not the device's:
the device holds about 2.2 n_pad^2 doubles (n_pad = n rounded up to 64), a few 10^5 unknowns in 288 GB.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ptr, reason_buffer
from .eikonal import Device
from .grid import DEG2RAD, EARTH_RADIUS_KM, GridDefinition


@dataclass
class FormalResolutionResult:
    """FormalResolutionResult (FormalResolution.cs:22): per kept column its R_ii, sqrt(C_ii) and
    resolution length (km, NaN without a place)."""

    columns: np.ndarray
    diagonal: np.ndarray
    standard_deviation: np.ndarray
    length_km: np.ndarray
    seconds: float


@dataclass
class FormalResolutionMaps:
    """FormalResolutionMaps (FormalResolution.cs:62): one block's values on the grid nodes."""

    resolution: np.ndarray
    sigma: np.ndarray
    length_km: np.ndarray


def _declare():
    L = _lib.load()
    if getattr(L, "_formal_declared", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.ts_formal_resolution.restype = i
    L.ts_formal_resolution.argtypes = [vp, i, i, vp, vp, vp, i, i, vp, vp, i, vp, vp, vp, vp, vp, vp, vp, vp,
                                       C.c_char_p, i]
    L._formal_declared = True
    return L


def to_cartesian(lon, lat, depth):
    """GeoMath.ToCartesian (Geo/GeoMath.cs:49-56): geocentric km."""
    r = EARTH_RADIUS_KM - np.asarray(depth, np.float64)
    la, lo = np.asarray(lat, np.float64) * DEG2RAD, np.asarray(lon, np.float64) * DEG2RAD
    cl = np.cos(la)
    return np.stack([r * cl * np.cos(lo), r * cl * np.sin(lo), r * np.sin(la)], axis=-1)


def add_places(places, grid: GridDefinition, offset: int, block: int, column_of_node=None):
    """AddPlaces (FormalResolution.cs:83-99): every column of a block at the centroid of the grid nodes
    it stands for. places = (xyz[cols, 3], block[cols]) is filled in place."""
    xyz, blk = places
    k, j, i = np.meshgrid(np.arange(grid.nz), np.arange(grid.ny), np.arange(grid.nx), indexing="ij")
    p = to_cartesian(grid.lon()[i.ravel()], grid.lat()[j.ravel()], grid.depth()[k.ravel()])
    col = offset + (np.arange(grid.count) if column_of_node is None
                    else np.array([column_of_node(q) for q in range(grid.count)]))
    cols = np.unique(col)
    for a in range(3):
        xyz[cols, a] = np.bincount(col, weights=p[:, a], minlength=xyz.shape[0])[cols] / \
            np.bincount(col, minlength=xyz.shape[0])[cols]
    blk[cols] = block
    return places


def new_places(columns: int):
    return np.full((columns, 3), np.nan), np.full(columns, -1, np.int32)


def on_nodes(r: FormalResolutionResult, column_count: int, nodes: int, offset: int,
             column_of_node=None) -> FormalResolutionMaps:
    """OnNodes (FormalResolution.cs:66-81): node i takes the values of column offset + columnOfNode(i)."""
    at = np.full(column_count, -1)
    at[r.columns] = np.arange(r.columns.size)
    col = offset + (np.arange(nodes) if column_of_node is None else np.array([column_of_node(q) for q in range(nodes)]))
    k = at[col]
    ok = k >= 0

    def take(v):
        return np.where(ok, v[np.maximum(k, 0)], np.nan)

    return FormalResolutionMaps(take(r.diagonal), take(r.standard_deviation), take(r.length_km))


def compute(device: Device, rows: int, cols: int, row_start, column, values, data_rows: int, eliminate=(),
            max_unknowns: int = 20000, places=None) -> FormalResolutionResult:
    """FormalResolution.Compute (:108-333) for the weighted system (CSR, FP64 values): rows below
    data_rows are data, the rest regularisation; eliminate: groups of columns separated out (the 4
    hypocentre columns of each event); places: (xyz[cols, 3], block[cols]) from ``add_places`` or None.
    Raises HipError with the reference's messages on the declines (too many unknowns, a row coupling
    two groups, a regularisation row coupling model and hypocentre columns, a matrix that is not
    positive definite)."""
    L = _declare()
    rs = np.ascontiguousarray(row_start, np.int32)
    ci = np.ascontiguousarray(column, np.int32)
    va = np.ascontiguousarray(values, np.float64)
    gs = np.zeros(len(eliminate) + 1, np.int32)
    gs[1:] = np.cumsum([len(g) for g in eliminate])
    gc = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32) for g in eliminate]) if len(eliminate)
                              else np.zeros(1, np.int32), np.int32)
    kept = np.zeros(max(1, cols), np.int32)
    nk = C.c_int32()
    diag, sd, ln = np.zeros(max(1, cols)), np.zeros(max(1, cols)), np.zeros(max(1, cols))
    pxyz = pblk = None
    if places is not None:
        pxyz = np.ascontiguousarray(places[0], np.float64).reshape(cols, 3)
        pblk = np.ascontiguousarray(places[1], np.int32)
    secs = C.c_double()
    r = reason_buffer()
    rc = L.ts_formal_resolution(device.handle, int(rows), int(cols), ptr(rs), ptr(ci), ptr(va), int(data_rows),
                                len(eliminate), ptr(gs), ptr(gc), int(max_unknowns),
                                None if pxyz is None else ptr(pxyz), None if pblk is None else ptr(pblk), ptr(kept),
                                C.byref(nk), ptr(diag), ptr(sd), ptr(ln), C.byref(secs), r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    n = nk.value
    return FormalResolutionResult(kept[:n].copy(), diag[:n].copy(), sd[:n].copy(), ln[:n].copy(), secs.value)


def velocity_formal(device: Device, system, grid: GridDefinition, nevents: int, max_unknowns: int = 20000):
    """TravelTimeTomography.Formal (Tomography/TravelTimeTomography.cs:525-548) on a plain node grid: the
    system of the last iteration (``system.build_system``'s ``System``, FP64 values kept), the 4
    hypocentre columns of every event separated out, the P and S blocks placed (blocks 0 and 1), the
    results mapped to the nodes. Returns (P maps or None, S maps or None, the result)."""
    lay = system.layout
    rs, ci, va = system.read()
    cols, rows = system.matrix.cols, system.matrix.rows
    groups = [] if lay["offH"] < 0 else [list(range(lay["offH"] + 4 * e, lay["offH"] + 4 * e + 4))
                                         for e in range(nevents)]
    places = new_places(cols)
    if lay["offP"] >= 0:
        add_places(places, grid, lay["offP"], 0)
    if lay["offS"] >= 0:
        add_places(places, grid, lay["offS"], 1)
    fr = compute(device, rows, cols, rs, ci, va, lay["data_rows"], groups, max_unknowns, places)
    p = on_nodes(fr, cols, grid.count, lay["offP"]) if lay["offP"] >= 0 else None
    s = on_nodes(fr, cols, grid.count, lay["offS"]) if lay["offS"] >= 0 else None
    return p, s, fr
