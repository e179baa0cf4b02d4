"""Event location's grid search on the device (EventLocator.GridSearch, Location/EventLocator.cs:
114-175) over the tables of a solved table set, every event at once. The Geiger refinement, the
outlier rounds and the error ellipse that follow it in EventLocator.Locate stay with the caller.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr, reason_buffer
from .grid import GridDefinition


@dataclass
class LocatorSettings:
    """The LocatorSettings fields the grid search reads (EventLocator.cs:10-36)."""
    global_search: bool = True
    coarse_candidates: int = 8000
    fix_depth: bool = False


@dataclass
class Picks:
    """One event's usable picks: table (field index of the table set), observed time relative to the
    event's reference (s), weight, station correction; and the event's start position (lon, lat, depth)."""
    field: np.ndarray
    obs: np.ndarray
    weight: np.ndarray
    correction: np.ndarray
    start: tuple


def _declare():
    L = _lib.load()
    if getattr(L, "_loc_declared", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.ts_locate_grid_search.restype = i
    L.ts_locate_grid_search.argtypes = [vp, i, vp, vp, vp, vp, vp, i, vp, vp, vp, vp, vp, C.c_char_p, i]
    L._loc_declared = True
    return L


def _round_half_even(x: float) -> int:
    return int(round(x))  # Math.Round's default, as Python's round


def grid_search(tables, grid: GridDefinition, events: list[Picks], settings: LocatorSettings | None = None):
    """GridSearch for every event: arrays (lon, lat, depth, t0, node, misfit). The start position
    places a local search (global_search False) and a fixed depth (fix_depth), as
    Math.Round(Grid.Fractional(start)) does in the reference."""
    settings = settings or LocatorSettings()
    L = _declare()
    ne = len(events)
    starts = np.zeros(ne + 1, np.int32)
    starts[1:] = np.cumsum([len(e.field) for e in events])
    cat = lambda name, dt: np.ascontiguousarray(np.concatenate([np.asarray(getattr(e, name), dt) for e in events])
                                                if ne else np.zeros(0, dt))
    field, obs, w, corr = cat("field", np.int32), cat("obs", np.float64), cat("weight", np.float64), \
        cat("correction", np.float64)
    center = None
    kfix = None
    if not settings.global_search or settings.fix_depth:
        frac = np.array([grid.fractional(*e.start) for e in events]).reshape(ne, 3)
        rnd = np.array([[_round_half_even(v) for v in row] for row in frac], np.int32).reshape(ne, 3)
        if not settings.global_search:
            center = np.ascontiguousarray(rnd)
        if settings.fix_depth:
            kfix = np.ascontiguousarray(np.clip(rnd[:, 2], 0, grid.nz - 1).astype(np.int32))
    node = np.zeros(ne, np.int32)
    t0 = np.zeros(ne)
    mis = np.zeros(ne)
    r = reason_buffer()
    rc = L.ts_locate_grid_search(tables.handle, ne, ptr(starts), ptr(field), ptr(obs), ptr(w), ptr(corr),
                                 int(settings.coarse_candidates), ptr(center) if center is not None else None,
                                 ptr(kfix) if kfix is not None else None, ptr(node), ptr(t0), ptr(mis), r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    i, j, k = node % grid.nx, (node // grid.nx) % grid.ny, node // (grid.nx * grid.ny)
    lon, lat, dep = grid.lon()[i], grid.lat()[j], grid.depth()[k]
    if settings.fix_depth:
        dep = np.array([e.start[2] for e in events], np.float64)
    return lon, lat, dep, t0, node, mis
