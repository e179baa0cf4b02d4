"""The regular spherical grid of the reference (Geo/SphericalGrid.cs) on the host side.

Nodes sit on the bounds; node i of Nx is at MinLon + i*(MaxLon - MinLon)/(Nx - 1); the index is
(k*Ny + j)*Nx + i, longitude fastest (SphericalGrid.cs:135).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import TsGrid

EARTH_RADIUS_KM = 6371.0  # Geo/GeoMath.cs:43
DEG2RAD = math.pi / 180.0


_TRIG = {"sin": 0, "cos": 1, "asin": 2, "atan2": 3}


def trig(func: str, a, b=None) -> np.ndarray:
    """The library's correctly rounded sin / cos / asin of a, or atan2(a, b), evaluated on the host
    (ts_crmath_eval): the trigonometry every position of the library is computed with."""
    import ctypes as C

    L = _lib.load()
    if not getattr(L, "_crmath_declared", False):
        L.ts_crmath_eval.restype = C.c_int
        L.ts_crmath_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L._crmath_declared = True
    a = np.ascontiguousarray(a, np.float64)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    out = np.empty_like(a)
    _lib.check(L.ts_crmath_eval(None, _TRIG[func], a.size, _lib.ptr(a), _lib.ptr(bb), _lib.ptr(out)))
    return out


@dataclass(frozen=True)
class GridDefinition:
    """GridDefinition (SphericalGrid.cs:13-56)."""

    min_lon: float
    max_lon: float
    min_lat: float
    max_lat: float
    min_depth: float
    max_depth: float
    nx: int
    ny: int
    nz: int

    @property
    def dlon(self) -> float:
        return (self.max_lon - self.min_lon) / max(1, self.nx - 1)

    @property
    def dlat(self) -> float:
        return (self.max_lat - self.min_lat) / max(1, self.ny - 1)

    @property
    def ddepth(self) -> float:
        return (self.max_depth - self.min_depth) / max(1, self.nz - 1)

    @property
    def count(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        """numpy shape of a node field: (nz, ny, nx)."""
        return (self.nz, self.ny, self.nx)

    def lon(self) -> np.ndarray:
        return self.min_lon + np.arange(self.nx) * self.dlon

    def lat(self) -> np.ndarray:
        return self.min_lat + np.arange(self.ny) * self.dlat

    def depth(self) -> np.ndarray:
        return self.min_depth + np.arange(self.nz) * self.ddepth

    def refined(self, factor: int) -> "GridDefinition":
        """SphericalGrid.cs:51-56."""
        if factor <= 1:
            return self
        return GridDefinition(self.min_lon, self.max_lon, self.min_lat, self.max_lat, self.min_depth,
                              self.max_depth, (self.nx - 1) * factor + 1, (self.ny - 1) * factor + 1,
                              (self.nz - 1) * factor + 1)

    @staticmethod
    def from_spacing(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth, horizontal_km,
                     vertical_km) -> "GridDefinition":
        """GridDefinition.FromSpacing (SphericalGrid.cs:59-74); Math.Round is half-to-even, as round()."""
        mid = 0.5 * (min_lat + max_lat) * DEG2RAD
        width = (max_lon - min_lon) * DEG2RAD * EARTH_RADIUS_KM * math.cos(mid)
        height = (max_lat - min_lat) * DEG2RAD * EARTH_RADIUS_KM
        return GridDefinition(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth,
                              max(2, int(round(width / horizontal_km)) + 1),
                              max(2, int(round(height / horizontal_km)) + 1),
                              max(2, int(round((max_depth - min_depth) / vertical_km)) + 1))

    def to_ts(self) -> TsGrid:
        return TsGrid(self.nx, self.ny, self.nz, self.min_lon, self.max_lon, self.min_lat, self.max_lat,
                      self.min_depth, self.max_depth)

    def min_spacing_km(self) -> float:
        """SphericalGrid.MinSpacingKm (SphericalGrid.cs:159-164)."""
        rmin = EARTH_RADIUS_KM - self.depth()[-1]
        cosmin = float(np.min(trig("cos", self.lat() * DEG2RAD)))
        return min(self.ddepth, min(rmin * (self.dlat * DEG2RAD), rmin * cosmin * (self.dlon * DEG2RAD)))

    def fractional(self, lon, lat, depth):
        """SphericalGrid.Fractional (SphericalGrid.cs:179-184), vectorised; no longitude unwrap needed
        inside ±180° of the centre."""
        return ((np.asarray(lon) - self.min_lon) / self.dlon, (np.asarray(lat) - self.min_lat) / self.dlat,
                (np.asarray(depth) - self.min_depth) / self.ddepth)
