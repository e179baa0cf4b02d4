"""Seeded synthetic inputs of the benchmark configurations.

Velocity models, stations and events are generated here with numpy's PCG64 generator (the
reference's own generators use .NET System.Random, which is not reproducible outside .NET). The
anomaly patterns restate the reference's resolution-test patterns (Tomography/ResolutionTests.cs:
22-94): the sine-product checkerboard, Gaussian spikes and tabular bodies with super-Gaussian edges.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .grid import DEG2RAD, EARTH_RADIUS_KM, GridDefinition


def surface_distance_km(lon1, lat1, lon2, lat2):
    """GeoMath.SurfaceDistanceKm (Geo/GeoMath.cs:83-95), haversine, vectorised."""
    p1, p2 = lat1 * DEG2RAD, lat2 * DEG2RAD
    dp, dl = p2 - p1, (lon2 - lon1) * DEG2RAD
    a = np.sin(dp / 2) ** 2 + np.cos(p1) * np.cos(p2) * np.sin(dl / 2) ** 2
    return 2 * np.arcsin(np.minimum(1, np.sqrt(a))) * EARTH_RADIUS_KM


def checkerboard(grid: GridDefinition, cell_h_km: float, cell_v_km: float, amplitude: float) -> np.ndarray:
    """Sine-product checkerboard (ResolutionTests.cs:71-80) as a fraction, shape (nz, ny, nx)."""
    lon, lat, dep = grid.lon(), grid.lat(), grid.depth()
    x = surface_distance_km(grid.min_lon, lat[:, None], lon[None, :], lat[:, None])        # (ny, nx)
    y = surface_distance_km(lon[None, :], grid.min_lat, lon[None, :], lat[:, None])         # (ny, nx)
    z = dep - grid.min_depth
    h = np.sin(math.pi * x / cell_h_km) * np.sin(math.pi * y / cell_h_km)
    v = np.sin(math.pi * (z + 0.5 * cell_v_km) / cell_v_km)
    return amplitude * v[:, None, None] * h[None, :, :]


def _local_km(grid: GridDefinition, lon0: float, lat0: float):
    """East, north (km) of every node column from (lon0, lat0), the flat-Earth offsets of
    TabularBody.At (ResolutionTests.cs:26-28)."""
    east = (grid.lon()[None, :] - lon0) * DEG2RAD * EARTH_RADIUS_KM * math.cos(lat0 * DEG2RAD)
    north = (grid.lat()[:, None] - lat0) * DEG2RAD * EARTH_RADIUS_KM
    return np.broadcast_to(east, (grid.ny, grid.nx)), np.broadcast_to(north, (grid.ny, grid.nx))


def ellipsoid_body(grid: GridDefinition, lon, lat, depth, a_km, b_km, c_km, amplitude) -> np.ndarray:
    """A smooth ellipsoidal anomaly: super-Gaussian edge of order 6 (ResolutionTests.cs:39-40), 1 inside,
    half at the nominal semi-axes, 0 well outside."""
    east, north = _local_km(grid, lon, lat)
    down = grid.depth() - depth
    rho2 = (east / a_km) ** 2 + (north / b_km) ** 2
    r = np.sqrt(rho2[None, :, :] + (down[:, None, None] / c_km) ** 2)
    return amplitude * np.exp(-math.log(2) * r ** 6)


def spike(grid: GridDefinition, lon, lat, depth, radius_km, amplitude) -> np.ndarray:
    """Gaussian spike (ResolutionTests.cs:85-89), with flat-Earth distances."""
    east, north = _local_km(grid, lon, lat)
    down = grid.depth() - depth
    d2 = east[None] ** 2 + north[None] ** 2 + down[:, None, None] ** 2
    return amplitude * np.exp(-d2 / (2 * radius_km ** 2))


def gaussian_random_field(grid: GridDefinition, rng: np.random.Generator, sigma: float, corr_km: float,
                          spacing_km: float) -> np.ndarray:
    """White noise filtered to a Gaussian covariance exp(-r^2/L^2) (filter width L/2), scaled to std sigma."""
    from scipy.ndimage import gaussian_filter

    w = rng.standard_normal(grid.shape)
    f = gaussian_filter(w, sigma=0.5 * corr_km / spacing_km, mode="reflect")
    f *= sigma / f.std()
    return f


@dataclass
class Workload:
    """Grid, P/S slowness and the stations/events of one configuration."""

    name: str
    grid: GridDefinition
    vp: np.ndarray  # (nz, ny, nx)
    vs: np.ndarray
    stations: np.ndarray  # (S, 3) lon, lat, depth
    events: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))

    @property
    def slowness_p(self) -> np.ndarray:
        return (1.0 / self.vp).ravel()

    @property
    def slowness_s(self) -> np.ndarray:
        return (1.0 / self.vs).ravel()


def _uniform_points(rng, grid, n, dmin, dmax, margin_deg=0.0):
    lon = rng.uniform(grid.min_lon + margin_deg, grid.max_lon - margin_deg, n)
    lat = rng.uniform(grid.min_lat + margin_deg, grid.max_lat - margin_deg, n)
    dep = rng.uniform(dmin, dmax, n) if dmax > dmin else np.full(n, dmin)
    return np.stack([lon, lat, dep], axis=1)


def one_iteration_case(seed: int = 0, stations: int = 25, events: int = 150) -> Workload:
    """One tomographic iteration on the 61x61x31 1-km grid of NumericsTests.cs:91: Vp = 5.0 + 0.04 z,
    Vs = Vp/1.73, +-5% sine-product checkerboard of 10-km cells; stations at z = 0, events at 3-15 km."""
    g = GridDefinition.from_spacing(13.0, 13.74, 43.0, 43.54, 0, 30, 1, 1)
    rng = np.random.default_rng(seed)
    z = g.depth()[:, None, None]
    vp = (5.0 + 0.04 * z) * (1 + checkerboard(g, 10.0, 10.0, 0.05))
    vp = np.ascontiguousarray(np.broadcast_to(vp, g.shape), dtype=np.float64)
    st = _uniform_points(rng, g, stations, 0.0, 0.0)
    ev = _uniform_points(rng, g, events, 3.0, 15.0)
    return Workload("61x61x31 at 1 km, 25 stations x {P,S}", g, vp, vp / 1.73, st, ev)


def box_grid(min_lon: float, min_lat: float, width_km: float, depth_km: float, spacing_km: float) -> GridDefinition:
    """A square box `width_km` on a side and `depth_km` deep with nodes every `spacing_km`."""
    dlat = width_km / (EARTH_RADIUS_KM * DEG2RAD)
    max_lat = min_lat + dlat
    mid = 0.5 * (min_lat + max_lat) * DEG2RAD
    dlon = width_km / (EARTH_RADIUS_KM * DEG2RAD * math.cos(mid))
    n = int(round(width_km / spacing_km)) + 1
    nz = int(round(depth_km / spacing_km)) + 1
    return GridDefinition(min_lon, min_lon + dlon, min_lat, max_lat, 0.0, depth_km, n, n, nz)


def batched_eikonal_case(seed: int = 1, stations: int = 200, width_km: float = 100.0, depth_km: float = 50.0,
                         spacing_km: float = 0.5) -> Workload:
    """Batched eikonal: 201x201x101 at 0.5 km (100 x 100 x 50 km). Vp = 5.0 + 0.04 z times
    (1 + Gaussian random field, std 4 %, correlation length 5 km + three smooth ellipsoidal bodies of
    +-10 %); Vs = Vp/1.73; stations uniform at z = 0 (one P and one S field each)."""
    g = box_grid(13.0, 42.5, width_km, depth_km, spacing_km)
    rng = np.random.default_rng(seed)
    z = g.depth()[:, None, None]
    pert = gaussian_random_field(g, rng, 0.04, 5.0, spacing_km)
    for amp in (0.10, -0.10, 0.10):
        lon = rng.uniform(g.min_lon + 0.2 * (g.max_lon - g.min_lon), g.max_lon - 0.2 * (g.max_lon - g.min_lon))
        lat = rng.uniform(g.min_lat + 0.2 * (g.max_lat - g.min_lat), g.max_lat - 0.2 * (g.max_lat - g.min_lat))
        dep = rng.uniform(0.2 * depth_km, 0.7 * depth_km)
        a, b, c = rng.uniform(0.08, 0.2, 3) * np.array([width_km, width_km, depth_km])
        pert += ellipsoid_body(g, lon, lat, dep, a, b, c, amp)
    vp = np.ascontiguousarray((5.0 + 0.04 * z) * (1 + pert), dtype=np.float64)
    st = _uniform_points(rng, g, stations, 0.0, 0.0)
    return Workload(f"{g.nx}x{g.ny}x{g.nz} at {spacing_km} km, {stations} stations x {{P,S}}", g, vp, vp / 1.73, st)


def station_seed_sets(wl: Workload, point_source, phases=("P", "S")):
    """Seed sets and model indices of every station x phase field: fields ordered station-major,
    phase 0 = P (model 0), 1 = S (model 1)."""
    models = {"P": wl.slowness_p, "S": wl.slowness_s}
    seeds, mof = [], []
    for st in wl.stations:
        for ph in phases:
            seeds.append(point_source(wl.grid, models[ph], st[0], st[1], st[2]))
            mof.append(0 if ph == "P" else 1)
    return seeds, np.asarray(mof, np.int32)


def scale_arrivals(wl: Workload, events: int = 20000, rays: int = 1_000_000, seed: int = 2):
    """The LSQR-at-scale arrivals: `events` hypocentres uniform in the box at 3-15 km, and the first `rays`
    station x event x phase triples of a seeded shuffle. Returns (event positions (E, 3), station,
    event and phase (0 P, 1 S) of each arrival, in arrival order)."""
    rng = np.random.default_rng(seed)
    ev = _uniform_points(rng, wl.grid, events, 3.0, 15.0)
    ns = len(wl.stations)
    total = ns * events * 2
    pick = np.sort(rng.permutation(total)[:min(rays, total)])
    # triple index = (event * stations + station) * 2 + phase: arrivals grouped by event, as picks are
    phase = (pick % 2).astype(np.int32)
    station = ((pick // 2) % ns).astype(np.int32)
    event = (pick // (2 * ns)).astype(np.int32)
    return ev, station, event, phase


@dataclass
class UnevenProblem:
    grid: GridDefinition
    slowness: np.ndarray  # P slowness, nodes of grid
    source: tuple         # lon, lat, depth
    pattern: str
    spacing_km: float


def uneven_batch_case(seed: int = 3, problems: int = 64, min_h: int = 41, max_h: int = 161, min_v: int = 21,
                      max_v: int = 81):
    """The uneven batch: per problem t ~ U[0, 1] sets Nx = Ny = round(min_h + (max_h - min_h) t) and
    Nz = round(min_v + (max_v - min_v) t) (the axes are drawn jointly), spacing U[0.5, 2] km; Vp =
    5.0 + 0.04 z times (1 + one of a checkerboard, a Gaussian spike or a tabular body of +-U[5, 10] %,
    ResolutionTests.cs:22-94 patterns); the source is uniform in the interior, at least 3 cells from
    every face. One P field per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(problems):
        t = rng.random()
        nh = int(round(min_h + (max_h - min_h) * t))
        nv = int(round(min_v + (max_v - min_v) * t))
        h = rng.uniform(0.5, 2.0)
        g0 = box_grid(13.0 + 0.5 * rng.random(), 42.5 + 0.5 * rng.random(), (nh - 1) * h, (nv - 1) * h, h)
        g = GridDefinition(g0.min_lon, g0.max_lon, g0.min_lat, g0.max_lat, g0.min_depth, g0.max_depth, nh, nh, nv)
        z = g.depth()[:, None, None]
        amp = rng.uniform(0.05, 0.10) * rng.choice([-1.0, 1.0])
        kind = ["checkerboard", "spike", "tabular"][q % 3]
        if kind == "checkerboard":
            cell = rng.uniform(5.0, 15.0) * max(1.0, h)
            pert = checkerboard(g, cell, cell, amp)
        elif kind == "spike":
            pert = spike(g, rng.uniform(g.min_lon, g.max_lon), rng.uniform(g.min_lat, g.max_lat),
                         rng.uniform(0.2, 0.8) * g.max_depth, 0.1 * (nh - 1) * h, amp)
        else:
            pert = ellipsoid_body(g, rng.uniform(g.min_lon, g.max_lon), rng.uniform(g.min_lat, g.max_lat),
                                  rng.uniform(0.3, 0.7) * g.max_depth, 0.25 * (nh - 1) * h, 0.15 * (nh - 1) * h,
                                  0.05 * (nv - 1) * h, amp)
        vp = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + pert), g.shape), dtype=np.float64)
        lon = g.min_lon + (g.max_lon - g.min_lon) * rng.uniform(3 / (nh - 1), 1 - 3 / (nh - 1))
        lat = g.min_lat + (g.max_lat - g.min_lat) * rng.uniform(3 / (nh - 1), 1 - 3 / (nh - 1))
        dep = g.min_depth + (g.max_depth - g.min_depth) * rng.uniform(3 / (nv - 1), 1 - 3 / (nv - 1))
        out.append(UnevenProblem(g, (1.0 / vp).ravel(), (lon, lat, dep), kind, h))
    return out
