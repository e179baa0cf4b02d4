"""Trade-off curves on the device (TravelTimeTomography.TradeOff, Tomography/TravelTimeTomography.cs:
566-622): the first linearised step solved for every (damping, smoothing) pair with the forward done
once - tables and rays on the device, then per pair the system assembled in HBM, LSQR (damp 0), the
data misfit from the device product G x, and the model's norms.

Differences from the reference, all of scope: a plain node grid (NodeUpdate is the identity column
map, TravelTimeTomography.cs:329-336; no parameter lattice or adaptive mesh), residuals from the
device TCalc, and the mesh update (UpdateMesh) omitted. The pairs run one after another, as in the
reference; each is one device assembly and one device LSQR solve.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import eikonal as E
from . import iteration as It
from . import lsqr as Q
from . import rays as Ry
from . import system as Sy
from .grid import DEG2RAD, EARTH_RADIUS_KM, GridDefinition
from .workloads import Workload


@dataclass
class TradeOffPoint:
    """TradeOffPoint (TravelTimeTomography.cs:557-558)."""
    damping: float
    smoothing: float
    residual_norm: float
    model_norm: float
    roughness: float
    data_variance: float
    model_variance: float
    vp: np.ndarray
    vs: np.ndarray


def axis_weights(g: GridDefinition, vertical: float, in_nodes: bool):
    """AxisWeights (Observations.cs:390-397): spacings at (Nx/2, Ny/2, 0) made dimensionless by their
    geometric horizontal mean, the vertical one times the vertical weight; in node units (1, 1, w)."""
    if in_nodes:
        return 1.0, 1.0, vertical
    r = EARTH_RADIUS_KM - g.min_depth
    lat = g.min_lat + (g.ny // 2) * g.dlat
    hx = max(1e-9, r * math.cos(lat * DEG2RAD) * (g.dlon * DEG2RAD))
    hy = r * (g.dlat * DEG2RAD)
    hz = g.ddepth
    h = math.sqrt(hx * hy)
    return (h / hx) ** 2, (h / hy) ** 2, vertical * (h / hz) ** 2


def roughness(g: GridDefinition, x: np.ndarray, vertical: float, in_nodes: bool = False) -> float:
    """Regularization.Roughness (Observations.cs:488-516): ||L x|| with the Laplacian of the axis
    weights over the in-range neighbours of every node."""
    wx, wy, wz = axis_weights(g, vertical, in_nodes)
    f = np.asarray(x, np.float64).reshape(g.nz, g.ny, g.nx)
    lap = np.zeros_like(f)
    diag = np.zeros_like(f)
    for ax, w in ((2, wx), (1, wy), (0, wz)):
        for step in (-1, 1):
            nb = np.roll(f, -step, axis=ax)
            ok = np.ones_like(f, bool)
            idx = [slice(None)] * 3
            idx[ax] = slice(-1, None) if step == 1 else slice(0, 1)
            ok[tuple(idx)] = False
            lap -= np.where(ok, w * nb, 0.0)
            diag += np.where(ok, w, 0.0)
    lap += diag * f
    return float(math.sqrt(float(np.sum(lap * lap))))


def trade_off_point(g: GridDefinition, settings: Sy.SystemSettings, sysd: Sy.System, x: np.ndarray, damping: float,
                    smoothing: float, sigma: float, sP: np.ndarray, sS: np.ndarray, kRef: np.ndarray) -> TradeOffPoint:
    """One point of TradeOff from the pair's system and LSQR solution (TravelTimeTomography.cs:598-617):
    the data misfit ||W (G x - b)|| over the data rows from the device product, the data variance with
    the arrivals' sigma (one value here), the node updates of the velocity blocks, their norm, the
    roughness of the P update, the variance of the model and the updated velocities."""
    n = g.count
    lay = sysd.layout
    nd = lay["data_rows"]
    r = sysd.matrix.multiply(x)[:nd] - sysd.rhs[:nd]
    rr = float(np.dot(r, r))
    dv = rr * sigma * sigma
    mP = x[lay["offP"]:lay["offP"] + n] if lay["offP"] >= 0 else np.zeros(n)
    mS = x[lay["offS"]:lay["offS"] + n] if lay["offS"] >= 0 else np.zeros(n)
    model = np.concatenate([mP, mS])
    mean = model.mean()
    vp = 1.0 / (sP * (1 + mP))
    vs = vp / (kRef * (1 + mS)) if settings.vpvs else 1.0 / (sS * (1 + mS))
    return TradeOffPoint(damping, smoothing, math.sqrt(rr), float(np.linalg.norm(model)),
                         roughness(g, mP, settings.vertical_weight, settings.smoothing_scale == "nodes"),
                         dv / nd if nd > 0 else 0.0, float(np.mean((model - mean) ** 2)), vp, vs)


def trade_off(dev: E.Device, wl: Workload, t_obs: np.ndarray, pairs, settings: Sy.SystemSettings | None = None,
              lsqr_iterations: int = 100, lsqr_tolerance: float = 1e-6, sigma: float = 0.05,
              lsqr_mode: str = "fast_fp64", eik_mode: str = "fast_fp32") -> list[TradeOffPoint]:
    """TradeOff for the (damping, smoothing) pairs from the workload's 1-D starting model: LSQR with
    damp 0, LsqrIterations, atol = btol = LsqrTolerance, conlim 1e8 per pair (:597)."""
    settings = settings or Sy.SystemSettings(vpvs=True)
    g = wl.grid
    n = g.count
    sP, sS = It.starting_model(wl)
    seeds = [E.point_source(g, sP if ph == 0 else sS, *st) for st in wl.stations for ph in (0, 1)]
    mof = np.tile(np.array([0, 1], np.int32), len(wl.stations))
    ts = E.TableSet(dev, g, [sP, sS], seeds, mof)
    rc, why = ts.solve(eik_mode)
    if rc != 0:
        raise RuntimeError(why)
    station, event, phase = It.all_arrivals(wl)
    batch = Ry.RayBatch(ts, g, np.repeat(wl.stations, 2, axis=0), 2 * station + phase, wl.events[event])
    ts.release()
    used = np.flatnonzero((batch.counts >= 0) & np.isfinite(t_obs)).astype(np.int32)
    residual = t_obs[used] - batch.tcalc_grad[used, 0]
    sig = np.full(used.size, sigma)
    kRef = sS / sP
    points = []
    for d, sm in pairs:
        sysd = Sy.build_system(dev, g, settings, batch, used, phase[used], event[used], station[used], residual, sig,
                               np.zeros(len(wl.events), np.uint8), np.zeros(len(wl.events), np.uint8), len(wl.stations),
                               sP, sS, sP, sS, kRef, damping=d, smoothing=sm, keep_fp64=True)
        rc, why, res = Q.solve_many(sysd.matrix, sysd.rhs[None, :], [0.0], lsqr_iterations, lsqr_tolerance,
                                    lsqr_tolerance, 1e8, mode=lsqr_mode)
        if rc != 0:
            raise RuntimeError(why)
        points.append(trade_off_point(g, settings, sysd, res[0].solution, d, sm, sigma, sP, sS, kRef))
        sysd.matrix.release()
    batch.release()
    return points
