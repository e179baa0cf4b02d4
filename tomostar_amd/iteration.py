"""One tomographic iteration on the device, end to end: the caller side of the hot path
(TravelTimeTomography's iteration, Tomography/TravelTimeTomography.cs:405-477 and :860-1034).

tables (every station x phase, one resident set) -> rays to every event -> the linear system
assembled in HBM -> LSQR. Observed times come from the tables of the true model (the workload's
velocities), calculated times from a 1-D starting model, so the system has the structure and the
right-hand side of a real first iteration.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import eikonal as E
from . import lsqr as Q
from . import rays as Ry
from . import system as Sy
from .workloads import Workload, station_seed_sets


def all_arrivals(wl: Workload):
    """Every event x station x phase, event-major: (station, event, phase) arrays in arrival order."""
    ns, ne = len(wl.stations), len(wl.events)
    event = np.repeat(np.arange(ne), 2 * ns).astype(np.int32)
    station = np.tile(np.repeat(np.arange(ns), 2), ne).astype(np.int32)
    phase = np.tile(np.array([0, 1], np.int32), ne * ns)
    return station, event, phase


def starting_model(wl: Workload):
    """The 1-D start: the horizontal mean of the workload's P velocity at every depth; Vs = Vp / 1.73."""
    vp1 = wl.vp.reshape(wl.grid.shape).mean(axis=(1, 2))
    vp = np.ascontiguousarray(np.broadcast_to(vp1[:, None, None], wl.grid.shape))
    return 1.0 / vp.ravel(), 1.73 / vp.ravel()


@dataclass
class IterationResult:
    stage_seconds: dict = field(default_factory=dict)
    rows: int = 0
    columns: int = 0
    nnz: int = 0
    data_rows: int = 0
    x: np.ndarray | None = None
    iterations: int = 0
    stop_code: int = 0
    layout: dict = field(default_factory=dict)
    rhs: np.ndarray | None = None
    system: Sy.System | None = None
    batch: Ry.RayBatch | None = None
    arrivals: tuple = ()
    used: np.ndarray | None = None


def observed_times(dev: E.Device, wl: Workload, eik_mode: str = "fast_fp64") -> np.ndarray:
    """Travel times of every arrival in the true model (TCalc on the true tables; +inf unreachable)."""
    seeds, mof = station_seed_sets(wl, E.point_source)
    ts = E.TableSet(dev, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve(eik_mode)
    if rc != 0:
        raise RuntimeError(why)
    station, event, phase = all_arrivals(wl)
    b = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), 2 * station + phase, wl.events[event])
    t = np.where(b.counts >= 0, b.tcalc_grad[:, 0], np.inf)
    b.release()
    ts.release()
    return t


def run(dev: E.Device, wl: Workload, t_obs: np.ndarray, settings: Sy.SystemSettings | None = None,
        damping: float = 1.0, smoothing: float = 5.0, lsqr_iterations: int = 100, lsqr_mode: str = "fast_fp64",
        eik_mode: str = "fast_fp32", sigma: float = 0.05, keep_fp64: bool = True) -> IterationResult:
    """One iteration from the 1-D starting model; LSQR with damp 0 and a fixed iteration budget
    (atol = btol = conlim = 0). Stage wall times in stage_seconds."""
    settings = settings or Sy.SystemSettings(vpvs=True)
    out = IterationResult()
    g = wl.grid
    sP, sS = starting_model(wl)
    t0 = time.perf_counter()
    seeds = [E.point_source(g, sP if ph == 0 else sS, *st) for st in wl.stations for ph in (0, 1)]
    mof = np.tile(np.array([0, 1], np.int32), len(wl.stations))
    ts = E.TableSet(dev, g, [sP, sS], seeds, mof)
    rc, why = ts.solve(eik_mode)
    if rc != 0:
        raise RuntimeError(why)
    out.stage_seconds["tables"] = time.perf_counter() - t0
    station, event, phase = all_arrivals(wl)
    t0 = time.perf_counter()
    batch = Ry.RayBatch(ts, g, np.repeat(wl.stations, 2, axis=0), 2 * station + phase, wl.events[event])
    out.stage_seconds["rays"] = time.perf_counter() - t0
    used = np.flatnonzero((batch.counts >= 0) & np.isfinite(t_obs)).astype(np.int32)
    residual = t_obs[used] - batch.tcalc_grad[used, 0]
    t0 = time.perf_counter()
    kRef = sS / sP
    sysd = Sy.build_system(dev, g, settings, batch, used, phase[used], event[used], station[used], residual,
                           np.full(used.size, sigma), np.zeros(len(wl.events), np.uint8),
                           np.zeros(len(wl.events), np.uint8), len(wl.stations), sP, sS, sP, sS, kRef,
                           damping=damping, smoothing=smoothing, keep_fp64=keep_fp64)
    out.stage_seconds["system"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc, why, res = Q.solve_many(sysd.matrix, sysd.rhs[None, :], [0.0], lsqr_iterations, 0.0, 0.0, 0.0,
                                mode=lsqr_mode)
    if rc != 0:
        raise RuntimeError(why)
    out.stage_seconds["lsqr"] = time.perf_counter() - t0
    ts.release()
    out.x, out.iterations, out.stop_code = res[0].solution, res[0].iterations, res[0].stop_code
    out.rows, out.columns, out.nnz = sysd.matrix.rows, sysd.matrix.cols, sysd.matrix.nnz
    out.layout, out.rhs, out.data_rows = sysd.layout, sysd.rhs, sysd.layout["data_rows"]
    out.system, out.batch, out.arrivals, out.used = sysd, batch, (station, event, phase), used
    return out
