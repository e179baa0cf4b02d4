"""Attenuation (Q) tomography from t* on the device (QTomography, Attenuation/QTomography.cs:126-554).

Along a ray t* = sum L s q with q = 1/Q on the nodes (:123-131): the rows of the velocity problem's
rays with other values. The device holds the rays (a ``RayBatch``); per least-squares solve it
assembles the system in HBM (``ts_qsystem_build``: the t* rows at the current q, the Gauss-Newton
right-hand side, the sensitivity-scaled damping and smoothing, the station terms), forms the
correction right-hand side with the device product, runs LSQR, and predicts every t* again
(``ts_q_predict``). The host keeps the reference's control flow: the reference Q from the data
(Prepare, :199-300), the Gauss-Newton iterations on ln(q/q0) (Solve, :252-289), the outlier rule
(RejectOutliers, :412-432), Run (:434-505) and TradeOff (:520-554).

Scope: a plain node grid (no adaptive mesh or parameter lattice), Laplacian smoothing or unweighted
first differences (no reweighting passes). Rays traced on an extended forward domain carry their
OutsideTime (the background time of the pieces outside the inversion grid), which enters Predict and the
Q0 estimate's path time on the device. The formal
resolution of the last system (:480-497) runs on the device when the settings ask for it.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import lsqr as Q
from ._lib import TsGrid, check, ptr, reason_buffer
from .eikonal import Device
from .grid import GridDefinition
from .rays import RayBatch
from .system import _AdoptedMatrix
from .tradeoff import roughness


class QSystemOpts(C.Structure):
    _fields_ = [("station_terms", C.c_int32), ("gradient_smoothing", C.c_int32), ("smoothing_in_nodes", C.c_int32),
                ("nstations", C.c_int32), ("fixed_scales", C.c_int32), ("damping", C.c_double),
                ("smoothing", C.c_double), ("vertical_weight", C.c_double), ("damping_station", C.c_double),
                ("q0", C.c_double)]


def _declare():
    L = _lib.load()
    if getattr(L, "_q_declared", False):
        return L
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.ts_q_predict.restype = i
    L.ts_q_predict.argtypes = [vp, i, vp, vp, vp, d, vp, vp, C.c_char_p, i]
    L.ts_qsystem_build.restype = i
    L.ts_qsystem_build.argtypes = [vp, C.POINTER(TsGrid), C.POINTER(QSystemOpts), vp, i, vp, vp, vp, vp, vp, vp, vp,
                                   i, vp, C.POINTER(vp), vp, C.c_char_p, i]
    L._q_declared = True
    return L


@dataclass
class QTomographySettings:
    """QTomographySettings (QTomography.cs:12-90): the fields of a plain node grid, reference defaults."""

    q0: float = 300.0                  # Q0: starting and reference Q, replaced by the data's when estimate_q0
    estimate_q0: bool = True
    damping: float = 2.0               # relative to the model block's sensitivity scale
    smoothing: float = 10.0
    vertical_weight: float = 0.1
    smoothing_scale: str = "km"        # or "nodes"
    smoothing_method: str = "laplacian"  # or "gradient" (first differences, no reweighting)
    station_terms: bool = True
    damping_station: float = 3.0       # relative to the station block's scale
    lsqr_iterations: int = 400
    log_parameterisation: bool = True  # unknowns ln(q/q0) by Gauss-Newton, else q/q0 - 1
    nonlinear_iterations: int = 20
    q_min: float = 10.0
    q_max: float = 5000.0
    outlier_mads: float = 4.0
    formal_resolution: bool = False    # FormalResolutionSettings.Enabled, on the last system
    formal_max_unknowns: int = 20000   # FormalResolutionSettings.MaxUnknowns


@dataclass
class QData:
    """The t* observations of one phase (QTomography.FromCatalogue, :133-178): per arrival its station
    (index into the stations), t* (s) and sigma (s; FromCatalogue keeps at least 0.002)."""

    station: np.ndarray
    tstar: np.ndarray
    sigma: np.ndarray
    nstations: int


@dataclass
class QSystem:
    matrix: _AdoptedMatrix
    rhs: np.ndarray
    layout: dict
    scales: tuple


@dataclass
class QPass:
    """One least-squares solve (SolveOnce's pass, :296-405)."""

    solution: np.ndarray
    q: np.ndarray
    terms: np.ndarray
    iterations: int
    stop_code: int
    rows: int
    columns: int
    nnz: int
    seconds: dict = field(default_factory=dict)


@dataclass
class QTomographyResult:
    """QTomographyResult (QTomography.cs:92-121), the fields of a plain node grid."""

    q: np.ndarray            # Q on the nodes
    dws: np.ndarray
    rms_before: float
    rms_after: float
    reference_q: float
    station_terms: np.ndarray
    rejected: int
    lsqr_iterations: int
    stop_code: int
    gauss_newton: tuple      # (iterations, RMS change of ln q at the last)
    residual: np.ndarray     # per arrival (NaN: no ray)
    seconds: dict
    formal: object = None    # FormalResolutionMaps of the unknowns on the nodes, when asked for and possible
    formal_reason: str = ""  # why it is missing when asked for


@dataclass
class QTradeOffPoint:
    """QTradeOffPoint (QTomography.cs:507-509)."""

    damping: float
    smoothing: float
    residual_norm: float
    model_norm: float
    roughness: float
    data_variance: float
    model_variance: float
    q: np.ndarray


def _seq_sum(a) -> float:
    a = np.asarray(a, np.float64)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


def _seq_mean(a) -> float:
    a = np.asarray(a, np.float64)
    return _seq_sum(a) / a.size


def q_predict(rays: RayBatch, ray_of_row, slowness, q, q0: float):
    """Predict (:307-312: q0 * OutsideTime + sum L s q) and the path time OutsideTime + sum L s (:231-233)
    of the given rows on the device; q None: q0 at every node. Returns (pred, tt)."""
    L = _declare()
    rows = np.ascontiguousarray(ray_of_row, np.int32)
    pred, tt = np.zeros(max(1, rows.size)), np.zeros(max(1, rows.size))
    s = np.ascontiguousarray(slowness, np.float64)
    qa = None if q is None else np.ascontiguousarray(q, np.float64)
    r = reason_buffer()
    rc = L.ts_q_predict(rays.handle, rows.size, ptr(rows), ptr(s), None if qa is None else ptr(qa), float(q0),
                        ptr(pred), ptr(tt), r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    return pred[:rows.size], tt[:rows.size]


def build_qsystem(device: Device, grid: GridDefinition, settings: QTomographySettings, rays: RayBatch, ray_of_row,
                  station, tstar, sigma, nstations: int, slowness, q0: float, damping: float, smoothing: float,
                  q_old=None, m_old=None, scales=None, keep_fp64: bool = True) -> QSystem:
    """SolveOnce's system (:318-392) assembled in HBM for the used rows (arrival ray_of_row[r] with its
    station, t* and sigma). scales None: the sensitivity scales of these rows, else fixed."""
    L = _declare()
    rows = np.ascontiguousarray(ray_of_row, np.int32)
    st = np.ascontiguousarray(station, np.int32)
    ts = np.ascontiguousarray(tstar, np.float64)
    sg = np.ascontiguousarray(sigma, np.float64)
    s = np.ascontiguousarray(slowness, np.float64).ravel()
    if s.size != grid.count:
        raise ValueError("slowness size differs from the grid's node count")
    qo = None if q_old is None else np.ascontiguousarray(q_old, np.float64)
    mo = None if m_old is None else np.ascontiguousarray(m_old, np.float64)
    o = QSystemOpts(station_terms=int(settings.station_terms),
                    gradient_smoothing=int(settings.smoothing_method == "gradient"),
                    smoothing_in_nodes=int(settings.smoothing_scale == "nodes"), nstations=int(nstations),
                    fixed_scales=int(scales is not None), damping=damping, smoothing=smoothing,
                    vertical_weight=settings.vertical_weight, damping_station=settings.damping_station, q0=q0)
    sc = np.array(scales if scales is not None else (1.0, 1.0), np.float64)
    layout = np.zeros(4, np.int64)
    h = C.c_void_p()
    r = reason_buffer()
    g = grid.to_ts()
    rc = L.ts_qsystem_build(device.handle, C.byref(g), C.byref(o), rays.handle, rows.size, ptr(rows), ptr(st),
                            ptr(ts), ptr(sg), ptr(s), None if qo is None else ptr(qo), None if mo is None else ptr(mo),
                            1 if keep_fp64 else 0, ptr(sc), C.byref(h), ptr(layout), r, len(r))
    if rc != 0:
        raise _lib.HipError(rc, r.value.decode(errors="replace"))
    mat = _AdoptedMatrix(device, h)
    mat.has_fp64 = bool(keep_fp64)
    rhs = np.zeros(max(1, mat.rows))
    check(_lib.load().ts_matrix_read(h, None, None, None, ptr(rhs)))
    names = ["off_station", "columns", "rows", "data_rows"]
    return QSystem(mat, rhs[:mat.rows], dict(zip(names, (int(x) for x in layout))), (float(sc[0]), float(sc[1])))


class QTomography:
    """QTomography (QTomography.cs:126) over resident rays; lsqr_mode picks the device LSQR form
    (plain_fp64 is LsqrOpenCl's arithmetic; fast_mixed keeps FP32 matrix values and FP64 vectors)."""

    def __init__(self, device: Device, grid: GridDefinition, settings: QTomographySettings | None = None,
                 lsqr_mode: str = "plain_fp64"):
        if settings is not None and settings.smoothing_method not in ("laplacian", "gradient"):
            raise ValueError("smoothing is the Laplacian or unweighted first differences")
        self.device, self.grid = device, grid
        self.s = settings or QTomographySettings()
        self.lsqr_mode = lsqr_mode
        self.seconds = {"system": 0.0, "lsqr": 0.0, "predict": 0.0}

    # ---- Prepare (:199-300)
    def prepare(self, rays: RayBatch, data: QData, slowness) -> None:
        import time
        s = self.s
        self.rays, self.data = rays, data
        self.slow = np.ascontiguousarray(slowness, np.float64).ravel()
        self.rows = np.flatnonzero(rays.counts >= 0).astype(np.int32)  # RayKernels' rows, arrival order
        q0 = 1 / max(1.0, s.q0)
        t0 = time.perf_counter()
        pred, tt = q_predict(rays, self.rows, self.slow, None, q0)
        if s.estimate_q0 and self.rows.size:
            ts, sg = data.tstar[self.rows], data.sigma[self.rows]
            ok = (tt > 0) & np.isfinite(ts)
            ratios = np.sort(ts[ok] / tt[ok])
            w = 1 / (sg[ok] * sg[ok])
            num = _seq_sum(w * ts[ok] * tt[ok])
            den = _seq_sum(w * tt[ok] * tt[ok])
            est = 0.0
            if s.station_terms:
                if den > 0 and num > 0:
                    est = num / den  # the uniform q fitting all t*
            elif ratios.size:
                m = ratios.size // 2  # the median of the path averages t* / T
                est = ratios[m] if ratios.size % 2 == 1 else 0.5 * (ratios[m - 1] + ratios[m])
            if est > 0:
                q0 = min(max(float(est), 1 / s.q_max), 1 / s.q_min)
            if q0 != 1 / max(1.0, s.q0):
                pred, _ = q_predict(rays, self.rows, self.slow, None, q0)
        self.seconds["predict"] += time.perf_counter() - t0
        self.q0 = q0
        na = data.tstar.size
        self.rejected = np.ones(na, bool)
        self.residual = np.full(na, np.nan)
        self.residual[self.rows] = data.tstar[self.rows] - pred
        self.rejected[self.rows] = False
        self.rms_before = self.rms()

    def rms(self) -> float:
        used = ~self.rejected & ~np.isnan(self.residual)
        return math.sqrt(_seq_mean(self.residual[used] ** 2)) if used.any() else 0.0

    # ---- SolveOnce (:296-405)
    def solve_once(self, damping, smoothing, m_old, fixed_scales, x_old) -> tuple[QPass, np.ndarray]:
        import time
        s, n, d = self.s, self.grid.count, self.data
        q0 = self.q0
        lo, hi = math.log(1 / (s.q_max * q0)), math.log(1 / (s.q_min * q0))
        q_old = None if m_old is None else q0 * np.exp(np.clip(m_old, lo, hi))
        used = self.rows[~self.rejected[self.rows]]
        t0 = time.perf_counter()
        sysd = build_qsystem(self.device, self.grid, s, self.rays, used, d.station[used], d.tstar[used], d.sigma[used],
                             d.nstations, self.slow, q0, damping, smoothing, q_old, m_old, fixed_scales)
        self.scales = sysd.scales
        rhs = sysd.rhs
        if x_old is not None:
            rhs = rhs - sysd.matrix.multiply(x_old)  # LSQR solves for the change from x_old (:393-397)
        t1 = time.perf_counter()
        rc, why, res = Q.solve_many(sysd.matrix, rhs[None, :], [0.0], s.lsqr_iterations, 1e-6, 1e-6, 1e8,
                                    mode=self.lsqr_mode)
        if rc != 0:
            raise _lib.HipError(rc, why)
        t2 = time.perf_counter()
        x = res[0].solution
        if x_old is not None:
            x = x + x_old
        if s.log_parameterisation:
            m_nodes = np.clip(x[:n], lo, hi)
            q = q0 * np.exp(m_nodes)
        else:
            q = np.clip(q0 * (1 + x[:n]), 1 / s.q_max, 1 / s.q_min)
            m_nodes = np.log(q / q0)
        terms = x[n:n + d.nstations] if s.station_terms else np.zeros(0)
        pred, _ = q_predict(self.rays, self.rows, self.slow, q, q0)
        self.residual[self.rows] = d.tstar[self.rows] - pred - (terms[d.station[self.rows]] if terms.size else 0.0)
        t3 = time.perf_counter()
        self.seconds["system"] += t1 - t0
        self.seconds["lsqr"] += t2 - t1
        self.seconds["predict"] += t3 - t2
        p = QPass(x, q, terms, res[0].iterations, res[0].stop_code, sysd.matrix.rows, sysd.matrix.cols,
                  sysd.matrix.nnz)
        if s.formal_resolution:  # the last pass's system, for the formal resolution (:480-497)
            from .system import System
            self.last_system = (System(sysd.matrix, sysd.rhs).read(), sysd.matrix.rows, sysd.matrix.cols,
                                sysd.layout["data_rows"])
        sysd.matrix.release()
        return p, m_nodes

    # ---- Solve (:252-289): Gauss-Newton on m = ln(q/q0), the regularisation at the first iteration's scale
    def solve(self, damping, smoothing) -> QPass:
        s = self.s
        if not s.log_parameterisation:
            self.gauss_newton = (1, float("nan"))
            return self.solve_once(damping, smoothing, None, None, None)[0]
        m = scales = x = None
        it, change, out = 0, float("nan"), None
        while it < max(1, s.nonlinear_iterations):
            out, m_nodes = self.solve_once(damping, smoothing, m, scales, x)
            x = out.solution
            scales = scales or self.scales
            change = math.inf if m is None else math.sqrt(_seq_mean((m_nodes - m) ** 2))
            m = m_nodes
            it += 1
            if change < 1e-3:
                break
        self.gauss_newton = (it, change)
        return out

    # ---- RejectOutliers (:412-432): beyond OutlierMads scaled MADs of the fitted residuals
    def reject_outliers(self) -> int:
        keep = ~self.rejected
        fitted = np.sort(self.residual[keep])
        med = fitted[fitted.size // 2] if fitted.size else 0.0
        mad = 1.4826 * (np.sort(np.abs(fitted - med))[fitted.size // 2] if fitted.size else 0.0)
        sig = self.data.sigma[keep]
        limit = max(self.s.outlier_mads * mad, 2 * (_seq_mean(sig) if sig.size else 0.002))
        bad = keep & (np.abs(self.residual - med) > limit)
        self.rejected |= bad
        return int(bad.sum())

    def _dws(self) -> np.ndarray:
        """Derivative weight sum: the rays' lengths per node over the used rows (:470-472)."""
        rays = self.rays
        if rays.nodes.size != rays.nnz:
            return np.zeros(self.grid.count)
        used = self.rows[~self.rejected[self.rows]]
        take = np.concatenate([np.arange(rays.row_start[a], rays.row_start[a + 1]) for a in used]) \
            if used.size else np.zeros(0, np.int64)
        return np.bincount(rays.nodes[take], weights=rays.lengths[take].astype(np.float64),
                           minlength=self.grid.count)

    # ---- Run (:434-505)
    def run(self, rays: RayBatch, data: QData, slowness) -> QTomographyResult:
        self.prepare(rays, data, slowness)
        p = self.solve(self.s.damping, self.s.smoothing)
        rejected = self.reject_outliers()
        if rejected > 0:
            p = self.solve(self.s.damping, self.s.smoothing)
        out = QTomographyResult(1 / p.q, self._dws(), self.rms_before, self.rms(), 1 / self.q0, p.terms, rejected,
                                p.iterations, p.stop_code, self.gauss_newton, self.residual.copy(), dict(self.seconds))
        if self.s.formal_resolution:
            self._formal(out)
        return out

    def _formal(self, out: QTomographyResult) -> None:
        """The formal resolution of the last system (:480-497): no groups to separate, the node block
        placed; the reference's declines are logged, not raised."""
        from . import formal as F
        (rs, ci, va), rows, cols, data_rows = self.last_system
        places = F.add_places(F.new_places(cols), self.grid, 0, 0)
        try:
            fr = F.compute(self.device, rows, cols, rs, ci, va, data_rows, (), self.s.formal_max_unknowns, places)
        except _lib.HipError as e:
            out.formal_reason = str(e)
            return
        out.formal = F.on_nodes(fr, cols, self.grid.count, 0)
        out.seconds["formal"] = fr.seconds

    # ---- TradeOff (:520-554): every pair a complete solve on the same data
    def trade_off(self, rays: RayBatch, data: QData, slowness, pairs) -> list[QTradeOffPoint]:
        n = self.grid.count
        self.prepare(rays, data, slowness)
        self.solve(self.s.damping, self.s.smoothing)
        self.reject_outliers()
        points = []
        for d, sm in pairs:
            p = self.solve(d, sm)
            used = ~self.rejected & ~np.isnan(self.residual)
            res, sig = self.residual[used], data.sigma[used]
            model = p.solution[:n]
            mean = _seq_mean(model)
            points.append(QTradeOffPoint(
                d, sm, math.sqrt(_seq_sum((res / sig) ** 2)), float(np.linalg.norm(model)),
                roughness(self.grid, model, self.s.vertical_weight, self.s.smoothing_scale == "nodes"),
                _seq_mean(res * res) if res.size else 0.0, _seq_mean((model - mean) ** 2), 1 / p.q))
        return points


# ------------------------------------------------------------------------------ synthetic case

@dataclass
class QCase:
    """P rays of every event x station pair traced in a workload's true velocity model (the model of the
    velocity inversion, as QTomography.Prepare traces them, :214-225), a Q model and its t*."""

    rays: RayBatch
    data: QData
    slowness: np.ndarray  # on the inversion grid
    q_true: np.ndarray    # Q on the inversion grid's nodes
    grid: GridDefinition | None = None  # the inversion grid


def q_model(grid: GridDefinition, q_mean: float = 250.0) -> np.ndarray:
    """Q on the nodes: q_mean times exp(+-0.3) in a 10-km checkerboard of ln Q, and a low-Q body (Q/3)
    at the grid's centre, 8 km deep."""
    from .workloads import checkerboard, ellipsoid_body
    lnq = 0.3 * checkerboard(grid, 10.0, 10.0, 1.0).ravel()
    lon, lat = 0.5 * (grid.min_lon + grid.max_lon), 0.5 * (grid.min_lat + grid.max_lat)
    body = ellipsoid_body(grid, lon, lat, 8.0, 8.0, 8.0, 5.0, 1.0).ravel()
    lnq = lnq - math.log(3.0) * np.clip(body, 0.0, 1.0)
    return q_mean * np.exp(lnq)


def q_case(device: Device, wl, sigma: float = 0.004, seed: int = 0, q_mean: float = 250.0,
           eik_mode: str = "fast_fp64", step: int = 1) -> QCase:
    """t* = Predict(1/Q) + N(0, sigma) of the P rays of every event x station pair of ``wl``; step > 1
    inverts on every step-th node of the forward grid along each axis (the slowness taken there)."""
    from . import eikonal as E
    g = wl.grid
    sP = wl.slowness_p
    seeds = [E.point_source(g, sP, *st) for st in wl.stations]
    ts = E.TableSet(device, g, [sP], seeds, np.zeros(len(wl.stations), np.int32))
    rc, why = ts.solve(eik_mode)
    if rc != 0:
        raise RuntimeError(why)
    ns, ne = len(wl.stations), len(wl.events)
    station = np.tile(np.arange(ns, dtype=np.int32), ne)
    event = np.repeat(np.arange(ne), ns)
    inv, s_inv = g, sP
    if step > 1:
        if any((m - 1) % step for m in (g.nx, g.ny, g.nz)):
            raise ValueError("the step must divide every axis' intervals")
        inv = GridDefinition(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth,
                             (g.nx - 1) // step + 1, (g.ny - 1) // step + 1, (g.nz - 1) // step + 1)
        s_inv = np.ascontiguousarray(sP.reshape(g.shape)[::step, ::step, ::step]).ravel()
    rays = RayBatch(ts, inv, wl.stations, station, wl.events[event])
    ts.release()
    Qn = q_model(inv, q_mean)
    rows = np.flatnonzero(rays.counts >= 0).astype(np.int32)
    pred, _ = q_predict(rays, rows, s_inv, 1 / Qn, 1.0)
    tstar = np.full(station.size, np.nan)
    rng = np.random.default_rng(seed)
    tstar[rows] = pred + sigma * rng.standard_normal(rows.size)
    return QCase(rays, QData(station, tstar, np.full(station.size, sigma), ns), s_inv, Qn, inv)
