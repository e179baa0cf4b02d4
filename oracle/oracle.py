"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (the plain-C FP64 restatement of the reference's CPU solvers) and
wraps it for numpy. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module, always as the checker, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class GridDef(C.Structure):
    """GridDefinition (Geo/SphericalGrid.cs:13-31): bounds in degrees / km and node counts."""

    _fields_ = [("min_lon", C.c_double), ("max_lon", C.c_double), ("min_lat", C.c_double),
                ("max_lat", C.c_double), ("min_depth", C.c_double), ("max_depth", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)]

    @property
    def count(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def as_tuple(self):
        return (self.min_lon, self.max_lon, self.min_lat, self.max_lat, self.min_depth,
                self.max_depth, self.nx, self.ny, self.nz)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_G = C.POINTER(GridDef)


def _declare(L):
    d = C.c_double
    L.or_trig_array.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_void_p, _dp]
    L.or_to_cartesian.argtypes = [d, d, d, _dp]
    L.or_from_cartesian.argtypes = [_dp, _dp]
    L.or_local_frame.argtypes = [d, d, _dp, _dp, _dp]
    L.or_grid_from_spacing.argtypes = [d, d, d, d, d, d, d, d, _G]
    L.or_fractional.argtypes = [_G, d, d, d, _dp]
    L.or_trilinear_weights_fractional.argtypes = [_G, d, d, d, _ip, _dp]
    L.or_interpolate.argtypes = [_G, _dp, d, d, d]
    L.or_interpolate.restype = d
    L.or_min_spacing_km.argtypes = [_G]
    L.or_min_spacing_km.restype = d
    L.or_node_coords.argtypes = [_G, _dp, _dp, _dp]
    L.or_surface_distance_km.argtypes = [d, d, d, d]
    L.or_surface_distance_km.restype = d
    L.or_metric.argtypes = [_G, _dp, _dp, C.POINTER(d)]
    L.or_point_source.argtypes = [_G, _dp, d, d, d, d, d, _ip, _dp, C.c_int]
    L.or_point_source.restype = C.c_int
    L.or_relax.argtypes = [_G, _dp, C.c_int, _ip, _dp, C.c_int, _dp, _ip]
    L.or_relax.restype = C.c_int
    L.or_fmm.argtypes = [_G, _dp, C.c_int, _ip, _dp, _dp]
    L.or_fmm.restype = C.c_int
    L.or_fmm_batch.argtypes = [_G, C.c_int, C.POINTER(C.c_void_p), _ip, _ip, _dp, _dp, C.c_int]
    L.or_fmm_batch.restype = C.c_int
    if hasattr(L, "or_table_time"):
        _declare_rays(L)
    if hasattr(L, "or_lsqr"):
        _declare_lsqr(L)


TRIG = {"sin": 0, "cos": 1, "asin": 2, "atan2": 3}


def trig(func: str, a, b=None, correct: bool = True) -> np.ndarray:
    """sin / cos / asin of a, or atan2(a, b): the correctly rounded values (libquadmath's 113-bit
    functions rounded to double) or, with correct=False, the platform C library's."""
    L = lib()
    a = np.ascontiguousarray(a, np.float64)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    out = np.empty_like(a)
    L.or_trig_array(TRIG[func], 1 if correct else 0, a.size, a, None if bb is None else bb.ctypes.data, out)
    return out


def _declare_rays(L):
    d = C.c_double
    L.or_table_time_fractional.argtypes = [_G, _fp, d, d, d]
    L.or_table_time_fractional.restype = d
    L.or_table_time.argtypes = [_G, _fp, d, d, d]
    L.or_table_time.restype = d
    L.or_table_gradient.argtypes = [_G, _fp, d, d, d, _dp]
    L.or_backtrack.argtypes = [_G, _fp, d, d, d, d, d, d, d, C.c_int, _dp, _dp, _dp]
    L.or_backtrack.restype = C.c_int
    L.or_ray_kernel.argtypes = [_G, C.c_int, _dp, _dp, _dp, C.c_int, _ip, _dp]
    L.or_ray_kernel.restype = C.c_int
    L.or_ray_row.argtypes = [_G, _fp, d, d, d, _G, d, d, d, d, C.c_int, _ip, _fp, _dp, C.c_void_p]
    L.or_ray_row.restype = C.c_int
    L.or_ray_rows_batch.argtypes = [_G, C.POINTER(C.c_void_p), _dp, _G, C.c_int, _ip, _dp, d,
                                    C.c_int, C.c_int, _ip, _ip, _fp, _dp]
    L.or_ray_rows_batch.restype = C.c_int


def _declare_lsqr(L):
    d = C.c_double
    L.or_csr_transpose.argtypes = [C.c_int, C.c_int, _lp, _ip, _dp, _lp, _ip, _dp]
    L.or_spmv.argtypes = [C.c_int, _lp, _ip, _dp, _dp, _dp, C.c_int]
    L.or_dot.argtypes = [C.c_int, _dp, _dp]
    L.or_dot.restype = d
    L.or_norm.argtypes = [C.c_int, _dp]
    L.or_norm.restype = d
    L.or_lsqr.argtypes = [C.c_int, C.c_int, _lp, _ip, _dp, _lp, _ip, _dp, _dp, d, C.c_int, d, d, d,
                          C.c_int, _dp, _dp, _dp]
    L.or_lsqr.restype = C.c_int
    L.or_lsqr_dot_mode.argtypes = [C.c_int]
    L.or_dot_device.argtypes = [C.c_int, _dp, _dp]
    L.or_dot_device.restype = d


# ---------------------------------------------------------------- geometry
def griddef(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth, nx, ny, nz) -> GridDef:
    return GridDef(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth, nx, ny, nz)


def from_spacing(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth, h_km, v_km) -> GridDef:
    g = GridDef()
    lib().or_grid_from_spacing(min_lon, max_lon, min_lat, max_lat, min_depth, max_depth, h_km, v_km,
                               C.byref(g))
    return g


def to_cartesian(lon, lat, depth):
    out = np.zeros(3)
    lib().or_to_cartesian(lon, lat, depth, out)
    return out


def from_cartesian(v):
    out = np.zeros(3)
    lib().or_from_cartesian(np.ascontiguousarray(v, dtype=np.float64), out)
    return out


def node_coords(g: GridDef):
    lon, lat, dep = np.zeros(g.nx), np.zeros(g.ny), np.zeros(g.nz)
    lib().or_node_coords(C.byref(g), lon, lat, dep)
    return lon, lat, dep


def trilinear_weights(g: GridDef, lon, lat, depth):
    f = np.zeros(3)
    lib().or_fractional(C.byref(g), lon, lat, depth, f)
    nodes, w = np.zeros(8, np.int32), np.zeros(8)
    lib().or_trilinear_weights_fractional(C.byref(g), f[0], f[1], f[2], nodes, w)
    return nodes, w


def interpolate(g: GridDef, field, lon, lat, depth) -> float:
    return lib().or_interpolate(C.byref(g), np.ascontiguousarray(field, dtype=np.float64), lon, lat, depth)


def min_spacing_km(g: GridDef) -> float:
    return lib().or_min_spacing_km(C.byref(g))


def surface_distance_km(lon1, lat1, lon2, lat2) -> float:
    return lib().or_surface_distance_km(lon1, lat1, lon2, lat2)


def cartesian_nodes(g: GridDef):
    """(N, 3) geocentric positions of every node, via the oracle's ToCartesian."""
    lon, lat, dep = node_coords(g)
    out = np.zeros((g.count, 3))
    n = 0
    for k in range(g.nz):
        for j in range(g.ny):
            for i in range(g.nx):
                out[n] = to_cartesian(lon[i], lat[j], dep[k])
                n += 1
    return out


# ---------------------------------------------------------------- eikonal
def metric(g: GridDef):
    hlon, hlat, hr = np.zeros(g.ny * g.nz), np.zeros(g.nz), C.c_double()
    lib().or_metric(C.byref(g), hlon, hlat, C.byref(hr))
    return hlon, hlat, hr.value


def point_source(g: GridDef, slowness, lon, lat, depth, origin_time=0.0, radius_cells=3.0):
    s = np.ascontiguousarray(slowness, dtype=np.float64).ravel()
    cap = 4096
    while True:
        idx, t = np.zeros(cap, np.int32), np.zeros(cap)
        n = lib().or_point_source(C.byref(g), s, lon, lat, depth, origin_time, radius_cells, idx, t, cap)
        if n <= cap:
            return idx[:n].copy(), t[:n].copy()
        cap = n


def relax(g: GridDef, slowness, seed_idx, seed_time, max_rounds=0):
    out = np.zeros(g.count)
    rounds = np.zeros(2, np.int32)
    s = np.ascontiguousarray(slowness, dtype=np.float64).ravel()
    ok = lib().or_relax(C.byref(g), s, len(seed_idx), np.ascontiguousarray(seed_idx, np.int32),
                        np.ascontiguousarray(seed_time, np.float64), max_rounds, out, rounds)
    return out, bool(ok), tuple(int(r) for r in rounds)


def fmm(g: GridDef, slowness, seed_idx, seed_time):
    out = np.zeros(g.count)
    s = np.ascontiguousarray(slowness, dtype=np.float64).ravel()
    rc = lib().or_fmm(C.byref(g), s, len(seed_idx), np.ascontiguousarray(seed_idx, np.int32),
                      np.ascontiguousarray(seed_time, np.float64), out)
    if rc != 0:
        raise ValueError("no seed nodes")
    return out


def fmm_batch(g: GridDef, slowness_list, seed_sets, threads=0):
    """One FMM table per field, fields spread over `threads` cores (0 = all)."""
    nf = len(seed_sets)
    offs = np.zeros(nf + 1, np.int32)
    for f, (idx, _) in enumerate(seed_sets):
        offs[f + 1] = offs[f] + len(idx)
    idx = np.ascontiguousarray(np.concatenate([s[0] for s in seed_sets]).astype(np.int32))
    tim = np.ascontiguousarray(np.concatenate([s[1] for s in seed_sets]).astype(np.float64))
    slows = [np.ascontiguousarray(s, dtype=np.float64).ravel() for s in slowness_list]
    ptrs = (C.c_void_p * nf)(*[s.ctypes.data for s in slows])
    out = np.zeros(nf * g.count)
    rc = lib().or_fmm_batch(C.byref(g), nf, ptrs, offs, idx, tim, out, threads)
    if rc != 0:
        raise ValueError("a field has no seed nodes")
    return out.reshape(nf, g.count)


# ---------------------------------------------------------------- sparse algebra and LSQR
LSQR_STOP = {0: "iteration budget exhausted", 1: "residual reduced to tolerance (btol)",
             2: "normal equations satisfied (atol)", 3: "condition-number limit reached (conlim)",
             4: "residual at machine precision", 5: "normal equations at machine precision",
             6: "condition number at machine precision", 7: "bidiagonalisation terminated",
             8: "empty system", 9: "zero right-hand side", 10: "A^T b = 0"}


def csr_transpose(rows, cols, rs, ci, va):
    rs = np.ascontiguousarray(rs, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    trs = np.zeros(cols + 1, np.int64)
    tci = np.zeros(len(ci), np.int32)
    tva = np.zeros(len(va), np.float64)
    lib().or_csr_transpose(rows, cols, rs, ci, va, trs, tci, tva)
    return trs, tci, tva


def spmv(rs, ci, va, x, threads=0):
    rs = np.ascontiguousarray(rs, np.int64)
    y = np.zeros(len(rs) - 1)
    lib().or_spmv(len(rs) - 1, rs, np.ascontiguousarray(ci, np.int32), np.ascontiguousarray(va, np.float64),
                  np.ascontiguousarray(x, np.float64), y, threads)
    return y


def lsqr(m, n, rs, ci, va, b, damp=0.0, max_it=500, atol=1e-8, btol=1e-8, conlim=1e8, threads=0,
         transpose=None, device_dots=False):
    """Lsqr.Solve (Numerics/Lsqr.cs:34-152). Returns (x, stop code, stats dict, rnorm history).
    device_dots: the dot products and norms of LsqrOpenCl.Run (LsqrOpenCl.cs:79-94: 64 partials of
    tree-reduced item sums, the norm their square root) instead of SimdVector's."""
    rs = np.ascontiguousarray(rs, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    va = np.ascontiguousarray(va, np.float64)
    trs, tci, tva = transpose if transpose is not None else csr_transpose(m, n, rs, ci, va)
    x = np.zeros(n)
    stats = np.zeros(4)
    hist = np.zeros(max(1, max_it))
    L = lib()
    L.or_lsqr_dot_mode(1 if device_dots else 0)
    try:
        stop = L.or_lsqr(m, n, rs, ci, va, trs, tci, tva, np.ascontiguousarray(b, np.float64), damp, max_it,
                         atol, btol, conlim, threads, x, stats, hist)
    finally:
        L.or_lsqr_dot_mode(0)
    it = int(stats[0])
    return x, stop, {"iterations": it, "bnorm": stats[1], "rnorm": stats[2], "acond": stats[3]}, hist[:it].copy()


def resolution_diagonal(m, n, rs, ci, va, data_rows, signs, lsqr_iterations, tolerance, threads=0):
    """ResolutionDiagonal.Estimate (Tomography/ResolutionDiagonal.cs:31-73) with the probes given:
    signs[probes][n] of +-1. Per probe: az = A z (:48), rhs = az on the first data_rows rows and 0 on
    the rest (:50-51), Lsqr.Solve(A, rhs, damp 0, lsqr_iterations, tolerance, tolerance, 1e8) (:52),
    e = z * x summed and squared (:54-59); then the mean, and the standard error
    sqrt(max(0, (sum2 - p mean^2) / (p - 1)) / p) (:61-69). Returns (diagonal, stderr, mean iterations)."""
    signs = np.asarray(signs, np.float64)
    probes = signs.shape[0]
    tr = csr_transpose(m, n, rs, ci, va)
    s1 = np.zeros(n)
    s2 = np.zeros(n)
    its = 0
    for z in signs:
        az = spmv(rs, ci, va, z, threads)
        rhs = np.zeros(m)
        rhs[:data_rows] = az[:data_rows]
        x, _, st, _ = lsqr(m, n, rs, ci, va, rhs, 0.0, lsqr_iterations, tolerance, tolerance, 1e8, threads, tr)
        its += st["iterations"]
        e = z * x
        s1 += e
        s2 += e * e
    mean = s1 / probes
    var = np.maximum(0.0, (s2 - probes * mean * mean) / (probes - 1))
    return mean, np.sqrt(var / probes), its // probes


def weighted_median(values, weights):
    """EventLocator.WeightedMedian (Location/EventLocator.cs:311-323): the finite values in a stable sort,
    their weights summed in that order, the first value whose running weight reaches half; 0 when none."""
    v = np.asarray(values, np.float64)
    w = np.asarray(weights, np.float64)
    keep = ~np.isnan(v)
    v, w = v[keep], w[keep]
    if v.size == 0:
        return 0.0
    o = np.argsort(v, kind="stable")
    acc = np.cumsum(w[o])  # sequential, in sorted order
    half = acc[-1] / 2
    hit = np.flatnonzero(acc >= half)
    return float(v[o][hit[0]]) if hit.size else float(v[o][-1])


def locate_grid_search(tables, nx, ny, nz, picks, coarse_candidates=8000, center=None, k_fixed=-1):
    """EventLocator.GridSearch (EventLocator.cs:114-175) for one event: tables[field] float32 [N];
    picks = (field, obs, weight, correction) arrays. Misfit per node: r = obs - t - correction (NaN where
    the table is not finite), t0 = weighted median of r, m = sum in pick order of 10 w (NaN) or
    w |r - t0|. A global search scans the stride grid (stride = max(1, ceil(cbrt(N / candidates))))
    then every node within two strides of the best; a local search (center = (i, j, k)) every node
    within 6; k_fixed >= 0 fixes the depth index. Returns (node, t0, misfit)."""
    field, obs, w, corr = (np.asarray(a) for a in picks)
    N = nx * ny * nz

    def misfits(idx):
        t = np.stack([tables[f][idx] for f in field], axis=1).astype(np.float64)  # [cand][pick]
        r = np.where(np.isfinite(t), (obs[None, :] - t) - corr[None, :], np.nan)
        t0 = np.array([weighted_median(row, w) for row in r])
        terms = np.where(np.isnan(r), 10 * w[None, :], w[None, :] * np.abs(r - t0[:, None]))
        return np.cumsum(terms, axis=1)[:, -1] if terms.shape[1] else np.zeros(len(idx)), t0

    def search(i0, i1, j0, j1, k0, k1, stride):
        i0, j0, k0 = max(0, i0), max(0, j0), max(0, k0)
        i1, j1, k1 = min(nx - 1, i1), min(ny - 1, j1), min(nz - 1, k1)
        if k_fixed >= 0:
            k0 = k1 = k_fixed
        kk, jj, ii = np.meshgrid(np.arange(k0, k1 + 1, stride), np.arange(j0, j1 + 1, stride),
                                 np.arange(i0, i1 + 1, stride), indexing="ij")
        idx = ((kk * ny + jj) * nx + ii).ravel()
        if idx.size == 0:
            return (k0 * ny + j0) * nx + i0
        m, _ = misfits(idx)
        return int(idx[int(np.argmin(m))])  # first minimum in k, j, i order

    if center is not None:
        ci, cj, ck = center
        best = search(ci - 6, ci + 6, cj - 6, cj + 6, ck - 6, ck + 6, 1)
    else:
        s = max(1, int(np.ceil(np.cbrt(N / max(1, coarse_candidates)))))
        best = search(0, nx - 1, 0, ny - 1, 0, nz - 1, s)
        bi, bj, bk = best % nx, (best // nx) % ny, best // (nx * ny)
        best = search(bi - 2 * s, bi + 2 * s, bj - 2 * s, bj + 2 * s, bk - 2 * s, bk + 2 * s, 1)
    m, t0 = misfits(np.array([best]))
    return best, float(t0[0]), float(m[0])


# ---------------------------------------------------------------- tables and rays
def table_time(g: GridDef, table, lon, lat, depth) -> float:
    return lib().or_table_time(C.byref(g), np.ascontiguousarray(table, np.float32), lon, lat, depth)


def table_gradient(g: GridDef, table, lon, lat, depth):
    out = np.zeros(3)
    lib().or_table_gradient(C.byref(g), np.ascontiguousarray(table, np.float32), lon, lat, depth, out)
    return out


def backtrack(g: GridDef, table, station, event, step_km, cap=200000):
    """RayTracing.Backtrack: (lon, lat, depth) arrays of the path, or None when the descent stalls."""
    pl, pa, pd = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    n = lib().or_backtrack(C.byref(g), np.ascontiguousarray(table, np.float32), station[0], station[1], station[2],
                           event[0], event[1], event[2], step_km, cap, pl, pa, pd)
    if n == -2:
        raise ValueError("path longer than cap")
    return None if n < 0 else (pl[:n].copy(), pa[:n].copy(), pd[:n].copy())


def ray_kernel(inv: GridDef, path, cap=100000):
    lon, lat, dep = (np.ascontiguousarray(a, np.float64) for a in path)
    nodes, lens = np.zeros(cap, np.int32), np.zeros(cap)
    n = lib().or_ray_kernel(C.byref(inv), len(lon), lon, lat, dep, cap, nodes, lens)
    if n < 0:
        raise ValueError("kernel longer than cap")
    return nodes[:n].copy(), lens[:n].copy()


class Background(C.Structure):
    """or_background: a 1-D model (Model/VelocityModel1D.cs), node depths and Vp, Vs."""

    _fields_ = [("count", C.c_int), ("depth_km", C.c_void_p), ("vp", C.c_void_p), ("vs", C.c_void_p)]


def background(depth_km, vp, vs):
    """(or_background, the arrays it points to: keep them alive while it is used)."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (depth_km, vp, vs)]
    return Background(len(arrs[0]), arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data), arrs


def background_velocity(bg, phase, depth):
    L = lib()
    L.or_background_velocity.argtypes = [C.POINTER(Background), C.c_int, C.c_double]
    L.or_background_velocity.restype = C.c_double
    return L.or_background_velocity(C.byref(bg[0]), phase, depth)


def ray_kernel_bg(inv: GridDef, path, bg, phase, cap=100000):
    """RayTracing.Kernel with a background: (nodes, lengths, outside time)."""
    L = lib()
    L.or_ray_kernel_bg.argtypes = [_G, C.c_int, _dp, _dp, _dp, C.c_void_p, C.c_int, C.c_int, _ip, _dp,
                                   C.POINTER(C.c_double)]
    lon, lat, dep = (np.ascontiguousarray(a, np.float64) for a in path)
    nodes, lens = np.zeros(cap, np.int32), np.zeros(cap)
    out = C.c_double()
    n = L.or_ray_kernel_bg(C.byref(inv), len(lon), lon, lat, dep, C.addressof(bg[0]) if bg else None, phase, cap,
                           nodes, lens, C.byref(out))
    if n < 0:
        raise ValueError("kernel longer than cap")
    return nodes[:n].copy(), lens[:n].copy(), out.value


def ray_rows_batch(fwd: GridDef, tables, stations, inv: GridDef, ray_table, events, step_km, cap_per_ray=4096,
                   threads=0, bg=None, phase_of_ray=None):
    """RayKernels.Compute's rows for many arrivals: list of (nodes, float32 lengths, tcalc, ge, gn, gz) or None;
    with a background (bg from background(), phase_of_ray 0 P / 1 S), the rows' outside times as well:
    (rows, outside[nrays])."""
    tabs = [np.ascontiguousarray(t, np.float32) for t in tables]
    ptrs = (C.c_void_p * len(tabs))(*[t.ctypes.data for t in tabs])
    st = np.ascontiguousarray(stations, np.float64).ravel()
    ev = np.ascontiguousarray(events, np.float64).ravel()
    rt = np.ascontiguousarray(ray_table, np.int32)
    nr = len(rt)
    counts = np.zeros(nr, np.int32)
    nodes = np.zeros(nr * cap_per_ray, np.int32)
    lens = np.zeros(nr * cap_per_ray, np.float32)
    tg = np.zeros(4 * nr)
    outside = np.zeros(max(1, nr))
    L = lib()
    L.or_ray_rows_batch_bg.argtypes = [_G, C.POINTER(C.c_void_p), _dp, _G, C.c_int, _ip, _dp, C.c_double,
                                       C.c_void_p, C.c_void_p, C.c_int, C.c_int, _ip, _ip, _fp, _dp, _dp]
    ph = None if phase_of_ray is None else np.ascontiguousarray(phase_of_ray, np.int32)
    L.or_ray_rows_batch_bg(C.byref(fwd), ptrs, st, C.byref(inv), nr, rt, ev, step_km,
                           C.addressof(bg[0]) if bg else None, None if ph is None else ph.ctypes.data,
                           cap_per_ray, threads, counts, nodes, lens, tg, outside)
    out = []
    for r in range(nr):
        c = counts[r]
        if c < 0:
            out.append(None)
        else:
            a = r * cap_per_ray
            out.append((nodes[a:a + c].copy(), lens[a:a + c].copy(), *tg[4 * r:4 * r + 4]))
    return (out, outside[:nr].copy()) if bg else out


class SystemOpts(C.Structure):
    """TomographySettings fields BuildSystem reads (Tomography/TravelTimeTomography.cs:30-110)."""

    _fields_ = [("vpvs", C.c_int), ("joint_hypocentres", C.c_int), ("station_corrections", C.c_int),
                ("invert_p", C.c_int), ("invert_s", C.c_int), ("gradient_smoothing", C.c_int),
                ("nevents", C.c_int), ("nstations", C.c_int), ("any_p", C.c_int), ("any_s", C.c_int),
                ("smoothing_in_nodes", C.c_int),
                ("damping", C.c_double), ("smoothing", C.c_double), ("vertical_weight", C.c_double),
                ("damping_hypocentre", C.c_double), ("damping_station", C.c_double), ("s_factor", C.c_double),
                ("absolute_weight", C.c_double)]


def build_system(g: GridDef, opts: dict, row_start, nodes, lens, tcalc_grad, phase, event, station, residual, sigma,
                 ev_fixed, ev_outside, sP, sS, sRefP, sRefS, kRef):
    """BuildSystem restated (plain node grid). Rows are the used data rows in order: row_start/nodes/lens
    their ray rows, tcalc_grad[r] = (TCalc, Ge, Gn, Gz). Returns (row_start, column, values, rhs, dims)
    with dims = (rows, cols, nnz, data rows)."""
    L = lib()
    if not getattr(L, "_sys_declared", False):
        L.or_build_system.restype = C.c_int
        L.or_build_system.argtypes = [_G, C.POINTER(SystemOpts), C.c_int, _lp, _ip, _fp, _dp, _ip, _ip, _ip, _dp, _dp,
                                      _u8, _u8, _dp, _dp, _dp, _dp, _dp, C.c_int64, C.c_int64, _lp, _ip, _dp, _dp, _lp]
        L._sys_declared = True
    o = SystemOpts(**{k: opts.get(k, 0) for k, _ in SystemOpts._fields_})
    n = g.count
    nrows = len(phase)
    row_start = np.ascontiguousarray(row_start, np.int64)
    nodes = np.ascontiguousarray(nodes, np.int32)
    cap_rows = nrows + 8 * n + 4 * o.nevents + 2 * o.nstations + 4
    cap_nnz = 2 * nodes.size + 6 * nrows + 30 * n + 4 * o.nevents + 4 * o.nstations + 8
    rs = np.zeros(cap_rows + 1, np.int64)
    ci = np.zeros(cap_nnz, np.int32)
    va = np.zeros(cap_nnz)
    rhs = np.zeros(cap_rows)
    dims = np.zeros(4, np.int64)
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    u8 = lambda a: np.ascontiguousarray(a, np.uint8) if len(a) else np.zeros(1, np.uint8)
    rc = L.or_build_system(C.byref(g), C.byref(o), nrows, row_start, nodes, np.ascontiguousarray(lens, np.float32),
                           f64(tcalc_grad).ravel() if nrows else np.zeros(4), i32(phase), i32(event), i32(station),
                           f64(residual), f64(sigma), u8(ev_fixed), u8(ev_outside), f64(sP), f64(sS), f64(sRefP),
                           f64(sRefS), f64(kRef), cap_rows, cap_nnz, rs, ci, va, rhs, dims)
    if rc != 0:
        raise RuntimeError("oracle system exceeds its capacity")
    m, cols, nnz, data_rows = (int(x) for x in dims)
    return rs[:m + 1].copy(), ci[:nnz].copy(), va[:nnz].copy(), rhs[:m].copy(), (m, cols, nnz, data_rows)


# ------------------------------------------------------------------------------ Q tomography

class QSystemOpts(C.Structure):
    """The QTomographySettings fields SolveOnce reads (Attenuation/QTomography.cs:12-90)."""

    _fields_ = [("station_terms", C.c_int), ("gradient_smoothing", C.c_int), ("smoothing_in_nodes", C.c_int),
                ("nstations", C.c_int), ("fixed_scales", C.c_int), ("damping", C.c_double),
                ("smoothing", C.c_double), ("vertical_weight", C.c_double), ("damping_station", C.c_double),
                ("q0", C.c_double)]


def _declare_q(L):
    if getattr(L, "_q_declared", False):
        return
    L.or_q_predict.restype = None
    L.or_q_predict.argtypes = [C.c_int, _lp, _ip, _fp, _dp, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                               C.c_void_p]
    L.or_build_qsystem.restype = C.c_int
    L.or_build_qsystem.argtypes = [_G, C.POINTER(QSystemOpts), C.c_int, _lp, _ip, _fp, C.c_void_p, _ip, _dp, _dp, _dp,
                                   C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, _lp, _ip, _dp, _dp, _lp, _dp]
    L._q_declared = True


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def q_predict(row_start, nodes, lens, slow, q, q0, outside=None):
    """QTomography.Predict (QTomography.cs:307-312) and the path time OutsideTime + sum L s (:231-233) of
    every row; outside: the rows' OutsideTime (None: 0); q None: q0 at every node. Returns (pred, tt)."""
    L = lib()
    _declare_q(L)
    rs = np.ascontiguousarray(row_start, np.int64)
    nr = rs.size - 1
    pred, tt = np.zeros(max(1, nr)), np.zeros(max(1, nr))
    qa = None if q is None else np.ascontiguousarray(q, np.float64)
    oa = None if outside is None else np.ascontiguousarray(outside, np.float64)
    L.or_q_predict(nr, rs, np.ascontiguousarray(nodes, np.int32), np.ascontiguousarray(lens, np.float32),
                   np.ascontiguousarray(slow, np.float64), _vp(qa), float(q0), _vp(oa), _vp(pred), _vp(tt))
    return pred[:nr], tt[:nr]


def build_qsystem(g: GridDef, opts: dict, row_start, nodes, lens, station, tstar, sigma, slow, q_old=None, m_old=None,
                  scales=None, outside=None):
    """QTomography.SolveOnce's system restated (QTomography.cs:318-392, plain node grid); the rows are the
    used ray rows in order. Returns (row_start, column, values, rhs, dims, scales) with dims = (rows,
    cols, nnz, data rows) and scales = (model, station)."""
    L = lib()
    _declare_q(L)
    o = QSystemOpts(**{k: opts.get(k, 0) for k, _ in QSystemOpts._fields_})
    o.fixed_scales = 1 if scales is not None else 0
    n = g.count
    rs_in = np.ascontiguousarray(row_start, np.int64)
    nr = rs_in.size - 1
    nodes = np.ascontiguousarray(nodes, np.int32)
    cap_rows = nr + 5 * n + o.nstations + 4
    cap_nnz = int(rs_in[-1]) + 2 * nr + 12 * n + 2 * o.nstations + 8
    rs = np.zeros(cap_rows + 1, np.int64)
    ci = np.zeros(cap_nnz, np.int32)
    va = np.zeros(cap_nnz)
    rhs = np.zeros(cap_rows)
    dims = np.zeros(4, np.int64)
    sc = np.array(scales if scales is not None else (1.0, 1.0), np.float64)
    qo = None if q_old is None else np.ascontiguousarray(q_old, np.float64)
    mo = None if m_old is None else np.ascontiguousarray(m_old, np.float64)
    st = np.ascontiguousarray(station if len(station) else np.zeros(1), np.int32)
    oa = None if outside is None else np.ascontiguousarray(outside, np.float64)
    rc = L.or_build_qsystem(C.byref(g), C.byref(o), nr, rs_in, nodes, np.ascontiguousarray(lens, np.float32), _vp(oa), st,
                            np.ascontiguousarray(tstar, np.float64) if nr else np.zeros(1),
                            np.ascontiguousarray(sigma, np.float64) if nr else np.ones(1),
                            np.ascontiguousarray(slow, np.float64), _vp(qo), _vp(mo), cap_rows, cap_nnz, rs, ci, va,
                            rhs, dims, sc)
    if rc != 0:
        raise RuntimeError("oracle Q system exceeds its capacity")
    m, cols, nnz, data_rows = (int(x) for x in dims)
    return rs[:m + 1].copy(), ci[:nnz].copy(), va[:nnz].copy(), rhs[:m].copy(), (m, cols, nnz, data_rows), \
        (float(sc[0]), float(sc[1]))


def seq_sum(a) -> float:
    """A running sum in index order (a foreach accumulation)."""
    a = np.asarray(a, np.float64)
    return float(np.cumsum(a)[-1]) if a.size else 0.0


def seq_mean(a) -> float:
    """Enumerable.Average: the running sum over the count."""
    a = np.asarray(a, np.float64)
    return seq_sum(a) / a.size


class QOracle:
    """QTomography's host steps (QTomography.cs:199-300 Prepare, :252-289 Solve, :296-405 SolveOnce, :412-432
    RejectOutliers, :434-505 Run, :520-554 TradeOff) over the C restatement of the system and Lsqr.Solve,
    for a plain node grid with Laplacian smoothing (or unweighted first differences), on rays given as
    rows in arrival order. settings: a dict of QTomographySettings fields (q0 as Q0, the rest by name)."""

    DEFAULTS = dict(q0=300.0, estimate_q0=True, damping=2.0, smoothing=10.0, vertical_weight=0.1, smoothing_in_nodes=0,
                    gradient_smoothing=0, station_terms=True, damping_station=3.0, lsqr_iterations=400,
                    log_parameterisation=True, nonlinear_iterations=20, q_min=10.0, q_max=5000.0, outlier_mads=4.0)

    def __init__(self, g: GridDef, settings: dict, threads: int = 0):
        self.g = g
        self.s = dict(self.DEFAULTS, **settings)
        self.threads = threads

    # rows: (row_start [nrows+1], nodes, lens) of the arrivals that have rays; arrival[r]: their arrival
    def prepare(self, row_start, nodes, lens, arrival, tstar, sigma, station, nstations, slow):
        s = self.s
        self.rs, self.nodes, self.lens = np.asarray(row_start, np.int64), nodes, lens
        self.arrival = np.asarray(arrival, np.int64)
        self.tstar, self.sigma = np.asarray(tstar, np.float64), np.asarray(sigma, np.float64)
        self.station, self.nstations, self.slow = np.asarray(station, np.int32), nstations, slow
        q0 = 1 / max(1.0, s["q0"])
        _, tt = q_predict(self.rs, nodes, lens, slow, None, q0)
        if s["estimate_q0"] and self.arrival.size:
            ratios, num, den = [], 0.0, 0.0
            for r, a in enumerate(self.arrival):
                t = tt[r]
                if not (t > 0) or not np.isfinite(self.tstar[a]):
                    continue
                ratios.append(self.tstar[a] / t)
                w = 1 / (self.sigma[a] * self.sigma[a])
                num += w * self.tstar[a] * t
                den += w * t * t
            est = 0.0
            if s["station_terms"]:
                if den > 0 and num > 0:
                    est = num / den
            elif ratios:
                ratios.sort()
                m = len(ratios) // 2
                est = ratios[m] if len(ratios) % 2 == 1 else 0.5 * (ratios[m - 1] + ratios[m])
            if est > 0:
                q0 = min(max(est, 1 / s["q_max"]), 1 / s["q_min"])
        self.q0 = q0
        na = self.tstar.size
        self.rejected = np.ones(na, bool)
        self.residual = np.full(na, np.nan)
        pred, _ = q_predict(self.rs, nodes, lens, slow, None, q0)
        self.start_residual = self.tstar[self.arrival] - pred
        self.residual[self.arrival] = self.start_residual
        self.rejected[self.arrival] = False
        self.rms_before = self.rms()

    def rms(self):
        used = ~self.rejected & ~np.isnan(self.residual)
        return math.sqrt(seq_mean(self.residual[used] ** 2)) if used.any() else 0.0

    def _used_rows(self):
        keep = ~self.rejected[self.arrival]
        idx = np.flatnonzero(keep)
        starts, ends = self.rs[idx], self.rs[idx + 1]
        lens = ends - starts
        rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        take = np.concatenate([np.arange(a, b) for a, b in zip(starts, ends)]) if idx.size else np.zeros(0, np.int64)
        return idx, rs, self.nodes[take], self.lens[take]

    def solve_once(self, damping, smoothing, m_old, fixed_scales, x_old):
        s, g, n = self.s, self.g, self.g.count
        q0 = self.q0
        lo, hi = math.log(1 / (s["q_max"] * q0)), math.log(1 / (s["q_min"] * q0))
        q_old = None if m_old is None else q0 * np.exp(np.clip(m_old, lo, hi))
        idx, rs, nd, ln = self._used_rows()
        a = self.arrival[idx]
        opts = dict(station_terms=int(s["station_terms"]), gradient_smoothing=s["gradient_smoothing"],
                    smoothing_in_nodes=s["smoothing_in_nodes"], nstations=self.nstations, damping=damping,
                    smoothing=smoothing, vertical_weight=s["vertical_weight"], damping_station=s["damping_station"],
                    q0=q0)
        grs, gci, gva, rhs, dims, scales = build_qsystem(g, opts, rs, nd, ln, self.station[a], self.tstar[a],
                                                         self.sigma[a], self.slow, q_old, m_old, fixed_scales)
        self.scales = scales
        m_rows, cols = dims[0], dims[1]
        if x_old is not None:
            rhs = rhs - spmv(grs, gci, gva, x_old, self.threads)
        x, stop, st, _ = lsqr(m_rows, cols, grs, gci, gva, rhs, 0.0, s["lsqr_iterations"], 1e-6, 1e-6, 1e8,
                              self.threads)
        if x_old is not None:
            x = x + x_old
        if s["log_parameterisation"]:
            m_nodes = np.clip(x[:n], lo, hi)
            q = q0 * np.exp(m_nodes)
        else:
            q = np.clip(q0 * (1 + x[:n]), 1 / s["q_max"], 1 / s["q_min"])
            m_nodes = np.log(q / q0)
        terms = x[n:n + self.nstations] if s["station_terms"] else np.zeros(0)
        pred, _ = q_predict(self.rs, self.nodes, self.lens, self.slow, q, q0)
        self.residual[self.arrival] = self.tstar[self.arrival] - pred - \
            (terms[self.station[self.arrival]] if terms.size else 0.0)
        return dict(x=x, q=q, terms=terms, iterations=st["iterations"], stop=stop, dims=dims), m_nodes

    def solve(self, damping, smoothing):
        s = self.s
        if not s["log_parameterisation"]:
            return self.solve_once(damping, smoothing, None, None, None)[0]
        m = scales = x = None
        it, change, out = 0, float("nan"), None
        while it < max(1, s["nonlinear_iterations"]):
            out, m_nodes = self.solve_once(damping, smoothing, m, scales, x)
            x = out["x"]
            scales = scales or self.scales
            change = math.inf if m is None else math.sqrt(seq_mean((m_nodes - m) ** 2))
            m = m_nodes
            it += 1
            if change < 1e-3:
                break
        self.gn = (it, change)
        return out

    def reject_outliers(self):
        keep = ~self.rejected
        fitted = np.sort(self.residual[keep])
        med = fitted[fitted.size // 2] if fitted.size else 0.0
        mad = 1.4826 * (np.sort(np.abs(fitted - med))[fitted.size // 2] if fitted.size else 0.0)
        sig = self.sigma[keep]
        limit = max(self.s["outlier_mads"] * mad, 2 * (seq_mean(sig) if sig.size else 0.002))
        bad = keep & (np.abs(self.residual - med) > limit)
        self.rejected |= bad
        return int(bad.sum())

    def run(self):
        out = self.solve(self.s["damping"], self.s["smoothing"])
        rejected = self.reject_outliers()
        if rejected > 0:
            out = self.solve(self.s["damping"], self.s["smoothing"])
        out["rejected"] = rejected
        out["rms_after"] = self.rms()
        return out


# ------------------------------------------------------------------------------ formal resolution

def formal_resolution(rows, cols, row_start, column, values, data_rows, eliminate=(), max_unknowns=20000,
                      places=None):
    """FormalResolution.Compute (Tomography/FormalResolution.cs:108-333) restated on dense numpy for small
    systems: N = A^T A, Gamma = the regularisation rows' part on the kept columns, the kept block reduced
    by the Schur complement of each eliminated group (its live columns, N_gg diagonal > 0; a group whose
    block is not positive definite is left out, as the reference's SmallCholesky failure does), S = L L^T,
    C_ii = (S^-1)_ii, R_ii = 1 - (S^-1 Gamma)_ii and, with places (xyz[cols, 3], block[cols]), the
    resolution length of each placed column from its row of R = I - S^-1 Gamma over the placed columns of
    its block. Returns (kept, diagonal, sd, length) or raises ValueError with the reference's reason."""
    rs = np.asarray(row_start, np.int64)
    ci = np.asarray(column, np.int64)
    va = np.asarray(values, np.float64)
    A = np.zeros((rows, cols))
    for r in range(rows):
        A[r, ci[rs[r]:rs[r + 1]]] += va[rs[r]:rs[r + 1]]
    group = np.full(cols, -1)
    for g, cs in enumerate(eliminate):
        group[np.asarray(cs)] = g
    kept = np.flatnonzero(group < 0)
    n = kept.size
    if n > max_unknowns:
        raise ValueError("too many unknowns")
    for r in range(rows):
        gs = set(group[ci[rs[r]:rs[r + 1]]].tolist()) - {-1}
        if len(gs) > 1:
            raise ValueError("a row couples the hypocentres of two events")
        if r >= data_rows and gs and np.any(group[ci[rs[r]:rs[r + 1]]] < 0):
            raise ValueError("a regularisation row couples model and hypocentre columns")
    Areg = A[data_rows:][:, kept]
    gamma = Areg.T @ Areg
    N = A.T @ A
    S = N[np.ix_(kept, kept)].copy()
    for cs in eliminate:
        cs = np.asarray(cs)
        live = cs[np.diag(N)[cs] > 0]
        if live.size == 0:
            continue
        Ngg = N[np.ix_(live, live)]
        try:
            np.linalg.cholesky(Ngg)
        except np.linalg.LinAlgError:
            continue
        Nkg = N[np.ix_(kept, live)]
        S -= Nkg @ np.linalg.solve(Ngg, Nkg.T)
    try:
        L = np.linalg.cholesky(S)
    except np.linalg.LinAlgError:
        raise ValueError("the reduced normal matrix is not positive definite")
    X = np.linalg.solve(L, np.eye(n))       # L^-1
    Sinv = X.T @ X
    sd = np.sqrt(np.diag(Sinv))
    R = np.eye(n) - Sinv @ gamma
    diag = np.diag(R).copy()
    length = np.full(n, np.nan)
    if places is not None:
        xyz = np.asarray(places[0], np.float64)[kept]
        blk = np.asarray(places[1])[kept]
        for i in range(n):
            if blk[i] < 0:
                continue
            same = blk == blk[i]
            r2 = R[i, same] ** 2
            d2 = np.sum((xyz[same] - xyz[i]) ** 2, axis=1)
            den = r2.sum()
            length[i] = math.sqrt((r2 * d2).sum() / den) if den > 0 else np.nan
    return kept, diag, sd, length
