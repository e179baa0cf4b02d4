"""Uneven batches: many single-field eikonal problems, each on its own grid, solved together on the
device (libtomostar_hip's ts_gridbatch). The reference solves such problems one grid per
EikonalOpenCl.TrySolve call (Forward/EikonalOpenCl.cs:138-343)."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import TS_EIK_FAST_FP32, TS_EIK_FAST_FP64, TsGrid, check, ptr, reason_buffer
from .eikonal import Device
from .grid import GridDefinition

MODES = {"fast_fp32": TS_EIK_FAST_FP32, "fast_fp64": TS_EIK_FAST_FP64}
METHODS = ["planes", "stream", "hybrid"]  # ts_gridbatch_schedule's method codes


def _declare():
    L = _lib.load()
    if getattr(L, "_gridbatch_declared", False):
        return L
    vp, i = C.c_void_p, C.c_int
    L.ts_gridbatch_create.restype = i
    L.ts_gridbatch_create.argtypes = [vp, i, vp, vp, vp, vp, vp, C.POINTER(vp), C.c_char_p, i]
    L.ts_gridbatch_tune.restype = i
    L.ts_gridbatch_tune.argtypes = [vp, vp]
    L.ts_gridbatch_solve.restype = i
    L.ts_gridbatch_solve.argtypes = [vp, i, vp, vp, C.c_char_p, i]
    L.ts_gridbatch_read.restype = i
    L.ts_gridbatch_read.argtypes = [vp, i, vp]
    L.ts_gridbatch_stats.restype = i
    L.ts_gridbatch_stats.argtypes = [vp, vp, vp, vp]
    L.ts_gridbatch_updates.restype = i
    L.ts_gridbatch_updates.argtypes = [vp, vp, vp]
    L.ts_gridbatch_schedule.restype = i
    L.ts_gridbatch_schedule.argtypes = [vp, vp, vp, vp, vp]
    L.ts_gridbatch_set_schedule.restype = i
    L.ts_gridbatch_set_schedule.argtypes = [vp, i, vp]
    L.ts_gridbatch_round_log.restype = i
    L.ts_gridbatch_round_log.argtypes = [vp, i, vp, vp]
    L.ts_gridbatch_release.restype = None
    L.ts_gridbatch_release.argtypes = [vp]
    L._gridbatch_declared = True
    return L


class GridBatch:
    """Problems q = 0..n-1: grids[q], slowness[q] (its nodes), seeds[q] = (node indices, times)."""

    def __init__(self, device: Device, grids: Sequence[GridDefinition], slowness: Sequence[np.ndarray],
                 seeds: Sequence[tuple]):
        L = _declare()
        n = len(grids)
        if len(slowness) != n or len(seeds) != n:
            raise ValueError("one slowness model and one seed set per problem")
        self.grids = list(grids)
        self._g = (TsGrid * n)(*[g.to_ts() for g in grids])
        self._s = [np.ascontiguousarray(s, np.float64).ravel() for s in slowness]
        for g, s in zip(grids, self._s):
            if s.size != g.count:
                raise ValueError("slowness size differs from its grid's node count")
        sp = (C.c_void_p * n)(*[s.ctypes.data for s in self._s])
        offs = np.zeros(n + 1, np.int64)
        for q, (idx, _) in enumerate(seeds):
            offs[q + 1] = offs[q] + len(idx)
        idx = np.ascontiguousarray(np.concatenate([np.asarray(a, np.int32) for a, _ in seeds]), np.int32)
        tm = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float64) for _, t in seeds]), np.float64)
        h = C.c_void_p()
        r = reason_buffer()
        rc = L.ts_gridbatch_create(device.handle, n, C.cast(self._g, C.c_void_p), C.cast(sp, C.c_void_p), ptr(offs),
                                   ptr(idx), ptr(tm), C.byref(h), r, len(r))
        if rc != 0:
            raise _lib.HipError(rc, r.value.decode(errors="replace"))
        self._h = h
        self.n = n

    def tune(self):
        """Runs the tuner; returns its pick of nodes per thread for problems of < 2e5, < 8e5 and >= 8e5 nodes
        (the streaming tile widths and the method it picked: schedule())."""
        out = np.zeros(3, np.int32)
        check(_declare().ts_gridbatch_tune(self._h, ptr(out)))
        return [int(x) for x in out]

    def schedule(self) -> dict:
        """The sweep in use: method ("planes": one launch per plane step; "stream": one persistent launch per
        ordering over every problem's column tiles), the streaming tile width and the plane method's nodes
        per thread per size class, and the tuner's trial round of the batch with each method (ms)."""
        m = C.c_int32()
        tw, npt, ms = np.zeros(3, np.int32), np.zeros(3, np.int32), np.zeros(3)
        check(_declare().ts_gridbatch_schedule(self._h, C.byref(m), ptr(tw), ptr(npt), ptr(ms)))
        return {"method": METHODS[m.value], "tile_width": [int(x) for x in tw],
                "nodes_per_thread": [int(x) for x in npt],
                "trial_ms": {name: float(ms[q]) for q, name in enumerate(METHODS)}}

    def set_schedule(self, method: str, tile_width: Sequence[int] | None = None):
        """Forces the sweep ("planes", "stream", or "hybrid": a run's first round in plane launches, the rest
        streaming) and, when given, the tile width (4, 8 or 16) per size class."""
        tw = None if tile_width is None else np.ascontiguousarray(tile_width, np.int32)
        check(_declare().ts_gridbatch_set_schedule(self._h, METHODS.index(method), ptr(tw)))

    def solve(self, mode: str = "fast_fp32", cancel: C.c_int32 | None = None):
        """(status, reason); rounds per problem in self.rounds (first order, second order)."""
        rounds = np.zeros(2 * self.n, np.int32)
        r = reason_buffer()
        rc = _declare().ts_gridbatch_solve(self._h, MODES[mode], C.byref(cancel) if cancel is not None else None,
                                           ptr(rounds), r, len(r))
        self.rounds = rounds.reshape(self.n, 2)
        return rc, r.value.decode(errors="replace")

    def round_log(self):
        """[(width, method, seconds)] per round of the last solve: width 32 or 64 (the times swept), the method
        that swept the round ("planes" or "stream")."""
        L = _declare()
        n = L.ts_gridbatch_round_log(self._h, 0, None, None)
        check(min(n, 0))
        kind, secs = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1))
        L.ts_gridbatch_round_log(self._h, n, ptr(kind), ptr(secs))
        return [(64 if k & 16 else 32, METHODS[k & 15], float(x)) for k, x in zip(kind[:n], secs[:n])]

    def read(self, q: int) -> np.ndarray:
        out = np.zeros(self.grids[q].count)
        check(_declare().ts_gridbatch_read(self._h, q, ptr(out)))
        return out

    def stats(self):
        n, s, k = C.c_int64(), C.c_double(), C.c_double()
        check(_declare().ts_gridbatch_stats(self._h, C.byref(n), C.byref(s), C.byref(k)))
        u32, u64 = C.c_double(), C.c_double()
        check(_declare().ts_gridbatch_updates(self._h, C.byref(u32), C.byref(u64)))
        return {"nodes": n.value, "seconds": s.value, "sweep_seconds": k.value, "updates_fp32": u32.value,
                "updates_fp64": u64.value}

    def release(self):
        if getattr(self, "_h", None) is not None:
            _declare().ts_gridbatch_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
