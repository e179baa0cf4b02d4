"""Diagonal of the model resolution matrix by random probes (ResolutionDiagonal.Estimate,
Tomography/ResolutionDiagonal.cs:31-73), with the probes' LSQR solves batched into one multi-RHS
solve on the device.

For a probe z of independent +-1 entries, R z is the LSQR solution of A m = [W G z; 0] (the data the
model z would produce on the data rows, zero on the regularisation rows); diag(R) is estimated by the
mean of z * R z over the probes, with the standard error of that mean (Bekas, Kokiopoulou & Saad 2007).
The reference solves the probes one after another; here up to 64 of them share every pass over A and
A^T (the fast multi-RHS form), since they share the matrix. The probes come from numpy's generator
with the given seed: the reference's are drawn from .NET System.Random, which is not reproduced.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import lsqr as Q


@dataclass
class ResolutionDiagonalResult:
    """ResolutionDiagonalResult (ResolutionDiagonal.cs:9)."""
    diagonal: np.ndarray
    standard_error: np.ndarray
    probes: int
    mean_lsqr_iterations: int


def probe_signs(n: int, probes: int, seed: int) -> np.ndarray:
    """probes x n entries of +-1 (the reference: rng.Next(2) == 0 ? -1 : 1, ResolutionDiagonal.cs:46)."""
    rng = np.random.default_rng(seed)
    return np.where(rng.integers(0, 2, size=(probes, n)) == 0, -1.0, 1.0)


def estimate(a: Q.Matrix, data_rows: int, probes: int, lsqr_iterations: int, tolerance: float, seed: int,
             mode: str | None = None, batch: int = 64, signs: np.ndarray | None = None) -> ResolutionDiagonalResult:
    """ResolutionDiagonal.Estimate on the device: probes >= 2 (as the reference clamps), LSQR with damp 0,
    atol = btol = tolerance, conlim 1e8 per probe, up to `batch` probes per multi-RHS solve. `signs`
    (probes x cols of +-1) replaces the generated probes."""
    if data_rows <= 0 or data_rows > a.rows:
        raise ValueError("data_rows out of range")
    probes = max(2, int(probes))
    n = a.cols
    Z = probe_signs(n, probes, seed) if signs is None else np.asarray(signs, np.float64)
    if Z.shape != (probes, n):
        raise ValueError("signs must be probes x cols")
    s1 = np.zeros(n)
    s2 = np.zeros(n)
    iterations = 0
    for p0 in range(0, probes, batch):
        z = np.ascontiguousarray(Z[p0:p0 + batch])
        rhs = Q.multiply(a, z)
        rhs[:, data_rows:] = 0.0  # regularisation rows ask for zero
        rc, why, res = Q.solve_many(a, rhs, np.zeros(z.shape[0]), lsqr_iterations, tolerance, tolerance, 1e8, mode=mode)
        if rc != 0:
            raise RuntimeError(f"LSQR declined: {why}")
        for j, r in enumerate(res):
            e = z[j] * r.solution
            s1 += e
            s2 += e * e
            iterations += r.iterations
    mean = s1 / probes
    variance = np.maximum(0.0, (s2 - probes * mean * mean) / (probes - 1))
    return ResolutionDiagonalResult(mean, np.sqrt(variance / probes), probes, iterations // probes)
