"""Writes tests/golden/lsqr_scipy.npz: LSQR solutions from SciPy's scipy.sparse.linalg.lsqr (SciPy 1.15,
the SOL algorithm the reference's Lsqr.Solve follows, Numerics/Lsqr.cs:7-30; the reference validates its
solver against it, validation/v02_lsqr.py). Stopping tests are disabled (atol = btol = 0, conlim = 0),
so every solve runs exactly `iters` iterations and the fixture pins the iterate itself.

Run: python tests/golden/make_lsqr_golden.py
"""
import os

import numpy as np
import scipy
import scipy.sparse as sp
from scipy.sparse.linalg import lsqr

HERE = os.path.dirname(os.path.abspath(__file__))


def system(rows, cols, per_row, seed):
    """A random sparse system with a few non-zeros per row, as a ray matrix has (the shape of
    LsqrOpenCl.TestSystem, LsqrOpenCl.cs:558-573, drawn with numpy's PCG64)."""
    rng = np.random.default_rng(seed)
    rs, ci, va = [0], [], []
    for _ in range(rows):
        c = np.sort(rng.choice(cols, per_row, replace=False))
        ci.extend(c.tolist())
        va.extend((rng.random(per_row) + 0.1).tolist())
        rs.append(len(ci))
    b = rng.random(rows) - 0.5
    return np.array(rs, np.int64), np.array(ci, np.int32), np.array(va), b


def main():
    out = {"scipy_version": np.array(scipy.__version__)}
    cases = []
    for name, (rows, cols, per_row, seed) in {"a": (300, 120, 12, 17), "b": (500, 400, 6, 3),
                                              "c": (90, 150, 9, 5)}.items():
        rs, ci, va, b = system(rows, cols, per_row, seed)
        A = sp.csr_matrix((va, ci, rs), shape=(rows, cols))
        out[f"{name}_rs"], out[f"{name}_ci"], out[f"{name}_va"], out[f"{name}_b"] = rs, ci, va, b
        out[f"{name}_shape"] = np.array([rows, cols])
        for damp in (0.0, 0.05, 1.0):
            for iters in (10, 40):
                r = lsqr(A, b, damp=damp, atol=0, btol=0, conlim=0, iter_lim=iters)
                key = f"{name}_d{damp}_i{iters}"
                out[key + "_x"] = r[0]
                out[key + "_r1norm"] = np.array(r[3])  # ||b - A x||
                out[key + "_r2norm"] = np.array(r[4])  # sqrt(||b - A x||^2 + damp^2 ||x||^2)
                cases.append(key)
    out["cases"] = np.array(cases)
    np.savez_compressed(os.path.join(HERE, "lsqr_scipy.npz"), **out)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
