"""Pins the oracle's CSR products and LSQR.

Against: SciPy's lsqr at fixed iteration counts (tests/golden/lsqr_scipy.npz, the SOL algorithm the
reference follows and validates against, validation/results/lsqr.json), the reference's closed-form
KATs (NumericsTests.cs:29-85: normal equations of a random 60 x 8 system; damped 2 x 2 gives
x = (2/3, 2/3)), the SIMD-vs-scalar check (NumericsTests.cs:14-27), and the invariant that the
Golub-Kahan vectors do not depend on `damp` (Lsqr.cs:79-92), so the residual history at damp = 0
is shared by construction only through them.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "lsqr_scipy.npz"), allow_pickle=False)


def test_oracle_matches_scipy_iterates(oracle, golden):
    for key in golden["cases"]:
        key = str(key)
        name, d, i = key.split("_")
        damp, iters = float(d[1:]), int(i[1:])
        m, n = golden[f"{name}_shape"]
        x, stop, st, hist = oracle.lsqr(int(m), int(n), golden[f"{name}_rs"], golden[f"{name}_ci"],
                                        golden[f"{name}_va"], golden[f"{name}_b"], damp=damp, max_it=iters,
                                        atol=0, btol=0, conlim=0)
        ref = golden[key + "_x"]
        assert st["iterations"] == iters and stop == 0
        # two FP64 implementations of the same recurrences round differently (SciPy's stable Givens
        # rotation vs sqrt(a^2 + b^2)) and Golub-Kahan amplifies it near convergence: the bar is the
        # reference's own device gate, 1e-8 (LsqrOpenCl.cs:33); the reference reached 2.5e-9 against
        # SciPy (validation/results/lsqr.json:7). Case c is underdetermined (90 x 150): with little or
        # no damping its iterate is ill-determined and two FP64 solvers part by ~1e-7 at 40 iterations.
        tol = 1e-6 if name == "c" and damp < 1 else 1e-8
        assert np.linalg.norm(x - ref) <= tol * np.linalg.norm(ref), key
        # the residual estimate of the recurrences equals the r2norm SciPy reports
        bn = float(np.linalg.norm(golden[f"{name}_b"]))
        assert abs(st["rnorm"] - float(golden[key + "_r2norm"])) <= (1e-7 if name == "c" else 1e-12) * bn


def test_normal_equations_of_a_random_system(oracle):
    # NumericsTests.cs:29-70: 60 x 8 dense system, LSQR reaches the least-squares solution
    rng = np.random.default_rng(7)
    m, n = 60, 8
    dense = rng.random((m, n)) - 0.5
    b = rng.random(m)
    rs = np.arange(0, m * n + 1, n, dtype=np.int64)
    ci = np.tile(np.arange(n, dtype=np.int32), m)
    x, stop, st, _ = oracle.lsqr(m, n, rs, ci, dense.ravel(), b, 0.0, 200, 1e-12, 1e-12)
    exact = np.linalg.solve(dense.T @ dense, dense.T @ b)
    assert np.allclose(x, exact, rtol=0, atol=1e-6)
    assert stop in (1, 2, 4, 5)


def test_damped_two_by_two(oracle):
    # NumericsTests.cs:72-85: (A'A + I) x = A'b with A'A = 2I gives x = (2/3, 2/3)
    rs = np.array([0, 2, 4], np.int64)
    ci = np.array([0, 1, 0, 1], np.int32)
    va = np.array([1.0, 1.0, 1.0, -1.0])
    b = np.array([2.0, 0.0])
    free = oracle.lsqr(2, 2, rs, ci, va, b, 0.0, 50, 1e-12, 1e-12)[0]
    damped = oracle.lsqr(2, 2, rs, ci, va, b, 1.0, 50, 1e-12, 1e-12)[0]
    assert abs(free[0] - 1) < 1e-8
    assert np.allclose(damped, [2 / 3, 2 / 3], atol=1e-6)


def test_simd_kernels_agree_with_scalar_loops(oracle):
    # NumericsTests.cs:14-27, length 1037 so the scalar tail is exercised
    rng = np.random.default_rng(3)
    a, b = rng.random(1037) - 0.5, rng.random(1037) - 0.5
    L = oracle.lib()
    assert abs(L.or_dot(1037, a, b) - float(np.sum(a * b))) < 1e-10
    assert abs(L.or_norm(1037, a) - np.sqrt(np.sum(a * a))) < 1e-10
    tiny = a * 1e-200
    assert abs(L.or_norm(1037, tiny) / 1e-200 - np.sqrt(np.sum(a * a))) < 1e-10


def test_transpose_and_products(oracle):
    rng = np.random.default_rng(11)
    m, n = 70, 40
    dense = np.where(rng.random((m, n)) < 0.15, rng.standard_normal((m, n)), 0.0)
    rows, cols = np.nonzero(dense)
    rs = np.zeros(m + 1, np.int64)
    np.add.at(rs, rows + 1, 1)
    rs = np.cumsum(rs)
    ci, va = cols.astype(np.int32), dense[rows, cols]
    trs, tci, tva = oracle.csr_transpose(m, n, rs, ci, va)
    x, y = rng.standard_normal(n), rng.standard_normal(m)
    assert np.allclose(oracle.spmv(rs, ci, va, x), dense @ x, atol=1e-12)
    assert np.allclose(oracle.spmv(trs, tci, tva, y), dense.T @ y, atol=1e-12)
    # rows of the transpose list the entries of A in increasing row order (SparseMatrix.cs:88-94)
    for c in range(n):
        r = tci[trs[c]:trs[c + 1]]
        assert np.all(np.diff(r) > 0)


def test_zero_rhs_and_empty_system(oracle):
    rs = np.array([0, 1], np.int64)
    x, stop, st, _ = oracle.lsqr(1, 1, rs, np.array([0], np.int32), np.array([2.0]), np.array([0.0]))
    assert stop == 9 and x[0] == 0


def test_planted_mistake_is_caught(oracle, golden):
    """Perturbing one matrix value by 1e-3 moves the iterate far beyond the fixture tolerance."""
    key = "a_d0.05_i40"
    va = golden["a_va"].copy()
    va[17] *= 1.001
    m, n = golden["a_shape"]
    x = oracle.lsqr(int(m), int(n), golden["a_rs"], golden["a_ci"], va, golden["a_b"], damp=0.05, max_it=40,
                    atol=0, btol=0, conlim=0)[0]
    ref = golden[key + "_x"]
    assert np.linalg.norm(x - ref) > 1e-8 * np.linalg.norm(ref)


def test_device_dot_product_order(oracle):
    """or_dot_device restates LsqrOpenCl's dot_partial (LsqrOpenCl.cs:79-94): exact on integers in any
    order, and on random values equal to a direct emulation of 64 groups x 256 strided item sums, each
    group halved in a tree, the group sums added in order - and different from SimdVector's order."""
    L = oracle.lib()
    rng = np.random.default_rng(5)
    a = rng.integers(-1000, 1000, 50_000).astype(np.float64)
    b = rng.integers(-1000, 1000, 50_000).astype(np.float64)
    assert L.or_dot_device(a.size, a, b) == float(np.dot(a.astype(np.int64), b.astype(np.int64)))
    x, y = rng.standard_normal(70_001), rng.standard_normal(70_001)
    part = []
    for g in range(64):
        red = [0.0] * 256
        for lane in range(256):
            s = 0.0
            for i in range(g * 256 + lane, x.size, 64 * 256):
                s += x[i] * y[i]
            red[lane] = s
        h = 128
        while h:
            for lane in range(h):
                red[lane] += red[lane + h]
            h >>= 1
        part.append(red[0])
    want = 0.0
    for p in part:
        want += p
    assert L.or_dot_device(x.size, x, y) == want
    assert L.or_dot(x.size, x, y) != want
