"""Parity of the device LSQR with the oracle's FP64 Lsqr.Solve (Numerics/Lsqr.cs:34-152) and SciPy's lsqr
fixtures, through the C ABI.

Bars: the plain FP64 form agrees with the oracle to 1e-8 relative (the reference's device gate,
LsqrOpenCl.cs:33) and reproduces its iteration counts and stop reasons; the fast FP32 form (k
right-hand sides, per-RHS damping) agrees with the FP64 oracle at equal iteration count to 1e-5
relative L2 in every model, the oracle solving the same FP32-valued matrix.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FAST_TOL = 1e-5


def ray_like_system(rows, cols, per_row, fd_block, seed):
    """Rows with `per_row` clustered non-zeros (a ray crossing nearby nodes) and, when fd_block, one
    first-difference row per column pair (a smoothing block), as the tomography matrix has."""
    rng = np.random.default_rng(seed)
    rs, ci, va = [0], [], []
    for _ in range(rows):
        start = rng.integers(0, cols - per_row * 3)
        c = np.unique(start + rng.choice(per_row * 3, per_row, replace=False))
        ci.extend(c.tolist())
        va.extend((rng.random(c.size) * 0.9 + 0.1).tolist())
        rs.append(len(ci))
    if fd_block:
        for c in range(cols - 1):
            ci.extend([c, c + 1])
            va.extend([-fd_block, fd_block])
            rs.append(len(ci))
    m = len(rs) - 1
    b = np.concatenate([rng.standard_normal(rows), np.zeros(m - rows)])
    return m, cols, np.array(rs, np.int32), np.array(ci, np.int32), np.array(va), b


def _oracle(oracle, m, n, rs, ci, va, b, damp, iters, **kw):
    return oracle.lsqr(m, n, rs.astype(np.int64), ci, va, b, damp=damp, max_it=iters, **kw)


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "lsqr_scipy.npz"), allow_pickle=False)


def test_selftest_gates(device):
    from tomostar_amd.lsqr import LsqrHip

    hip = LsqrHip(device)
    plain, fast = hip.self_test()
    assert plain <= 1e-8 and fast <= FAST_TOL


def test_plain_fp64_matches_oracle_and_scipy(device, oracle, golden):
    from tomostar_amd import lsqr as Q

    for name in ("a", "b", "c"):
        m, n = (int(v) for v in golden[f"{name}_shape"])
        rs, ci, va, b = golden[f"{name}_rs"], golden[f"{name}_ci"], golden[f"{name}_va"], golden[f"{name}_b"]
        A = Q.Matrix(device, m, n, rs.astype(np.int32), ci, va)
        for damp in (0.0, 0.05, 1.0):
            rc, why, res = Q.solve_many(A, b[None, :], [damp], 40, 0, 0, 0, mode="plain_fp64")
            assert rc == 0, why
            x = res[0].solution
            ref = _oracle(oracle, m, n, rs, ci, va, b, damp, 40, atol=0, btol=0, conlim=0)[0]
            tol = 1e-6 if name == "c" and damp < 1 else 1e-8
            assert np.linalg.norm(x - ref) <= tol * np.linalg.norm(ref)
            sc = golden[f"{name}_d{damp}_i40_x"]
            assert np.linalg.norm(x - sc) <= tol * np.linalg.norm(sc)


def test_plain_stop_rules_follow_the_oracle(device, oracle):
    from tomostar_amd import lsqr as Q

    m, n, rs, ci, va, b = ray_like_system(400, 150, 10, 0.5, 1)
    A = Q.Matrix(device, m, n, rs, ci, va)
    for damp, atol, conlim in ((0.0, 1e-8, 1e8), (0.3, 1e-6, 1e8), (0.0, 1e-12, 50.0)):
        rc, why, res = Q.solve_many(A, b[None, :], [damp], 300, atol, atol, conlim, mode="plain_fp64", history=True)
        assert rc == 0, why
        x, stop, st, hist = _oracle(oracle, m, n, rs, ci, va, b, damp, 300, atol=atol, btol=atol, conlim=conlim)
        r = res[0]
        # the same stop rule fires; a threshold crossed within rounding may be crossed one iteration
        # apart, as the device stores the rows in another order and reduces in another order
        assert r.stop_code == stop and abs(r.iterations - st["iterations"]) <= 1
        # dot products reduce in a different order on the device (64 partials) and on the host (SIMD
        # lanes): once the Golub-Kahan vectors lose orthogonality the two runs follow different (equally
        # valid) paths whose residual estimates part by up to ~1e-6 before both converge
        common = min(r.iterations, st["iterations"])
        assert np.allclose(r.residual_history[:common], hist[:common], rtol=1e-5, atol=0)
        # the two paths meet at the same solution to the accuracy the stop rule asked for, or to the
        # spread rounding alone causes: an early stop (conlim 50: 20 iterations) leaves an iterate that
        # moves by 1e-3 when b is perturbed by 1e-15 - measured here on the oracle itself (both at the
        # device's iteration count when the counts differ)
        bp = b * (1 + 1e-15 * np.random.default_rng(3).standard_normal(b.size))
        if r.iterations != st["iterations"]:
            x = _oracle(oracle, m, n, rs, ci, va, b, damp, r.iterations, atol=0, btol=0, conlim=0)[0]
            xp = _oracle(oracle, m, n, rs, ci, va, bp, damp, r.iterations, atol=0, btol=0, conlim=0)[0]
        else:
            xp = _oracle(oracle, m, n, rs, ci, va, bp, damp, 300, atol=atol, btol=atol, conlim=conlim)[0]
        spread = np.linalg.norm(xp - x) / np.linalg.norm(x)
        assert np.linalg.norm(r.solution - x) <= max(1e-7, 10 * atol, 20 * spread) * np.linalg.norm(x)


@pytest.mark.parametrize("mode,tol", [("fast_fp32", FAST_TOL), ("fast_mixed", 1e-8)])
@pytest.mark.parametrize("k", [1, 2, 3, 5, 32, 33])
def test_fast_many_rhs_against_fp64_oracle(device, oracle, k, mode, tol):
    """FP32 matrix values: the oracle runs on the same rounded values. FP32 vectors stay within 1e-5,
    FP64 vectors (mixed) within 1e-8 of the FP64 oracle at equal iteration counts."""
    from tomostar_amd import lsqr as Q

    m, n, rs, ci, va, b0 = ray_like_system(1500, 600, 24, 1.0, 7)
    rng = np.random.default_rng(k)
    B = b0[None, :] + 0.05 * rng.standard_normal((k, m)) * (np.arange(m) < 1500)
    damps = 10 ** (-1 + 3 * np.arange(k) / max(1, k - 1))  # log-spaced 0.1 - 100
    A = Q.Matrix(device, m, n, rs, ci, va, keep_fp64=False)
    rc, why, res = Q.solve_many(A, B, damps, 200, 0, 0, 0, mode=mode)
    assert rc == 0, why
    va32 = va.astype(np.float32).astype(np.float64)
    picks = sorted(set([0, k // 2, k - 1]))
    for j in picks:
        # with atol = btol = conlim = 0 only the machine-precision rules (Lsqr.cs:144-146) can stop a
        # solve; the FP32 form reaches them earlier than FP64 does, so the oracle runs the same count
        it = res[j].iterations
        assert it == 200 or res[j].stop_code in (4, 5, 6)
        ref = _oracle(oracle, m, n, rs, ci, va32, B[j], damps[j], it, atol=0, btol=0, conlim=0)[0]
        err = np.linalg.norm(res[j].solution - ref) / np.linalg.norm(ref)
        assert err <= tol, (j, damps[j], err)


def test_fast_stop_rules_and_frozen_rhs(device, oracle):
    """RHS that stop early keep their solution while the others go on."""
    from tomostar_amd import lsqr as Q

    m, n, rs, ci, va, b0 = ray_like_system(800, 300, 16, 0.8, 3)
    B = np.stack([b0, 2 * b0, b0[::-1].copy()])
    damps = [0.0, 5.0, 50.0]
    A = Q.Matrix(device, m, n, rs, ci, va)
    rc, why, res = Q.solve_many(A, B, damps, 400, 1e-6, 1e-6, 1e8, mode="fast_fp32")
    assert rc == 0, why
    va32 = va.astype(np.float32).astype(np.float64)
    for j in range(3):
        x, stop, st, _ = _oracle(oracle, m, n, rs, ci, va32, B[j], damps[j], 400, atol=1e-6, btol=1e-6, conlim=1e8)
        assert stop == res[j].stop_code
        # FP32 vectors take a few more iterations to meet atol = 1e-6 than FP64 ones
        assert abs(res[j].iterations - st["iterations"]) <= max(3, st["iterations"] // 5)
        assert np.linalg.norm(res[j].solution - x) <= 1e-4 * np.linalg.norm(x)
    # the RHS that stopped first kept its solution while the others went on
    order = np.argsort([r.iterations for r in res])
    assert res[order[0]].iterations < res[order[-1]].iterations


def test_degenerate_systems(device):
    from tomostar_amd import _lib
    from tomostar_amd import lsqr as Q

    m, n, rs, ci, va, b = ray_like_system(50, 40, 5, 0.0, 2)
    A = Q.Matrix(device, m, n, rs, ci, va)
    for mode in ("plain_fp64", "fast_fp32", "fast_mixed"):
        rc, why, res = Q.solve_many(A, np.zeros((2, m)), [0.0, 1.0], 20, mode=mode)
        assert rc == 0 and all(r.stop_code == 9 and r.iterations == 0 for r in res)
        assert all(np.all(r.solution == 0) for r in res)
    # A^T b = 0: b lives on a row with no entries
    rs2 = np.array([0, 2, 2], np.int32)
    A2 = Q.Matrix(device, 2, 3, rs2, np.array([0, 2], np.int32), np.array([1.0, 2.0]))
    for mode in ("plain_fp64", "fast_fp32", "fast_mixed"):
        rc, why, res = Q.solve_many(A2, np.array([[0.0, 3.0]]), [0.0], 20, mode=mode)
        assert rc == 0 and res[0].stop_code == 10 and res[0].residual_norm == 3.0
    # empty system
    A3 = Q.Matrix(device, 0, 5, np.array([0], np.int32), np.zeros(0, np.int32), np.zeros(0))
    rc, why, res = Q.solve_many(A3, np.zeros((1, 0)), [0.0], 20)
    assert rc == 0 and res[0].stop_code == 8
    # invalid: k > 64, cancellation
    rc, why, _ = Q.solve_many(A, np.ones((65, m)), np.zeros(65), 5)
    assert rc == _lib.TS_EINVAL
    flag = C.c_int32(1)
    rc, why, _ = Q.solve_many(A, b[None, :], [0.0], 50, cancel=flag)
    assert rc == _lib.TS_CANCELLED
    with pytest.raises(_lib.HipError):
        Q.Matrix(device, 2, 3, np.array([0, 1, 2], np.int32), np.array([0, 7], np.int32), np.array([1.0, 1.0]))


def test_two_by_two_damped_known_answer(device):
    from tomostar_amd import lsqr as Q

    A = Q.Matrix(device, 2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                 np.array([1.0, 1.0, 1.0, -1.0]))
    b = np.array([[2.0, 0.0], [2.0, 0.0]])
    for mode, tol in (("plain_fp64", 1e-8), ("fast_fp32", 1e-6), ("fast_mixed", 1e-8)):
        rc, why, res = Q.solve_many(A, b, [0.0, 1.0], 50, 1e-12, 1e-12, mode=mode)
        assert rc == 0, why
        assert abs(res[0].solution[0] - 1) < tol and np.allclose(res[1].solution, [2 / 3, 2 / 3], atol=tol)
