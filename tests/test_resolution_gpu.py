"""The device product A x / A^T x (ts_matrix_multiply, CsrMatrix.MultiplyInto) and the resolution-
diagonal estimate (ResolutionDiagonal.Estimate, Tomography/ResolutionDiagonal.cs:31-73) with its
probes' LSQR solves batched into multi-RHS solves.

Bars: A x equals the oracle's FP64 product to the last bit (same entries, same order; rows and
columns of an assembled system come back in the given order), A^T x within 1e-13 (its rows sum in the
stored row order); the estimate matches the closed form g^2 / (g^2 + lam^2) of a diagonal system to
1e-8, and the oracle's estimate with the same probes on the 61 x 61 x 31 iteration system within 1e-5
relative L2 at 20 LSQR iterations (before the iterates' rounding drift) - in batches of 3 probes as
in one batch, so batching changes nothing but the number of passes over A.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_csr(m, n, per_row, seed):
    rng = np.random.default_rng(seed)
    rs, ci, va = [0], [], []
    for _ in range(m):
        c = np.sort(rng.choice(n, per_row, replace=False))
        ci.extend(c.tolist())
        va.extend(rng.standard_normal(per_row).tolist())
        rs.append(len(ci))
    return np.array(rs, np.int32), np.array(ci, np.int32), np.array(va)


def test_multiply_matches_oracle(device, oracle):
    from tomostar_amd import lsqr as Q

    m, n = 700, 300
    rs, ci, va = _random_csr(m, n, 9, 4)
    A = Q.Matrix(device, m, n, rs, ci, va, keep_fp64=True)
    X = np.random.default_rng(5).standard_normal((3, n))
    Y = A.multiply(X)
    for j in range(3):
        assert np.array_equal(Y[j], oracle.spmv(rs, ci, va, X[j]))
    U = np.random.default_rng(6).standard_normal((2, m))
    trs, tci, tva = oracle.csr_transpose(m, n, rs.astype(np.int64), ci, va)
    V = A.multiply(U, transpose=True)
    for j in range(2):
        ref = oracle.spmv(trs, tci, tva, U[j])
        assert np.allclose(V[j], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    with pytest.raises(ValueError):
        A.multiply(np.zeros(n + 1))
    A.release()


@pytest.fixture(scope="module")
def iteration(device):
    from tomostar_amd import iteration as It
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    t_obs = It.observed_times(device, wl)
    return It.run(device, wl, t_obs, lsqr_iterations=5)


def test_multiply_on_an_assembled_system(device, oracle, iteration):
    """Rows and columns of an assembled system are stored reordered; the product comes back in the
    given orders, equal to the oracle's product over the system read back."""
    r = iteration
    rs, ci, va = r.system.read()
    x = np.random.default_rng(8).standard_normal(r.columns)
    assert np.array_equal(r.system.matrix.multiply(x), oracle.spmv(rs, ci, va, x))


def test_diagonal_system_closed_form(device):
    from tomostar_amd import lsqr as Q
    from tomostar_amd import resolution as R

    rng = np.random.default_rng(0)
    g = rng.uniform(0.2, 3.0, 50)
    lam = 0.7
    n = g.size
    rs = np.arange(2 * n + 1, dtype=np.int32)
    ci = np.concatenate([np.arange(n), np.arange(n)]).astype(np.int32)
    va = np.concatenate([g, np.full(n, lam)])
    A = Q.Matrix(device, 2 * n, n, rs, ci, va, keep_fp64=False)
    res = R.estimate(A, n, 8, 200, 1e-14, seed=3)
    va32 = va.astype(np.float32).astype(np.float64)
    g32, lam32 = va32[:n], va32[n]
    assert res.probes == 8
    assert np.allclose(res.diagonal, g32 * g32 / (g32 * g32 + lam32 * lam32), rtol=1e-8, atol=1e-10)
    assert np.all(res.standard_error < 3e-8)
    A.release()


def test_estimate_matches_oracle_on_the_iteration_system(device, oracle, iteration):
    from tomostar_amd import resolution as R

    r = iteration
    rs, ci, va = r.system.read()
    probes, its = 6, 20
    signs = R.probe_signs(r.columns, probes, seed=9)
    va32 = va.astype(np.float32).astype(np.float64)
    ref, ref_err, ref_its = oracle.resolution_diagonal(r.rows, r.columns, rs, ci, va32, r.data_rows, signs, its, 0.0)
    one = R.estimate(r.system.matrix, r.data_rows, probes, its, 0.0, seed=9, signs=signs)
    three = R.estimate(r.system.matrix, r.data_rows, probes, its, 0.0, seed=9, signs=signs, batch=3)
    for res in (one, three):
        assert res.mean_lsqr_iterations == its == ref_its
        d = np.linalg.norm(res.diagonal - ref) / np.linalg.norm(ref)
        e = np.linalg.norm(res.standard_error - ref_err) / np.linalg.norm(ref_err)
        print(f"resolution diagonal vs oracle: {d:.2e}, standard error {e:.2e}")
        assert d <= 1e-5 and e <= 1e-5
