"""The oracle's resolution-diagonal estimate (ResolutionDiagonal.Estimate restated) against closed
forms. For A = [G; lam I] with G = diag(g) the resolution matrix R = (G^2 + lam^2)^-1 G^2 is
diagonal, so every probe gives z * R z = diag(R) exactly: the estimate is g^2 / (g^2 + lam^2) with
zero standard error (up to the rounding of its formula). With no regularisation rows and a full-rank G, R = I. A planted mistake (the
regularisation rows fed the probe's data instead of zero) moves the first case off its closed form."""
import numpy as np


def _diag_system(g, lam):
    n = g.size
    rows = [([i], [g[i]]) for i in range(n)] + [([i], [lam]) for i in range(n)]
    rs = np.concatenate([[0], np.cumsum([len(c) for c, _ in rows])]).astype(np.int64)
    ci = np.concatenate([c for c, _ in rows]).astype(np.int32)
    va = np.concatenate([v for _, v in rows]).astype(np.float64)
    return 2 * n, n, rs, ci, va


def test_diagonal_system_closed_form(oracle):
    rng = np.random.default_rng(0)
    g = rng.uniform(0.2, 3.0, 40)
    lam = 0.7
    m, n, rs, ci, va = _diag_system(g, lam)
    signs = np.where(rng.integers(0, 2, (6, n)) == 0, -1.0, 1.0)
    d, err, its = oracle.resolution_diagonal(m, n, rs, ci, va, n, signs, 200, 1e-14)
    assert np.allclose(d, g * g / (g * g + lam * lam), rtol=1e-9, atol=1e-12)
    # zero up to the formula's own rounding: sqrt of a difference of equal sums of squares
    assert np.all(err < 3e-8)
    # planted mistake: the regularisation rows keep the probe's data
    tr = oracle.csr_transpose(m, n, rs, ci, va)
    x = oracle.lsqr(m, n, rs, ci, va, oracle.spmv(rs, ci, va, signs[0]), 0.0, 200, 1e-14, 1e-14, 1e8, 0, tr)[0]
    assert not np.allclose(signs[0] * x, g * g / (g * g + lam * lam), rtol=1e-3)


def test_full_rank_without_regularisation_is_identity(oracle):
    rng = np.random.default_rng(1)
    m, n = 60, 12
    G = rng.standard_normal((m, n))
    rs = np.arange(0, m * n + 1, n).astype(np.int64)
    ci = np.tile(np.arange(n), m).astype(np.int32)
    va = G.ravel()
    signs = np.where(rng.integers(0, 2, (3, n)) == 0, -1.0, 1.0)
    d, err, its = oracle.resolution_diagonal(m, n, rs, ci, va, m, signs, 200, 1e-14)
    assert np.allclose(d, 1.0, atol=1e-8)
    assert np.all(err < 3e-8)
