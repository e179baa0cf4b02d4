"""The formal resolution oracle (FormalResolution.Compute, Tomography/FormalResolution.cs:108-333) against
what the mathematics fixes:

- damped least squares A = [G; lambda I]: with G = U s V^T, R = V s^2/(s^2 + lambda^2) V^T and
  C = V 1/(s^2 + lambda^2) V^T (the SVD closed forms);
- parameter separation: with the hypocentre groups eliminated, C_kk is the kept block of the full
  inverse N^-1 and R_kk = I - (N^-1)_kk Gamma_kk (the block-inverse identity);
- resolution lengths: zero when R is diagonal; the two-column closed form d |R_12| / |R_1|;
- the declines: too many unknowns, rows coupling two groups or a regularisation row coupling model
  and hypocentre columns, a column with neither data nor damping.
"""
import math

import numpy as np
import pytest


def _csr(A):
    rs, ci, va = [0], [], []
    for row in A:
        nz = np.flatnonzero(row)
        ci.extend(nz.tolist())
        va.extend(row[nz].tolist())
        rs.append(len(ci))
    return np.array(rs, np.int32), np.array(ci, np.int32), np.array(va)


def test_damped_least_squares_closed_form(oracle):
    rng = np.random.default_rng(0)
    m, n, lam = 40, 25, 0.7
    G = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)
    A = np.vstack([G, lam * np.eye(n)])
    rs, ci, va = _csr(A)
    kept, diag, sd, _ = oracle.formal_resolution(A.shape[0], n, rs, ci, va, m)
    U, s, Vt = np.linalg.svd(G, full_matrices=True)
    s_full = np.zeros(n)
    s_full[:s.size] = s
    V = Vt.T
    R = (V * (s_full ** 2 / (s_full ** 2 + lam ** 2))) @ V.T
    Cm = (V * (1 / (s_full ** 2 + lam ** 2))) @ V.T
    assert np.array_equal(kept, np.arange(n))
    assert np.allclose(diag, np.diag(R), rtol=1e-10, atol=1e-13)
    assert np.allclose(sd, np.sqrt(np.diag(Cm)), rtol=1e-10)


def _separable(rng, nmodel=30, nev=4, lam=0.5, hyp_damp=0.2):
    """Data rows touching the model and one event's 4 columns; damping of the model; a Laplacian-like
    difference row on the model; damping of the hypocentres (a row that touches them only)."""
    ncols = nmodel + 4 * nev
    rows = []
    for e in range(nev):
        for _ in range(12):
            r = np.zeros(ncols)
            idx = rng.choice(nmodel, 6, replace=False)
            r[idx] = rng.random(6) + 0.1
            r[nmodel + 4 * e:nmodel + 4 * e + 4] = rng.standard_normal(4)
            rows.append(r)
    data = len(rows)
    for c in range(nmodel):
        r = np.zeros(ncols)
        r[c] = lam
        rows.append(r)
    for c in range(nmodel - 1):
        r = np.zeros(ncols)
        r[c], r[c + 1] = -0.3, 0.3
        rows.append(r)
    for c in range(nmodel, ncols):
        r = np.zeros(ncols)
        r[c] = hyp_damp
        rows.append(r)
    groups = [list(range(nmodel + 4 * e, nmodel + 4 * e + 4)) for e in range(nev)]
    return np.array(rows), data, groups, nmodel


def test_parameter_separation_is_the_block_inverse(oracle):
    rng = np.random.default_rng(1)
    A, data, groups, nmodel = _separable(rng)
    rs, ci, va = _csr(A)
    kept, diag, sd, _ = oracle.formal_resolution(A.shape[0], A.shape[1], rs, ci, va, data, groups)
    assert np.array_equal(kept, np.arange(nmodel))
    N = A.T @ A
    Ninv = np.linalg.inv(N)
    Areg = A[data:, :nmodel]
    gamma = Areg.T @ Areg
    assert np.allclose(sd, np.sqrt(np.diag(Ninv)[:nmodel]), rtol=1e-9)
    assert np.allclose(diag, np.diag(np.eye(nmodel) - Ninv[:nmodel, :nmodel] @ gamma), rtol=1e-9, atol=1e-12)


def test_resolution_lengths(oracle):
    # orthogonal columns: R is diagonal, every length is 0
    A = np.vstack([np.diag([1.0, 2.0, 3.0]), 0.5 * np.eye(3)])
    rs, ci, va = _csr(A)
    xyz = np.array([[0.0, 0, 0], [10.0, 0, 0], [0, 7.0, 0]])
    _, _, _, length = oracle.formal_resolution(6, 3, rs, ci, va, 3, places=(xyz, np.zeros(3, int)))
    assert np.allclose(length, 0.0, atol=1e-12)
    # two columns, one data row [1, 1], damping lam: N = [[1 + l2, 1], [1, 1 + l2]], Gamma = l2 I
    lam, d = 0.8, 5.0
    A = np.array([[1.0, 1.0], [lam, 0.0], [0.0, lam]])
    rs, ci, va = _csr(A)
    xyz = np.array([[0.0, 0, 0], [d, 0, 0]])
    _, diag, _, length = oracle.formal_resolution(3, 2, rs, ci, va, 1, places=(xyz, np.zeros(2, int)))
    l2 = lam * lam
    det = (1 + l2) ** 2 - 1
    r11 = 1 - l2 * (1 + l2) / det
    r12 = l2 / det
    assert math.isclose(diag[0], r11, rel_tol=1e-12)
    assert math.isclose(length[0], d * abs(r12) / math.hypot(r11, r12), rel_tol=1e-12)
    # columns of other blocks do not count
    _, _, _, length = oracle.formal_resolution(3, 2, rs, ci, va, 1, places=(xyz, np.array([0, 1])))
    assert np.allclose(length, 0.0)


def test_declines(oracle):
    rng = np.random.default_rng(2)
    A, data, groups, nmodel = _separable(rng)
    rs, ci, va = _csr(A)
    with pytest.raises(ValueError, match="too many"):
        oracle.formal_resolution(A.shape[0], A.shape[1], rs, ci, va, data, groups, max_unknowns=nmodel - 1)
    B = A.copy()
    B[0, nmodel + 4:nmodel + 8] = 1.0  # a data row touching two events
    rs2, ci2, va2 = _csr(B)
    with pytest.raises(ValueError, match="two events"):
        oracle.formal_resolution(B.shape[0], B.shape[1], rs2, ci2, va2, data, groups)
    B = A.copy()
    B[data, nmodel] = 1.0  # a damping row of the model touching a hypocentre column
    rs2, ci2, va2 = _csr(B)
    with pytest.raises(ValueError, match="couples model"):
        oracle.formal_resolution(B.shape[0], B.shape[1], rs2, ci2, va2, data, groups)
    C = np.vstack([np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]])])  # column 2: no data, no damping
    rs3, ci3, va3 = _csr(C)
    with pytest.raises(ValueError, match="positive definite"):
        oracle.formal_resolution(2, 3, rs3, ci3, va3, 2)
