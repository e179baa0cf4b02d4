"""The Q tomography oracle (Attenuation/QTomography.cs) pinned by what the mathematics fixes, on
synthetic ray rows (ascending node lists, float32 lengths) over small grids:

- Predict and the path time: closed forms in a uniform medium at uniform q, linearity in q;
- the reference Q from the data: on t* = q T exactly, the weighted fit (station terms) and the
  median of the path averages (without) both return q, clamped to [1/QMax, 1/QMin];
- the data rows: consistent with the model that made the data - at q0 for the unknowns q/q0 - 1 with
  station terms, and at the true model for the Gauss-Newton rows of ln(q/q0), whose right-hand side
  carries the J m term a dropped term would break;
- the sensitivity scales: exact powers of two in sigma and q0 scale them as the RMS of the weighted
  entries must;
- the regularisation: damping rows, Laplacian rows that annihilate constants, the zero-sum row;
- the Gauss-Newton solve recovers ln(q/q0) from noise-free data with negligible regularisation, the
  linear one recovers q/q0 - 1;
- the outlier rule removes exactly the planted outliers.
"""
import math

import numpy as np
import pytest


def _grid(oracle, nx=6, ny=5, nz=4):
    return oracle.griddef(13.0, 13.3, 43.0, 43.2, 0.0, 6.0, nx, ny, nz)


def _rows(n, nrows, per_row, seed, cover=False):
    """Ray-like rows: ascending unique nodes, float32 lengths in (0.05, 1.5) km."""
    rng = np.random.default_rng(seed)
    rs, nodes, lens = [0], [], []
    for r in range(nrows):
        k = min(n, per_row if not cover else max(per_row, 1))
        c = np.sort(rng.choice(n, k, replace=False))
        if cover:  # every node on some row
            c = np.unique(np.concatenate([c, [r % n]]))
        nodes.extend(c.tolist())
        lens.extend((0.05 + 1.45 * rng.random(c.size)).tolist())
        rs.append(len(nodes))
    return np.array(rs, np.int64), np.array(nodes, np.int32), np.array(lens, np.float32)


def _slowness(n, seed):
    rng = np.random.default_rng(seed)
    return 1.0 / (5.0 + rng.random(n))


def test_predict_closed_forms_and_linearity(oracle):
    n = 60
    rs, nodes, lens = _rows(n, 40, 9, 1)
    s = np.full(n, 0.2)
    pred, tt = oracle.q_predict(rs, nodes, lens, s, None, 1 / 250)
    total = np.add.reduceat(lens.astype(np.float64), rs[:-1])
    assert np.allclose(tt, 0.2 * total, rtol=1e-14, atol=0)
    assert np.allclose(pred, 0.2 * total / 250, rtol=1e-14, atol=0)
    rng = np.random.default_rng(2)
    s = _slowness(n, 3)
    q1, q2 = rng.random(n) * 1e-2, rng.random(n) * 1e-2
    p1, _ = oracle.q_predict(rs, nodes, lens, s, q1, 1.0)
    p2, _ = oracle.q_predict(rs, nodes, lens, s, q2, 1.0)
    p12, _ = oracle.q_predict(rs, nodes, lens, s, q1 + q2, 1.0)
    assert np.allclose(p12, p1 + p2, rtol=1e-13, atol=0)


def _setup(oracle, g, nrows=80, per_row=10, nst=6, seed=0, cover=False):
    n = g.count
    rs, nodes, lens = _rows(n, nrows, per_row, seed, cover)
    rng = np.random.default_rng(seed + 10)
    station = rng.integers(0, nst, nrows).astype(np.int32)
    sigma = 0.002 + 0.01 * rng.random(nrows)
    return n, rs, nodes, lens, station, sigma, _slowness(n, seed + 20)


@pytest.mark.parametrize("station_terms", [True, False])
def test_reference_q_from_the_data(oracle, station_terms):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g)
    for q_true, expect in ((1 / 180.0, 1 / 180.0), (1 / 5.0, 1 / 10.0), (1 / 9000.0, 1 / 5000.0)):
        _, tt = oracle.q_predict(rs, nodes, lens, s, None, 1.0)
        tstar = q_true * tt
        O = oracle.QOracle(g, dict(station_terms=station_terms))
        O.prepare(rs, nodes, lens, np.arange(tt.size), tstar, sigma, station, 6, s)
        assert math.isclose(O.q0, expect, rel_tol=1e-13)
    # an even number of path averages: the mean of the middle two
    tstar = np.array([1.0, 2.0, 4.0, 3.0]) * 1e-3 * tt[:4]
    O = oracle.QOracle(g, dict(station_terms=False))
    O.prepare(rs[:5], nodes, lens, np.arange(4), tstar, sigma[:4], station[:4], 6, s)
    assert math.isclose(O.q0, 2.5e-3, rel_tol=1e-13)


def _data_block(oracle, g, opts, rs, nodes, lens, station, tstar, sigma, s, **kw):
    grs, gci, gva, rhs, dims, scales = oracle.build_qsystem(g, opts, rs, nodes, lens, station, tstar, sigma, s, **kw)
    nd = dims[3]
    A = np.zeros((nd, dims[1]))
    for r in range(nd):
        A[r, gci[grs[r]:grs[r + 1]]] = gva[grs[r]:grs[r + 1]]
    return A, rhs[:nd], (grs, gci, gva, rhs, dims, scales)


def test_linear_rows_are_consistent_with_the_true_model(oracle):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g)
    q0 = 1 / 300
    rng = np.random.default_rng(5)
    x_true = 0.3 * rng.standard_normal(n)
    c_true = 0.01 * rng.standard_normal(6)
    pred, _ = oracle.q_predict(rs, nodes, lens, s, q0 * (1 + x_true), q0)
    tstar = pred + c_true[station]
    opts = dict(station_terms=1, nstations=6, damping=2.0, smoothing=10.0, vertical_weight=0.1, damping_station=3.0,
                q0=q0)
    A, b, _ = _data_block(oracle, g, opts, rs, nodes, lens, station, tstar, sigma, s)
    assert np.linalg.norm(A @ np.concatenate([x_true, c_true]) - b) <= 1e-12 * np.linalg.norm(b)
    # a wrong sign of the right-hand side or a missing station column would not be consistent
    assert np.linalg.norm(A @ np.concatenate([x_true, 0 * c_true]) - b) > 1e-3 * np.linalg.norm(b)


def test_gauss_newton_rows_carry_the_linearisation(oracle):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g, seed=3)
    q0 = 1 / 300
    rng = np.random.default_rng(7)
    m_true = 0.4 * rng.standard_normal(n)
    q_true = q0 * np.exp(m_true)
    tstar, _ = oracle.q_predict(rs, nodes, lens, s, q_true, q0)
    opts = dict(station_terms=0, nstations=6, damping=2.0, smoothing=10.0, vertical_weight=0.1, q0=q0)
    A, b, _ = _data_block(oracle, g, opts, rs, nodes, lens, station, tstar, sigma, s, q_old=q_true, m_old=m_true,
                          scales=(1.0, 1.0))
    # at the true model, J m_true = t* - Predict(q_true) + J m_true exactly
    assert np.linalg.norm(A @ m_true - b) <= 1e-12 * np.linalg.norm(b)
    # the first iteration's rows (q0) are not consistent with m_true: the problem is nonlinear
    A0, b0, _ = _data_block(oracle, g, opts, rs, nodes, lens, station, tstar, sigma, s)
    assert np.linalg.norm(A0 @ m_true - b0) > 1e-2 * np.linalg.norm(b0)


def test_sensitivity_scales(oracle):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g)
    tstar = np.full(station.size, 0.02)
    opts = dict(station_terms=1, nstations=6, damping=2.0, smoothing=10.0, vertical_weight=0.1, damping_station=3.0,
                q0=1 / 300)
    *_, base = oracle.build_qsystem(g, opts, rs, nodes, lens, station, tstar, sigma, s)
    *_, half = oracle.build_qsystem(g, opts, rs, nodes, lens, station, tstar, 2 * sigma, s)
    assert half == (base[0] / 2, base[1] / 2)
    opts2 = dict(opts, q0=2 / 300)
    *_, dq = oracle.build_qsystem(g, opts2, rs, nodes, lens, station, tstar, sigma, s)
    assert dq == (2 * base[0], base[1])
    # one row with one node: the RMS is that entry
    rs1, nodes1, lens1 = np.array([0, 1], np.int64), np.array([3], np.int32), np.array([0.5], np.float32)
    *_, one = oracle.build_qsystem(g, opts, rs1, nodes1, lens1, station[:1], tstar[:1], sigma[:1], s)
    assert math.isclose(one[0], 0.5 * s[3] / 300 / sigma[0], rel_tol=1e-14)
    assert math.isclose(one[1], 1 / sigma[0], rel_tol=1e-14)


def test_regularisation_rows(oracle):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g)
    tstar = np.full(station.size, 0.02)
    opts = dict(station_terms=1, nstations=6, damping=2.0, smoothing=10.0, vertical_weight=0.1, damping_station=3.0,
                q0=1 / 300)
    grs, gci, gva, rhs, dims, (sm, ss) = oracle.build_qsystem(g, opts, rs, nodes, lens, station, tstar, sigma, s)
    m, cols, nnz, nd = dims
    assert cols == n + 6 and m == nd + n + n + 6 + 1
    d = slice(nd, nd + n)  # damping: one entry damping * scale per node
    assert np.all(np.diff(grs[nd:nd + n + 1]) == 1)
    assert np.array_equal(gci[grs[nd]:grs[nd + n]], np.arange(n))
    assert np.all(gva[grs[nd]:grs[nd + n]] == 2.0 * sm) and np.all(rhs[d] == 0)
    ones = np.ones(n)  # the Laplacian block annihilates a constant field
    lap = np.zeros(n)
    for r in range(nd + n, nd + 2 * n):
        lap[r - nd - n] = gva[grs[r]:grs[r + 1]] @ ones[gci[grs[r]:grs[r + 1]]]
    assert np.max(np.abs(lap)) <= 1e-12 * 10 * sm
    st = slice(nd + 2 * n, nd + 2 * n + 6)  # station damping
    assert np.all(gva[grs[st.start]:grs[st.stop]] == 3.0 * ss)
    last = m - 1  # zero-sum row over the station columns, weight 10
    assert np.array_equal(gci[grs[last]:grs[last + 1]], n + np.arange(6))
    assert np.all(gva[grs[last]:grs[last + 1]] == 10 * ss) and rhs[last] == 0


def test_gauss_newton_recovers_ln_q(oracle):
    g = _grid(oracle, 3, 3, 2)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g, nrows=200, per_row=6, cover=True, seed=11)
    q0 = 1 / 300
    rng = np.random.default_rng(12)
    m_true = 0.5 * rng.standard_normal(n)
    tstar, _ = oracle.q_predict(rs, nodes, lens, s, q0 * np.exp(m_true), q0)
    settings = dict(q0=300.0, estimate_q0=False, damping=1e-7, smoothing=1e-7, station_terms=False,
                    lsqr_iterations=400, nonlinear_iterations=20)
    O = oracle.QOracle(g, settings)
    O.prepare(rs, nodes, lens, np.arange(tstar.size), tstar, sigma, station, 6, s)
    out = O.solve(settings["damping"], settings["smoothing"])
    assert O.gn[0] < 20
    assert np.allclose(np.log(out["q"] / q0), m_true, atol=1e-5)
    # the unknowns q/q0 - 1 on data linear in q
    x_true = 0.3 * rng.standard_normal(n)
    tstar, _ = oracle.q_predict(rs, nodes, lens, s, q0 * (1 + x_true), q0)
    O = oracle.QOracle(g, dict(settings, log_parameterisation=False))
    O.prepare(rs, nodes, lens, np.arange(tstar.size), tstar, sigma, station, 6, s)
    out = O.solve(1e-7, 1e-7)
    # one LSQR solve, stopped by btol = 1e-6 (QTomography.cs:386-389): accurate to about that
    assert np.allclose(out["q"] / q0 - 1, x_true, atol=1e-5)


def test_outliers_are_rejected(oracle):
    g = _grid(oracle)
    n, rs, nodes, lens, station, sigma, s = _setup(oracle, g, nrows=120)
    rng = np.random.default_rng(4)
    tstar, _ = oracle.q_predict(rs, nodes, lens, s, np.full(n, 1 / 200), 1 / 300)
    sigma = np.full(tstar.size, 0.002)
    tstar = tstar + 0.002 * rng.standard_normal(tstar.size)
    planted = [5, 40, 77]
    tstar[planted] += 0.2
    O = oracle.QOracle(g, dict(station_terms=True, lsqr_iterations=50, nonlinear_iterations=2))
    O.prepare(rs, nodes, lens, np.arange(tstar.size), tstar, sigma, station, 6, s)
    O.solve(2.0, 10.0)
    assert O.reject_outliers() == 3
    assert np.array_equal(np.flatnonzero(O.rejected), planted)
