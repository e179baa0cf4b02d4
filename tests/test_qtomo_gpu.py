"""Q tomography on the device (Attenuation/QTomography.cs) against the oracle, on the P rays of the
61 x 61 x 31 case traced in its true model (25 stations x 150 events) with t* of a Q model (a
checkerboard in ln Q and a low-Q body) plus noise:

- ts_q_predict equals the oracle's Predict and path times to the bit;
- the assembled system equals the oracle's: the data rows to the bit at the first Gauss-Newton
  iteration (q0) and at a later one (q_old, m_old); the sensitivity scales, and with them the
  regularisation rows, to 1e-13 when computed (the device sums each row's squared entries and then
  the rows, the reference keeps one running sum), to the bit when fixed;
- Run with LsqrOpenCl's FP64 arithmetic follows the oracle's Run: the same reference Q, the same
  rejected arrivals, the same Gauss-Newton count, and Q and the final RMS within 20 times the spread
  a 1e-15 perturbation of the t* causes in the oracle itself (LSQR stops on atol = 1e-6 after some
  30 iterations, and the iteration it stops at moves the iterate); the fast mixed form likewise;
- TradeOff's points follow Tikhonov's trade-off and agree among themselves.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def qcase(device):
    from tomostar_amd import qtomography as QT
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    c = QT.q_case(device, wl)
    yield wl, c
    c.rays.release()


def _host_rows(rays, arrivals):
    starts, ends = rays.row_start[arrivals], rays.row_start[arrivals + 1]
    rs = np.concatenate([[0], np.cumsum(ends - starts)]).astype(np.int64)
    take = np.concatenate([np.arange(a, b) for a, b in zip(starts, ends)])
    return rs, rays.nodes[take], rays.lengths[take]


def test_predict_is_the_oracles(device, oracle, qcase):
    from tomostar_amd import qtomography as QT

    wl, c = qcase
    rows = np.flatnonzero(c.rays.counts >= 0).astype(np.int32)
    assert rows.size > 3500
    rs, nd, ln = _host_rows(c.rays, rows)
    for q, q0 in ((None, 1 / 300), (1 / c.q_true, 1 / 300)):
        pred, tt = QT.q_predict(c.rays, rows, c.slowness, q, q0)
        opred, ott = oracle.q_predict(rs, nd, ln, c.slowness, q, q0)
        assert np.array_equal(pred, opred) and np.array_equal(tt, ott)


@pytest.mark.parametrize("later", [False, True])
def test_system_is_the_oracles(device, oracle, qcase, later):
    from tomostar_amd import qtomography as QT
    from tomostar_amd.system import System

    wl, c = qcase
    g = wl.grid
    og = oracle.griddef(*dataclasses.astuple(g))
    rows = np.flatnonzero(c.rays.counts >= 0).astype(np.int32)[::2]
    d = c.data
    s = QT.QTomographySettings()
    q0 = 1 / 250
    q_old = m_old = scales = None
    if later:
        rng = np.random.default_rng(3)
        m_old = 0.2 * rng.standard_normal(g.count)
        q_old = q0 * np.exp(m_old)
        scales = (0.37, 123.0)
    sysd = QT.build_qsystem(device, g, s, c.rays, rows, d.station[rows], d.tstar[rows], d.sigma[rows], d.nstations,
                            c.slowness, q0, s.damping, s.smoothing, q_old, m_old, scales)
    rs, ci, va = System(sysd.matrix, sysd.rhs).read()
    hrs, hnd, hln = _host_rows(c.rays, rows)
    opts = dict(station_terms=1, nstations=d.nstations, damping=s.damping, smoothing=s.smoothing,
                vertical_weight=s.vertical_weight, damping_station=s.damping_station, q0=q0)
    ors, oci, ova, orhs, dims, oscales = oracle.build_qsystem(og, opts, hrs, hnd, hln, d.station[rows], d.tstar[rows],
                                                              d.sigma[rows], c.slowness, q_old, m_old, scales)
    assert (sysd.matrix.rows, sysd.matrix.cols, sysd.matrix.nnz) == dims[:3]
    assert sysd.layout == {"off_station": g.count, "columns": dims[1], "rows": dims[0], "data_rows": dims[3]}
    assert np.array_equal(rs, ors) and np.array_equal(ci, oci)
    nd = dims[3]
    e = ors[nd]
    assert np.array_equal(va[:e], ova[:e]) and np.array_equal(sysd.rhs[:nd], orhs[:nd])
    if later:
        assert sysd.scales == oscales
        assert np.array_equal(va, ova) and np.array_equal(sysd.rhs, orhs)
    else:
        assert np.allclose(sysd.scales, oscales, rtol=1e-13, atol=0)
        assert np.allclose(va[e:], ova[e:], rtol=1e-13, atol=0)
        assert np.allclose(sysd.rhs, orhs, rtol=1e-13, atol=0)
    sysd.matrix.release()


def _run_oracle(oracle, wl, c, settings, tstar):
    O = oracle.QOracle(oracle.griddef(*dataclasses.astuple(wl.grid)), settings, threads=16)
    rows = np.flatnonzero(c.rays.counts >= 0)
    rs, nd, ln = _host_rows(c.rays, rows)
    O.prepare(rs, nd, ln, rows, tstar, c.data.sigma, c.data.station, c.data.nstations, c.slowness)
    out = O.run()
    return O, out


def test_run_follows_the_oracle(device, oracle, qcase):
    from tomostar_amd import qtomography as QT

    wl, c = qcase
    settings = dict(lsqr_iterations=100, nonlinear_iterations=3)
    s = QT.QTomographySettings(**settings)
    res = QT.QTomography(device, wl.grid, s, lsqr_mode="plain_fp64").run(c.rays, c.data, c.slowness)
    O, out = _run_oracle(oracle, wl, c, settings, c.data.tstar)
    assert res.reference_q == 1 / O.q0
    assert res.rejected == out["rejected"]
    assert res.gauss_newton[0] == O.gn[0]
    t2 = c.data.tstar * (1 + 1e-15 * np.random.default_rng(5).standard_normal(c.data.tstar.size))
    _, out2 = _run_oracle(oracle, wl, c, settings, t2)
    spread = np.max(np.abs(out2["q"] / out["q"] - 1))
    err = np.max(np.abs(res.q * out["q"] - 1))
    print(f"Q run: max relative error {err:.3e}, oracle spread {spread:.3e}, GN {res.gauss_newton}, "
          f"rejected {res.rejected}, reference Q {res.reference_q:.1f}")
    assert err <= max(1e-9, 20 * spread)
    rms_spread = abs(out2["rms_after"] / out["rms_after"] - 1)
    assert abs(res.rms_after / out["rms_after"] - 1) <= max(1e-9, 20 * rms_spread)
    assert res.rms_after < 0.5 * res.rms_before
    fast = QT.QTomography(device, wl.grid, s, lsqr_mode="fast_mixed").run(c.rays, c.data, c.slowness)
    assert fast.rejected == out["rejected"]
    assert np.max(np.abs(fast.q * out["q"] - 1)) <= max(1e-3, 20 * spread)
    # the recovered Q correlates with the true one where the rays sample it
    ok = res.dws > 1.0
    assert np.corrcoef(np.log(res.q[ok]), np.log(c.q_true[ok]))[0, 1] > 0.5


def test_trade_off_points(device, qcase):
    """Unknowns q/q0 - 1 (a linear problem): less damping fits the data better with a larger model, as
    Tikhonov regularisation must; the misfit, data variance and roughness agree with one another."""
    from tomostar_amd import qtomography as QT
    from tomostar_amd import tradeoff as T

    wl, c = qcase
    s = QT.QTomographySettings(lsqr_iterations=200, log_parameterisation=False)
    qt = QT.QTomography(device, wl.grid, s, lsqr_mode="fast_mixed")
    pairs = [(20.0, 10.0), (0.5, 10.0)]
    pts = qt.trade_off(c.rays, c.data, c.slowness, pairs)
    assert [(p.damping, p.smoothing) for p in pts] == pairs
    assert pts[1].residual_norm < pts[0].residual_norm
    assert pts[1].model_norm > pts[0].model_norm
    used = int(np.sum(~qt.rejected & ~np.isnan(qt.residual)))
    sigma = c.data.sigma[0]
    for p in pts:
        assert np.isclose(p.residual_norm ** 2 * sigma ** 2 / used, p.data_variance, rtol=1e-12)
        assert np.all((p.q >= s.q_min * (1 - 1e-12)) & (p.q <= s.q_max * (1 + 1e-12)))
        # away from the bounds the node unknowns are q/q0 - 1 with q = 1 / Q
        if np.all((p.q > 1.001 * s.q_min) & (p.q < 0.999 * s.q_max)):
            x = 1 / (p.q * qt.q0) - 1
            assert np.isclose(p.roughness, T.roughness(wl.grid, x, 0.1), rtol=1e-9)


def test_edges_and_invalid_arguments(device, oracle, qcase):
    """No rows (the regularisation alone, unit scales); no station terms with first differences, against
    the oracle; q_old without m_old, a bad station and a non-positive sigma are refused."""
    from tomostar_amd import _lib
    from tomostar_amd import qtomography as QT
    from tomostar_amd.system import System

    wl, c = qcase
    g = wl.grid
    og = oracle.griddef(*dataclasses.astuple(g))
    d = c.data
    s = QT.QTomographySettings()
    pred, tt = QT.q_predict(c.rays, np.zeros(0, np.int32), c.slowness, None, 1 / 300)
    assert pred.size == 0 and tt.size == 0
    sysd = QT.build_qsystem(device, g, s, c.rays, np.zeros(0, np.int32), [], [], [], d.nstations, c.slowness,
                            1 / 300, s.damping, s.smoothing)
    assert sysd.scales == (1.0, 1.0) and sysd.layout["data_rows"] == 0
    assert sysd.matrix.rows == 2 * g.count + d.nstations + 1
    sysd.matrix.release()
    rows = np.flatnonzero(c.rays.counts >= 0).astype(np.int32)[:500]
    s2 = QT.QTomographySettings(station_terms=False, smoothing_method="gradient", smoothing_scale="nodes")
    sysd = QT.build_qsystem(device, g, s2, c.rays, rows, d.station[rows], d.tstar[rows], d.sigma[rows], d.nstations,
                            c.slowness, 1 / 300, 1.5, 4.0)
    rs, ci, va = System(sysd.matrix, sysd.rhs).read()
    hrs, hnd, hln = _host_rows(c.rays, rows)
    opts = dict(station_terms=0, gradient_smoothing=1, smoothing_in_nodes=1, nstations=d.nstations, damping=1.5,
                smoothing=4.0, vertical_weight=s2.vertical_weight, q0=1 / 300)
    ors, oci, ova, orhs, dims, oscales = oracle.build_qsystem(og, opts, hrs, hnd, hln, d.station[rows], d.tstar[rows],
                                                              d.sigma[rows], c.slowness)
    assert sysd.layout["off_station"] == -1 and sysd.matrix.cols == g.count
    assert np.array_equal(rs, ors) and np.array_equal(ci, oci)
    assert np.allclose(va, ova, rtol=1e-13, atol=0) and np.allclose(sysd.rhs, orhs, rtol=1e-13, atol=0)
    sysd.matrix.release()
    with pytest.raises(_lib.HipError) as e:
        QT.build_qsystem(device, g, s, c.rays, rows, d.station[rows], d.tstar[rows], d.sigma[rows], d.nstations,
                         c.slowness, 1 / 300, 1.0, 1.0, q_old=np.full(g.count, 1 / 300))
    assert e.value.status == _lib.TS_EINVAL
    bad = d.station[rows].copy()
    bad[3] = d.nstations
    with pytest.raises(_lib.HipError, match="station"):
        QT.build_qsystem(device, g, s, c.rays, rows, bad, d.tstar[rows], d.sigma[rows], d.nstations, c.slowness,
                         1 / 300, 1.0, 1.0)
    sig = d.sigma[rows].copy()
    sig[0] = 0.0
    with pytest.raises(_lib.HipError, match="sigma"):
        QT.build_qsystem(device, g, s, c.rays, rows, d.station[rows], d.tstar[rows], sig, d.nstations, c.slowness,
                         1 / 300, 1.0, 1.0)
