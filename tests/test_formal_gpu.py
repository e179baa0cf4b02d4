"""The formal resolution on the device (FormalResolution.Compute, Tomography/FormalResolution.cs:108-333)
against the oracle: random weighted systems whose kept count is not a multiple of the 64-wide tiles
(padding), with eliminated groups and places; the closed-form cases of the oracle's own tests; the
declines with the reference's reasons; and the velocity system of a small grid (tables, rays and the
assembly on the device, joint hypocentres and station terms), its hypocentres separated out, its P and
S blocks placed. Bars: the diagonal and the standard deviations to 1e-9 relative, the lengths to 1e-7
(the factor is accumulated in another order than the oracle's LAPACK, and the lengths sum over rows of
R whose entries are differences of nearly equal terms)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr(A):
    rs, ci, va = [0], [], []
    for row in A:
        nz = np.flatnonzero(row)
        ci.extend(nz.tolist())
        va.extend(row[nz].tolist())
        rs.append(len(ci))
    return np.array(rs, np.int32), np.array(ci, np.int32), np.array(va)


def _random_system(rng, nmodel, nev, per_event=10):
    ncols = nmodel + 4 * nev
    rows = []
    for e in range(nev):
        for _ in range(per_event):
            r = np.zeros(ncols)
            idx = rng.choice(nmodel, 12, replace=False)
            r[idx] = rng.random(12) + 0.1
            r[nmodel + 4 * e:nmodel + 4 * e + 4] = rng.standard_normal(4)
            rows.append(r)
    data = len(rows)
    for c in range(nmodel):
        r = np.zeros(ncols)
        r[c] = 0.4
        rows.append(r)
    for c in range(nmodel - 2):
        r = np.zeros(ncols)
        r[c], r[c + 1], r[c + 2] = -0.5, 1.0, -0.5
        rows.append(r)
    for c in range(nmodel, ncols):
        r = np.zeros(ncols)
        r[c] = 0.1
        rows.append(r)
    groups = [list(range(nmodel + 4 * e, nmodel + 4 * e + 4)) for e in range(nev)]
    return np.array(rows), data, groups


def _compare(dev_res, ref, tol_len=1e-7):
    kept, diag, sd, length = ref
    assert np.array_equal(dev_res.columns, kept)
    assert np.allclose(dev_res.diagonal, diag, rtol=1e-9, atol=1e-12)
    assert np.allclose(dev_res.standard_deviation, sd, rtol=1e-9)
    both = ~np.isnan(length)
    assert np.array_equal(np.isnan(dev_res.length_km), np.isnan(length))
    assert np.allclose(dev_res.length_km[both], length[both], rtol=tol_len, atol=1e-9)


@pytest.mark.parametrize("nmodel,nev", [(37, 3), (150, 0), (700, 25)])
def test_random_systems(device, oracle, nmodel, nev):
    from tomostar_amd import formal as F

    rng = np.random.default_rng(nmodel)
    A, data, groups = _random_system(rng, nmodel, nev, per_event=max(10, 2 * nmodel // max(1, nev)) if nev else 0)
    if nev == 0:
        G = rng.standard_normal((200, nmodel)) * (rng.random((200, nmodel)) < 0.1)
        A = np.vstack([G, 0.6 * np.eye(nmodel)])
        data = 200
    rs, ci, va = _csr(A)
    cols = A.shape[1]
    xyz = rng.random((cols, 3)) * 50
    blk = np.where(np.arange(cols) < nmodel, np.arange(cols) % 2, -1).astype(np.int32)
    blk[::7] = -1
    res = F.compute(device, A.shape[0], cols, rs, ci, va, data, groups, places=(xyz, blk))
    _compare(res, oracle.formal_resolution(A.shape[0], cols, rs, ci, va, data, groups, places=(xyz, blk)))
    # without places the lengths are NaN and the rest unchanged
    res2 = F.compute(device, A.shape[0], cols, rs, ci, va, data, groups)
    assert np.all(np.isnan(res2.length_km)) and np.array_equal(res2.diagonal, res.diagonal)


def test_closed_forms_and_declines(device, oracle):
    from tomostar_amd import _lib
    from tomostar_amd import formal as F

    lam, d = 0.8, 5.0
    A = np.array([[1.0, 1.0], [lam, 0.0], [0.0, lam]])
    rs, ci, va = _csr(A)
    xyz = np.array([[0.0, 0, 0], [d, 0, 0]])
    res = F.compute(device, 3, 2, rs, ci, va, 1, places=(xyz, np.zeros(2, np.int32)))
    l2 = lam * lam
    det = (1 + l2) ** 2 - 1
    r11, r12 = 1 - l2 * (1 + l2) / det, l2 / det
    assert np.isclose(res.diagonal[0], r11, rtol=1e-12)
    assert np.isclose(res.length_km[0], d * abs(r12) / np.hypot(r11, r12), rtol=1e-12)
    rng = np.random.default_rng(2)
    A, data, groups = _random_system(rng, 30, 4)
    rs, ci, va = _csr(A)
    with pytest.raises(_lib.HipError, match="above the limit") as e:
        F.compute(device, A.shape[0], A.shape[1], rs, ci, va, data, groups, max_unknowns=29)
    assert e.value.status == _lib.TS_DECLINED
    B = A.copy()
    B[0, 34:38] = 1.0
    with pytest.raises(_lib.HipError, match="two events"):
        F.compute(device, B.shape[0], B.shape[1], *_csr(B), data, groups)
    B = A.copy()
    B[data, 30] = 1.0
    with pytest.raises(_lib.HipError, match="couples model and hypocentre"):
        F.compute(device, B.shape[0], B.shape[1], *_csr(B), data, groups)
    C = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    with pytest.raises(_lib.HipError, match="not positive definite"):
        F.compute(device, 2, 3, *_csr(C), 2)


def test_velocity_system_of_a_small_grid(device, oracle):
    """Tables, rays and BuildSystem on the device for a 12 x 12 x 8 grid (8 stations x P,S, 40 events,
    Vp/Vs, joint hypocentres, station terms); the hypocentres separated out, P and S blocks placed."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import formal as F
    from tomostar_amd import iteration as It
    from tomostar_amd import rays as Ry
    from tomostar_amd import system as Sy
    from tomostar_amd import workloads as W
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition.from_spacing(13.0, 13.15, 43.0, 43.1, 0, 10.5, 1.2, 1.5)
    rng = np.random.default_rng(5)
    z = g.depth()[:, None, None]
    vp = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + W.checkerboard(g, 5.0, 5.0, 0.05)), g.shape))
    st = np.stack([rng.uniform(g.min_lon, g.max_lon, 8), rng.uniform(g.min_lat, g.max_lat, 8), np.zeros(8)], 1)
    ev = np.stack([rng.uniform(g.min_lon, g.max_lon, 40), rng.uniform(g.min_lat, g.max_lat, 40),
                   rng.uniform(2, 8, 40)], 1)
    wl = W.Workload("small", g, vp, vp / 1.73, st, ev)
    t_obs = It.observed_times(device, wl)
    sP, sS = It.starting_model(wl)
    seeds = [E.point_source(g, sP if ph == 0 else sS, *s) for s in wl.stations for ph in (0, 1)]
    ts = E.TableSet(device, g, [sP, sS], seeds, np.tile(np.array([0, 1], np.int32), len(wl.stations)))
    assert ts.solve("fast_fp64")[0] == 0
    station, event, phase = It.all_arrivals(wl)
    batch = Ry.RayBatch(ts, g, np.repeat(wl.stations, 2, axis=0), 2 * station + phase, wl.events[event])
    ts.release()
    used = np.flatnonzero((batch.counts >= 0) & np.isfinite(t_obs)).astype(np.int32)
    residual = t_obs[used] - batch.tcalc_grad[used, 0]
    settings = Sy.SystemSettings(vpvs=True)
    sysd = Sy.build_system(device, g, settings, batch, used, phase[used], event[used], station[used], residual,
                           np.full(used.size, 0.05), np.zeros(40, np.uint8), np.zeros(40, np.uint8), 8, sP, sS, sP,
                           sS, sS / sP, damping=1.0, smoothing=5.0, keep_fp64=True)
    rs, ci, va = sysd.read()
    lay = sysd.layout
    cols = lay["columns"]
    groups = [list(range(lay["offH"] + 4 * e, lay["offH"] + 4 * e + 4)) for e in range(40)]
    places = F.new_places(cols)
    F.add_places(places, g, lay["offP"], 0)
    F.add_places(places, g, lay["offS"], 1)
    res = F.compute(device, sysd.matrix.rows, cols, rs, ci, va, lay["data_rows"], groups, places=places)
    ref = oracle.formal_resolution(sysd.matrix.rows, cols, rs, ci, va, lay["data_rows"], groups, places=places)
    print(f"formal resolution: {res.columns.size} unknowns in {res.seconds:.3f} s, "
          f"median R_ii {np.median(res.diagonal[:g.count]):.3f}")
    _compare(res, ref)
    assert np.all(res.diagonal < 1 + 1e-12) and np.all(res.standard_deviation > 0)
    maps = F.on_nodes(res, cols, g.count, lay["offP"])
    assert np.array_equal(maps.resolution, res.diagonal[:g.count])
    # TravelTimeTomography.Formal's wrapper: the same groups and places, the blocks mapped to the nodes
    pm, sm, fr = F.velocity_formal(device, sysd, g, 40)
    assert np.array_equal(fr.diagonal, res.diagonal) and np.array_equal(pm.resolution, maps.resolution)
    assert np.array_equal(sm.sigma, F.on_nodes(res, cols, g.count, lay["offS"]).sigma)
    sysd.matrix.release()
    batch.release()


def test_edges(device, oracle):
    """Every column separated out (nothing kept), no columns at all, a single kept column, and a kept
    count of exactly one tile."""
    from tomostar_amd import formal as F

    A = np.array([[1.0, 2.0], [0.5, 0.0], [0.0, 0.5]])
    rs, ci, va = _csr(A)
    res = F.compute(device, 3, 2, rs, ci, va, 1, [[0, 1]])
    assert res.columns.size == 0
    res = F.compute(device, 0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), 0)
    assert res.columns.size == 0
    A = np.array([[2.0], [0.5]])
    rs, ci, va = _csr(A)
    res = F.compute(device, 2, 1, rs, ci, va, 1)
    assert np.isclose(res.diagonal[0], 4.0 / 4.25, rtol=1e-14) and np.isclose(res.standard_deviation[0],
                                                                              (1 / 4.25) ** 0.5, rtol=1e-14)
    rng = np.random.default_rng(9)
    G = rng.standard_normal((100, 64)) * (rng.random((100, 64)) < 0.2)
    A = np.vstack([G, 0.5 * np.eye(64)])
    rs, ci, va = _csr(A)
    _compare(F.compute(device, A.shape[0], 64, rs, ci, va, 100),
             oracle.formal_resolution(A.shape[0], 64, rs, ci, va, 100))


def test_q_run_with_formal_resolution(device, oracle):
    """QTomography.Run with the formal resolution on (QTomography.cs:480-497): the 61 x 61 x 31 case's P
    t* inverted on every third node (21 x 21 x 11 nodes and 25 station terms), the last pass's system
    against the oracle; too small a limit is logged, not raised."""
    from tomostar_amd import formal as F
    from tomostar_amd import qtomography as QT
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    c = QT.q_case(device, wl, step=3)
    s = QT.QTomographySettings(lsqr_iterations=60, nonlinear_iterations=2, formal_resolution=True)
    qt = QT.QTomography(device, c.grid, s)
    res = qt.run(c.rays, c.data, c.slowness)
    assert res.formal is not None and res.formal_reason == ""
    (rs, ci, va), rows, cols, nd = qt.last_system
    places = F.add_places(F.new_places(cols), c.grid, 0, 0)
    kept, diag, sd, length = oracle.formal_resolution(rows, cols, rs, ci, va, nd, (), 20000, places)
    n = c.grid.count
    assert np.allclose(res.formal.resolution, diag[:n], rtol=1e-9, atol=1e-12)
    assert np.allclose(res.formal.sigma, sd[:n], rtol=1e-9)
    assert np.allclose(res.formal.length_km, length[:n], rtol=1e-7, atol=1e-9)
    s2 = QT.QTomographySettings(lsqr_iterations=20, nonlinear_iterations=1, formal_resolution=True,
                                formal_max_unknowns=1000)
    res2 = QT.QTomography(device, c.grid, s2).run(c.rays, c.data, c.slowness)
    assert res2.formal is None and "above the limit" in res2.formal_reason
    c.rays.release()
