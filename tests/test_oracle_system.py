"""The oracle's restatement of BuildSystem (Tomography/TravelTimeTomography.cs:860-1034, plain node
grid) pinned by what the mathematics fixes, independently of how the oracle forms the rows:

- data rows are the linearisation of the travel time t = sum L s along the ray in the inversion
  unknowns (x = ds / sRef, and Vp/Vs in the VpVpVs form): finite differences of t reproduce them;
- AxisWeights (Observations.cs:390-397) follow from the grid metric (or_metric, pinned elsewhere);
- a Laplacian row annihilates constant and linear fields away from the edges, first-difference rows
  return the axis weights on a linear ramp, and every smoothing row's right-hand side is minus the row
  applied to the current deviation;
- damping rows are lambda I, the zero-sum row is 10 over the station block, rows and right-hand sides
  scale by absoluteWeight / sigma, entries below 1e-4 are dropped and empty data rows skipped.
"""
import numpy as np
import pytest
import scipy.sparse as sp


@pytest.fixture(scope="module")
def case(oracle):
    g = oracle.from_spacing(13.0, 13.3, 43.0, 43.2, 0.0, 10.0, 2.0, 2.0)
    n = g.count
    rng = np.random.default_rng(5)
    depth = np.repeat(np.linspace(g.min_depth, g.max_depth, g.nz), g.nx * g.ny)
    sRefP = 1.0 / (5.0 + 0.04 * depth)
    kRef = np.full(n, 1.73)
    sRefS = sRefP * kRef
    xP = 0.03 * rng.standard_normal(n)
    xK = 0.02 * rng.standard_normal(n)
    sP = sRefP * (1 + xP)
    sS = sP * kRef * (1 + xK)
    nev, nst, nrows = 6, 4, 40
    rows, lens = [], []
    for r in range(nrows):
        c = np.sort(rng.choice(n, rng.integers(3, 12), replace=False))
        rows.append(c)
        lens.append((rng.random(c.size) * 2 + 0.05).astype(np.float32))
    row_start = np.concatenate([[0], np.cumsum([c.size for c in rows])]).astype(np.int64)
    data = dict(row_start=row_start, nodes=np.concatenate(rows).astype(np.int32), lens=np.concatenate(lens),
                tcalc_grad=np.column_stack([rng.random(nrows) * 5, 0.1 * rng.standard_normal((nrows, 3))]),
                phase=(np.arange(nrows) % 2).astype(np.int32), event=rng.integers(0, nev, nrows).astype(np.int32),
                station=rng.integers(0, nst, nrows).astype(np.int32), residual=0.1 * rng.standard_normal(nrows),
                sigma=np.where(np.arange(nrows) % 3 == 0, 0.05, 0.1), ev_fixed=np.array([0, 0, 1, 0, 0, 0], np.uint8),
                ev_outside=np.array([0, 1, 0, 0, 0, 0], np.uint8))
    models = dict(sP=sP, sS=sS, sRefP=sRefP, sRefS=sRefS, kRef=kRef)
    return g, data, models, (nev, nst)


def _opts(nev, nst, **kw):
    o = dict(vpvs=1, joint_hypocentres=1, station_corrections=1, invert_p=1, invert_s=1, gradient_smoothing=0,
             nevents=nev, nstations=nst, any_p=1, any_s=1, damping=1.0, smoothing=5.0, vertical_weight=0.1,
             damping_hypocentre=1.0, damping_station=5.0, s_factor=1.0, absolute_weight=1.0)
    o.update(kw)
    return o


def _build(oracle, case, **kw):
    g, data, models, (nev, nst) = case
    o = _opts(nev, nst, **kw)
    rs, ci, va, rhs, dims = oracle.build_system(g, o, **data, **models)
    A = sp.csr_matrix((va, ci, rs), shape=(dims[0], dims[1]))
    return A, rhs, dims, o


def _layout(n, nev, nst):
    offP, offS, offH = 0, n, 2 * n
    offCP, offCS = offH + 4 * nev, offH + 4 * nev + nst
    return offP, offS, offH, offCP, offCS


def test_row_and_column_counts(oracle, case):
    g, data, models, (nev, nst) = case
    A, rhs, (m, cols, nnz, data_rows), _ = _build(oracle, case)
    n = g.count
    assert cols == 2 * n + 4 * nev + 2 * nst
    assert data_rows == len(data["phase"])
    # damping and Laplacian of P and S, hypocentre damping, two station blocks with their zero-sum rows
    assert m == data_rows + 4 * n + 4 * nev + 2 * (nst + 1)
    assert rhs.size == m


def test_data_rows_are_the_travel_time_linearisation(oracle, case):
    """VpVpVs: t_P = sum L sRefP (1 + xP); t_S = sum L sRefP (1 + xP) kRef (1 + xK). The P-block and
    S-block entries of a row are dt/dxP and dt/dxK at the current model (central differences)."""
    g, data, models, (nev, nst) = case
    A, rhs, (m, cols, nnz, drows), o = _build(oracle, case)
    n = g.count
    sRefP, kRef = models["sRefP"], models["kRef"]
    xP = models["sP"] / sRefP - 1
    xK = models["sS"] / models["sP"] / kRef - 1
    offP, offS, offH, offCP, offCS = _layout(n, nev, nst)
    rs = data["row_start"]
    for r in range(drows):
        nodes, L = data["nodes"][rs[r]:rs[r + 1]], data["lens"][rs[r]:rs[r + 1]].astype(np.float64)
        w = 1.0 / data["sigma"][r]
        row = A.getrow(r).toarray().ravel() / w

        def t(xp, xk):
            s = sRefP[nodes] * (1 + xp[nodes])
            return np.sum(L * s) if data["phase"][r] == 0 else np.sum(L * s * kRef[nodes] * (1 + xk[nodes]))

        for q, nd in enumerate(nodes):
            for off, x in ((offP, 0), (offS, 1)):
                h = 1e-6
                xa, xb = [xP.copy(), xK.copy()], [xP.copy(), xK.copy()]
                xa[x][nd] += h
                xb[x][nd] -= h
                d = (t(*xa) - t(*xb)) / (2 * h)
                if data["phase"][r] == 0 and off == offS:
                    assert row[off + nd] == 0
                    continue
                assert row[off + nd] == pytest.approx(d, rel=1e-7, abs=1e-4)
        # hypocentre columns: 1, Ge, Gn, Gz unless fixed; only the origin time when outside
        e = data["event"][r]
        hyp = row[offH + 4 * e: offH + 4 * e + 4]
        if data["ev_fixed"][e]:
            assert np.all(hyp == 0)
        elif data["ev_outside"][e]:
            assert hyp[0] == 1 and np.all(hyp[1:] == 0)
        else:
            tg = data["tcalc_grad"][r]
            exp = np.where(np.abs(tg[1:]) >= 1e-4, tg[1:], 0)
            assert hyp[0] == 1 and np.allclose(hyp[1:], exp, rtol=1e-15)
        # station term in its phase's block
        sc = offCP if data["phase"][r] == 0 else offCS
        assert row[sc + data["station"][r]] == 1
        assert rhs[r] == pytest.approx(data["residual"][r] * w, rel=1e-15)


def test_rows_scale_with_the_weight_and_small_entries_drop(oracle, case):
    g, data, models, (nev, nst) = case
    A1, rhs1, d1, _ = _build(oracle, case, absolute_weight=1.0)
    A2, rhs2, d2, _ = _build(oracle, case, absolute_weight=2.0)
    dr = d1[3]
    # an exact power-of-two weight scales every data entry and right-hand side exactly
    assert np.array_equal(A2[:dr].toarray(), 2 * A1[:dr].toarray())
    assert np.array_equal(rhs2[:dr], 2 * rhs1[:dr])
    assert np.array_equal(A2[dr:].toarray(), A1[dr:].toarray())
    # no stored entry of a data row is below 1e-4 before weighting
    sig = case[1]["sigma"]
    for r in range(dr):
        v = A1.getrow(r).data * sig[r]
        assert np.all(np.abs(v) >= 1e-4 * (1 - 1e-12))


def test_empty_rows_are_skipped(oracle, case):
    g, data, models, (nev, nst) = case
    d = dict(data)
    lens = d["lens"].copy()
    rs = d["row_start"]
    lens[rs[3]:rs[4]] = 1e-6  # every velocity entry of row 3 falls below 1e-4
    d["lens"] = lens
    ev = d["event"].copy()
    ev[3] = 2  # a fixed event: no hypocentre terms
    d["event"] = ev
    o = _opts(nev, nst, station_corrections=0)
    rs2, ci, va, rhs, dims = oracle.build_system(g, o, **d, **models)
    assert dims[3] == len(d["phase"]) - 1


def test_axis_weights_follow_the_metric(oracle, case):
    """wx hx^2 = wy hy^2 = h^2 and wz = vertical (h / hz)^2, with the spacings of the metric at
    (Nx/2, Ny/2, 0). Read back from the first-difference rows on a linear ramp."""
    g, data, models, (nev, nst) = case
    vw, lam = 0.3, 2.0
    o = _opts(nev, nst, gradient_smoothing=1, vertical_weight=vw, smoothing=lam, damping=0.0, vpvs=0)
    rs, ci, va, rhs, (m, cols, nnz, dr) = oracle.build_system(g, o, **data, **models)
    A = sp.csr_matrix((va, ci, rs), shape=(m, cols))
    hlon, hlat, hr = oracle.metric(g)
    hx, hy, hz = hlon[0 * g.ny + g.ny // 2], hlat[0], hr
    h2 = hx * hy
    wx, wy, wz = h2 / hx ** 2, h2 / hy ** 2, vw * h2 / hz ** 2
    n = g.count
    i = np.tile(np.arange(g.nx), g.ny * g.nz).astype(float)
    j = np.tile(np.repeat(np.arange(g.ny), g.nx), g.nz).astype(float)
    k = np.repeat(np.arange(g.nz), g.nx * g.ny).astype(float)
    reg = A[dr:dr + (g.nx - 1) * g.ny * g.nz + g.nx * (g.ny - 1) * g.nz + g.nx * g.ny * (g.nz - 1)]
    for ramp, w in ((i, wx), (j, wy), (k, wz)):
        x = np.zeros(cols)
        x[:n] = ramp
        y = reg @ x
        nz_ = y[np.abs(y) > 0]
        assert nz_.size > 0 and np.allclose(nz_, lam * np.sqrt(w), rtol=1e-12)
    # rows ordered by node, then axis x, y, z; each row is [-a, +a] at (n, n + axis step)
    first = reg.getrow(0)
    assert list(first.indices) == [0, 1]
    assert reg.getrow(1).indices.tolist() == [0, g.nx]


def test_laplacian_rows(oracle, case):
    g, data, models, (nev, nst) = case
    lam = 5.0
    A, rhs, (m, cols, nnz, dr), _ = _build(oracle, case, vpvs=0, damping=1.0, smoothing=lam)
    n = g.count
    # block order after the data rows: damping P, Laplacian P, damping S, Laplacian S
    dampP = A[dr:dr + n]
    assert np.array_equal(dampP.toarray()[:, :n], np.eye(n)) and np.all(rhs[dr:dr + n] == 0)
    lap = A[dr + n:dr + 2 * n]
    assert np.allclose(lap @ np.r_[np.ones(n), np.zeros(cols - n)], 0, atol=1e-12)
    # a linear ramp: zero at every node with both neighbours on each axis
    i = np.tile(np.arange(g.nx), g.ny * g.nz)
    j = np.tile(np.repeat(np.arange(g.ny), g.nx), g.nz)
    k = np.repeat(np.arange(g.nz), g.nx * g.ny)
    interior = (i > 0) & (i < g.nx - 1) & (j > 0) & (j < g.ny - 1) & (k > 0) & (k < g.nz - 1)
    ramp = 0.3 * i - 0.2 * j + 0.7 * k
    y = lap @ np.r_[ramp, np.zeros(cols - n)]
    assert np.allclose(y[interior], 0, atol=1e-9)
    assert np.any(np.abs(y[~interior]) > 1e-3)
    # the right-hand side is minus the row applied to the current deviation of P slowness
    dev = models["sP"] / models["sRefP"] - 1
    assert np.allclose(rhs[dr + n:dr + 2 * n], -(lap[:, :n] @ dev), rtol=1e-12, atol=1e-15)
    # the S block is smoothed on sS / sRefS - 1 with the S factor
    lapS = A[dr + 3 * n:dr + 4 * n]
    devS = models["sS"] / models["sRefS"] - 1
    assert np.allclose(rhs[dr + 3 * n:dr + 4 * n], -(lapS[:, n:2 * n] @ devS), rtol=1e-12, atol=1e-15)
    # diagonal at an interior node: lambda * 2 (wx + wy + wz); off-diagonals negative
    q = int(np.flatnonzero(interior)[0])
    row = lap.getrow(q)
    assert row.indices[-1] == q and row.data[-1] > 0 and np.all(row.data[:-1] < 0)
    assert row.data[-1] == pytest.approx(-row.data[:-1].sum(), rel=1e-14)


def test_hypocentre_and_station_blocks(oracle, case):
    g, data, models, (nev, nst) = case
    A, rhs, (m, cols, nnz, dr), _ = _build(oracle, case, damping_hypocentre=1.5, damping_station=5.0)
    n = g.count
    offP, offS, offH, offCP, offCS = _layout(n, nev, nst)
    tail = A[m - (4 * nev + 2 * (nst + 1)):].toarray()
    hyp = tail[:4 * nev]
    assert np.array_equal(hyp[:, offH:offH + 4 * nev], 1.5 * np.eye(4 * nev))
    st = tail[4 * nev:]
    assert np.array_equal(st[:nst, offCP:offCP + nst], 5.0 * np.eye(nst))
    assert np.array_equal(st[nst, offCP:offCP + nst], np.full(nst, 10.0)) and st[nst].sum() == 10.0 * nst
    assert np.array_equal(st[nst + 1:2 * nst + 1, offCS:offCS + nst], 5.0 * np.eye(nst))
    assert np.array_equal(st[2 * nst + 1, offCS:offCS + nst], np.full(nst, 10.0))
    assert np.all(rhs[m - (4 * nev + 2 * (nst + 1)):] == 0)


def test_layout_without_optional_blocks(oracle, case):
    g, data, models, (nev, nst) = case
    # P data only, no S inversion, no hypocentres, no station terms: P block alone
    d = dict(data)
    d["phase"] = np.zeros_like(d["phase"])
    o = _opts(nev, nst, any_s=0, joint_hypocentres=0, station_corrections=0, vpvs=0)
    rs, ci, va, rhs, (m, cols, nnz, dr) = oracle.build_system(g, o, **d, **models)
    assert cols == g.count and m == dr + 2 * g.count
    # VpVs with S data only still carries the P block (t_S depends on s_P)
    d["phase"] = np.ones_like(d["phase"])
    o = _opts(nev, nst, any_p=0, joint_hypocentres=0, station_corrections=0, vpvs=1)
    rs, ci, va, rhs, (m, cols, nnz, dr) = oracle.build_system(g, o, **d, **models)
    assert cols == 2 * g.count and m == dr + 4 * g.count
