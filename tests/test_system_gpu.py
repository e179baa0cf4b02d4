"""Parity of the device system assembly (ts_system_build) with the oracle's BuildSystem
(Tomography/TravelTimeTomography.cs:860-1034, plain node grid), through the C ABI.

Bar: the same matrix to the last bit - row starts, column indices and FP64 values - and the same
right-hand side, since every entry is the reference's product in the reference's order. The stored
transpose equals the counting-sort transpose (SparseMatrix.cs:79-97) of that matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


@pytest.fixture(scope="module")
def traced(device):
    """The 61 x 61 x 31 one-iteration case: tables, every station x event x phase ray (7,500)."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import rays as Ry
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    assert ts.solve("fast_fp32")[0] == 0
    ns, ne = len(wl.stations), len(wl.events)
    field = np.array([2 * s + ph for e in range(ne) for s in range(ns) for ph in (0, 1)], np.int32)
    events = np.array([wl.events[e] for e in range(ne) for s in range(ns) for ph in (0, 1)])
    batch = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), field, events)
    arrivals = dict(phase=field % 2, station=field // 2, event=np.repeat(np.arange(ne), 2 * ns).astype(np.int32))
    return wl, ts, batch, arrivals


def _models(wl, rng):
    n = wl.grid.count
    sRefP = wl.slowness_p.ravel() * (1 + 0.02 * rng.standard_normal(n))
    kRef = np.full(n, 1.73)
    return dict(sP=wl.slowness_p.ravel(), sS=wl.slowness_s.ravel(), sRefP=sRefP, sRefS=sRefP * kRef, kRef=kRef)


CASES = {
    # the reference defaults on a joint P/S inversion (VpVpVs, Laplacian in km, hypocentres, stations)
    "vpvs_laplacian": dict(settings=dict(vpvs=True), damping=1.0, smoothing=5.0, weight=1.0),
    "vpvp_gradient_nodes": dict(settings=dict(vpvs=False, smoothing_method="gradient", smoothing_scale="nodes",
                                              vertical_weight=0.5, s_factor=2.0), damping=0.7, smoothing=3.0,
                                weight=1.0),
    "p_only_no_hypocentres": dict(settings=dict(vpvs=False, joint_hypocentres=False, station_corrections=False,
                                                invert_s=False), damping=0.0, smoothing=2.0, weight=2.0,
                                  p_only=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_system_matches_the_oracle(device, oracle, traced, name):
    from tomostar_amd import system as Sy

    wl, ts, batch, arr = traced
    cfg = CASES[name]
    rng = np.random.default_rng(len(name))
    ne, ns = len(wl.events), len(wl.stations)
    # used rows: arrivals with a ray, minus a few "rejected" ones, in arrival order
    has = np.flatnonzero(batch.counts >= 0)
    keep = has[rng.random(has.size) > 0.05]
    if cfg.get("p_only"):
        keep = keep[arr["phase"][keep] == 0]
    rows = keep.astype(np.int32)
    phase, event, station = arr["phase"][rows], arr["event"][rows], arr["station"][rows]
    residual = 0.1 * rng.standard_normal(rows.size)
    sigma = np.where(phase == 0, 0.05, 0.1)
    ev_fixed = (rng.random(ne) < 0.1).astype(np.uint8)
    ev_outside = (rng.random(ne) < 0.1).astype(np.uint8)
    models = _models(wl, rng)
    st = Sy.SystemSettings(**cfg["settings"])
    sysd = Sy.build_system(device, wl.grid, st, batch, rows, phase, event, station, residual, sigma, ev_fixed,
                           ev_outside, ns, **models, damping=cfg["damping"], smoothing=cfg["smoothing"],
                           absolute_weight=cfg["weight"], keep_fp64=True)
    # the oracle on the same rows
    rs_rows = np.concatenate([[0], np.cumsum(batch.counts[rows])]).astype(np.int64)
    nodes = np.concatenate([batch.nodes[batch.row_start[r]:batch.row_start[r + 1]] for r in rows])
    lens = np.concatenate([batch.lengths[batch.row_start[r]:batch.row_start[r + 1]] for r in rows])
    o = dict(vpvs=int(st.vpvs), joint_hypocentres=int(st.joint_hypocentres),
             station_corrections=int(st.station_corrections), invert_p=int(st.invert_p), invert_s=int(st.invert_s),
             gradient_smoothing=int(st.smoothing_method == "gradient"),
             smoothing_in_nodes=int(st.smoothing_scale == "nodes"), nevents=ne, nstations=ns,
             any_p=int(np.any(phase == 0)), any_s=int(np.any(phase == 1)), damping=cfg["damping"],
             smoothing=cfg["smoothing"], vertical_weight=st.vertical_weight,
             damping_hypocentre=st.damping_hypocentre, damping_station=st.damping_station, s_factor=st.s_factor,
             absolute_weight=cfg["weight"])
    rs, ci, va, rhs, (m, cols, nnz, drows) = oracle.build_system(
        _og(oracle, wl.grid), o, rs_rows, nodes, lens, batch.tcalc_grad[rows], phase, event, station, residual,
        sigma, ev_fixed, ev_outside, **models)
    L = sysd.layout
    assert (L["rows"], L["columns"], L["data_rows"]) == (m, cols, drows)
    assert sysd.matrix.nnz == nnz
    drs, dci, dva = sysd.read()
    assert np.array_equal(drs, rs.astype(np.int32))
    assert np.array_equal(dci, ci)
    assert np.array_equal(dva, va)
    assert np.array_equal(sysd.rhs, rhs)
    # the stored transpose: column-ordered rows, FP32 values
    T = sp.csr_matrix((va, ci, rs), shape=(m, cols)).T.tocsr()
    T.has_sorted_indices = False
    trs, tci, tva = sysd.read_transpose()
    assert np.array_equal(trs, T.indptr) and np.array_equal(tci, T.indices)
    assert np.array_equal(tva, T.data.astype(np.float32))
    print(f"{name}: {m} rows, {cols} columns, {nnz} non-zeros, assembled in {sysd.build_seconds:.3f} s")


def test_invalid_rows_are_rejected(device, traced):
    from tomostar_amd import _lib
    from tomostar_amd import system as Sy

    wl, ts, batch, arr = traced
    n = wl.grid.count
    models = dict(sP=np.ones(n), sS=np.ones(n), sRefP=np.ones(n), sRefS=np.ones(n), kRef=np.ones(n))
    bad = np.flatnonzero(batch.counts < 0)
    rows = np.array([bad[0]] if bad.size else [batch.nrays + 5], np.int32)
    with pytest.raises(_lib.HipError) as e:
        Sy.build_system(device, wl.grid, Sy.SystemSettings(), batch, rows, [0], [0], [0], [0.0], [0.1],
                        np.zeros(len(wl.events), np.uint8), np.zeros(len(wl.events), np.uint8), len(wl.stations),
                        **models, damping=1.0, smoothing=1.0)
    assert e.value.status == _lib.TS_EINVAL
