"""The serving loop (ts_server): table and LSQR requests on resident models and matrices, admitted by
a device-memory budget and run concurrently on worker streams.

Bars: every request's result is identical (bit for bit) run in a concurrent mix and run alone; the
peak of admitted memory never exceeds the budget; a request larger than the budget is declined at
submission; results equal the direct entry points' (ts_tables_solve, ts_lsqr_solve).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(seed, n, nz):
    from tomostar_amd import workloads as W
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(13.0, 13.0 + 0.012 * (n - 1), 43.0, 43.0 + 0.009 * (n - 1), 0.0, 1.0 * (nz - 1), n, n, nz)
    z = g.depth()[:, None, None]
    vp = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + W.checkerboard(g, 6.0, 6.0, 0.05)), g.shape))
    return g, (1.0 / vp).ravel(), (1.73 / vp).ravel()


@pytest.fixture(scope="module")
def setup(device):
    from tomostar_amd import lsqr as Q
    from test_lsqr_gpu import ray_like_system

    models = [_model(1, 31, 16), _model(2, 41, 21)]
    rng = np.random.default_rng(9)
    reqs = []
    for q in range(8):
        if q % 3 == 2:
            k = int(rng.integers(1, 5))
            reqs.append(("lsqr", k, 10 ** rng.uniform(-1, 2, k), ["fast_mixed", "fast_fp32"][q % 2]))
        else:
            mi = q % 2
            g = models[mi][0]
            ns = int(rng.integers(1, 6))
            src = np.stack([rng.uniform(g.min_lon, g.max_lon, ns), rng.uniform(g.min_lat, g.max_lat, ns),
                            rng.uniform(0, g.max_depth * 0.5, ns)], 1)
            reqs.append(("tables", mi, src, ["fast_fp32", "fast_fp64"][q % 2]))
    m, n, rs, ci, va, b0 = ray_like_system(600, 250, 12, 0.7, 4)
    A = Q.Matrix(device, m, n, rs, ci, va)
    B = {i: b0[None, :] + 0.1 * np.random.default_rng(i).standard_normal((r[1], m)) for i, r in enumerate(reqs)
         if r[0] == "lsqr"}
    return models, reqs, A, B


def _run(device, setup, workers, concurrent, budget_gb=8.0):
    from tomostar_amd.serving import Server

    models, reqs, A, B = setup
    srv = Server(device, workers=workers, memory_budget_gb=budget_gb)
    ids = [(srv.add_model(g, sp), srv.add_model(g, ss)) for g, sp, ss in models]
    out = {}

    def submit(i):
        r = reqs[i]
        if r[0] == "tables":
            return srv.submit_tables(ids[r[1]][0], ids[r[1]][1], r[2], r[3])
        return srv.submit_lsqr(A, B[i], r[2], 60, mode=r[3])

    def collect(i, rid):
        st, why, t = srv.wait(rid)
        assert st == 0, why
        r = reqs[i]
        if r[0] == "tables":
            g = models[r[1]][0]
            out[i] = srv.tables(rid, 0, 2 * len(r[2]), g.count)
        else:
            out[i] = srv.lsqr_result(rid, r[1], A.cols)[0]
        srv.release_request(rid)

    if concurrent:
        rids = [submit(i) for i in range(len(reqs))]
        for i, rid in enumerate(rids):
            collect(i, rid)
    else:
        for i in range(len(reqs)):
            collect(i, submit(i))
    stats = srv.stats()
    srv.close()
    return out, stats


def test_same_result_alone_and_in_the_mix(device, setup):
    mixed, st = _run(device, setup, workers=4, concurrent=True)
    alone, _ = _run(device, setup, workers=1, concurrent=False)
    assert st["peak_inflight"] >= 2
    for i in mixed:
        assert np.array_equal(mixed[i], alone[i]), i


def test_results_equal_the_direct_entry_points(device, setup):
    from tomostar_amd import eikonal as E
    from tomostar_amd import lsqr as Q

    models, reqs, A, B = setup
    mixed, _ = _run(device, setup, workers=3, concurrent=True)
    for i, r in enumerate(reqs):
        if r[0] == "tables":
            g, sp, ss = models[r[1]]
            seeds = []
            for p in r[2]:
                seeds.append(E.point_source(g, sp, *p))
                seeds.append(E.point_source(g, ss, *p))
            ts = E.TableSet(device, g, [sp, ss], seeds, np.tile([0, 1], len(r[2])).astype(np.int32))
            assert ts.solve(r[3])[0] == 0
            assert np.array_equal(ts.read_f32(), mixed[i])
            ts.release()
        else:
            rc, why, res = Q.solve_many(A, B[i], r[2], 60, 0.0, 0.0, 0.0, mode=r[3])
            assert rc == 0, why
            assert np.array_equal(np.stack([x.solution for x in res]), mixed[i])


def test_memory_budget_admission(device, setup):
    from tomostar_amd import _lib
    from tomostar_amd.serving import Server

    models, reqs, A, B = setup
    out, st = _run(device, setup, workers=4, concurrent=True, budget_gb=0.3)
    assert st["peak_bytes"] <= st["budget_bytes"]
    srv = Server(device, workers=2, memory_budget_gb=0.05)
    g, sp, ss = models[1]
    p, s = srv.add_model(g, sp), srv.add_model(g, ss)
    with pytest.raises(_lib.HipError) as e:
        srv.submit_tables(p, s, np.array([[13.1, 43.1, 1.0]] * 40), "fast_fp32")
    assert e.value.status == _lib.TS_DECLINED
    srv.close()
