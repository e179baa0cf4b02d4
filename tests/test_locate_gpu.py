"""Event location's grid search on the device (ts_locate_grid_search) against the oracle's restatement
of EventLocator.GridSearch on the same float32 tables: on the 61 x 61 x 31 case (25 stations x P, S)
with 40 events' noisy picks, quarter/half/full weights and station corrections, the chosen node,
origin time and misfit are equal to the last bit - global search, local search around a start, and
fixed depth. (Every node of this case is reachable; the cost of an unreachable table value is pinned
on the oracle, test_oracle_locate.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(device):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    g = wl.grid
    seeds = [E.point_source(g, wl.slowness_p if ph == 0 else wl.slowness_s, *st) for st in wl.stations for ph in (0, 1)]
    mof = np.tile(np.array([0, 1], np.int32), len(wl.stations))
    ts = E.TableSet(device, g, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    tabs = ts.read_f32()
    yield wl, ts, tabs
    ts.release()


def _events(wl, tabs, n, seed):
    from tomostar_amd.locate import Picks

    g = wl.grid
    rng = np.random.default_rng(seed)
    out = []
    for e in range(n):
        node = int(rng.integers(0, g.count))
        field = rng.choice(len(tabs), size=int(rng.integers(6, 40)), replace=False).astype(np.int32)
        t0 = rng.uniform(-2, 2)
        obs = np.array([t0 + float(tabs[f][node]) for f in field]) + rng.normal(0, 0.05, field.size)
        w = rng.choice([1.0, 0.5, 0.25], field.size)
        corr = rng.normal(0, 0.02, field.size)
        k, j, i = node // (g.nx * g.ny), (node // g.nx) % g.ny, node % g.nx
        start = (g.lon()[i] + 0.03, g.lat()[j] - 0.02, g.depth()[k] + 2.2)
        out.append(Picks(field, obs, w, corr, start))
    return out


def _oracle_all(oracle, g, tabs, events, settings):
    from tomostar_amd import locate as Lc

    res = []
    for e in events:
        center, kf = None, -1
        if not settings.global_search or settings.fix_depth:
            fr = g.fractional(*e.start)
            rnd = [int(round(float(v))) for v in fr]
            if not settings.global_search:
                center = tuple(rnd)
            if settings.fix_depth:
                kf = int(np.clip(rnd[2], 0, g.nz - 1))
        res.append(oracle.locate_grid_search(tabs, g.nx, g.ny, g.nz, (e.field, e.obs, e.weight, e.correction),
                                             settings.coarse_candidates, center, kf))
    return res


@pytest.mark.parametrize("glob,fix", [(True, False), (False, False), (True, True)])
def test_grid_search_matches_oracle(device, oracle, case, glob, fix):
    from tomostar_amd import locate as Lc

    wl, ts, tabs = case
    g = wl.grid
    events = _events(wl, tabs, 40, seed=5)
    s = Lc.LocatorSettings(global_search=glob, coarse_candidates=8000, fix_depth=fix)
    lon, lat, dep, t0, node, mis = Lc.grid_search(ts, g, events, s)
    ref = _oracle_all(oracle, g, tabs, events, s)
    for e, (n, t, m) in enumerate(ref):
        assert node[e] == n and t0[e] == t and mis[e] == m, (e, node[e], n, t0[e], t, mis[e], m)
    if fix:
        assert np.array_equal(dep, [e.start[2] for e in events])


def test_picks_beyond_the_limit_are_declined(device, case):
    from tomostar_amd import _lib
    from tomostar_amd import locate as Lc

    wl, ts, tabs = case
    many = Lc.Picks(np.zeros(600, np.int32), np.zeros(600), np.ones(600), np.zeros(600), (13.3, 43.2, 5.0))
    with pytest.raises(_lib.HipError):
        Lc.grid_search(ts, wl.grid, [many])
