"""Parity of the device ray forward step with the oracle's Backtrack + Kernel (Forward/RayTracing.cs:73-159,
Tomography/RayKernels.cs:62-105), through the C ABI, on the same float32 tables.

Bars (north-star): ray cell lists identical for >= 99.9 % of rays, path lengths within 1e-4; TCalc and
the event gradient equal to the oracle's (the same trilinear arithmetic on the same table). Closed
form: in a homogeneous medium the ray is the chord and sum(kernel * s) is the time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


@pytest.fixture(scope="module")
def iteration_case(device):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    return wl, ts, mof


def _arrivals(wl):
    """every station x event x phase: field 2*s + phase"""
    ns, ne = len(wl.stations), len(wl.events)
    field = np.array([2 * s + ph for e in range(ne) for s in range(ns) for ph in (0, 1)], np.int32)
    events = np.array([wl.events[e] for e in range(ne) for s in range(ns) for ph in (0, 1)])
    return field, events


def test_rows_match_the_oracle(device, oracle, iteration_case):
    from tomostar_amd import rays as Ry

    wl, ts, mof = iteration_case
    field, events = _arrivals(wl)
    stations = np.repeat(wl.stations, 2, axis=0)
    batch = Ry.RayBatch(ts, wl.grid, stations, field, events)
    tables = ts.read_f32()
    og = _og(oracle, wl.grid)
    step = Ry.ray_step(wl.grid, wl.grid)
    ref = oracle.ray_rows_batch(og, list(tables), stations, og, field, events, step)
    same_cells, compared, worst_len = 0, 0, 0.0
    for r, rr in enumerate(ref):
        dev = batch.row(r)
        assert (rr is None) == (dev is None), r
        if rr is None:
            continue
        compared += 1
        nodes, lens, tcalc, ge, gn, gz = rr
        assert dev.tcalc == tcalc and (dev.ge, dev.gn, dev.gz) == (ge, gn, gz)
        if np.array_equal(nodes, dev.nodes):
            same_cells += 1
            worst_len = max(worst_len, float(np.max(np.abs(dev.length - lens) / np.maximum(lens, 1e-30))))
        total_ref, total_dev = float(np.sum(lens, dtype=np.float64)), float(np.sum(dev.length, dtype=np.float64))
        assert abs(total_dev - total_ref) <= 1e-4 * total_ref
    assert compared > 0.99 * len(ref)
    assert same_cells >= 0.999 * compared
    assert worst_len <= 1e-4


def test_homogeneous_rays_are_chords(device, oracle):
    from tomostar_amd import eikonal as E
    from tomostar_amd import rays as Ry
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition.from_spacing(13, 13.74, 43, 43.54, 0, 30, 1, 1)
    s = np.full(g.count, 1 / 6.0)
    st = np.array([[13.2, 43.1, 0.0], [13.6, 43.4, 0.0]])
    seeds = [E.point_source(g, s, *p) for p in st]
    ts = E.TableSet(device, g, [s], seeds)
    assert ts.solve("fast_fp32")[0] == 0
    ev = np.array([[13.6, 43.45, 12.0], [13.05, 43.05, 5.0], [13.7, 43.0, 3.0], [13.3, 43.3, 25.0]])
    field = np.array([0, 0, 1, 1], np.int32)
    batch = Ry.RayBatch(ts, g, st, field, ev)
    for r in range(4):
        row = batch.row(r)
        chord = np.linalg.norm(oracle.to_cartesian(*st[field[r]]) - oracle.to_cartesian(*ev[r]))
        length = float(np.sum(row.length, dtype=np.float64))
        assert abs(length - chord) <= 2e-3 * chord
        assert abs(float(np.sum(row.length * s[row.nodes])) - chord / 6) <= 2e-3 * chord / 6
        assert np.all(np.diff(row.nodes) > 0) and np.all(row.length > 0)


def test_unreached_events_give_no_row(device):
    from tomostar_amd import eikonal as E
    from tomostar_amd import rays as Ry
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(13, 13.3, 43, 43.3, 0, 20, 21, 21, 11)
    s = np.full(g.count, 1 / 6.0).reshape(g.shape)
    s[:, :, 12] = 0.0  # a wall of invalid slowness: the far side is never reached
    s = s.ravel()
    seeds = [E.point_source(g, s, 13.05, 43.1, 0.0)]
    ts = E.TableSet(device, g, [s], seeds)
    assert ts.solve("fast_fp32")[0] == 0
    ev = np.array([[13.1, 43.2, 5.0], [13.25, 43.2, 5.0]])
    batch = Ry.RayBatch(ts, g, np.array([[13.05, 43.1, 0.0]]), np.zeros(2, np.int32), ev)
    assert batch.counts[0] > 0 and batch.counts[1] == -1
    assert len(batch.rows()) == 1
