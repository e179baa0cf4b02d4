"""Parity of the device eikonal with the oracle's Relax (Forward/SphericalEikonal.cs:284-396), through
the C ABI.

Bars: the FP64 forms reproduce Relax to the last bit (the same update arithmetic in the same order,
fields solved in the reference's plane schedule); the mixed form (FP32 first order, FP64 second order) stays within 1e-4 relative error of
Relax at every computed node (measured errors are far below it); the plain
form meets the reference's device gate, <= 1e-6 of the field range (EikonalOpenCl.cs:45).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


def _selftest_case():
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(10, 12, 42, 44, -2, 120, 22, 20, 16)
    k, j, i = np.meshgrid(np.arange(g.nz), np.arange(g.ny), np.arange(g.nx), indexing="ij")
    s = 1 / (5.8 + 0.02 * g.depth()[k] + 0.2 * np.sin(0.7 * i) + 0.15 * np.cos(0.5 * j))
    return g, np.ascontiguousarray(s.ravel())


def _checker_case(seed, nx=31, ny=29, nz=17, nst=6):
    from tomostar_amd import workloads as W
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(13.0, 13.4, 43.0, 43.3, -1.0, 20.0, nx, ny, nz)
    rng = np.random.default_rng(seed)
    z = g.depth()[:, None, None]
    vp = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + W.checkerboard(g, 8.0, 6.0, 0.08)), g.shape))
    st = np.stack([rng.uniform(g.min_lon, g.max_lon, nst), rng.uniform(g.min_lat, g.max_lat, nst),
                   rng.uniform(-1.0, 12.0, nst)], axis=1)
    return W.Workload("checker", g, vp, vp / 1.73, st)


def _relax_fields(oracle, g, models, seeds, mof):
    og = _og(oracle, g)
    out = []
    for (idx, t), m in zip(seeds, mof):
        f, ok, _ = oracle.relax(og, models[m], idx, t)
        assert ok
        out.append(f)
    return np.array(out)


def _rel_err(dev, ref, seeds):
    mask = np.isfinite(ref) & (ref > 0)
    for f, (idx, _) in enumerate(seeds):
        mask[f, idx] = False
    return np.max(np.abs(dev[mask] - ref[mask]) / ref[mask])


def test_device_opens_and_passes_the_self_test(device):
    eik, lsqr = device.self_test_errors()
    assert eik <= 1e-6
    assert "gfx950" in device.info


@pytest.mark.parametrize("mode", ["fast_fp64", "fast_fp32", "plain_fp64"])
def test_selftest_field_against_relax(device, oracle, mode):
    from tomostar_amd import eikonal as E

    g, s = _selftest_case()
    seeds = [E.point_source(g, s, 10.37, 42.61, 8.3)]
    ts = E.TableSet(device, g, [s], seeds)
    rc, why = ts.solve(mode)
    assert rc == 0, why
    dev = ts.read()
    ref = _relax_fields(oracle, g, [s], seeds, [0])
    if mode == "fast_fp32":
        assert _rel_err(dev, ref, seeds) < FP32_TOL
    else:
        assert np.array_equal(dev, ref)
    ref_rounds = oracle.relax(_og(oracle, g), s, *seeds[0])[2]
    assert tuple(ts.field_rounds()[0]) == ref_rounds or mode == "fast_fp32"


def test_mixed_models_many_groups_fp64_bit_identical(device, oracle):
    """70 fields (two 64-lane groups, the second partly empty), P and S models interleaved."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = _checker_case(5, nx=21, ny=19, nz=11, nst=35)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    models = [wl.slowness_p, wl.slowness_s]
    ts = E.TableSet(device, wl.grid, models, seeds, mof)
    rc, why = ts.solve("fast_fp64")
    assert rc == 0, why
    dev = ts.read()
    ref = _relax_fields(oracle, wl.grid, models, seeds, mof)
    assert np.array_equal(dev, ref)
    # the FP32 form on the same table set
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    assert _rel_err(ts.read(), ref, seeds) < FP32_TOL
    # stored tables are the float32 rounding of the times (TravelTimeTable.cs:73)
    t32 = ts.read_f32()
    assert np.array_equal(t32, ts.read().astype(np.float32))


@pytest.mark.parametrize("seed", [11, 12])
def test_checkerboard_fields(device, oracle, seed):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = _checker_case(seed)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    models = [wl.slowness_p, wl.slowness_s]
    ref = _relax_fields(oracle, wl.grid, models, seeds, mof)
    ts = E.TableSet(device, wl.grid, models, seeds, mof)
    assert ts.solve("fast_fp64")[0] == 0
    assert np.array_equal(ts.read(), ref)
    assert ts.solve("fast_fp32")[0] == 0
    assert _rel_err(ts.read(), ref, seeds) < FP32_TOL
    assert ts.solve("plain_fp64")[0] == 0
    plain = ts.read()
    rng_ = np.nanmax(np.where(np.isfinite(ref), ref, np.nan), axis=1) - np.nanmin(ref, axis=1)
    assert np.max(np.abs(plain - ref).max(axis=1) / rng_) <= 1e-6


def test_plain_single_field_bit_identical(device, oracle):
    from tomostar_amd import eikonal as E

    g, s = _selftest_case()
    seeds = [E.point_source(g, s, 11.2, 43.1, 30.0)]
    hip = E.EikonalHip(device, mode="plain_fp64")
    out = np.zeros(g.count)
    ok, why = hip.try_solve(g, s, seeds, out)
    assert ok, why
    ref = _relax_fields(oracle, g, [s], seeds, [0])[0]
    assert np.array_equal(out, ref)
    assert hip.last_stats.endswith("rounds of 8 sweeps")


def test_unreachable_region_and_minimal_grid(device, oracle):
    from tomostar_amd import eikonal as E
    from tomostar_amd.grid import GridDefinition

    # a slab of zero slowness (not a valid medium) cuts the grid: nodes there are never updated
    g, s = _selftest_case()
    s = s.copy().reshape(g.shape)
    s[:, :, 10] = 0.0
    s = s.ravel()
    seeds = [E.point_source(g, s, 10.37, 42.61, 8.3)]
    ts = E.TableSet(device, g, [s], seeds)
    ref = _relax_fields(oracle, g, [s], seeds, [0])
    for mode in ("fast_fp64", "plain_fp64"):
        assert ts.solve(mode)[0] == 0
        assert np.array_equal(ts.read(), ref)
    # 2 x 2 x 2 grid, one seed
    g2 = GridDefinition(10, 10.01, 40, 40.01, 0, 1, 2, 2, 2)
    s2 = np.full(8, 0.2)
    seeds2 = [(np.array([0], np.int32), np.array([0.0]))]
    ts2 = E.TableSet(device, g2, [s2], seeds2)
    assert ts2.solve("fast_fp64")[0] == 0
    ref2 = _relax_fields(oracle, g2, [s2], seeds2, [0])
    assert np.array_equal(ts2.read(), ref2)


def test_declines_and_errors(device):
    from tomostar_amd import _lib
    from tomostar_amd import eikonal as E

    g, s = _selftest_case()
    # seeds outside the grid only: no seed nodes -> decline at creation
    with pytest.raises(_lib.HipError) as e:
        E.TableSet(device, g, [s], [(np.array([-5, g.count + 3], np.int32), np.array([0.0, 0.0]))])
    assert e.value.status == _lib.TS_DECLINED
    # model index out of range
    with pytest.raises(_lib.HipError) as e:
        E.TableSet(device, g, [s], [E.point_source(g, s, 11, 43, 5)], model_of_field=[3])
    assert e.value.status == _lib.TS_EINVAL
    # cancellation before the first round
    ts = E.TableSet(device, g, [s], [E.point_source(g, s, 11, 43, 5)])
    flag = C.c_int32(1)
    rc, why = ts.solve("fast_fp32", cancel=flag)
    assert rc == _lib.TS_CANCELLED
    rc, why = ts.solve("fast_fp32")
    assert rc == 0


def test_duplicate_seeds_keep_the_smallest_time(device, oracle):
    from tomostar_amd import eikonal as E

    g, s = _selftest_case()
    idx, t = E.point_source(g, s, 10.37, 42.61, 8.3)
    dup_idx = np.concatenate([idx, idx[:5]]).astype(np.int32)
    dup_t = np.concatenate([t, t[:5] - 0.01])
    ts = E.TableSet(device, g, [s], [(dup_idx, dup_t)])
    assert ts.solve("fast_fp64")[0] == 0
    ref = oracle.relax(_og(oracle, g), s, dup_idx, dup_t)[0]
    assert np.array_equal(ts.read()[0], ref)


def test_full_size_batched_mixed_against_fp64(device):
    """The 400-field, 201 x 201 x 101 batch: the mixed form against the FP64 form (which equals Relax to
    the last bit, as the tests above show on every field they check), on every tenth field."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = W.batched_eikonal_case(seed=1, stations=200)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    picks = list(range(0, len(seeds), 10))
    assert ts.solve("fast_fp64")[0] == 0
    ref = np.stack([ts.read(f, 1)[0] for f in picks])
    assert ts.solve("fast_fp32")[0] == 0
    dev = np.stack([ts.read(f, 1)[0] for f in picks])
    err = _rel_err(dev, ref, [seeds[f] for f in picks])
    print(f"max relative error of the mixed form over {len(picks)} fields: {err:.3e}")
    assert err < FP32_TOL
    # reads large enough to go through the pinned staging buffers (whole chunks and a partial one)
    # return what field-by-field reads return
    all32 = ts.read_f32(0, len(seeds))
    for f in (0, 63, 64, 200, 399):
        assert np.array_equal(all32[f], ts.read_f32(f, 1)[0], equal_nan=True), f
    del all32
    some64 = ts.read(0, 40)
    for f in (0, 17, 39):
        assert np.array_equal(some64[f], ts.read(f, 1)[0], equal_nan=True), f
    ts.release()


def test_mixed_schedule_leaves_the_times_unchanged(device, monkeypatch):
    """How a round reads the dirty flags (every load at once, the flag first, a few or more flags per wave) and
    how many streams the groups sweep on change when loads are issued, not which nodes update nor their
    arithmetic: no node of a plane reads or wakes another node of the same plane, and the groups are
    independent fields. The mixed form's times are bit-identical under every schedule, also when the
    stream count changes between solves of one table set (its graphs then cover other groups)."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = _checker_case(7, nx=41, ny=37, nz=21, nst=70)  # 140 fields: three groups of 64 lanes
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    base = ts.read()
    assert np.isfinite(base).any()
    schedules = [
        {"TOMOSTAR_SPARSE_FROM": "99", "TOMOSTAR_SECOND_SPARSE_FROM": "99"},  # flag first, no sparse rounds
        {"TOMOSTAR_FP64_FLAG_FIRST": "0"},  # second-order phase sparse from its round 0 (no all-loads round)
        {"TOMOSTAR_SECOND_SPARSE_FROM": "3"},  # second-order rounds 1-2 flag first
        {"TOMOSTAR_WIDE_FROM": "1", "TOMOSTAR_SECOND_WIDE_FROM": "1"},  # more flags per wave from round 1
        {"TOMOSTAR_SECOND_WIDE_FROM": "99"},  # never more flags per wave
        {"TOMOSTAR_GROUP_STREAMS": "1"},
        {"TOMOSTAR_GROUP_STREAMS": "2"},  # chunk 1 now holds groups 1-2 (with 3 streams: group 1)
        {},
    ]
    for env in schedules:
        for k in ("TOMOSTAR_SPARSE_FROM", "TOMOSTAR_SECOND_SPARSE_FROM", "TOMOSTAR_FP64_FLAG_FIRST",
                  "TOMOSTAR_GROUP_STREAMS", "TOMOSTAR_WIDE_FROM", "TOMOSTAR_SECOND_WIDE_FROM"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        print(f"schedule {env}", flush=True)
        rc, why = ts.solve("fast_fp32")
        assert rc == 0, why
        assert np.array_equal(ts.read(), base, equal_nan=True), env
        if "TOMOSTAR_WIDE_FROM" in env:  # both phases' rounds from 1 on read more flags per wave
            orders = [(w, o) for w, o, _ in ts.round_log()]
            assert (32, 4) in orders and (64, 4) in orders, orders
    ts.release()


@pytest.mark.parametrize("nst", [70, 40, 50])
def test_narrow_group_four_nodes_per_wave(device, monkeypatch, nst):
    """A last group of 1-16 fields (140 fields: 12 over two groups; 80: 16 over one) updates four plane
    nodes per wave, 16 lanes each, in the first-order phase and the second-order phase's rounds
    before the late four-flag ones; 100 fields leave 36, a full-width group. The
    times of the mixed form are the same to the bit as with one node per wave (TOMOSTAR_NARROW=0), and
    so are the counted updates."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = _checker_case(11, nx=37, ny=33, nz=19, nst=nst)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    out = {}
    for narrow in ("0", "1"):
        monkeypatch.setenv("TOMOSTAR_NARROW", narrow)
        print(f"fields {2 * nst} narrow {narrow}", flush=True)
        ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
        rc, why = ts.solve("fast_fp32")
        assert rc == 0, why
        st = ts.stats()
        out[narrow] = (ts.read(), st["updates_fp32"], st["updates_fp64"])
        ts.release()
    assert np.isfinite(out["1"][0]).any()
    assert np.array_equal(out["0"][0], out["1"][0], equal_nan=True)
    assert out["0"][1:] == out["1"][1:]


def test_first_order_wake_threshold(device, oracle, monkeypatch):
    """The mixed form's first-order phase wakes a node's readers only on drops above
    TOMOSTAR_FP32_WAKE_ULPS ulps of its time (16 by default): the times stay within the bar of Relax
    whatever the threshold, a threshold of 0 (every drop) gives the same times as one of a quarter ulp
    (no drop of an FP32 time is smaller than half an ulp of it), and the FP64 form ignores the setting (Relax to the bit under every threshold)."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = _checker_case(11, nx=37, ny=33, nz=19, nst=36)  # 72 fields: two groups, the second of 8 fields
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    models = [wl.slowness_p, wl.slowness_s]
    ref = _relax_fields(oracle, wl.grid, models, seeds, mof)
    ts = E.TableSet(device, wl.grid, models, seeds, mof)
    out = {}
    for ulps in ("0", "0.25", "16", "256"):
        monkeypatch.setenv("TOMOSTAR_FP32_WAKE_ULPS", ulps)
        print(f"wake threshold {ulps} ulps", flush=True)
        rc, why = ts.solve("fast_fp32")
        assert rc == 0, why
        out[ulps] = ts.read()
        err = _rel_err(out[ulps], ref, seeds)
        print(f"  max relative error against Relax {err:.3e}")
        assert err <= FP32_TOL
        rc, why = ts.solve("fast_fp64")
        assert rc == 0, why
        assert np.array_equal(ts.read(), ref), ulps
    assert np.array_equal(out["0"], out["0.25"])
    ts.release()


def _suggest_batch_rule(nodes, global_mem):
    """EikonalOpenCl.SuggestBatch (EikonalOpenCl.cs:121-132) with Real = 8 (an FP64 device) and MaxAlloc =
    the device's memory (HIP sets no separate allocation limit); C#'s long division truncates."""
    real = 8
    per_field = nodes * real + 2 * nodes
    budget = int(global_mem * 0.5) - 2 * nodes * real
    by_alloc = global_mem // max(1, nodes * real)
    by_host = (768 << 20) // max(1, nodes * 8)
    enough = max(4, (128 << 20) // max(1, nodes))
    return int(np.clip(min(min(budget // max(1, per_field), by_alloc), min(by_host, enough)), 0, 32))


def test_suggest_batch_follows_the_reference_rule(device):
    import re

    from tomostar_amd import _lib

    L = _lib.load()
    m = re.search(r"([0-9.]+) GB", device.info)
    assert m, device.info
    global_mem = float(m.group(1)) * 1e9
    assert global_mem > 100e9  # on this device the host and batch terms bind before the memory ones
    for nodes in (1, 1000, 115_351, 3_000_000, 4_080_501, 10_000_000, 33_554_432, 50_000_000, 100_663_296,
                  200_000_000, 2_000_000_000):
        want = _suggest_batch_rule(nodes, int(global_mem))
        # the memory terms are far from binding, so the device's exact byte count cannot change the answer
        assert want == _suggest_batch_rule(nodes, int(global_mem * 1.001)) == _suggest_batch_rule(nodes, int(global_mem * 0.999))
        assert L.ts_eik_suggest_batch(device.handle, nodes) == want, nodes
    assert L.ts_eik_suggest_batch(device.handle, 4_080_501) == 24  # 768 MiB of FP64 host buffers: 24 fields of 201 x 201 x 101


def test_host_seed_and_metric_rows_beside_the_device(oracle):
    """GridMetric and PointSource against the oracle to the last bit, run where the GPU tests run."""
    import test_abi

    test_abi.test_grid_metric_and_point_source_match_oracle(oracle)
    test_abi.test_point_source_at_faces_and_corners(oracle)
