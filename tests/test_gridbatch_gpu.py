"""Parity of the uneven batch (ts_gridbatch: many problems, each on its own grid, swept together)
with the oracle's Relax (Forward/SphericalEikonal.cs:284-396) per problem, through the C ABI.

Bars: the FP64 form equals Relax to the last bit on every problem; the mixed form (first order on
FP32 times, second order on FP64 times) stays within 1e-4 relative error; a problem's result is the
same solved alone and in the batch, and the same after the tuner changes the launch shape. The
streaming sweep (one persistent launch per ordering over every problem's column tiles) updates the
same nodes in a topological order of each ordering, so it equals the plane sweep to the bit, at every
tile width.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


@pytest.fixture(scope="module")
def small_batch():
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    probs = W.uneven_batch_case(seed=5, problems=7, min_h=11, max_h=31, min_v=7, max_v=17)
    seeds = [E.point_source(p.grid, p.slowness, *p.source) for p in probs]
    return probs, seeds


@pytest.fixture(scope="module")
def relax_ref(oracle, small_batch):
    probs, seeds = small_batch
    out = []
    for p, (idx, t) in zip(probs, seeds):
        f, ok, _ = oracle.relax(_og(oracle, p.grid), p.slowness, idx, t)
        assert ok
        out.append(f)
    return out


def test_fp64_batch_equals_relax_on_every_problem(device, small_batch, relax_ref):
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    rc, why = b.solve("fast_fp64")
    assert rc == 0, why
    for q in range(len(probs)):
        assert np.array_equal(b.read(q), relax_ref[q]), q


def test_mixed_batch_within_1e4_and_tuner(device, small_batch, relax_ref):
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    rc, why = b.solve("fast_fp32")
    assert rc == 0, why
    first = [b.read(q) for q in range(len(probs))]
    for q, (idx, _) in enumerate(seeds):
        ref = relax_ref[q]
        mask = np.isfinite(ref) & (ref > 0)
        mask[idx] = False
        err = np.max(np.abs(first[q][mask] - ref[mask]) / ref[mask])
        assert err < 1e-4, (q, err)
    choice = b.tune()
    assert all(c in (1, 2, 4) for c in choice)
    rc, why = b.solve("fast_fp32")
    assert rc == 0, why
    for q in range(len(probs)):
        assert np.array_equal(b.read(q), first[q])


def test_alone_and_in_the_batch_agree(device, small_batch):
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    assert b.solve("fast_fp32")[0] == 0
    for q in (0, len(probs) - 1):
        alone = GridBatch(device, [probs[q].grid], [probs[q].slowness], [seeds[q]])
        assert alone.solve("fast_fp32")[0] == 0
        assert np.array_equal(alone.read(0), b.read(q))
        alone.release()


def test_declines_and_invalid(device, small_batch):
    from tomostar_amd import _lib
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    with pytest.raises(_lib.HipError) as e:
        GridBatch(device, [probs[0].grid], [probs[0].slowness], [(np.array([-1], np.int32), np.array([0.0]))])
    assert e.value.status == _lib.TS_DECLINED
    with pytest.raises(ValueError):
        GridBatch(device, [probs[0].grid], [probs[0].slowness[:-1]], [seeds[0]])


@pytest.mark.parametrize("width", [4, 8, 16])
def test_streaming_fp64_equals_relax(device, small_batch, relax_ref, width):
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    b.set_schedule("stream", [width] * 3)
    assert b.schedule()["method"] == "stream"
    print(f"streaming sweep, tile width {width}", flush=True)
    rc, why = b.solve("fast_fp64")
    assert rc == 0, why
    for q in range(len(probs)):
        assert np.array_equal(b.read(q), relax_ref[q]), q
    b.release()


def test_streaming_mixed_equals_planes_and_tuner(device, small_batch):
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    b.set_schedule("planes")
    assert b.solve("fast_fp32")[0] == 0
    planes = [b.read(q) for q in range(len(probs))]
    for method, widths in (("stream", [4, 8, 16]), ("stream", [16, 4, 8]), ("hybrid", [8, 8, 8])):
        b.set_schedule(method, widths)
        print(f"{method} mixed form, widths {widths}", flush=True)
        rc, why = b.solve("fast_fp32")
        assert rc == 0, why
        for q in range(len(probs)):
            assert np.array_equal(b.read(q), planes[q]), (method, widths, q)
        if method == "hybrid":
            methods = [m for _, m, _ in b.round_log()]
            assert methods[0] == "planes" and "stream" in methods, methods
    b.tune()
    sch = b.schedule()
    print(f"tuner: {sch}", flush=True)
    assert sch["method"] in ("planes", "stream", "hybrid") and all(w in (4, 8, 16) for w in sch["tile_width"])
    assert all(ms > 0 for ms in sch["trial_ms"].values())
    assert b.solve("fast_fp32")[0] == 0
    for q in range(len(probs)):
        assert np.array_equal(b.read(q), planes[q]), q
    b.release()


def test_streaming_wait_limit_raises_an_error(device, small_batch, monkeypatch):
    """A wait beyond the spin limit (forced by a limit of zero polls, read when the batch is made) is a
    device error, not a hang; a batch made with the normal limit then solves correctly."""
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    monkeypatch.setenv("TOMOSTAR_STREAM_SPIN", "0")
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    b.set_schedule("stream", [4, 4, 4])
    rc, why = b.solve("fast_fp64")
    assert rc != 0 and "spin limit" in why, (rc, why)
    b.release()
    monkeypatch.delenv("TOMOSTAR_STREAM_SPIN")
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    b.set_schedule("stream", [4, 4, 4])
    assert b.solve("fast_fp64")[0] == 0
    b.release()


def test_set_schedule_rejects_bad_widths(device, small_batch):
    from tomostar_amd import _lib
    from tomostar_amd.gridbatch import GridBatch

    probs, seeds = small_batch
    b = GridBatch(device, [probs[0].grid], [probs[0].slowness], [seeds[0]])
    with pytest.raises(_lib.HipError) as e:
        b.set_schedule("stream", [4, 5, 8])
    assert e.value.status == _lib.TS_EINVAL
    b.release()
