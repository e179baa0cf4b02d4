"""The streaming sweep (one persistent launch per ordering over column tiles, csrc/eikonal_stream.hpp)
against Relax (Forward/SphericalEikonal.cs:284-396) and against the plane sweep, through the C ABI.

A tile walks the diagonal planes of its columns in order and waits for its upwind tiles, so every node
is updated after its upwind neighbours and before its downwind ones, with the same arithmetic and the
same set of dirty nodes as the plane sweep: the FP64 form equals Relax to the last bit, and every form
equals its plane-sweep result to the last bit, whichever rounds stream. The cases cover tiles cut by
the grid's edges, several 64-field groups (the last partly empty), both phases and both widths.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KNOBS = ("TOMOSTAR_STREAM_ROUNDS", "TOMOSTAR_STREAM_ROUNDS2", "TOMOSTAR_STREAM_SPIN")


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


def _case(seed, nx, ny, nz, nst):
    from tomostar_amd import workloads as W
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(13.0, 13.4, 43.0, 43.3, -1.0, 20.0, nx, ny, nz)
    rng = np.random.default_rng(seed)
    z = g.depth()[:, None, None]
    vp = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + W.checkerboard(g, 8.0, 6.0, 0.08)), g.shape))
    st = np.stack([rng.uniform(g.min_lon, g.max_lon, nst), rng.uniform(g.min_lat, g.max_lat, nst),
                   rng.uniform(-1.0, 12.0, nst)], axis=1)
    return W.Workload("checker", g, vp, vp / 1.73, st)


def _tables(device, wl):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    seeds, mof = W.station_seed_sets(wl, E.point_source)
    models = [wl.slowness_p, wl.slowness_s]
    return E.TableSet(device, wl.grid, models, seeds, mof), models, seeds, mof


def _set(monkeypatch, env):
    for k in KNOBS:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)


def _solve(ts, mode):
    rc, why = ts.solve(mode)
    assert rc == 0, why
    return ts.read()


@pytest.mark.parametrize("shape", [(21, 19, 11, 35), (18, 27, 9, 33)])
def test_streaming_fp64_form_equals_relax(device, oracle, monkeypatch, shape):
    """Every round streaming; grids whose x and y extents are not multiples of the tile's."""
    nx, ny, nz, nst = shape
    wl = _case(5, nx, ny, nz, nst)
    ts, models, seeds, mof = _tables(device, wl)
    _set(monkeypatch, {"TOMOSTAR_STREAM_ROUNDS": "999"})
    print(f"streaming FP64 form on {nx}x{ny}x{nz}, {len(seeds)} fields", flush=True)
    dev = _solve(ts, "fast_fp64")
    og = _og(oracle, wl.grid)
    for f, ((idx, t0), m) in enumerate(zip(seeds, mof)):
        ref, ok, _ = oracle.relax(og, models[m], idx, t0)
        assert ok
        assert np.array_equal(dev[f], ref), f
    ts.release()


def test_streaming_equals_the_plane_sweep(device, monkeypatch):
    """Both forms, streaming in every round, in the first rounds only, or not at all: bit-identical.
    140 fields (three groups), 45 x 39 x 23 nodes."""
    wl = _case(7, 45, 39, 23, 70)
    ts, _, _, _ = _tables(device, wl)
    for mode in ("fast_fp64", "fast_fp32"):
        _set(monkeypatch, {})
        base = _solve(ts, mode)
        assert np.isfinite(base).any()
        for env in ({"TOMOSTAR_STREAM_ROUNDS": "999", "TOMOSTAR_STREAM_ROUNDS2": "999"},
                    {"TOMOSTAR_STREAM_ROUNDS": "1", "TOMOSTAR_STREAM_ROUNDS2": "2"},
                    {"TOMOSTAR_STREAM_ROUNDS2": "1"}):
            _set(monkeypatch, env)
            print(f"{mode} {env}", flush=True)
            assert np.array_equal(_solve(ts, mode), base, equal_nan=True), (mode, env)
    ts.release()


def test_streaming_larger_grid_equals_the_plane_sweep(device, monkeypatch):
    """Four groups on 97 x 83 x 41 nodes (many tiles in flight at once, long upwind chains)."""
    wl = _case(3, 97, 83, 41, 125)
    ts, _, _, _ = _tables(device, wl)
    _set(monkeypatch, {})
    base64 = _solve(ts, "fast_fp64")
    base32 = _solve(ts, "fast_fp32")
    _set(monkeypatch, {"TOMOSTAR_STREAM_ROUNDS": "999", "TOMOSTAR_STREAM_ROUNDS2": "999"})
    print("streaming, FP64 form", flush=True)
    assert np.array_equal(_solve(ts, "fast_fp64"), base64, equal_nan=True)
    print("streaming, mixed form", flush=True)
    assert np.array_equal(_solve(ts, "fast_fp32"), base32, equal_nan=True)
    ts.release()


def test_streaming_wait_limit_raises_an_error(device, monkeypatch):
    """A wait beyond the spin limit (forced here by a limit of zero polls) is reported as a device error,
    not a hang; the next solve with the normal limit is correct again."""
    wl = _case(9, 33, 29, 13, 20)
    _set(monkeypatch, {"TOMOSTAR_STREAM_ROUNDS": "999", "TOMOSTAR_STREAM_SPIN": "0"})
    ts, _, _, _ = _tables(device, wl)  # the spin limit is read when the table set is made
    rc, why = ts.solve("fast_fp64")
    assert rc != 0 and "spin limit" in why, (rc, why)
    ts.release()
    _set(monkeypatch, {})
    ts, _, _, _ = _tables(device, wl)
    base = _solve(ts, "fast_fp64")
    _set(monkeypatch, {"TOMOSTAR_STREAM_ROUNDS": "999"})
    assert np.array_equal(_solve(ts, "fast_fp64"), base, equal_nan=True)
    ts.release()
