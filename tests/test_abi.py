"""The C ABI library: it loads, it exports every entry point include/tomostar_hip.h declares, and its
host-side helpers (GridMetric, PointSource) agree with the oracle to the last bit. No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(ts_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from tomostar_amd import _lib

    L = _lib.load()
    declared = _declared()
    assert len(declared) >= 15
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from tomostar_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def _cases():
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=3)
    return wl


def test_grid_metric_and_point_source_match_oracle(oracle):
    from tomostar_amd import eikonal as E

    wl = _cases()
    g = wl.grid
    og = oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)
    a, b = E.metric(g), oracle.metric(og)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    for st in list(wl.stations[:6]) + list(wl.events[:6]):
        for s in (wl.slowness_p, wl.slowness_s):
            x, y = E.point_source(g, s, *st), oracle.point_source(og, s, *st)
            assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_point_source_at_faces_and_corners(oracle):
    from tomostar_amd import eikonal as E
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(10, 11, 40, 41, -1, 20, 12, 13, 9)
    og = oracle.griddef(10, 11, 40, 41, -1, 20, 12, 13, 9)
    s = np.linspace(0.15, 0.2, g.count)
    for p in [(10, 40, -1), (11, 41, 20), (10.5, 40.5, 50), (9.9, 40.2, 3), (10.99999, 40.5, 7.5)]:
        x, y = E.point_source(g, s, *p), oracle.point_source(og, s, *p)
        assert len(x[0]) > 0 and np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_invalid_arguments_are_rejected():
    from tomostar_amd import _lib

    L = _lib.load()
    bad = _lib.TsGrid(1, 5, 5, 0, 1, 0, 1, 0, 1)  # one node on an axis
    out = np.zeros(4)
    assert L.ts_grid_metric(C.byref(bad), out.ctypes.data, out.ctypes.data, C.byref(C.c_double())) == _lib.TS_EINVAL
    assert L.ts_eik_suggest_batch(None, 1000) == _lib.TS_EINVAL
    r = _lib.reason_buffer()
    assert L.ts_tables_create(None, None, 1, None, 1, None, None, None, None, None, r, len(r)) == _lib.TS_EINVAL
    from tomostar_amd import lsqr as Q

    Q._declare()
    assert L.ts_matrix_multiply(None, 1, None, None, 0, r, len(r)) == _lib.TS_EINVAL
    assert b"invalid argument" in r.value


def test_system_build_marshalling_without_a_device():
    """The Python front end of ts_system_build passes every argument of the C signature: with no
    context the library rejects the call before touching a GPU."""
    from types import SimpleNamespace

    from tomostar_amd import _lib
    from tomostar_amd import system as Sy
    from tomostar_amd.grid import GridDefinition

    g = GridDefinition(13.0, 13.2, 43.0, 43.2, 0.0, 10.0, 5, 5, 4)
    n = g.count
    nothing = SimpleNamespace(handle=None)
    models = dict(sP=np.ones(n), sS=np.ones(n), sRefP=np.ones(n), sRefS=np.ones(n), kRef=np.ones(n))
    with pytest.raises(_lib.HipError) as e:
        Sy.build_system(nothing, g, Sy.SystemSettings(), nothing, [0], [0], [0], [0], [0.0], [0.1],
                        np.zeros(1, np.uint8), np.zeros(1, np.uint8), 1, **models, damping=1.0, smoothing=1.0)
    assert e.value.status == _lib.TS_EINVAL
    L = Sy._declare()
    assert L.ts_matrix_shape(None, None, None, None, None) == _lib.TS_EINVAL
    assert L.ts_matrix_read(None, None, None, None, None) == _lib.TS_EINVAL
    assert L.ts_matrix_read_transpose(None, None, None, None) == _lib.TS_EINVAL


def test_gridbatch_rejects_without_a_device():
    from tomostar_amd import _lib
    from tomostar_amd import gridbatch as Gb

    L = Gb._declare()
    r = _lib.reason_buffer()
    assert L.ts_gridbatch_create(None, 1, None, None, None, None, None, None, r, len(r)) == _lib.TS_EINVAL
    assert L.ts_gridbatch_solve(None, 1, None, None, r, len(r)) == _lib.TS_EINVAL
    assert L.ts_gridbatch_read(None, 0, None) == _lib.TS_EINVAL
    assert L.ts_gridbatch_tune(None, None) == _lib.TS_EINVAL
