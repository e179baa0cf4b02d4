"""TravelTimeTableSet.Build on the device (tomostar_amd.tables.build): the tables of 3 stations x P, S on
the 61 x 61 x 31 case solved in one device table set and written as .qttb files; each file's times
are the device's FP32 tables, within 1e-4 relative of the FP64 form (Relax to the bit, test_eikonal_gpu);
a second build finds every table current and computes none; a changed model computes all again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_build_writes_and_reuses(device, tmp_path):
    from tomostar_amd import eikonal as E
    from tomostar_amd import tables as T
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    g = wl.grid
    src = [T.TableSource(f"ST{q:02d}", *wl.stations[q]) for q in range(3)]
    rep = T.build(device, g, wl.slowness_p, wl.slowness_s, src, str(tmp_path))
    assert len(rep.computed) == 6 and not rep.reused
    seeds = [E.point_source(g, wl.slowness_p if ph == "P" else wl.slowness_s, s.lon, s.lat, s.depth_km)
             for ph in ("P", "S") for s in src]
    ts = E.TableSet(device, g, [wl.slowness_p, wl.slowness_s], seeds, np.array([0, 0, 0, 1, 1, 1], np.int32))
    rc, why = ts.solve("fast_fp64")
    assert rc == 0, why
    ref = ts.read()
    ts.release()
    for q, (ph, s) in enumerate([(ph, s) for ph in ("P", "S") for s in src]):
        h, data = T.open_table(str(tmp_path / f"{s.station_id}.{ph}.qttb"))
        assert h.station_id == s.station_id and h.phase == ph and h.grid == g
        assert h.model_key == T.model_key(g, wl.slowness_p if ph == "P" else wl.slowness_s)
        d = np.asarray(data, np.float64)
        fin = np.isfinite(ref[q])
        assert np.array_equal(np.isfinite(d), fin)
        rel = np.abs(d[fin] - ref[q][fin]) / np.maximum(ref[q][fin], 1e-3)
        assert rel.max() < 1e-4
    again = T.build(device, g, wl.slowness_p, wl.slowness_s, src, str(tmp_path))
    assert not again.computed and len(again.reused) == 6
    changed = T.build(device, g, wl.slowness_p * 1.01, wl.slowness_s, src, str(tmp_path), phases=("P",))
    assert len(changed.computed) == 3
