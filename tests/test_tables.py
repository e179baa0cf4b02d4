"""The .qttb table files (TravelTimeTable.Write / Open / ReadHeader, Forward/TravelTimeTable.cs:57-112),
the model key (ForwardModel.Key, :173-180) and the reuse rule of TravelTimeTableSet.Build (:270-282).

The byte-exact header cannot be compared with a file the reference wrote (none is at hand): the
tests pin the format's rules - magic, little-endian length, 64-byte alignment of the data, float32
times with +inf for unreachable nodes, System.Text.Json's indented layout and number forms - and the
round trip."""
import hashlib
import os
import struct

import numpy as np

from tomostar_amd import tables as T
from tomostar_amd.grid import GridDefinition

G = GridDefinition(13.0, 13.74, 43.0, 43.54, 0.0, 30.0, 5, 4, 3)


def test_number_forms():
    assert [T._number(x) for x in (13.0, 13.74, -3.0, 0.0, 6371.0, 0.1, 1 / 3)] == \
        ["13", "13.74", "-3", "0", "6371", "0.1", "0.3333333333333333"]
    assert T._number(1e-5) == "1E-05" and T._number(1.5e-5) == "1.5E-05" and T._number(0.0001) == "0.0001"
    assert T._number(1e15) == "1E+15" and T._number(123456789012345.0) == "123456789012345"
    assert T._number(7) == "7"


def test_header_layout_and_escaping():
    h = T.TableHeader("A+B<1>", "S", 13.37, 43.27, 0.0, G, "K", "HIP fast sweeping")
    js = T.header_json(h).decode()
    lines = js.split("\n")
    assert lines[0] == "{" and lines[-1] == "}"
    assert lines[1] == '  "StationId": "A\\u002BB\\u003C1\\u003E",'
    assert '  "Grid": {' in lines and '    "Nx": 5,' in lines and '  },' in lines
    assert lines[-2] == '  "Solver": "HIP fast sweeping"'
    assert [l.split(":")[0].strip() for l in lines if l.startswith('  "')] == \
        ['"StationId"', '"Phase"', '"Lon"', '"Lat"', '"DepthKm"', '"Grid"', '"ModelKey"', '"Solver"']


def test_write_open_round_trip(tmp_path):
    n = G.nx * G.ny * G.nz
    t = np.linspace(0, 12.3456789, n)
    t[3] = np.inf
    t[5] = np.nan
    path = str(tmp_path / "ST1.P.qttb")
    h = T.TableHeader("ST1", "P", 13.1, 43.2, 1.5, G, T.model_key(G, np.full(n, 0.2)), "x")
    T.write(path, h, t)
    raw = open(path, "rb").read()
    assert raw[:8] == b"QTTB\x01\x00\x00\x00"
    n_js = struct.unpack("<i", raw[8:12])[0]
    off = 12 + n_js + (64 - (12 + n_js) % 64) % 64
    assert off % 64 == 0 and len(raw) == off + 4 * n
    assert raw[12 + n_js:off] == b"\x00" * (off - 12 - n_js)
    h2, data = T.open_table(path)
    assert h2 == h
    want = np.where(np.isfinite(t), t, np.inf).astype(np.float32)
    assert np.array_equal(np.asarray(data), want)
    assert T.read_header(str(tmp_path / "missing.qttb")) is None
    open(tmp_path / "junk.qttb", "wb").write(b"not a table")
    assert T.read_header(str(tmp_path / "junk.qttb")) is None


def test_model_key_is_sha256_of_slowness_and_grid():
    s = np.linspace(0.15, 0.25, G.nx * G.ny * G.nz)
    key = T.model_key(G, s)
    ref = hashlib.sha256(s.astype("<f8").tobytes() + T.grid_json(G).encode()).hexdigest()[:32].upper()
    assert key == ref and len(key) == 32
    s2 = s.copy()
    s2[7] = np.nextafter(s2[7], 1)
    assert T.model_key(G, s2) != key


def test_reuse_rule(tmp_path):
    n = G.nx * G.ny * G.nz
    slow = {"P": np.full(n, 0.2), "S": np.full(n, 0.35)}
    src = [T.TableSource("ST1", 13.2, 43.1, 0.0), T.TableSource("ST 2", 13.5, 43.3, 0.0)]
    jobs, reused = T.pending(str(tmp_path), G, slow, src)
    assert len(jobs) == 4 and not reused
    assert jobs[1][2].endswith("ST_2.P.qttb")
    s, ph, path, key = jobs[0]
    T.write(path, T.TableHeader(s.station_id, ph, s.lon, s.lat, s.depth_km, G, key, "x"), np.zeros(n))
    jobs, reused = T.pending(str(tmp_path), G, slow, src)
    assert len(jobs) == 3 and reused == [path]
    moved = [T.TableSource("ST1", 13.2 + 1e-6, 43.1, 0.0), src[1]]  # another position: computed again
    assert len(T.pending(str(tmp_path), G, slow, moved)[0]) == 4
    slow["P"] = slow["P"] * 1.01  # another model: computed again
    assert len(T.pending(str(tmp_path), G, slow, src)[0]) == 4
