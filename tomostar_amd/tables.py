"""Travel-time table files and table sets (Forward/TravelTimeTable.cs): the .qttb format the
reference writes and memory-maps, its model key, and TravelTimeTableSet.Build with the tables solved
on the device.

File layout (TravelTimeTable.Write / Open, TravelTimeTable.cs:57-112): the magic "QTTB" 01 00 00 00, the
header's JSON length as little-endian int32, the header as indented JSON (System.Text.Json with the
reference's options: PascalCase names, enums as strings, two-space indentation), zero padding to a
multiple of 64 bytes, then the N times as little-endian float32 (+inf where not finite).

The JSON is written the way System.Text.Json writes these types (numbers in their shortest round-trip
form, integral values without a fraction, HTML-sensitive and non-ASCII characters escaped as \\uXXXX);
no file written by the reference is at hand to compare bytes with, so that equality is unpinned and the
tests check the format's rules and the round trip instead.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from .grid import GridDefinition

MAGIC = b"QTTB\x01\x00\x00\x00"
PHASES = ("P", "S")


@dataclass
class TableHeader:
    """TravelTimeTableHeader (TravelTimeTable.cs:14-27)."""
    station_id: str
    phase: str
    lon: float
    lat: float
    depth_km: float
    grid: GridDefinition
    model_key: str = ""
    solver: str = ""


def _number(x) -> str:
    """A double as System.Text.Json writes it: the shortest string that reads back to the same value,
    integral values without a fraction, E notation with a signed two-digit exponent beyond 1e15 and
    below 1e-4."""
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return str(int(x))
    x = float(x)
    if math.isnan(x):
        return '"NaN"'
    if math.isinf(x):
        return '"Infinity"' if x > 0 else '"-Infinity"'
    if x == 0:
        return "-0" if math.copysign(1.0, x) < 0 else "0"
    r = repr(x)
    mant, _, exp = r.partition("e")
    e10 = int(exp) if exp else None
    mag = abs(x)
    if 1e-4 <= mag < 1e15:
        if e10 is not None:  # repr chose E notation for |x| in [1e15, 1e16): never here, kept for safety
            r = np.format_float_positional(x, trim="-")
        if r.endswith(".0"):
            r = r[:-2]
        return r
    # E notation: mantissa digits of the shortest form, exponent with a sign and at least two digits
    digits = np.format_float_scientific(x, trim="-", exp_digits=2)
    m, _, e = digits.partition("e")
    if m.endswith("."):
        m = m[:-1]
    return f"{m}E{e[0]}{e[1:]}"


def _string(s: str) -> str:
    """A JSON string as System.Text.Json's default encoder escapes it."""
    out = ['"']
    for ch in s:
        c = ord(ch)
        if ch == '"':
            out.append('\\"')
        elif ch == "\\":
            out.append("\\\\")
        elif ch in "\b\f\n\r\t":
            out.append({"\b": "\\b", "\f": "\\f", "\n": "\\n", "\r": "\\r", "\t": "\\t"}[ch])
        elif c < 0x20 or c > 0x7E or ch in "<>&'+`":
            if c > 0xFFFF:
                c -= 0x10000
                out.append(f"\\u{0xD800 + (c >> 10):04X}\\u{0xDC00 + (c & 0x3FF):04X}")
            else:
                out.append(f"\\u{c:04X}")
        else:
            out.append(ch)
    out.append('"')
    return "".join(out)


def _object(items, indent: int) -> str:
    pad = "  " * (indent + 1)
    body = ",\n".join(f"{pad}{_string(k)}: {v}" for k, v in items)
    return "{\n" + body + "\n" + "  " * indent + "}"


def grid_json(g: GridDefinition, indent: int = 0) -> str:
    """GridDefinition's JSON (SphericalGrid.cs:13-24; the derived spacings are [JsonIgnore])."""
    return _object([("MinLon", _number(g.min_lon)), ("MaxLon", _number(g.max_lon)), ("MinLat", _number(g.min_lat)),
                    ("MaxLat", _number(g.max_lat)), ("MinDepthKm", _number(g.min_depth)),
                    ("MaxDepthKm", _number(g.max_depth)), ("Nx", _number(int(g.nx))), ("Ny", _number(int(g.ny))),
                    ("Nz", _number(int(g.nz)))], indent)


def header_json(h: TableHeader) -> bytes:
    return _object([("StationId", _string(h.station_id)), ("Phase", _string(h.phase)), ("Lon", _number(h.lon)),
                    ("Lat", _number(h.lat)), ("DepthKm", _number(h.depth_km)), ("Grid", grid_json(h.grid, 1)),
                    ("ModelKey", _string(h.model_key)), ("Solver", _string(h.solver))], 0).encode()


def model_key(grid: GridDefinition, slowness: np.ndarray) -> str:
    """ForwardModel.Key (TravelTimeTable.cs:173-180): the first 32 hex digits (upper case) of SHA-256 over
    the slowness as little-endian float64 followed by the grid definition's JSON."""
    s = np.ascontiguousarray(slowness, "<f8")
    return hashlib.sha256(s.tobytes() + grid_json(grid).encode()).hexdigest()[:32].upper()


def safe_name(s: str) -> str:
    """TravelTimeTableSet.Safe (TravelTimeTable.cs:425): letters, digits, '.', '_' and '-' kept, the rest '_'."""
    return "".join(c if c.isalnum() or c in "._-" else "_" for c in s)


def write(path: str, header: TableHeader, times: np.ndarray) -> None:
    """TravelTimeTable.Write: through a temporary file renamed over the target."""
    js = header_json(header)
    pad = (64 - (12 + len(js)) % 64) % 64
    t = np.asarray(times)
    data = np.where(np.isfinite(t), t, np.inf).astype("<f4")
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<i", len(js)))
        f.write(js)
        f.write(b"\x00" * pad)
        f.write(data.tobytes())
    os.replace(tmp, path)


def read_header(path: str) -> TableHeader | None:
    """TravelTimeTable.ReadHeader: None when the file is not a table or its header does not parse."""
    try:
        with open(path, "rb") as f:
            head = f.read(12)
            if len(head) != 12 or head[:8] != MAGIC:
                return None
            n = struct.unpack("<i", head[8:])[0]
            d = json.loads(f.read(n))
    except (OSError, ValueError):
        return None
    lower = {k.lower(): v for k, v in d.items()}  # PropertyNameCaseInsensitive
    gd = {k.lower(): v for k, v in lower.get("grid", {}).items()}
    g = GridDefinition(float(gd.get("minlon", 0)), float(gd.get("maxlon", 0)), float(gd.get("minlat", 0)),
                       float(gd.get("maxlat", 0)), float(gd.get("mindepthkm", 0)), float(gd.get("maxdepthkm", 0)),
                       int(gd.get("nx", 2)), int(gd.get("ny", 2)), int(gd.get("nz", 2)))
    return TableHeader(lower.get("stationid", ""), lower.get("phase", "P"), float(lower.get("lon", 0)),
                       float(lower.get("lat", 0)), float(lower.get("depthkm", 0)), g, lower.get("modelkey", ""),
                       lower.get("solver", ""))


def open_table(path: str):
    """TravelTimeTable.Open: (header, the times as a read-only float32 memory map)."""
    with open(path, "rb") as f:
        head = f.read(12)
        if head[:8] != MAGIC:
            raise ValueError("Not a travel-time table.")
        n = struct.unpack("<i", head[8:])[0]
    h = read_header(path)
    offset = 12 + n + (64 - (12 + n) % 64) % 64
    count = h.grid.nx * h.grid.ny * h.grid.nz
    return h, np.memmap(path, dtype="<f4", mode="r", offset=offset, shape=(count,))


@dataclass
class TableSource:
    """TableSource (TravelTimeTable.cs:245)."""
    station_id: str
    lon: float
    lat: float
    depth_km: float


@dataclass
class BuildReport:
    computed: list = field(default_factory=list)
    reused: list = field(default_factory=list)
    seconds: float = 0.0


def pending(folder: str, grid: GridDefinition, slowness: dict, sources, phases=PHASES):
    """The jobs TravelTimeTableSet.Build computes (TravelTimeTable.cs:270-282): every (source, phase)
    whose file is missing or was written for another model key or source position."""
    jobs, reused = [], []
    for phase in phases:
        key = model_key(grid, slowness[phase])
        for s in sources:
            path = os.path.join(folder, f"{safe_name(s.station_id)}.{phase}.qttb")
            h = read_header(path)
            if (h is not None and h.model_key == key and abs(h.lon - s.lon) < 1e-9 and abs(h.lat - s.lat) < 1e-9
                    and abs(h.depth_km - s.depth_km) < 1e-9):
                reused.append(path)
                continue
            jobs.append((s, phase, path, key))
    return jobs, reused


def build(device, grid: GridDefinition, slowness_p: np.ndarray, slowness_s: np.ndarray, sources, folder: str,
          phases=PHASES, mode: str = "fast_fp32") -> BuildReport:
    """TravelTimeTableSet.Build with every table the folder lacks solved on the device in one table set
    (both phases' fields together, one slowness model each) and written as .qttb files."""
    import time

    from . import eikonal as E

    t0 = time.perf_counter()
    os.makedirs(folder, exist_ok=True)
    slow = {"P": np.asarray(slowness_p, np.float64), "S": np.asarray(slowness_s, np.float64)}
    jobs, reused = pending(folder, grid, slow, sources, phases)
    rep = BuildReport(reused=reused)
    if jobs:
        models = [slow["P"], slow["S"]]
        seeds = [E.point_source(grid, slow[ph], s.lon, s.lat, s.depth_km) for s, ph, _, _ in jobs]
        mof = np.array([0 if ph == "P" else 1 for _, ph, _, _ in jobs], np.int32)
        ts = E.TableSet(device, grid, models, seeds, mof)
        try:
            rc, why = ts.solve(mode)
            if rc != 0:
                raise RuntimeError(f"the device declined the tables: {why}")
            for q, (s, ph, path, key) in enumerate(jobs):
                times = ts.read_f32(q, 1)[0]
                write(path, TableHeader(s.station_id, ph, s.lon, s.lat, s.depth_km, grid, key, "HIP fast sweeping"),
                      times)
                rep.computed.append(path)
        finally:
            ts.release()
    rep.seconds = time.perf_counter() - t0
    return rep
