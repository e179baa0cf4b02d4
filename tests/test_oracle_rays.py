"""Pins the oracle's tables, back-trace and ray kernel (Forward/TravelTimeTable.cs:116-157,
Forward/RayTracing.cs:73-159, Tomography/RayKernels.cs:62-105).

Against what the mathematics fixes: trilinear time is exact on linear fields and its gradient is the
field's gradient; in a homogeneous medium the ray is the straight chord, its kernel sums to the path
length (the trilinear weights sum to one) and sum(kernel * s) is the travel time; a ray in a
vertical gradient bends (it is longer than the chord) and sum(kernel * s) still matches the table's
time; rows are ascending with positive lengths. No reference-held ray fixtures exist.
"""
import numpy as np
import pytest


def _homogeneous(oracle, v=6.0):
    g = oracle.from_spacing(13, 13.74, 43, 43.54, 0, 30, 1, 1)
    s = np.full(g.count, 1 / v)
    return g, s


def _table(oracle, g, s, station):
    idx, t0 = oracle.point_source(g, s, *station)
    return oracle.fmm(g, s, idx, t0).astype(np.float32)


def test_table_time_and_gradient_exact_on_linear_fields(oracle):
    g = oracle.griddef(10, 11, 40, 41, 0, 20, 9, 8, 7)
    lon, lat, dep = oracle.node_coords(g)
    field = (0.5 * lon[None, None, :] + 2.0 * lat[None, :, None] + 0.25 * dep[:, None, None]).astype(np.float32)
    t = oracle.table_time(g, field.ravel(), 10.41, 40.37, 7.7)
    assert abs(t - (0.5 * 10.41 + 2.0 * 40.37 + 0.25 * 7.7)) < 1e-4
    gr = oracle.table_gradient(g, field.ravel(), 10.41, 40.37, 7.7)
    # dT/dlon = 0.5 s/deg over r cos(lat) deg->km; dT/dlat = 2 s/deg over r deg->km; dT/dz = 0.25
    r = 6371.0 - 7.7
    km_lon = r * np.cos(40.37 * np.pi / 180) * np.pi / 180
    km_lat = r * np.pi / 180
    assert np.allclose(gr, [0.5 / km_lon, 2.0 / km_lat, 0.25], rtol=1e-4)


def test_unreached_corner_makes_time_infinite(oracle):
    g = oracle.griddef(10, 11, 40, 41, 0, 20, 5, 5, 5)
    t = np.ones(g.count, np.float32)
    t[(2 * 5 + 2) * 5 + 2] = np.inf
    lon, lat, dep = oracle.node_coords(g)
    assert np.isinf(oracle.table_time(g, t, (lon[1] + lon[2]) / 2, (lat[1] + lat[2]) / 2, (dep[1] + dep[2]) / 2))
    assert oracle.table_time(g, t, lon[0], lat[0], dep[0]) == 1.0  # zero weights skip the unreached corner


@pytest.mark.parametrize("event", [(13.6, 43.45, 12.0), (13.05, 43.05, 5.0), (13.7, 43.0, 3.0), (13.3, 43.3, 25.0)])
def test_homogeneous_ray_is_the_chord(oracle, event):
    g, s = _homogeneous(oracle)
    st = (13.2, 43.1, 0.0)
    T = _table(oracle, g, s, st)
    step = 0.5 * oracle.min_spacing_km(g)
    path = oracle.backtrack(g, T, st, event, step)
    assert path is not None
    assert np.allclose([path[0][0], path[1][0], path[2][0]], event)
    assert np.allclose([path[0][-1], path[1][-1], path[2][-1]], st)
    xyz = np.array([oracle.to_cartesian(*p) for p in zip(*path)])
    plen = np.sum(np.linalg.norm(np.diff(xyz, axis=0), axis=1))
    chord = np.linalg.norm(oracle.to_cartesian(*st) - oracle.to_cartesian(*event))
    assert abs(plen - chord) <= 2e-3 * chord
    nodes, lens = oracle.ray_kernel(g, path)
    assert np.all(np.diff(nodes) > 0) and np.all(lens > 0)
    assert abs(lens.sum() - plen) <= 1e-9 * plen           # weights sum to one on every piece
    assert abs(np.sum(lens * s[nodes]) - chord / 6) <= 2e-3 * chord / 6


def test_gradient_medium_ray_bends_and_integrates_to_the_time(oracle):
    g = oracle.from_spacing(13, 13.74, 43, 43.54, 0, 30, 1, 1)
    lon, lat, dep = oracle.node_coords(g)
    v = 4.0 + 0.08 * dep
    s = np.ascontiguousarray(np.broadcast_to((1 / v)[:, None, None], (g.nz, g.ny, g.nx))).ravel()
    st, ev = (13.05, 43.05, 0.0), (13.7, 43.5, 4.0)
    T = _table(oracle, g, s, st)
    step = 0.5 * oracle.min_spacing_km(g)
    path = oracle.backtrack(g, T, st, ev, step)
    xyz = np.array([oracle.to_cartesian(*p) for p in zip(*path)])
    plen = np.sum(np.linalg.norm(np.diff(xyz, axis=0), axis=1))
    chord = np.linalg.norm(oracle.to_cartesian(*st) - oracle.to_cartesian(*ev))
    assert plen > 1.005 * chord                             # the ray dives into the faster layers
    assert max(path[2]) > 8.0
    nodes, lens = oracle.ray_kernel(g, path)
    tt = oracle.table_time(g, T, *ev)
    # the reference reports <= 2.13 % on real data (validation/results/rays.json:20)
    assert abs(np.sum(lens * s[nodes]) - tt) <= 0.01 * tt


def test_rows_batch_matches_single_rows(oracle):
    g, s = _homogeneous(oracle, 5.5)
    sts = np.array([[13.2, 43.1, 0.0], [13.5, 43.4, 0.0]])
    tabs = [_table(oracle, g, s, st) for st in sts]
    ev = np.array([[13.6, 43.45, 12.0], [13.05, 43.05, 5.0], [13.7, 43.0, 3.0]])
    rt = np.array([0, 1, 1], np.int32)
    step = 0.5 * oracle.min_spacing_km(g)
    rows = oracle.ray_rows_batch(g, tabs, sts, g, rt, ev, step, threads=2)
    for r in range(3):
        path = oracle.backtrack(g, tabs[rt[r]], sts[rt[r]], ev[r], step)
        nodes, lens = oracle.ray_kernel(g, path)
        assert np.array_equal(rows[r][0], nodes) and np.array_equal(rows[r][1], lens.astype(np.float32))
        assert rows[r][2] == oracle.table_time(g, tabs[rt[r]], *ev[r])


# a 1-D background with a step at 10 km (VelocityModel1D: a repeated depth is a discontinuity)
BG = ([0.0, 10.0, 10.0, 35.0, 100.0], [5.0, 6.0, 6.5, 7.0, 8.0], [2.9, 3.5, 3.7, 4.0, 4.5])


def test_background_velocity_is_piecewise_linear_with_steps(oracle):
    """VelocityModel1D.Evaluate (Model/VelocityModel1D.cs:31-45) on values fixed by its definition."""
    bg = oracle.background(*BG)

    def v(ph, d):
        return oracle.background_velocity(bg, ph, d)

    assert v(0, -2.0) == 5.0 and v(1, -2.0) == 2.9      # above the first node: its value
    assert v(0, 5.0) == 5.5 and v(1, 5.0) == 3.2         # halfway between the first two nodes
    assert v(0, 10.0) == 6.0 and v(1, 10.0) == 3.5       # on the interface: the upper value
    assert abs(v(0, 10.0 + 1e-9) - 6.5) < 1e-9           # just below it: the lower branch
    assert v(0, 22.5) == 6.75                            # halfway 6.5 .. 7.0
    assert v(1, 100.0) == 4.5 and v(1, 400.0) == 4.5     # at and below the last node


def _vertical_path(oracle, lon, lat, depths, per_km=8):
    """A radial path through the given depths, `per_km` points per km (Backtrack's step is smaller than
    the pieces; the kernel only sees the polyline)."""
    d = [depths[0]]
    for a, b in zip(depths[:-1], depths[1:]):
        k = max(1, int(round(abs(b - a) * per_km)))
        d.extend(a + (b - a) * np.arange(1, k + 1) / k)
    d = np.array(d)
    return np.full(d.size, lon), np.full(d.size, lat), d


def test_pieces_outside_the_grid_are_timed_not_attributed(oracle):
    """RayTracing.Kernel with an outside slowness (RayTracing.cs:126-147, RayKernels.cs:62-64): a radial
    ray from 5 km down to the grid's bottom at 20 km and on to 40.2 km. The inside part keeps its
    kernel (lengths summing to 15 km, the rows of the inside part alone); the outside part adds
    int dz / v(z) of the phase's background, which for v linear in depth is ln(v1 / v0) / (dv/dz): the
    midpoint rule over pieces <= 0.5 km long is within 2e-5 of it. Without a background the whole
    35.2 km is attributed to the grid (clamped to the bottom nodes)."""
    g = oracle.griddef(13.0, 13.5, 43.0, 43.4, 0.0, 20.0, 11, 9, 21)  # 1 km in depth, ~4 km laterally
    assert abs(0.5 * oracle.min_spacing_km(g) - 0.5) < 1e-9
    lin = oracle.background([0.0, 20.0, 60.0], [6.0, 6.0, 8.0], [3.5, 3.5, 4.5])  # 0.05 and 0.025 /s below 20 km
    path = _vertical_path(oracle, 13.23, 43.17, [5.0, 20.0, 40.2])
    inside = _vertical_path(oracle, 13.23, 43.17, [5.0, 20.0])
    n_in, l_in = oracle.ray_kernel(g, inside)
    for phase, (v0, grad) in enumerate([(6.0, 0.05), (3.5, 0.025)]):
        nodes, lens, out = oracle.ray_kernel_bg(g, path, lin, phase)
        assert np.array_equal(nodes, n_in) and np.array_equal(lens, l_in)
        assert abs(lens.sum() - 15.0) < 1e-9
        exact = np.log((v0 + grad * 20.2) / v0) / grad
        assert abs(out - exact) <= 2e-5 * exact, (phase, out, exact)
    nodes, lens, out = oracle.ray_kernel_bg(g, path, None, 0)
    assert out == 0 and abs(lens.sum() - 35.2) < 1e-9
    bottom = nodes >= (g.nz - 1) * g.nx * g.ny
    assert lens[bottom].sum() > 20.2


def test_outside_time_sums_the_leaving_length_on_every_side(oracle):
    """A constant background v: length inside + v * OutsideTime is the path length whichever face the
    ray leaves by (east, north, bottom, across a corner), and a ray that stays inside is unchanged."""
    g = oracle.griddef(13.0, 13.5, 43.0, 43.4, 0.0, 20.0, 11, 9, 21)
    const = oracle.background([0.0], [7.5], [4.2])
    cases = [((13.25, 43.2, 5.0), (13.7, 43.2, 5.0)),      # east
             ((13.25, 43.2, 5.0), (13.25, 43.6, 8.0)),     # north
             ((13.25, 43.2, 5.0), (13.3, 43.25, 33.0)),    # bottom
             ((13.1, 43.1, 2.0), (12.8, 42.9, 26.0)),      # south-west, below
             ((13.1, 43.1, 2.0), (13.4, 43.3, 15.0))]      # inside
    for a, b in cases:
        n = 400
        t = np.arange(n + 1) / n
        pa, pb = np.array(a), np.array(b)
        path = tuple((pa[None, :] + (pb - pa)[None, :] * t[:, None]).T.copy())
        xyz = np.array([oracle.to_cartesian(*p) for p in zip(*path)])
        plen = float(np.sum(np.linalg.norm(np.diff(xyz, axis=0), axis=1)))
        for phase, v in ((0, 7.5), (1, 4.2)):
            nodes, lens, out = oracle.ray_kernel_bg(g, path, const, phase)
            assert abs(lens.sum() + v * out - plen) <= 1e-9 * plen, (a, b, phase)
            if b == (13.4, 43.3, 15.0):
                n0, l0 = oracle.ray_kernel(g, path)
                assert out == 0 and np.array_equal(nodes, n0) and np.array_equal(lens, l0)
            else:
                assert out > 0.05


def test_rows_batch_with_a_background_matches_single_rows(oracle):
    g, s = _homogeneous(oracle, 5.5)
    inv = oracle.griddef(13.2, 13.5, 43.1, 43.4, 0.0, 12.0, 9, 9, 13)
    bg = oracle.background(*BG)
    sts = np.array([[13.3, 43.2, 0.0], [13.05, 43.45, 0.0]])
    tabs = [_table(oracle, g, s, st) for st in sts]
    ev = np.array([[13.6, 43.45, 12.0], [13.3, 43.25, 20.0], [13.35, 43.3, 6.0], [13.7, 43.05, 25.0]])
    rt = np.array([0, 0, 1, 1], np.int32)
    ph = np.array([0, 1, 1, 0], np.int32)
    step = 0.5 * min(oracle.min_spacing_km(g), oracle.min_spacing_km(inv))
    rows, outside = oracle.ray_rows_batch(g, tabs, sts, inv, rt, ev, step, threads=2, bg=bg, phase_of_ray=ph)
    for r in range(len(ev)):
        path = oracle.backtrack(g, tabs[rt[r]], sts[rt[r]], ev[r], step)
        nodes, lens, out = oracle.ray_kernel_bg(inv, path, bg, int(ph[r]))
        assert np.array_equal(rows[r][0], nodes) and np.array_equal(rows[r][1], lens.astype(np.float32))
        assert outside[r] == out
    assert np.count_nonzero(outside) >= 3
