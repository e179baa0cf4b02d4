"""The oracle's grid search (EventLocator.GridSearch and WeightedMedian restated) against what the
mathematics fixes: picks made from the tables at a node X with origin time t0 put the minimum (zero
misfit) at X and return t0; the weighted median of known sets (ties kept in order, NaN ignored, the
running weight reaching exactly half); a non-finite table value costs 10 x its weight; the stride
rule max(1, ceil(cbrt(N / candidates)))."""
import numpy as np


def test_weighted_median_cases(oracle):
    wm = oracle.weighted_median
    assert wm([3.0, 1.0, 2.0], [1, 1, 1]) == 2.0
    assert wm([1.0, 2.0, 3.0, 4.0], [1, 1, 1, 1]) == 2.0  # running weight reaches exactly half at 2
    assert wm([5.0, np.nan, 1.0], [1, 100, 3]) == 1.0     # NaN ignored, 1 carries 3 of 4
    assert wm([np.nan], [1]) == 0.0 and wm([], []) == 0.0
    assert wm([2.0, 1.0, 2.0], [0.1, 0.1, 5]) == 2.0


def _tables(rng, nx, ny, nz, nf):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    tabs = []
    for f in range(nf):
        c = rng.uniform([0, 0, 0], [nx, ny, nz])
        tabs.append((0.2 * np.sqrt((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)).ravel().astype(np.float32))
    return tabs


def test_exact_picks_recover_the_node(oracle):
    rng = np.random.default_rng(1)
    nx, ny, nz = 21, 17, 9
    tabs = _tables(rng, nx, ny, nz, 12)
    X = (4 * ny + 11) * nx + 7
    t0 = 3.25
    field = np.arange(12)
    obs = np.array([t0 + float(tabs[f][X]) for f in field])
    w = rng.choice([1.0, 0.5, 0.25], 12)
    node, t, m = oracle.locate_grid_search(tabs, nx, ny, nz, (field, obs, w, np.zeros(12)), coarse_candidates=300)
    assert node == X and abs(t - t0) < 1e-12 and m < 1e-9
    node, t, m = oracle.locate_grid_search(tabs, nx, ny, nz, (field, obs, w, np.zeros(12)), center=(9, 10, 3))
    assert node == X
    # a fixed depth away from X: the best node lies in that layer
    node, t, m = oracle.locate_grid_search(tabs, nx, ny, nz, (field, obs, w, np.zeros(12)), coarse_candidates=300,
                                           k_fixed=2)
    assert node // (nx * ny) == 2 and m > 0


def test_unreachable_pick_costs_ten_times_its_weight(oracle):
    nx, ny, nz = 6, 5, 4
    tabs = [np.full(nx * ny * nz, 1.0, np.float32), np.full(nx * ny * nz, np.inf, np.float32)]
    node, t, m = oracle.locate_grid_search(tabs, nx, ny, nz, (np.array([0, 1]), np.array([4.0, 4.0]),
                                                               np.array([1.0, 0.7]), np.zeros(2)))
    assert node == 0 and t == 3.0 and abs(m - 7.0) < 1e-12
