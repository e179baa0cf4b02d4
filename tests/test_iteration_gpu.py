"""One tomographic iteration end to end on the device (the 61 x 61 x 31 case: 25 stations x P,S,
150 events, Vp/Vs parameterisation, damping 1, Laplacian 5, 100 LSQR iterations at damp 0).

Bars: the system has the reference's shape (its assembly is bit-exact, test_system_gpu); the mixed
LSQR model is within 1e-5 relative L2 of the FP64 oracle at equal iteration count while the iterates
are still reproducible (20 iterations), and after (100) within 20x the rounding spread of the oracle
run on the same FP32-rounded matrix values; the plain FP64 device LSQR equal to the last bit to the
oracle's restatement of the reference's device loop in the device's storage order. Closed check on the physics: the
recovered P slowness perturbation correlates with the true checkerboard.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def iteration(device):
    from tomostar_amd import iteration as It
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    t_obs = It.observed_times(device, wl)
    r = It.run(device, wl, t_obs, lsqr_mode="fast_mixed")
    return wl, t_obs, r


def test_system_shape(iteration):
    wl, t_obs, r = iteration
    n = wl.grid.count
    ne, ns = len(wl.events), len(wl.stations)
    assert r.columns == 2 * n + 4 * ne + 2 * ns
    # data rows + damping and Laplacian of both blocks + hypocentre damping + 2 x (station damping + zero sum)
    assert r.rows == r.data_rows + 4 * n + 4 * ne + 2 * (ns + 1)
    assert r.data_rows >= 0.99 * 2 * ns * ne


def _oracle_spread(oracle, r, rs, ci, va, its):
    """How far the FP64 oracle's own iterate moves under rounding alone: LSQR's iterates on this system
    separate exponentially from rounding after ~30 iterations (loss of orthogonality), so no
    implementation reproduces them more closely at equal iteration count. FP64 runs that differ only in
    rounding: b perturbed by 1e-15 (relative, three draws), the rows taken in another order (three
    orders: the same least-squares problem, every row sum and dot product rounded differently), and the
    dot products and norms of the reference's device loop (LsqrOpenCl.cs:79-94) instead of
    SimdVector's. The separations of such runs spread over two orders of magnitude from draw to draw
    (at 60 iterations three row orders gave 4.6e-6, 2.6e-5 and 7.8e-4), so one draw says little.
    Returns the oracle's iterate, the largest of the seven distances of the iterates and the largest
    relative change of the residual norm ||A x - b||."""
    def run(rs_, ci_, va_, b_, device_dots=False):
        return oracle.lsqr(r.rows, r.columns, rs_, ci_, va_, b_, damp=0.0, max_it=its, atol=0.0, btol=0.0,
                           conlim=0.0, threads=16, device_dots=device_dots)[0]

    x = run(rs, ci, va, r.rhs)
    res_x = _residual(oracle, rs, ci, va, x, r.rhs)
    d, dr = [], []

    def note(y):
        d.append(float(np.linalg.norm(y - x) / np.linalg.norm(x)))
        dr.append(abs(_residual(oracle, rs, ci, va, y, r.rhs) / res_x - 1))

    for seed in (11, 21, 31):
        note(run(rs, ci, va, r.rhs * (1 + 1e-15 * np.random.default_rng(seed).standard_normal(r.rhs.size))))
    for seed in (12, 22, 32):
        perm = np.random.default_rng(seed).permutation(r.rows)
        lens = np.diff(rs)[perm]
        prs = np.concatenate([[0], np.cumsum(lens)]).astype(rs.dtype)
        take = np.concatenate([np.arange(rs[q], rs[q + 1]) for q in perm]) if rs[-1] else np.zeros(0, np.int64)
        y = run(prs, ci[take], va[take], r.rhs[perm])
        note(y)
    note(run(rs, ci, va, r.rhs, device_dots=True))
    return x, max(d), max(dr)


def _residual(oracle, rs, ci, va, x, b):
    return float(np.linalg.norm(oracle.spmv(rs, ci, va, x, 16) - b))


def test_fast_lsqr_model_matches_fp64_oracle(device, oracle, iteration):
    """The mixed form (FP32 matrix values, FP64 vectors, the pipeline's default): within 1e-5 of the
    FP64 oracle at 20 iterations, before rounding separates the iterates. At 100 iterations the FP32
    rounding of the values alone moves the iterates of this system by ~1e-3 (the same separation
    that a 1e-15 change of b starts), so the mixed form is held to the FP64 oracle run on the
    FP32-rounded values: within 20x that oracle's own rounding spread (1e-15 in b, another row order). The distance to the oracle
    on the exact values and the FP32-vector form are reported beside it."""
    from tomostar_amd import lsqr as Q

    wl, t_obs, r = iteration
    rs, ci, va = r.system.read()
    x20 = oracle.lsqr(r.rows, r.columns, rs, ci, va, r.rhs, damp=0.0, max_it=20, atol=0.0, btol=0.0, conlim=0.0)[0]
    rc, why, res = Q.solve_many(r.system.matrix, r.rhs[None, :], [0.0], 20, 0.0, 0.0, 0.0, mode="fast_mixed")
    assert rc == 0, why
    err20 = np.linalg.norm(res[0].solution - x20) / np.linalg.norm(x20)
    assert r.iterations == 100
    va32 = va.astype(np.float32).astype(np.float64)
    x100_32, spread32, rspread32 = _oracle_spread(oracle, r, rs, ci, va32, 100)
    err100 = np.linalg.norm(r.x - x100_32) / np.linalg.norm(x100_32)
    x100 = oracle.lsqr(r.rows, r.columns, rs, ci, va, r.rhs, damp=0.0, max_it=100, atol=0.0, btol=0.0, conlim=0.0)[0]
    exact100 = np.linalg.norm(r.x - x100) / np.linalg.norm(x100)
    rc, why, r32 = Q.solve_many(r.system.matrix, r.rhs[None, :], [0.0], 100, 0.0, 0.0, 0.0, mode="fast_fp32")
    err32 = np.linalg.norm(r32[0].solution - x100) / np.linalg.norm(x100)
    res_dev = _residual(oracle, rs, ci, va32, r.x, r.rhs)
    res_ora = _residual(oracle, rs, ci, va32, x100_32, r.rhs)
    print(f"mixed LSQR vs FP64 oracle: {err20:.2e} at 20 iterations; at 100 {err100:.2e} against the oracle on "
          f"the FP32-rounded values (its own rounding spread {spread32:.2e}), {exact100:.2e} against the oracle on "
          f"the exact values; FP32-vector form {err32:.2e} at 100; residual norms {res_dev:.6e} / {res_ora:.6e}")
    assert err20 <= 1e-5
    # at 100 iterations the iterate and its residual norm within the reach of the oracle's own rounding
    # spread (one draw of which varies a hundredfold)
    assert abs(res_dev / res_ora - 1) <= max(1e-9, 100 * rspread32)
    assert err100 <= max(1e-5, 100 * spread32)


def test_plain_lsqr_equals_the_reference_device_loop(device, oracle, iteration):
    """The plain FP64 form is the reference's device loop (LsqrOpenCl.Run, LsqrOpenCl.cs:272-533): the
    oracle's Lsqr.Solve restatement with LsqrOpenCl's dot products and norms (64 partials of
    tree-reduced item sums, the norm their square root), run on the matrix in the order the device
    stores it (rows sorted, node columns tiled; each row's entries in the given order), reproduces it
    to the last bit at every iteration count - before the iterates separate under rounding (20) and
    well after (40, 100), where any other rounding moves them by up to 1e-3. Against the oracle in the
    given order with SimdVector's dot products the distance is reported beside its rounding spread."""
    from tomostar_amd import lsqr as Q

    wl, t_obs, r = iteration
    rs, ci, va = r.system.read()
    rp, cp = r.system.stored_order()
    stored_of = np.empty_like(cp)
    stored_of[cp] = np.arange(cp.size, dtype=cp.dtype)
    lens = np.diff(rs)[rp]
    srs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    take = np.concatenate([np.arange(rs[q], rs[q + 1]) for q in rp])
    sci, sva, sb = stored_of[ci[take]], va[take], r.rhs[rp]
    for its in (20, 40, 100):
        rc, why, res = Q.solve_many(r.system.matrix, r.rhs[None, :], [0.0], its, 0.0, 0.0, 0.0, mode="plain_fp64")
        assert rc == 0, why
        xs = oracle.lsqr(r.rows, r.columns, srs, sci, sva, sb, damp=0.0, max_it=its, atol=0.0, btol=0.0,
                         conlim=0.0, threads=16, device_dots=True)[0]
        want = np.empty_like(xs)
        want[cp] = xs
        x, spread, _ = _oracle_spread(oracle, r, rs, ci, va, its)
        err = np.linalg.norm(res[0].solution - x) / np.linalg.norm(x)
        print(f"plain FP64 LSQR at {its} iterations: max |difference| from the device loop restated "
              f"{np.max(np.abs(res[0].solution - want)):.1e}; from Lsqr.Solve in the given order {err:.2e} "
              f"(the oracle's own rounding spread {spread:.2e})")
        assert np.array_equal(res[0].solution, want)


def test_recovered_model_correlates_with_the_checkerboard(iteration):
    from tomostar_amd import iteration as It

    wl, t_obs, r = iteration
    n = wl.grid.count
    sP0, _ = It.starting_model(wl)
    true = wl.slowness_p / sP0 - 1  # the slowness perturbation the data see
    xp = r.x[:n]
    # nodes the rays sample: depths above the deepest event, away from the edges
    k = np.repeat(np.arange(wl.grid.nz), wl.grid.nx * wl.grid.ny)
    mask = (k >= 1) & (k <= 12)
    c = np.corrcoef(xp[mask], true[mask])[0, 1]
    print(f"correlation of the recovered and true P slowness perturbations: {c:.3f}")
    assert c > 0.3
