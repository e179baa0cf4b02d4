"""LSQR on a system with the structure of the LSQR-at-scale configuration (curved rays through a 3-D
model, P and S velocity columns, 4 hypocentre columns per event and a station term per station and
phase, a first-difference smoothing block, 32 right-hand sides with damping log-spaced 0.1-100 passed
as LSQR damp, 200 iterations), at a size the CPU oracle solves in seconds: 40,000 arrivals on the
61 x 61 x 31 grid. Lsqr.Solve (Numerics/Lsqr.cs:34-152) is the oracle.

At 200 iterations the weakly damped right-hand sides are past the point where LSQR's iterates separate
under rounding alone (loss of orthogonality in the Golub-Kahan process): any change of summation order,
made at every iteration, moves them by up to ~1e-2. Each comparison is therefore held to the oracle's
own spread: the largest distance between the oracle's iterate and three oracle runs that differ from it
only in rounding - b perturbed by 1e-15 (relative), the dot products and norms of the reference's device
loop (LsqrOpenCl.cs:79-94), the rows in another order. A single 1e-15 perturbation of b understates the
spread badly here (1.7e-5 against 1e-2 on the weakest damping). Bars (relative L2 of the model at equal
iteration count): <= max(1e-5, 5 x spread), with
- the FP64-valued form (TS_LSQR_FAST_FP64) against the oracle on the exact values;
- the mixed form (FP32 matrix values, FP64 vectors) against the oracle on the same FP32-rounded values;
and <= 1e-5 outright for the most strongly damped right-hand side (damp 100), which converges well
before the iterates can separate. The mixed form's distance to the oracle on the exact values (what
FP32 values cost) is reported beside.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPREAD_FACTOR = 5
K = 32
ITERATIONS = 200


@pytest.fixture(scope="module")
def scale_system(device):
    from tomostar_amd import eikonal as E
    from tomostar_amd import rays as Ry
    from tomostar_amd import system as Sy
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=4)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    ev, st_idx, ev_idx, ph = W.scale_arrivals(wl, events=2000, rays=40_000, seed=5)
    ns, ne = len(wl.stations), len(ev)
    batch = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), 2 * st_idx + ph, ev[ev_idx], host_rows=False)
    used = np.flatnonzero(batch.counts >= 0).astype(np.int32)
    rng = np.random.default_rng(6)
    sigma = 0.05
    settings = Sy.SystemSettings(vpvs=False, smoothing_method="gradient", damping_hypocentre=0.0, damping_station=0.0)
    sP, sS = wl.slowness_p, wl.slowness_s
    sysd = Sy.build_system(device, wl.grid, settings, batch, used, ph[used], ev_idx[used], st_idx[used],
                           sigma * rng.standard_normal(used.size), np.full(used.size, sigma),
                           np.zeros(ne, np.uint8), np.zeros(ne, np.uint8), ns, sP, sS, sP, sS, sS / sP,
                           damping=0.0, smoothing=1.0, keep_fp64=True)
    batch.release()
    ts.release()
    dr = sysd.layout["data_rows"]
    B = np.repeat(sysd.rhs[None, :], K, axis=0)
    B[:, :dr] += rng.standard_normal((K, dr))
    damps = 10.0 ** (-1 + 3 * np.arange(K) / (K - 1))
    rs, ci, va = sysd.read()
    yield sysd, rs.astype(np.int64), ci, va, B, damps
    sysd.matrix.release()


def _oracle_x(oracle, sysd, rs, ci, va, b, damp, its=ITERATIONS, device_dots=False):
    """The oracle's model after `its` iterations (atol = btol = conlim = 0: only the machine-precision
    tests stop it earlier, as they stop the device's)."""
    A = sysd.matrix
    x, _, st, _ = oracle.lsqr(A.rows, A.cols, rs, ci, va, b, damp=damp, max_it=its, atol=0.0, btol=0.0, conlim=0.0,
                              threads=16, device_dots=device_dots)
    return x, st["iterations"]


def _rel(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def _spread(oracle, sysd, rs, ci, va, b, damp, its, x, rng):
    """The largest distance from x (the oracle's iterate) of three oracle runs that round differently."""
    d = [_rel(_oracle_x(oracle, sysd, rs, ci, va, b * (1 + 1e-15 * rng.standard_normal(b.size)), damp, its)[0], x),
         _rel(_oracle_x(oracle, sysd, rs, ci, va, b, damp, its, device_dots=True)[0], x)]
    perm = rng.permutation(sysd.matrix.rows)
    lens = np.diff(rs)[perm]
    prs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    take = np.concatenate([np.arange(rs[q], rs[q + 1]) for q in perm])
    d.append(_rel(_oracle_x(oracle, sysd, prs, ci[take], va[take], b[perm], damp, its)[0], x))
    return max(d), d


def test_scale_structured_system_against_the_fp64_oracle(scale_system, oracle):
    from tomostar_amd import lsqr as Q

    sysd, rs, ci, va, B, damps = scale_system
    A = sysd.matrix
    assert A.nnz > 2_000_000 and sysd.layout["data_rows"] > 35_000
    runs = {}
    for mode in ("fast_fp64", "fast_mixed"):
        rc, why, res = Q.solve_many(A, B, damps, ITERATIONS, 0.0, 0.0, 0.0, mode=mode)
        assert rc == 0, why
        runs[mode] = res
    va32 = va.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(8)
    print(f"\nsystem {A.rows} x {A.cols}, {A.nnz} non-zeros, {K} RHS, {ITERATIONS} iterations")
    for j in (0, 16, K - 1):
        b = B[j]
        r64, r32 = runs["fast_fp64"][j], runs["fast_mixed"][j]
        # each device run against the oracle at the device run's own iteration count (a large damping
        # converges before 200 iterations and stops on a machine-precision test); the oracle stops there too
        x64, it64 = _oracle_x(oracle, sysd, rs, ci, va, b, damps[j], r64.iterations)
        spread64, parts = _spread(oracle, sysd, rs, ci, va, b, damps[j], r64.iterations, x64, rng)
        x32, it32 = _oracle_x(oracle, sysd, rs, ci, va32, b, damps[j], r32.iterations)
        spread32, parts32 = _spread(oracle, sysd, rs, ci, va32, b, damps[j], r32.iterations, x32, rng)
        e64 = _rel(r64.solution, x64)
        e32 = _rel(r32.solution, x32)
        cost = _rel(r32.solution, _oracle_x(oracle, sysd, rs, ci, va, b, damps[j], r32.iterations)[0])
        print(f"RHS {j:2d} damp {damps[j]:8.3f}, {r64.iterations}/{r32.iterations} iterations: FP64-valued form "
              f"{e64:.2e} (oracle spread {spread64:.2e}: b {parts[0]:.1e}, dots {parts[1]:.1e}, rows {parts[2]:.1e}); "
              f"mixed form {e32:.2e} against the oracle on FP32 values (spread {spread32:.2e}), {cost:.2e} "
              f"against the oracle on exact values")
        assert it64 == r64.iterations and it32 == r32.iterations
        assert e64 <= max(1e-5, SPREAD_FACTOR * spread64)
        assert e32 <= max(1e-5, SPREAD_FACTOR * spread32)
        if j == K - 1:
            assert e64 <= 1e-5 and e32 <= 1e-5
