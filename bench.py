"""Benchmark of the TomoSTAR hot path on one MI355X.

Workload (the batched-eikonal configuration): 400 travel-time fields (200 stations x {P, S}) on a
201 x 201 x 101 grid at 0.5 km, Vp = 5.0 + 0.04 z with a 4 % Gaussian random field (correlation
length 5 km) and three +-10 % ellipsoidal bodies, Vs = Vp / 1.73. A step is one full solve of all
400 fields (both phases to convergence) with the inputs resident in HBM, through the C ABI.

Reported: solved field-nodes per second (value, Gnode/s), the node updates per second actually
executed, the speed-up over the plain FP64 one-plane-per-launch form (timed on a bounded sample of
fields), the p50/p99 step latency, the HBM roofline of the sweep kernel, and the CPU baseline (the
oracle's fast-marching restatement of the reference's CPU path, one table per core).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one node update: read and write its own time, read its slowness (neighbours
# are counted as served from cache): FP32 times 4+4+4, FP64 times 8+8+4
BYTES_PER_UPDATE_FP32 = 12.0
BYTES_PER_UPDATE_FP64 = 20.0
# HBM bytes per launch (one plane step: the three chunks' kernels of one ordering's plane) of the sweep
# kernels on this workload: FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc (KiB per dispatch, summed over
# the solve's dispatches of each width and divided by its plane steps), scaled by what the same
# counters read for an elementwise copy of 1 GiB (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000). The runs and
# the figures are summarised in profiles/eik_fast_plane_batched_400x201x201x101.md.
TRAFFIC_BYTES_PER_LAUNCH: dict = {
    "eik_fast_plane<double>": 8.113e7,
    "eik_fast_plane<float>": 3.728e7,
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """The host's CPUs: nproc, the CPUs this process may run on (sched_getaffinity), the threads the
    CPU baselines use, and the CPU model. The threads are every CPU of the affinity set, capped by
    OMP_NUM_THREADS when it is set: a host shared by several GPUs' jobs gives each its share of the
    cores through it."""
    avail = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(share)) if share.isdigit() and int(share) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count() or 1, "affinity": avail, "threads": threads, "model": model,
            "omp_num_threads": share or None}


def _field_errors(dev, ref, seed_idx):
    """max and RMS relative error over the computed nodes outside the seeds (t > 0), and the device
    gate's range-normalised metric max |dt| / (max t - min t) (EikonalOpenCl.cs:371)."""
    mask = np.isfinite(ref) & (ref > 0)
    mask[seed_idx] = False
    r = np.abs(dev[mask] - ref[mask]) / ref[mask]
    fin = np.isfinite(ref)
    span = float(ref[fin].max() - ref[fin].min())
    return {"max_rel": float(r.max()), "rms_rel": float(np.sqrt(np.mean(r * r))),
            "range_normalised": float(np.max(np.abs(dev[fin] - ref[fin])) / span)}


def cpu_baseline(wl, seeds, mof, budget_s: float = 20.0, compare=None):
    """The reference's CPU path restated (fast marching, one table per core) on a bounded sample. With
    compare = {field: (mixed table, FP64-form table)}, the fast-marching tables of those fields (within
    the sample) are also the reference the device's tables are checked against: FMM is what the
    reference's CPU workers produce (TravelTimeTable.cs:341-354), while the device forms follow Relax."""
    from oracle import oracle as O

    g = wl.grid
    og = O.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)
    host = host_cpu()
    cores = host["threads"]
    models = [wl.slowness_p, wl.slowness_s]
    # probe one field on one core, then size the sample to ~budget_s of wall time on all cores
    t0 = time.perf_counter()
    O.fmm_batch(og, [models[mof[0]]], [seeds[0]], threads=1)
    one = time.perf_counter() - t0
    nf = int(max(cores, min(len(seeds), cores * max(1, round(budget_s / max(one, 1e-3))))))
    nf = min(nf, len(seeds))
    t0 = time.perf_counter()
    tabs = O.fmm_batch(og, [models[m] for m in mof[:nf]], seeds[:nf], threads=cores)
    wall = time.perf_counter() - t0
    against_fmm = None
    if compare:
        tabs = np.asarray(tabs).reshape(nf, -1)
        against_fmm = {}
        for f, (mixed, fp64) in compare.items():
            if f < nf:
                against_fmm[str(f)] = {"mixed_vs_fmm": _field_errors(mixed, tabs[f], seeds[f][0]),
                                       "relax_vs_fmm": _field_errors(fp64, tabs[f], seeds[f][0])}
    return {"value": nf * g.count / wall / 1e9, "unit": "Gnode/s", "cores": cores, "kind": "port",
            "nproc": host["nproc"], "affinity_cpus": host["affinity"], "omp_num_threads": host["omp_num_threads"],
            "cpu_model": host["model"],
            "sample": f"{nf} of {len(seeds)} fields, fast marching (SphericalEikonal.Solve restated in C), "
                      f"one field per core, {wall:.1f} s",
            "device_tables_vs_fmm": against_fmm}


def lsqr_bytes_per_iteration(m: int, n: int, nnz: int, k: int, s: int = 4, v: int = 4) -> float:
    """Algorithmic HBM bytes of one LSQR iteration over A and its stored transpose for k right-hand sides:
    both matrices streamed once (a v-byte value + int32 index per non-zero, int32 row starts), and per RHS
    u read + written, the gathered v and u, v, w, x read + written: 2((v + 4) nnz + 4(m+1)) + k s (3m + 7n)."""
    return 2 * ((v + 4) * nnz + 4 * (m + 1)) + k * s * (3 * m + 7 * n)


def lsqr_leg(dev, wl, ts, args):
    """LSQR at scale: G from curved rays through the batched-eikonal model (20,000 events, the 200
    stations, first `lsqr_rays` station x event x phase triples of a seeded shuffle), a first-difference
    smoothing block, k right-hand sides (data residual + independent noise) with damping log-spaced
    0.1 - 100, a fixed iteration budget. The fast multi-RHS form against the plain FP64 form."""
    from tomostar_amd import lsqr as Q
    from tomostar_amd import rays as Ry
    from tomostar_amd import system as Sy
    from tomostar_amd import workloads as W

    ev, st_idx, ev_idx, ph = W.scale_arrivals(wl, events=20000, rays=args.lsqr_rays, seed=2)
    ns, ne = len(wl.stations), len(ev)
    t0 = time.perf_counter()
    batch = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), 2 * st_idx + ph, ev[ev_idx], host_rows=False)
    t_rays = time.perf_counter() - t0
    used = np.flatnonzero(batch.counts >= 0).astype(np.int32)
    log(f"lsqr leg: {args.lsqr_rays} arrivals traced in {t_rays:.2f} s, {used.size} rows, {batch.nnz} non-zeros")
    rng = np.random.default_rng(3)
    sigma = 0.05
    residual = sigma * rng.standard_normal(used.size)
    sP, sS = wl.slowness_p, wl.slowness_s
    settings = Sy.SystemSettings(vpvs=False, smoothing_method="gradient", damping_hypocentre=0.0, damping_station=0.0)
    t0 = time.perf_counter()
    sysd = Sy.build_system(dev, wl.grid, settings, batch, used, ph[used], ev_idx[used], st_idx[used], residual,
                           np.full(used.size, sigma), np.zeros(ne, np.uint8), np.zeros(ne, np.uint8), ns, sP, sS,
                           sP, sS, sS / sP, damping=0.0, smoothing=1.0, keep_fp64=True)
    t_sys = time.perf_counter() - t0
    batch.release()
    A = sysd.matrix
    m, n, nnz = A.rows, A.cols, A.nnz
    log(f"lsqr leg: system {m} x {n}, {nnz} non-zeros, assembled in {t_sys:.2f} s")
    k = args.lsqr_rhs
    dr = sysd.layout["data_rows"]
    B = np.repeat(sysd.rhs[None, :], k, axis=0)
    B[:, :dr] += rng.standard_normal((k, dr))  # N(0, sigma) noise in seconds is N(0, 1) on weighted rows
    damps = 10.0 ** (-1 + 3 * np.arange(k) / max(1, k - 1))
    its = args.lsqr_iterations
    runs = {}
    for mode in ("fast_mixed", "fast_fp32", "fast_fp64"):
        t0 = time.perf_counter()
        rc, why, res = Q.solve_many(A, B, damps, its, 0.0, 0.0, 0.0, mode=mode)
        wall = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"{mode} LSQR declined: {why}")
        runs[mode] = (res, wall, A.last_seconds())
        log(f"lsqr leg: {mode}, {k} RHS, {max(r.iterations for r in res)} iterations in {wall:.2f} s "
            f"(device {runs[mode][2]:.3f} s)")
    # plain FP64 on sampled right-hand sides, each for the iterations its mixed run took, and the
    # spread rounding alone causes in FP64 (b perturbed by 1e-15) on the first of them
    picks = sorted(set([0, k // 3, (2 * k) // 3, k - 1]))
    plain_x, plain_s, plain_its = {}, 0.0, 0
    for j in picks:
        t0 = time.perf_counter()
        rc, why, pr = Q.solve_many(A, B[j:j + 1], [damps[j]], runs["fast_mixed"][0][j].iterations, 0.0, 0.0, 0.0,
                                   mode="plain_fp64")
        plain_s += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"plain LSQR declined: {why}")
        plain_its += pr[0].iterations
        plain_x[j] = pr[0].solution
    bp = B[picks[0]] * (1 + 1e-15 * rng.standard_normal(m))
    rc, why, pp = Q.solve_many(A, bp[None, :], [damps[picks[0]]], runs["fast_mixed"][0][picks[0]].iterations,
                               0.0, 0.0, 0.0, mode="plain_fp64")
    x0 = plain_x[picks[0]]
    spread = float(np.linalg.norm(pp[0].solution - x0) / np.linalg.norm(x0))
    # the FP64 solver's own rounding spread per sampled RHS: the distance between two FP64 solvers that
    # differ only in summation order (the plain one-RHS form and the FP64-valued SpMM form)
    fp64_pair = {str(j): float(np.linalg.norm(runs["fast_fp64"][0][j].solution - plain_x[j]) /
                               max(1e-300, np.linalg.norm(plain_x[j]))) for j in picks}
    plain_rate = plain_its / plain_s                       # RHS-iterations per second
    forms = {}
    for mode, (res, wall, dev_s) in runs.items():
        errs = {str(j): float(np.linalg.norm(res[j].solution - plain_x[j]) / max(1e-300, np.linalg.norm(plain_x[j])))
                for j in picks}
        it_max = max(r.iterations for r in res)
        bpi = lsqr_bytes_per_iteration(m, n, nnz, k, 4 if mode == "fast_fp32" else 8,
                                       8 if mode == "fast_fp64" else 4)
        rate = sum(r.iterations for r in res) / wall
        forms[mode] = {
            "rhs_iterations_per_s": round(rate, 2), "seconds": round(wall, 3), "device_seconds": round(dev_s, 3),
            "iterations": [int(r.iterations) for r in res], "stop_codes": sorted(set(int(r.stop_code) for r in res)),
            "speedup_vs_plain_fp64_gpu": round(rate / plain_rate, 2),
            "model_rel_l2_vs_fp64": {"max": max(errs.values()), "per_rhs": errs},
            "roofline": {"bound": "hbm", "bytes_per_iteration": bpi,
                         "achieved": round(bpi * it_max / dev_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(bpi * it_max / dev_s / 1e9 / HBM_PEAK_GBS, 4)},
        }
    main = forms["fast_fp64"]
    out = {
        "workload": f"G of {used.size} curved rays ({args.lsqr_rays} arrivals, 20,000 events x {ns} stations x P,S) "
                    f"+ first differences on {wl.grid.nx}x{wl.grid.ny}x{wl.grid.nz} (P, S), {k} RHS, "
                    f"damping 0.1-100, {its} iterations",
        "rows": m, "columns": n, "nnz": nnz,
        "rays_seconds": round(t_rays, 3), "system_seconds": round(t_sys, 3),
        "rhs_iterations_per_s": main["rhs_iterations_per_s"], "unit": "RHS-iterations/s",
        "form": "fast_fp64 (FP64 matrix values, vectors and scalars: the FP64 solver's arithmetic in the SpMM "
                "summation order); fast_mixed (FP32 values) and fast_fp32 beside it",
        "speedup_vs_plain_fp64_gpu": main["speedup_vs_plain_fp64_gpu"],
        "model_rel_l2_vs_fp64": main["model_rel_l2_vs_fp64"]["max"],
        "fp64_rounding_spread": {"value": spread, "rhs": picks[0],
                                 "what": "plain FP64 iterate moved by a 1e-15 relative perturbation of b"},
        "fp64_solver_pair_spread": {"per_rhs": fp64_pair, "max": max(fp64_pair.values()),
                                    "what": "distance between two FP64 solvers that differ only in summation "
                                            "order: the plain one-RHS form and the FP64-valued SpMM form"},
        "damping_per_rhs": {str(j): float(damps[j]) for j in picks},
        "plain_fp64_gpu": {"rhs_iterations_per_s": round(plain_rate, 2), "sample": f"RHS {picks}"},
        "forms": forms,
        "roofline": main["roofline"],
    }
    errs = [main["model_rel_l2_vs_fp64"]["max"]]
    out["rhs_iterations_per_s"] = main["rhs_iterations_per_s"]
    log(f"lsqr leg: {out['rhs_iterations_per_s']} RHS-it/s, {out['speedup_vs_plain_fp64_gpu']}x plain, "
        f"max model error {max(errs):.2e} (FP64 spread {spread:.2e}, FP64 solver pair {max(fp64_pair.values()):.2e}), "
        f"{out['roofline']['achieved']} GB/s")
    A.release()
    return out


def iteration_leg(dev, cpu: bool):
    """One tomographic iteration on the 61 x 61 x 31 grid (25 stations x P,S, 150 events): tables,
    rays, system, 100 LSQR iterations, each stage timed on the GPU and, as the CPU baseline, with the
    oracle's restatement of the reference's CPU path on the host cores."""
    from tomostar_amd import iteration as It
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    t_obs = It.observed_times(dev, wl)
    It.run(dev, wl, t_obs)  # warm-up: code objects and graphs
    r = It.run(dev, wl, t_obs)
    out = {"workload": "61x61x31 at 1 km, 25 stations x P,S, 150 events, Vp/Vs parameterisation, damping 1, "
                       "Laplacian 5, 100 LSQR iterations",
           "rows": r.rows, "columns": r.columns, "nnz": r.nnz, "data_rows": r.data_rows,
           "gpu_seconds": {k: round(v, 4) for k, v in r.stage_seconds.items()}}
    out["gpu_seconds"]["total"] = round(sum(r.stage_seconds.values()), 4)
    out["consumers"] = consumers_leg(dev, wl, t_obs, r)
    if cpu:
        out["cpu_seconds"] = cpu_baseline_iteration(wl, r)
        out["consumers"]["cpu_seconds"] = cpu_baseline_consumers(r, out["consumers"])
    out["q_tomography"] = q_leg(dev, wl, cpu)
    log(f"iteration leg: {out}")
    return out


def q_leg(dev, wl, cpu: bool):
    """Q tomography from t* on the same case: P rays of the 3750 event x station pairs traced in the
    true model, t* of a Q model (ln Q checkerboard +-0.3 around 250, a Q/3 body) with 4 ms noise, and
    QTomography.Run with the reference's defaults (Q0 from the data, ln(q/q0) by up to 20 Gauss-Newton
    iterations, 400 LSQR iterations with atol = btol = 1e-6, outliers rejected and solved again). The
    device run in the fast mixed LSQR form and in the plain FP64 one; the oracle's Run on the host cores."""
    from tomostar_amd import qtomography as QT

    c = QT.q_case(dev, wl)
    s = QT.QTomographySettings()
    QT.QTomography(dev, wl.grid, QT.QTomographySettings(lsqr_iterations=20, nonlinear_iterations=1),
                   lsqr_mode="fast_mixed").run(c.rays, c.data, c.slowness)  # warm-up
    res = {}
    for mode in ("fast_mixed", "plain_fp64"):
        t0 = time.perf_counter()
        res[mode] = QT.QTomography(dev, wl.grid, s, lsqr_mode=mode).run(c.rays, c.data, c.slowness)
        res[mode].seconds["total"] = time.perf_counter() - t0
    f, p = res["fast_mixed"], res["plain_fp64"]
    out = {"workload": "P t* of 3750 paths on 61x61x31, reference defaults (up to 20 Gauss-Newton iterations, "
                       "400 LSQR iterations, atol = btol = 1e-6)",
           "gpu_seconds": {k: round(v, 4) for k, v in f.seconds.items()},
           "gauss_newton_iterations": f.gauss_newton[0], "rejected": f.rejected,
           "reference_q": round(f.reference_q, 2), "rms_ms": [round(f.rms_before * 1e3, 3), round(f.rms_after * 1e3, 3)],
           "plain_fp64_gpu_seconds": round(p.seconds["total"], 4),
           "max_rel_q_difference_fast_vs_plain": float(np.max(np.abs(f.q / p.q - 1)))}
    if cpu:
        from oracle import oracle as O

        g = wl.grid
        og = O.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)
        cores = host_cpu()["threads"]
        rows = np.flatnonzero(c.rays.counts >= 0)
        starts, ends = c.rays.row_start[rows], c.rays.row_start[rows + 1]
        rs = np.concatenate([[0], np.cumsum(ends - starts)]).astype(np.int64)
        take = np.concatenate([np.arange(a, b) for a, b in zip(starts, ends)])
        t0 = time.perf_counter()
        orc = O.QOracle(og, {}, threads=cores)
        orc.prepare(rs, c.rays.nodes[take], c.rays.lengths[take], rows, c.data.tstar, c.data.sigma, c.data.station,
                    c.data.nstations, c.slowness)
        orc.run()
        out["cpu_seconds"] = round(time.perf_counter() - t0, 3)
        out["cpu_cores"] = cores
        out["cpu_sample"] = "the whole Run, the oracle's restatement (C system, LSQR with 16-thread SpMV)"
    c.rays.release()
    return out


def formal_leg(dev, args, cpu: bool):
    """The exact formal resolution (FormalResolution.Compute) of the Q system of the 61 x 61 x 31 case
    inverted on every second node (31 x 31 x 16 = 15,376 nodes and 25 station terms: 15,401 unknowns,
    under the reference's 20,000 limit): S = A^T A, its Cholesky factor, the inverse, the diagonal, the
    standard deviations and the resolution length of every node. The CPU baseline is LAPACK's Cholesky
    and triangular inverse (numpy / scipy on the host cores) of the same S: the n^3 part only."""
    from tomostar_amd import formal as F
    from tomostar_amd import qtomography as QT
    from tomostar_amd import workloads as W

    wl = W.one_iteration_case(seed=0)
    c = QT.q_case(dev, wl, step=2)
    inv = c.grid
    rows = np.flatnonzero(c.rays.counts >= 0).astype(np.int32)
    s = QT.QTomographySettings()
    d = c.data
    sysd = QT.build_qsystem(dev, inv, s, c.rays, rows, d.station[rows], d.tstar[rows], d.sigma[rows], d.nstations,
                            c.slowness, 1 / 250, s.damping, s.smoothing)
    from tomostar_amd.system import System
    rs, ci, va = System(sysd.matrix, sysd.rhs).read()
    m, cols, nd = sysd.matrix.rows, sysd.matrix.cols, sysd.layout["data_rows"]
    sysd.matrix.release()
    c.rays.release()
    places = F.add_places(F.new_places(cols), inv, 0, 0)
    F.compute(dev, m, cols, rs, ci, va, nd, places=places)  # warm-up: code objects
    t0 = time.perf_counter()
    res = F.compute(dev, m, cols, rs, ci, va, nd, places=places)
    sec = time.perf_counter() - t0
    n = res.columns.size
    out = {"workload": f"Q system of 3750 P t* on 31x31x16 nodes + 25 station terms ({n} unknowns), resolution "
                       "diagonal, standard deviations and resolution lengths",
           "unknowns": int(n), "gpu_seconds": round(sec, 3),
           "median_diagonal": round(float(np.nanmedian(res.diagonal[:inv.count])), 4),
           "median_length_km": round(float(np.nanmedian(res.length_km[:inv.count])), 2)}
    if cpu:
        import scipy.linalg as SL
        import scipy.sparse as SP

        A = SP.csr_matrix((va, ci, rs), shape=(m, cols))
        S = (A.T @ A).toarray()
        t0 = time.perf_counter()
        Lf = np.linalg.cholesky(S)
        SL.lapack.dtrtri(Lf, lower=1)
        out["cpu_seconds_cholesky_and_inverse"] = round(time.perf_counter() - t0, 3)
        out["cpu_sample"] = "LAPACK dpotrf + dtrtri of the same S (numpy / scipy, host cores); lengths not timed"
    log(f"formal leg: {out}")
    return out


def location_leg(wl, ts, args, cpu: bool):
    """Event location's grid search over the 400 resident tables of the headline solve: events at random
    nodes with 40 picks each (random station x phase tables, origin time, 0.05 s noise, quarter/half/
    full weights, station corrections), global search with 8000 coarse candidates (stride 8 on this
    grid, then every node within 16). The oracle's restatement timed on a sample of the events."""
    from tomostar_amd import locate as Lc

    g = wl.grid
    rng = np.random.default_rng(17)
    nf = 2 * len(wl.stations)
    nodes = rng.integers(0, g.count, args.locate_events)
    events = []
    for node in nodes:
        field = rng.choice(nf, size=40, replace=False).astype(np.int32)
        events.append(Lc.Picks(field, np.zeros(40), rng.choice([1.0, 0.5, 0.25], 40), rng.normal(0, 0.02, 40),
                               (0.0, 0.0, 0.0)))
    tabs = ts.read_f32()  # [field][node] float32, 6.5 GB: the observed times come from the tables
    for e, node in zip(events, nodes):
        e.obs = rng.uniform(-2, 2) + tabs[e.field, int(node)].astype(np.float64) + rng.normal(0, 0.05, 40)
    Lc.grid_search(ts, g, events[:4])  # warm-up
    t0 = time.perf_counter()
    lon, lat, dep, torg, node, mis = Lc.grid_search(ts, g, events)
    sec = time.perf_counter() - t0
    hit = float(np.mean(node == nodes))
    out = {"workload": f"{args.locate_events} events x 40 picks over {nf} tables of {g.nx}x{g.ny}x{g.nz}, global "
                       "search (8000 coarse candidates, then every node within two strides)",
           "gpu_seconds": round(sec, 4), "events_per_s": round(args.locate_events / sec, 1),
           "best_node_is_the_source": round(hit, 4)}
    if cpu:
        from oracle import oracle as O

        sample = 3
        t0 = time.perf_counter()
        for e in events[:sample]:
            O.locate_grid_search(tabs, g.nx, g.ny, g.nz, (e.field, e.obs, e.weight, e.correction), 8000)
        per = (time.perf_counter() - t0) / sample
        out["cpu_seconds_per_event"] = round(per, 3)
        out["cpu_sample"] = (f"{sample} events, the oracle's numpy restatement on one core (a checker, not an "
                             "optimised CPU locator)")
    del tabs
    log(f"location leg: {out}")
    return out


def consumers_leg(dev, wl, t_obs, r):
    """The multi-RHS consumers of the iteration's system: the resolution diagonal (32 probes, 100 LSQR
    iterations each, one batched solve; the plain FP64 form one probe at a time, timed on 2 probes)
    and a trade-off curve of 4 (damping, smoothing) pairs (assembly + 100 LSQR iterations per pair)."""
    from tomostar_amd import lsqr as Q
    from tomostar_amd import resolution as R
    from tomostar_amd import tradeoff as T

    probes, its = 32, 100
    R.estimate(r.system.matrix, r.data_rows, 2, 5, 0.0, seed=1)  # warm-up
    t0 = time.perf_counter()
    res = R.estimate(r.system.matrix, r.data_rows, probes, its, 0.0, seed=1)
    t_fast = time.perf_counter() - t0
    z = R.probe_signs(r.columns, 2, seed=1)
    rhs = Q.multiply(r.system.matrix, z)
    rhs[:, r.data_rows:] = 0.0
    t0 = time.perf_counter()
    for j in range(2):
        Q.solve_many(r.system.matrix, rhs[j][None, :], [0.0], its, 0.0, 0.0, 1e8, mode="plain_fp64")
    t_plain = (time.perf_counter() - t0) / 2
    pairs = [(10.0, 5.0), (3.0, 5.0), (1.0, 5.0), (0.3, 5.0)]
    t0 = time.perf_counter()
    pts = T.trade_off(dev, wl, t_obs, pairs, lsqr_iterations=its, lsqr_tolerance=0.0)
    t_trade = time.perf_counter() - t0
    return {"resolution_diagonal": {"probes": probes, "lsqr_iterations": its, "gpu_seconds": round(t_fast, 4),
                                    "mean_diagonal": round(float(res.diagonal.mean()), 6),
                                    "plain_fp64_gpu_seconds_per_probe": round(t_plain, 4),
                                    "speedup_vs_plain_fp64_gpu": round(t_plain * probes / t_fast, 2)},
            "trade_off": {"pairs": pairs, "lsqr_iterations": its, "gpu_seconds": round(t_trade, 4),
                          "residual_norms": [round(p.residual_norm, 4) for p in pts],
                          "model_norms": [round(p.model_norm, 4) for p in pts]}}


def cpu_baseline_consumers(r, cons):
    """The oracle's resolution diagonal on 2 of the probes (scaled to 32) and its LSQR per trade-off
    pair (the assembly left out), on the host cores."""
    from oracle import oracle as O
    from tomostar_amd import resolution as R

    cores = host_cpu()["threads"]
    rs, ci, va = r.system.read()
    z = R.probe_signs(r.columns, 2, seed=1)
    t0 = time.perf_counter()
    O.resolution_diagonal(r.rows, r.columns, rs, ci, va, r.data_rows, z, 100, 0.0, threads=cores)
    t_probe = (time.perf_counter() - t0) / 2
    t0 = time.perf_counter()
    O.lsqr(r.rows, r.columns, rs, ci, va, r.rhs, damp=0.0, max_it=100, atol=0, btol=0, conlim=1e8, threads=cores)
    t_pair = time.perf_counter() - t0
    probes = cons["resolution_diagonal"]["probes"]
    return {"resolution_diagonal": round(t_probe * probes, 3), "trade_off_lsqr": round(t_pair * len(cons["trade_off"]["pairs"]), 3),
            "cores": cores, "sample": "2 probes scaled to 32; one LSQR per pair"}


def cpu_baseline_iteration(wl, r):
    """The same iteration on the oracle (the reference's CPU solvers restated: fast marching one table
    per core, rays in parallel over arrivals, the system, LSQR with 16-thread SpMV)."""
    from oracle import oracle as O
    from tomostar_amd import iteration as It

    g = wl.grid
    og = O.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)
    cores = host_cpu()["threads"]
    sP, sS = It.starting_model(wl)
    sec = {}
    t0 = time.perf_counter()
    seeds = [O.point_source(og, sP if ph == 0 else sS, *st) for st in wl.stations for ph in (0, 1)]
    tabs = O.fmm_batch(og, [sP if q % 2 == 0 else sS for q in range(len(seeds))], seeds, threads=cores)
    sec["tables"] = time.perf_counter() - t0
    station, event, phase = r.arrivals
    t0 = time.perf_counter()
    O.ray_rows_batch(og, [t.astype(np.float32) for t in tabs], np.repeat(wl.stations, 2, axis=0), og,
                     2 * station + phase, wl.events[event], 0.5 * O.min_spacing_km(og), threads=cores)
    sec["rays"] = time.perf_counter() - t0
    rs, ci, va = r.system.read()
    t0 = time.perf_counter()
    O.lsqr(r.rows, r.columns, rs, ci, va, r.rhs, damp=0.0, max_it=100, atol=0, btol=0, conlim=0, threads=cores)
    sec["lsqr"] = time.perf_counter() - t0
    sec = {k: round(v, 4) for k, v in sec.items()}
    sec["cores"] = cores
    return sec


def uneven_leg(dev, args):
    """The uneven batch: 64 single-field problems on grids from 41 x 41 x 21 to 161 x 161 x 81 at
    0.5-2 km, swept together (plane launches over all problems, or one persistent launch per ordering
    over every problem's column tiles, or the hybrid); the tuner's choices (sweep method, tile width
    and nodes per thread per size class) by timing; the mixed form against the FP64 form (Relax to the
    last bit) per problem; the plain FP64 one-problem-per-solve form timed on a sample of the problems."""
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W
    from tomostar_amd.gridbatch import GridBatch

    t0 = time.perf_counter()
    probs = W.uneven_batch_case(seed=3, problems=args.uneven_problems)
    seeds = [E.point_source(p.grid, p.slowness, *p.source) for p in probs]
    total = sum(p.grid.count for p in probs)
    log(f"uneven leg: {len(probs)} problems, {total} nodes ({time.perf_counter() - t0:.1f} s)")
    b = GridBatch(dev, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    choice = b.tune()
    rc, why = b.solve("fast_fp32")  # warm-up
    if rc != 0:
        raise RuntimeError(f"uneven batch declined: {why}")
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        rc, why = b.solve("fast_fp32")
        times.append(time.perf_counter() - t0)
    mixed = [b.read(q) for q in range(len(probs))]
    rounds = b.rounds.copy()
    st = b.stats()
    sched = b.schedule()
    round_log = b.round_log()
    t0 = time.perf_counter()
    rc, why = b.solve("fast_fp64")
    t64 = time.perf_counter() - t0
    errs = []
    for q, (idx, _) in enumerate(seeds):
        ref = b.read(q)
        mask = np.isfinite(ref) & (ref > 0)
        mask[idx] = False
        errs.append(float(np.max(np.abs(mixed[q][mask] - ref[mask]) / ref[mask])))
    b.release()
    # the plain FP64 form, one problem per solve, on every eighth problem
    sample = list(range(0, len(probs), 8))
    ps, pn = 0.0, 0
    for q in sample:
        p = probs[q]
        ts = E.TableSet(dev, p.grid, [p.slowness], [seeds[q]])
        t0 = time.perf_counter()
        rc, why = ts.solve("plain_fp64")
        ps += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"plain solve declined: {why}")
        pn += p.grid.count
        ts.release()
    value = total / float(np.mean(times)) / 1e9
    plain = pn / ps / 1e9
    out = {"workload": f"{len(probs)} problems, grids 41x41x21 - 161x161x81 (axes drawn jointly), 0.5-2 km, "
                       f"checkerboard / spike / tabular body +-5-10 %, one interior source each",
           "nodes": total, "value": round(value, 4), "unit": "Gnode/s (solved nodes per second)",
           "seconds": round(float(np.mean(times)), 3), "fp64_seconds": round(t64, 3),
           "tuner_nodes_per_thread": {"<2e5 nodes": choice[0], "2e5-8e5": choice[1], ">=8e5": choice[2]},
           "tuner_schedule": {"method": sched["method"],
                              "tile_width": {"<2e5 nodes": sched["tile_width"][0], "2e5-8e5": sched["tile_width"][1],
                                             ">=8e5": sched["tile_width"][2]},
                              "trial_mixed_solve_ms": {k: round(v, 1) for k, v in sched["trial_ms"].items()}},
           "round_ms": [[w, m, round(x * 1e3, 1)] for w, m, x in round_log],
           "max_rel_error_vs_fp64": max(errs), "rounds_max": [int(rounds[:, 0].max()), int(rounds[:, 1].max())],
           "plain_fp64_gpu": {"value": round(plain, 4), "unit": "Gnode/s", "sample": f"problems {sample}"},
           "speedup_vs_plain_fp64_gpu": round(value / plain, 2)}
    # the sweep kernel's roofline: algorithmic bytes of the executed updates (12 B on FP32 times, 20 B on
    # FP64 times) over the sweep time of the last timed solve (HIP events around the rounds)
    alg = BYTES_PER_UPDATE_FP32 * st["updates_fp32"] + BYTES_PER_UPDATE_FP64 * st["updates_fp64"]
    if st["sweep_seconds"] > 0:
        ach = alg / st["sweep_seconds"] / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": {"planes": "mg_plane", "stream": "mg_stream", "hybrid": "mg_plane + mg_stream"}[sched["method"]],
                           "units_per_step": {"fp32_updates": st["updates_fp32"], "fp64_updates": st["updates_fp64"]},
                           "sweep_seconds": round(st["sweep_seconds"], 4)}
    log(f"uneven leg: {out}")
    return out


def serving_leg(dev, wl, ts, args):
    """Serving: seeded Poisson arrivals, 70 % table requests (grids 61^3 - 201 x 201 x 101, 1-50 sources,
    P and S) and 30 % LSQR requests (G of 1e5 - 1e6 rays through the batched-eikonal model, 1-8 RHS,
    damping log-uniform 0.1-100, 50 iterations); models and matrices resident; admission by a device
    memory budget. The plain FP64 form serves the same requests one at a time (one worker, live at the
    same arrivals in real time). The arrival rate puts the plain form at 80 % utilisation, from its
    service times on every request in a first pass."""
    from tomostar_amd import lsqr as Q
    from tomostar_amd import rays as Ry
    from tomostar_amd import system as Sy
    from tomostar_amd import workloads as W
    from tomostar_amd.serving import Server

    rng = np.random.default_rng(4)
    budget = args.serving_budget_gb
    # resident models: four grid sizes of the batched-eikonal kind
    sizes = [(60.0, 30.0, 1.0), (100.0, 50.0, 1.0), (100.0, 50.0, 0.667), (100.0, 50.0, 0.5)]
    grids, models_p, models_s = [], [], []
    for w, d, h in sizes[:-1]:
        m = W.batched_eikonal_case(seed=5, stations=1, width_km=w, depth_km=d, spacing_km=h)
        grids.append(m.grid); models_p.append(m.slowness_p); models_s.append(m.slowness_s)
    grids.append(wl.grid); models_p.append(wl.slowness_p); models_s.append(wl.slowness_s)
    # resident matrices: G of 1e5, 3e5 and 1e6 rays through the model's tables
    mats = []
    for nr in (100_000, 300_000, 1_000_000):
        ev, st_idx, ev_idx, ph = W.scale_arrivals(wl, events=20000, rays=nr, seed=6 + nr % 7)
        b = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), 2 * st_idx + ph, ev[ev_idx], host_rows=False)
        used = np.flatnonzero(b.counts >= 0).astype(np.int32)
        sysd = Sy.build_system(dev, wl.grid, Sy.SystemSettings(vpvs=False, smoothing_method="gradient",
                                                               damping_hypocentre=0.0, damping_station=0.0),
                               b, used, ph[used], ev_idx[used], st_idx[used], 0.05 * rng.standard_normal(used.size),
                               np.full(used.size, 0.05), np.zeros(len(ev), np.uint8), np.zeros(len(ev), np.uint8),
                               len(wl.stations), wl.slowness_p, wl.slowness_s, wl.slowness_p, wl.slowness_s,
                               wl.slowness_s / wl.slowness_p, damping=0.0, smoothing=1.0, keep_fp64=True)
        b.release()
        mats.append((sysd.matrix, sysd.rhs))
    log(f"serving leg: models {[g.count for g in grids]}, matrices {[a.nnz for a, _ in mats]} non-zeros")
    # the request stream
    nreq = args.serving_requests
    reqs = []
    for q in range(nreq):
        if rng.random() < 0.7:
            gi = int(rng.integers(0, len(grids)))
            g = grids[gi]
            ns = int(rng.integers(1, 51))
            src = np.stack([rng.uniform(g.min_lon, g.max_lon, ns), rng.uniform(g.min_lat, g.max_lat, ns),
                            rng.uniform(0.0, 0.3 * g.max_depth, ns)], 1)
            reqs.append(("tables", gi, src))
        else:
            mi = int(rng.integers(0, len(mats)))
            k = int(rng.integers(1, 9))
            a, rhs = mats[mi]
            B = rhs[None, :] + 0.05 * rng.standard_normal((k, rhs.size))
            reqs.append(("lsqr", mi, B, 10.0 ** rng.uniform(-1, 2, k)))

    def serve(server, ids, modes, arrivals=None):
        """Submits every request (at its arrival time when given), waits for all; per request
        (submitted, started, finished)."""
        t0 = time.perf_counter()
        rids = []
        for q, r in enumerate(reqs):
            if arrivals is not None:
                delay = arrivals[q] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            if r[0] == "tables":
                rids.append(server.submit_tables(ids[r[1]][0], ids[r[1]][1], r[2], modes[0]))
            else:
                rids.append(server.submit_lsqr(mats[r[1]][0], r[2], r[3], 50, mode=modes[1]))
        times = []
        for rid in rids:
            st, why, t = server.wait(rid, 1200.0)
            if st != 0:
                raise RuntimeError(f"request declined: {why}")
            times.append(t)
            server.release_request(rid)
        return np.array(times)

    def open_server(workers):
        srv = Server(dev, workers=workers, memory_budget_gb=budget)
        ids = [(srv.add_model(g, p), srv.add_model(g, s_)) for g, p, s_ in zip(grids, models_p, models_s)]
        return srv, ids

    # plain FP64, one request at a time (one worker: the reference's one solve at a time per device): the
    # arrival rate from its service times on every request (80 % utilisation; a subset of the requests
    # misjudges the mean), then every request live at those Poisson arrivals
    srv, ids = open_server(1)
    t_cal = serve(srv, ids, ("plain_fp64", "plain_fp64"))
    lam = 0.8 / float(np.mean(t_cal[:, 2] - t_cal[:, 1]))
    arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
    t_plain = serve(srv, ids, ("plain_fp64", "plain_fp64"), arrivals=arrivals)
    srv.close()
    service = t_plain[:, 2] - t_plain[:, 1]
    plain_lat = t_plain[:, 2] - t_plain[:, 0]
    plain_span = float(t_plain[:, 2].max() - t_plain[:, 0].min())
    # the fast forms, four workers, the same arrivals in real time
    srv, ids = open_server(args.serving_workers)
    serve(srv, ids, ("fast_fp32", "fast_mixed"))  # warm
    t_fast = serve(srv, ids, ("fast_fp32", "fast_mixed"), arrivals=arrivals)
    stats = srv.stats()
    # identity: the first table and LSQR requests again, alone
    ident = True
    for r_i, r in enumerate(reqs):
        if r[0] != "tables":
            continue
        rid = srv.submit_tables(ids[r[1]][0], ids[r[1]][1], r[2], "fast_fp32")
        srv.wait(rid)
        alone = srv.tables(rid, 0, min(4, 2 * len(r[2])), grids[r[1]].count)
        srv.release_request(rid)
        rid = srv.submit_tables(ids[r[1]][0], ids[r[1]][1], r[2], "fast_fp32")
        rid2 = srv.submit_lsqr(mats[0][0], mats[0][1][None, :], [1.0], 50, mode="fast_mixed")
        srv.wait(rid)
        srv.wait(rid2)
        mixed = srv.tables(rid, 0, min(4, 2 * len(r[2])), grids[r[1]].count)
        srv.release_request(rid)
        srv.release_request(rid2)
        ident = ident and bool(np.array_equal(alone, mixed))
        break
    srv.close()
    for a, _ in mats:
        a.release()
    fast_lat = t_fast[:, 2] - t_fast[:, 0]
    makespan = float(t_fast[:, 2].max() - t_fast[:, 0].min())
    out = {"workload": f"{nreq} requests, Poisson arrivals at {lam:.3f}/s (80 % utilisation of the plain FP64 "
                       f"form), 70 % tables (grids {[g.count for g in grids]} nodes, 1-50 sources, P and S), 30 % "
                       f"LSQR (G of 1e5 / 3e5 / 1e6 rays, 1-8 RHS, damping 0.1-100, 50 iterations)",
           "memory_budget_gb": budget, "workers": args.serving_workers,
           "latency_s": {"p50": float(np.percentile(fast_lat, 50)), "p99": float(np.percentile(fast_lat, 99))},
           "throughput_req_per_s": round(nreq / makespan, 4),
           "plain_fp64_gpu": {"latency_s": {"p50": float(np.percentile(plain_lat, 50)),
                                            "p99": float(np.percentile(plain_lat, 99))},
                              "throughput_req_per_s": round(nreq / plain_span, 4),
                              "mean_service_s": float(np.mean(service)),
                              "utilisation": round(float(np.sum(service)) / plain_span, 3),
                              "how": "run live: the same requests at the same Poisson arrivals in real time, "
                                     "one worker (one solve at a time), plain FP64 forms"},
           "peak_admitted_gb": round(stats["peak_bytes"] / 1e9, 2), "peak_inflight": stats["peak_inflight"],
           "same_result_alone_and_in_mix": ident}
    log(f"serving leg: {out}")
    return out


def calibrate_copy(mib: int) -> int:
    """A plain device-to-device copy of known size: the reference launch that hardware byte counters
    (FETCH_SIZE, WRITE_SIZE) are scaled by."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel that reads n and writes n floats
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "ok": ok}), flush=True)
    return 0 if ok else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    # two warm-up solves: the first counts node updates (its own graphs), the second captures the graphs
    # the timed solves launch
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--stations", type=int, default=200)
    ap.add_argument("--plain-fields", type=int, default=23, help="fields of the plain-FP64 sample (one batch); 0 skips it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lsqr-rays", type=int, default=1_000_000, help="arrivals of the LSQR-at-scale leg (0 skips it)")
    ap.add_argument("--lsqr-rhs", type=int, default=32)
    ap.add_argument("--lsqr-iterations", type=int, default=200)
    ap.add_argument("--no-iteration", action="store_true", help="skip the one-iteration pipeline leg")
    ap.add_argument("--uneven-problems", type=int, default=64, help="problems of the uneven-batch leg (0 skips it)")
    ap.add_argument("--serving-requests", type=int, default=16, help="requests of the serving leg (0 skips it)")
    ap.add_argument("--locate-events", type=int, default=2000, help="events of the location leg (0 skips it)")
    ap.add_argument("--no-formal", action="store_true", help="skip the formal resolution leg")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64-form solve that checks the timed tables (profiler runs)")
    ap.add_argument("--serving-workers", type=int, default=4)
    ap.add_argument("--serving-budget-gb", type=float, default=96.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path is single-GPU, "
              "fields shard across GPUs without exchange)", file=sys.stderr)
        return 2

    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    t0 = time.perf_counter()
    wl = W.batched_eikonal_case(seed=1, stations=args.stations)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    log(f"workload {wl.name}: {len(seeds)} fields, {wl.grid.count} nodes each ({time.perf_counter() - t0:.1f} s)")
    dev = E.Device.open(0, self_test=True)
    log(f"device {dev.info}")
    models = [wl.slowness_p, wl.slowness_s]
    ts = E.TableSet(dev, wl.grid, models, seeds, mof)
    nf, N = len(seeds), wl.grid.count

    for w in range(args.warmup):
        rc, why = ts.solve("fast_fp32")
        if rc != 0:
            raise RuntimeError(f"solve declined: {why}")
        log(f"warmup {w}: rounds {ts.rounds}, {ts.stats()['seconds']:.3f} s")
    lat, sweep, sw32, sw64 = [], [], [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        rc, why = ts.solve("fast_fp32")
        if rc != 0:
            raise RuntimeError(f"solve declined: {why}")
        st = ts.stats()
        lat.append(st["seconds"])
        sweep.append(st["sweep_seconds"])
        sw32.append(st["sweep_seconds_fp32"])
        sw64.append(st["sweep_seconds_fp64"])
        log(f"step {k}: {st['seconds']:.3f} s (sweep graphs {st['sweep_seconds']:.3f} s: FP32 "
            f"{st['sweep_seconds_fp32']:.3f} s in {st['rounds_fp32']} rounds, FP64 {st['sweep_seconds_fp64']:.3f} s "
            f"in {st['rounds_fp64']} rounds), rounds {ts.rounds}")
    total = time.perf_counter() - t_start
    # executed node updates are counted on the set's first fast solve (warm-up or timed)
    e32, e64 = st["updates_fp32"], st["updates_fp64"]
    evals = e32 + e64
    ms_per_step = 1e3 * total / args.steps
    value = nf * N / (total / args.steps) / 1e9
    upd_rate = evals / (total / args.steps) / 1e9
    # the roofline of the dominant kernel: per launch, algorithmic bytes / average launch duration
    # (HIP events around the rounds on the solve's stream; a round is the plane launches plus two
    # small bookkeeping kernels, whose time is charged to the plane kernel here). A launch is one plane
    # step: with the groups split over streams, the chunks' kernels of a step run side by side.
    launches32 = st["rounds_fp32"] * st["plane_launches_per_round"]
    launches64 = st["rounds_fp64"] * st["plane_launches_per_round"]
    kern = {
        "eik_fast_plane<float>": (e32 * BYTES_PER_UPDATE_FP32, launches32, float(np.mean(sw32))),
        "eik_fast_plane<double>": (e64 * BYTES_PER_UPDATE_FP64, launches64, float(np.mean(sw64))),
    }
    dom = max(kern, key=lambda k: kern[k][2])
    per_kernel = {}
    for name, (byts, launches, secs) in kern.items():
        if launches == 0:
            continue
        per_kernel[name] = {"launches": launches, "avg_launch_us": round(1e6 * secs / launches, 3),
                            "alg_bytes_per_launch": round(byts / launches, 1),
                            "achieved_gbs": round(byts / secs / 1e9, 1)}
    dom_bytes, dom_launches, dom_secs = kern[dom]
    achieved = dom_bytes / dom_secs / 1e9
    traffic = TRAFFIC_BYTES_PER_LAUNCH.get(dom)

    # accuracy of the timed (mixed) form per field: against the FP64 form, which is Relax's result to the
    # last bit, on the first 16 fields (8 stations x P, S); the table set is then solved again in the
    # timed form for the legs below
    acc_fields = [] if args.no_accuracy else list(range(min(16, nf)))
    mixed_tab = {f: ts.read(f, 1)[0] for f in acc_fields}
    rc, why = ts.solve("fast_fp64") if acc_fields else (0, "")
    if rc != 0:
        raise RuntimeError(f"FP64 solve declined: {why}")
    fp64_tab = {f: ts.read(f, 1)[0] for f in acc_fields}
    accuracy = None
    if acc_fields:
        per_field = {str(f): _field_errors(mixed_tab[f], fp64_tab[f], seeds[f][0]) for f in acc_fields}
        accuracy = {"reference": "the FP64 form (SphericalEikonal.Relax's arithmetic and schedule, bit for bit)",
                    "fields": len(acc_fields), "per_field": per_field,
                    "max_rel": max(v["max_rel"] for v in per_field.values()),
                    "rms_rel": float(np.sqrt(np.mean([v["rms_rel"] ** 2 for v in per_field.values()]))),
                    "range_normalised": max(v["range_normalised"] for v in per_field.values())}
        log(f"accuracy: mixed vs FP64 form over {len(acc_fields)} fields: max {accuracy['max_rel']:.2e}, "
            f"rms {accuracy['rms_rel']:.2e}, range-normalised {accuracy['range_normalised']:.2e}")
        rc, why = ts.solve("fast_fp32")
        if rc != 0:
            raise RuntimeError(f"solve declined: {why}")

    # plain FP64 one-plane-per-launch form on a bounded sample (one of its batches)
    pf = min(args.plain_fields, nf)
    plain_rate, plain_s = None, None
    if pf > 0:
        plain = E.TableSet(dev, wl.grid, models, seeds[:pf], mof[:pf])
        t0 = time.perf_counter()
        rc, why = plain.solve("plain_fp64")
        plain_s = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"plain solve declined: {why}")
        plain_rate = pf * N / plain_s / 1e9
        log(f"plain FP64: {pf} fields in {plain_s:.2f} s ({plain_rate:.3f} Gnode/s)")
        plain.release()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, seeds, mof, compare={f: (mixed_tab[f], fp64_tab[f]) for f in acc_fields})
        log(f"cpu baseline: {cpu}")

    lsqr = lsqr_leg(dev, wl, ts, args) if args.lsqr_rays > 0 else None
    location = location_leg(wl, ts, args, cpu=not args.no_cpu) if args.locate_events > 0 else None
    serving = serving_leg(dev, wl, ts, args) if args.serving_requests > 0 else None
    ts.release()
    it_leg = None if args.no_iteration else iteration_leg(dev, cpu=not args.no_cpu)
    uneven = uneven_leg(dev, args) if args.uneven_problems > 0 else None
    formal = None if args.no_formal else formal_leg(dev, args, cpu=not args.no_cpu)

    out = {
        "metric": "eikonal Gnode-upd/s + LSQR iters/s vs plain FP64 GPU; p50/p99 solve latency",
        "value": round(value, 4),
        "unit": "Gnode/s (solved field-nodes per second)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (first-order phase on FP32 times, second-order phase on FP64 times)",
        "data": "synthetic (seeded Gaussian random field + ellipsoidal bodies)",
        "config": {"workload": f"batched eikonal: {nf} fields ({len(wl.stations)} stations x P,S) on "
                               f"{wl.grid.nx}x{wl.grid.ny}x{wl.grid.nz} at 0.5 km, one solve per step",
                   "fields": nf, "nodes_per_field": N},
        "node_updates_per_s": round(upd_rate, 4),
        "node_updates_per_solve": evals,
        "speedup_vs_plain_fp64_gpu": round(value / plain_rate, 2) if plain_rate else None,
        "plain_fp64_gpu": {"value": round(plain_rate, 4), "unit": "Gnode/s", "sample": f"{pf} fields, one batch"}
                          if plain_rate else None,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "accuracy": accuracy,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom,
                     "bytes_per_unit": {"fp32_update": BYTES_PER_UPDATE_FP32, "fp64_update": BYTES_PER_UPDATE_FP64},
                     "units_per_step": {"fp32_updates": e32, "fp64_updates": e64},
                     "kernels": per_kernel,
                     "sweep_seconds_per_step": float(np.mean(sweep))},
        "cpu_baseline": cpu,
        "lsqr": lsqr,
        "iteration": it_leg,
        "uneven": uneven,
        "serving": serving,
        "location": location,
        "formal_resolution": formal,
    }
    print(json.dumps(out), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
