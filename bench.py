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
# HBM bytes per launch of the sweep kernels on this workload: FETCH_SIZE + WRITE_SIZE from rocprofv3
# --pmc (KiB per dispatch, averaged over the solve's 24,048 launches of each kernel), scaled by what the
# same counters read for an elementwise copy of 1 GiB (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000). The
# runs and the figures are summarised in profiles/eik_fast_plane_batched_400x201x201x101.md.
TRAFFIC_BYTES_PER_LAUNCH: dict = {
    "eik_fast_plane<double>": 2.440e8,
    "eik_fast_plane<float>": 4.462e7,
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(wl, seeds, mof, budget_s: float = 20.0):
    """The reference's CPU path restated (fast marching, one table per core) on a bounded sample."""
    from oracle import oracle as O

    g = wl.grid
    og = O.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)
    cores = min(os.cpu_count() or 1, 16)
    models = [wl.slowness_p, wl.slowness_s]
    # probe one field on one core, then size the sample to ~budget_s of wall time on all cores
    t0 = time.perf_counter()
    O.fmm_batch(og, [models[mof[0]]], [seeds[0]], threads=1)
    one = time.perf_counter() - t0
    nf = int(max(cores, min(len(seeds), cores * max(1, round(budget_s / max(one, 1e-3))))))
    nf = min(nf, len(seeds))
    t0 = time.perf_counter()
    O.fmm_batch(og, [models[m] for m in mof[:nf]], seeds[:nf], threads=cores)
    wall = time.perf_counter() - t0
    return {"value": nf * g.count / wall / 1e9, "unit": "Gnode/s", "cores": cores, "kind": "port",
            "sample": f"{nf} of {len(seeds)} fields, fast marching (SphericalEikonal.Solve restated in C), "
                      f"one field per core, {wall:.1f} s"}


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
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stations", type=int, default=200)
    ap.add_argument("--plain-fields", type=int, default=23, help="fields of the plain-FP64 sample (one batch); 0 skips it")
    ap.add_argument("--no-cpu", action="store_true")
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
    # small bookkeeping kernels, whose time is charged to the plane kernel here)
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
        cpu = cpu_baseline(wl, seeds, mof)
        log(f"cpu baseline: {cpu}")

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
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": dom,
                     "bytes_per_unit": {"fp32_update": BYTES_PER_UPDATE_FP32, "fp64_update": BYTES_PER_UPDATE_FP64},
                     "units_per_step": {"fp32_updates": e32, "fp64_updates": e64},
                     "kernels": per_kernel,
                     "sweep_seconds_per_step": float(np.mean(sweep))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ts.release()
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
