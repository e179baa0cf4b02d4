"""The workload configurations at their full sizes, on the device, against the oracle.

- Uneven batch (64 eikonal problems of 41 x 41 x 21 to 161 x 161 x 81 nodes, spacing 0.5-2 km, one
  work-queue solve): the FP64 form equals the oracle's Relax (Forward/SphericalEikonal.cs:284-396) to the
  last bit on four problems spanning the size range, and the mixed form stays within 1e-4 relative of
  the FP64 form on all 64.
- Batched eikonal (400 fields of 201 x 201 x 101 at 0.5 km): the mixed form within 1e-4 relative of the
  oracle's Relax on two full fields; its distance to the oracle's fast marching (SphericalEikonal.Solve,
  :102-152, the reference's CPU production path) is reported and held to the FMM-vs-Relax distance
  plus 1e-4.
- Serving (table requests on resident 61 x 61 x 61 and 201 x 201 x 101 models, LSQR requests of 1-8
  right-hand sides on a resident G of >= 1e5 curved rays): every result is identical alone and in the
  concurrent mix, and the admitted memory never exceeds the budget.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _og(oracle, g):
    return oracle.griddef(g.min_lon, g.max_lon, g.min_lat, g.max_lat, g.min_depth, g.max_depth, g.nx, g.ny, g.nz)


def _rel(dev, ref, seed_idx):
    mask = np.isfinite(ref) & (ref > 0)
    mask[seed_idx] = False
    r = np.abs(dev[mask] - ref[mask]) / ref[mask]
    return float(r.max()), float(np.sqrt(np.mean(r * r)))


def _parallel(fns):
    """Runs the oracle calls side by side (ctypes releases the interpreter lock)."""
    out = [None] * len(fns)

    def run(q):
        out[q] = fns[q]()

    th = [threading.Thread(target=run, args=(q,)) for q in range(len(fns))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return out


def test_uneven_batch_full_size(device, oracle):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W
    from tomostar_amd.gridbatch import GridBatch

    probs = W.uneven_batch_case(seed=3, problems=64)
    sizes = [p.grid.count for p in probs]
    assert min(sizes) < 60_000 and max(sizes) > 1_900_000
    seeds = [E.point_source(p.grid, p.slowness, *p.source) for p in probs]
    b = GridBatch(device, [p.grid for p in probs], [p.slowness for p in probs], seeds)
    b.set_schedule("planes")
    rc, why = b.solve("fast_fp64")
    assert rc == 0, why
    fp64 = [b.read(q) for q in range(len(probs))]
    rc, why = b.solve("fast_fp32")
    assert rc == 0, why
    mixed = [b.read(q) for q in range(len(probs))]
    # the streaming sweep (one persistent launch per ordering over all 64 problems' tiles) and the hybrid
    # (a run's first round in plane launches) to the bit
    for method, mode, ref in (("stream", "fast_fp64", fp64), ("stream", "fast_fp32", mixed), ("hybrid", "fast_fp32", mixed)):
        b.set_schedule(method, [4, 8, 16])
        print(f"uneven batch, {method} sweep, {mode}", flush=True)
        rc, why = b.solve(mode)
        assert rc == 0, why
        for q in range(len(probs)):
            assert np.array_equal(b.read(q), ref[q]), (mode, q)
    b.release()
    order = np.argsort(sizes)
    picks = [int(order[0]), int(order[len(order) // 3]), int(order[(2 * len(order)) // 3]), int(order[-1])]
    print(f"\nuneven batch: {len(probs)} problems, {sum(sizes)} nodes; Relax on problems {picks} "
          f"({[probs[q].grid.shape for q in picks]})", flush=True)
    refs = _parallel([lambda q=q: oracle.relax(_og(oracle, probs[q].grid), probs[q].slowness, *seeds[q]) for q in picks])
    for q, (f, ok, _) in zip(picks, refs):
        assert ok
        assert np.array_equal(fp64[q], f), q
    worst = max(_rel(mixed[q], fp64[q], seeds[q][0])[0] for q in range(len(probs)))
    print(f"uneven batch: FP64 form equal to Relax on {len(picks)} problems; mixed form within {worst:.2e} "
          f"of the FP64 form over all {len(probs)}")
    assert worst < 1e-4


def test_batched_fields_full_size_against_relax_and_fmm(device, oracle):
    from tomostar_amd import eikonal as E
    from tomostar_amd import workloads as W

    wl = W.batched_eikonal_case(seed=1, stations=200)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    rc, why = ts.solve("fast_fp32")
    assert rc == 0, why
    picks = [37, 290]  # a P and an S field
    dev = {f: ts.read(f, 1)[0] for f in picks}
    ts.release()
    og = _og(oracle, wl.grid)
    models = [wl.slowness_p, wl.slowness_s]
    jobs = []
    for f in picks:
        s = models[mof[f]]
        jobs.append(lambda s=s, f=f: oracle.relax(og, s, *seeds[f]))
        jobs.append(lambda s=s, f=f: oracle.fmm(og, s, *seeds[f]))
    res = _parallel(jobs)
    for q, f in enumerate(picks):
        relax, ok, _ = res[2 * q]
        fmm = res[2 * q + 1]
        fmm = fmm[0] if isinstance(fmm, tuple) else fmm
        assert ok
        e_relax = _rel(dev[f], relax, seeds[f][0])
        e_fmm = _rel(dev[f], fmm, seeds[f][0])
        gap = _rel(relax, fmm, seeds[f][0])
        print(f"field {f}: mixed vs Relax max {e_relax[0]:.2e} rms {e_relax[1]:.2e}; mixed vs FMM max "
              f"{e_fmm[0]:.2e} rms {e_fmm[1]:.2e}; Relax vs FMM max {gap[0]:.2e} rms {gap[1]:.2e}")
        assert e_relax[0] < 1e-4
        assert e_fmm[0] <= gap[0] + 1e-4


def test_serving_at_full_sizes(device):
    from tomostar_amd import eikonal as E
    from tomostar_amd import lsqr as Q
    from tomostar_amd import rays as Ry
    from tomostar_amd import system as Sy
    from tomostar_amd import workloads as W
    from tomostar_amd.grid import GridDefinition
    from tomostar_amd.serving import Server

    big = W.batched_eikonal_case(seed=1, stations=4)
    g61 = GridDefinition(13.0, 13.74, 43.0, 43.54, 0.0, 60.0, 61, 61, 61)
    z = g61.depth()[:, None, None]
    vp61 = np.ascontiguousarray(np.broadcast_to((5.0 + 0.04 * z) * (1 + W.checkerboard(g61, 10.0, 10.0, 0.05)),
                                                g61.shape)).ravel()
    models = [(g61, 1 / vp61, 1.73 / vp61), (big.grid, big.slowness_p, big.slowness_s)]
    # a resident G of 1e5 curved rays (25 stations x 2000 events x P, S on the 61 x 61 x 31 case)
    wl = W.one_iteration_case(seed=7)
    seeds, mof = W.station_seed_sets(wl, E.point_source)
    ts = E.TableSet(device, wl.grid, [wl.slowness_p, wl.slowness_s], seeds, mof)
    assert ts.solve("fast_fp32")[0] == 0
    ev, st_idx, ev_idx, ph = W.scale_arrivals(wl, events=2000, rays=100_000, seed=9)
    batch = Ry.RayBatch(ts, wl.grid, np.repeat(wl.stations, 2, axis=0), 2 * st_idx + ph, ev[ev_idx], host_rows=False)
    used = np.flatnonzero(batch.counts >= 0).astype(np.int32)
    assert used.size >= 99_000
    rng = np.random.default_rng(10)
    sysd = Sy.build_system(device, wl.grid, Sy.SystemSettings(vpvs=False, smoothing_method="gradient"), batch, used,
                           ph[used], ev_idx[used], st_idx[used], 0.05 * rng.standard_normal(used.size),
                           np.full(used.size, 0.05), np.zeros(len(ev), np.uint8), np.zeros(len(ev), np.uint8),
                           len(wl.stations), wl.slowness_p, wl.slowness_s, wl.slowness_p, wl.slowness_s,
                           wl.slowness_s / wl.slowness_p, damping=0.0, smoothing=1.0, keep_fp64=True)
    batch.release()
    ts.release()
    A = sysd.matrix
    reqs = []
    for q in range(9):
        if q % 3 == 2:
            k = int(rng.integers(1, 9))
            B = sysd.rhs[None, :] + 0.1 * rng.standard_normal((k, A.rows))
            reqs.append(("lsqr", B, 10 ** rng.uniform(-1, 2, k), ["fast_fp64", "fast_mixed", "fast_fp32"][q % 3]))
        else:
            mi = (q // 2) % 2
            g = models[mi][0]
            ns = int(rng.integers(1, 8)) if mi == 1 else int(rng.integers(1, 50))
            src = np.stack([rng.uniform(g.min_lon, g.max_lon, ns), rng.uniform(g.min_lat, g.max_lat, ns),
                            rng.uniform(0, g.max_depth * 0.3, ns)], 1)
            reqs.append(("tables", mi, src, ["fast_fp32", "fast_fp64"][q % 2]))

    def serve(workers, concurrent, budget_gb):
        srv = Server(device, workers=workers, memory_budget_gb=budget_gb)
        ids = [(srv.add_model(g, sp), srv.add_model(g, ss)) for g, sp, ss in models]
        out = {}

        def submit(i):
            r = reqs[i]
            if r[0] == "tables":
                return srv.submit_tables(ids[r[1]][0], ids[r[1]][1], r[2], r[3])
            return srv.submit_lsqr(A, r[1], r[2], 100, mode=r[3])

        def collect(i, rid):
            st, why, _ = srv.wait(rid)
            assert st == 0, why
            r = reqs[i]
            if r[0] == "tables":
                out[i] = srv.tables(rid, 0, 2 * len(r[2]), models[r[1]][0].count)
            else:
                out[i] = srv.lsqr_result(rid, r[1].shape[0], A.cols)[0]
            srv.release_request(rid)

        if concurrent:
            rids = [submit(i) for i in range(len(reqs))]
            for i, rid in enumerate(rids):
                collect(i, rid)
        else:
            for i in range(len(reqs)):
                collect(i, submit(i))
        stats = srv.stats()
        srv.close()
        return out, stats

    print(f"\nserving: {len(reqs)} requests; models {[m[0].shape for m in models]}; G {A.rows} x {A.cols}, "
          f"{A.nnz} non-zeros from {used.size} rays", flush=True)
    mixed, st = serve(4, True, 40.0)
    alone, _ = serve(1, False, 40.0)
    print(f"serving: peak in flight {st['peak_inflight']}, peak admitted {st['peak_bytes'] / 1e9:.2f} GB of "
          f"{st['budget_bytes'] / 1e9:.0f} GB")
    assert st["peak_inflight"] >= 2
    assert st["peak_bytes"] <= st["budget_bytes"]
    for i in mixed:
        assert np.array_equal(mixed[i], alone[i]), i
    A.release()
