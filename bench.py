"""Benchmark of the GPU batch Delaunay path (one MI355X).

A step is one qdt_gpu_execute_device() call on 10,000,000 device-resident
FP64 points uniform in [0,1)^2 (splitmix64 seed 1): dedupe + median-cut tree +
merges + triangle extraction, results left in HBM.  `value` is points per
second over the K timed steps.

Extra fields:
  roofline      the dominant kernel, k_cell (one wavefront per LDS cell builds
                the cell's whole kd subtree: partition, leaves, merges; one
                launch per step): algorithmic bytes per launch = 100 B per
                distinct point (reads: 16 B coordinates, 4 + 4 B of the two
                frame orders; writes: three 24 B quad-edge records and a 4 B
                vertex-edge entry) + 16 B per cell (its result), over the
                launch's duration measured with HIP events on the engine's
                stream around the cell stage of every timed step (k_cell,
                which runs the floating-point filters alone, and the exact
                pass over the cells it lists: none on this input).  The
                k_merge launches (levels above the cells merged with one
                thread per node) are reported beside it.
  cpu_baseline  the reference's threaded qdt_execute_parallel, compiled from
                the reference sources into oracle/_ref/, on the same 10M input,
                on this host's cores (median of --cpu-reps runs).
  verify        the exact device verifier (qdt_gpu_verify) on the last
                step's result: its verdict and kernel time.
  dropin        qdt_execute() from host memory (H2D, pipeline, D2H, host
                result store): the PCIe-inclusive rate of the drop-in call.
  mesh          the half-edge mesh of the last step's result built on the
                device (qdt_gpu_build_mesh: half-edges, twins, local
                vertices), median kernel time of 3 builds, and as its CPU
                baseline the reference's own mesh build (ensure_mesh_parallel_,
                after its batch call on the same input).  Algorithmic bytes:
                12 B per triangle read, 16 B per half-edge, 8 B per local
                vertex and 4 B per input index written.

  batched       the batched form: 4,096 sets, sizes log-uniform in [64, 65536]
                (seed 3), each uniform / Gaussian (sigma 0.1) / 32-cluster
                mixture; one qdt_gpu_batch_execute_device call per step
                versus the same sets one qdt_gpu_execute_device call at a
                time, and the reference (qdt_execute per set, one context
                per host thread) as its CPU baseline.
  serve         the serving form: a trace of requests (sizes log-uniform in
                [1000, 2,000,000], the same mix) with Poisson arrivals at
                each QPS, through qdt_gpu_server (continuous batching, 200 GB
                budget): p50/p99 submit-to-done latency and points/s, beside
                the plain per-request path on the same trace (one request at
                a time in arrival order; each request's own qdt_execute time,
                queueing included in its latency).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def traffic_per_launch():
    """HBM bytes per k_cell launch from the PMC runs summarised in
    profiles/k_cell_10M_uniform.md (FETCH_SIZE and WRITE_SIZE, each scaled
    by a calibration copy); None when that summary is absent."""
    p = os.path.join(ROOT, "profiles", "k_cell_10M_uniform.md")
    try:
        for line in open(p):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def concat(sets):
    import numpy as np

    offs = np.zeros(len(sets) + 1, dtype=np.uint64)
    for i, p in enumerate(sets):
        offs[i + 1] = offs[i] + len(p)
    return np.ascontiguousarray(np.concatenate(sets)), offs


def bench_batched(args, Q, W, np, torch):
    """The batched form against one call per set (see the module docstring)."""
    sets = W.batch_sets(count=args.batch_sets, seed=3)
    pts, offs = concat(sets)
    n = len(pts)
    dev = torch.from_numpy(pts).to("cuda:0")
    torch.cuda.synchronize()
    b = Q.Batch()
    for _ in range(1):
        assert b.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
    ts = []
    for _ in range(max(args.steps // 2, 2)):
        t0 = time.perf_counter()
        assert b.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
        ts.append(time.perf_counter() - t0)
    t_batch = statistics.median(ts)
    tri_total = int(b.all_triangles().shape[0])
    b.close()
    # one call per set, device-resident input, same sets
    t = Q.Triangulation()
    for i in range(8):
        t.execute_device(dev.data_ptr() + 16 * int(offs[i]), len(sets[i]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(sets)):
        assert t.execute_device(dev.data_ptr() + 16 * int(offs[i]), len(sets[i])) == Q.QDT_OK
    t_single = time.perf_counter() - t0
    t.close()
    out = {
        "workload": "%d sets, sizes log-uniform [64, 65536] (seed 3), uniform / Gaussian / 32-cluster mix, "
                    "%d points in all, device-resident" % (len(sets), n),
        "sets_per_s": round(len(sets) / t_batch, 1),
        "mpts_per_s": round(n / t_batch / 1e6, 3),
        "ms_per_batch": round(t_batch * 1e3, 3),
        "triangles": tri_total,
        "one_call_per_set": {"sets_per_s": round(len(sets) / t_single, 1),
                             "mpts_per_s": round(n / t_single / 1e6, 3), "ms": round(t_single * 1e3, 1)},
        "speedup_vs_one_call_per_set": round(t_single / t_batch, 2),
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        try:
            import pyoracle

            if pyoracle.ref_available():
                thr = min(len(os.sched_getaffinity(0)), 64)
                dt, tri_ref = pyoracle.ref_batch_time(pts, offs, thr)
                out["cpu_baseline"] = {"sets_per_s": round(len(sets) / dt, 1), "mpts_per_s": round(n / dt / 1e6, 3),
                                       "cores": thr, "kind": "reference",
                                       "sample": "all %d sets, reference qdt_execute per set, one context per "
                                                 "host thread, %.2f s; triangles %d (GPU %d)"
                                                 % (len(sets), dt, tri_ref, tri_total)}
        except Exception as e:  # a report, never the measurement
            out["cpu_baseline"] = {"error": str(e)}
    del dev
    return out


def bench_serve(args, Q, W, np):
    """The serving form against the plain per-request path (module docstring)."""
    nreq = args.serve_requests
    sizes = W.log_uniform_sizes(nreq, 1000, 2_000_000, 5)
    reqs = [W.mixed_set(int(sz), 5_000_011 + i)[0] for i, sz in enumerate(sizes)]
    npts = int(sum(len(r) for r in reqs))
    # plain path: each request alone through qdt_execute (host input, host results)
    t = Q.Triangulation()
    t.execute(reqs[0])
    dur = []
    for r in reqs:
        t0 = time.perf_counter()
        assert t.execute(r) == Q.QDT_OK
        dur.append(time.perf_counter() - t0)
    t.close()
    rng = np.random.default_rng(17)
    gaps = rng.exponential(1.0, size=nreq)
    rows = []
    with Q.Server(hbm_budget_bytes=200 << 30, workers=args.serve_workers) as srv:
        # warm the workers' workspaces
        for rid in [srv.submit(reqs[i]) for i in range(min(4, nreq))]:
            srv.wait(rid)
            srv.release(rid)
        for qps in args.serve_qps:
            arr = np.cumsum(gaps) / qps
            arr -= arr[0]
            ids = []
            b0 = srv.counters()
            t0 = time.perf_counter()
            for i in range(nreq):
                delay = t0 + arr[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(reqs[i]))
            lat, first, last = [], None, None
            for rid in ids:
                st, l = srv.wait(rid)
                assert st == Q.QDT_OK, Q.status_string(st)
                sub, _, done = srv.times(rid)
                lat.append(l)
                first = sub if first is None else min(first, sub)
                last = done if last is None else max(last, done)
                srv.release(rid)
            b1 = srv.counters()
            # the plain path on the same arrivals: one request at a time, FIFO
            end, slat = 0.0, []
            for i in range(nreq):
                start = max(arr[i], end)
                end = start + dur[i]
                slat.append((end - arr[i]) * 1e3)
            rows.append({
                "qps": qps,
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mpts_per_s": round(npts / ((last - first) * 1e-3) / 1e6, 2),
                "batches": b1[0] - b0[0],
                "plain_p50_ms": round(float(np.percentile(slat, 50)), 2),
                "plain_p99_ms": round(float(np.percentile(slat, 99)), 2),
                "plain_mpts_per_s": round(float(npts / (end - arr[0]) / 1e6), 2),
            })
            print("serve qps %s: %s" % (qps, rows[-1]), file=sys.stderr, flush=True)
    return {
        "workload": "%d requests, sizes log-uniform [1000, 2000000] (seed 5), uniform / Gaussian / 32-cluster "
                    "mix, %d points in all (mean %d), Poisson arrivals, host input and results, 200 GB budget, "
                    "%d workers" % (nreq, npts, npts // nreq, args.serve_workers),
        "plain_path_ms_per_request_mean": round(1e3 * statistics.mean(dur), 2),
        "by_qps": rows,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--dist", default="uniform", choices=["uniform", "gaussian"])
    ap.add_argument("--cpu-reps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--batch-sets", type=int, default=4096)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-workers", type=int, default=2)
    ap.add_argument("--serve-qps", type=lambda v: [float(x) for x in v.split(",")],
                    default=[100, 250, 500, 1000, 2500, 5000])
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one plain device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import libqdt_amd as Q
    from libqdt_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    pts = W.uniform(n, 1) if args.dist == "uniform" else W.gaussian(n, 1)
    dev = torch.from_numpy(pts).to("cuda:0")
    torch.cuda.synchronize()

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ctx = Q.Triangulation()
    ctx.set_profiling(True)
    for _ in range(args.warmup):
        st = ctx.execute_device(dev.data_ptr(), n)
        assert st == Q.QDT_OK, Q.status_string(st)
    torch.cuda.synchronize()

    step_ms, merge_ms, merge_launches, merge_edges, cell_ms, stats = [], [], [], [], [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st = ctx.execute_device(dev.data_ptr(), n)
        step_ms.append((time.perf_counter() - ts) * 1e3)
        assert st == Q.QDT_OK, Q.status_string(st)
        stats = ctx.stats()
        merge_ms.append(stats["merge_launch_ms_sum"])
        merge_launches.append(stats["merge_launches"])
        merge_edges.append(stats["merge_edges"])
        cell_ms.append(stats["stage_ms"]["leaves"])
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    ms_per_step = total_s * 1e3 / args.steps
    value = n / (total_s / args.steps) / 1e6

    # the exact device verifier on the last step's result
    ver = ctx.verify()

    T = int(stats["triangles"])
    m = int(stats["unique_points"])
    E = 3 * m - 3 - (2 * m - 2 - T)  # Euler: E = 3m - 3 - b with b = 2m - 2 - T
    # k_cell roofline (module docstring): 100 B per distinct point + 16 B per cell
    cells = int(stats["cells"])
    cell_bytes = 100.0 * m + 16.0 * cells
    cell_avg_ms = statistics.mean(cell_ms)
    achieved = cell_bytes / (cell_avg_ms * 1e-3) / 1e9
    # k_merge: each edge a launch creates is a 24 B record (four onext + two
    # origins) written once and read once -> 48 B per edge
    launches = merge_launches[-1] or 1
    avg_launch_ms = statistics.mean(merge_ms) / launches
    bytes_per_launch = 48.0 * statistics.mean(merge_edges) / launches
    merge_achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9

    line = {
        "metric": "Mpts/s triangulated (10M uniform f64); batched sets/s; serve p50/p99 ms",
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "single-instance batch Delaunay, %d FP64 points %s (splitmix64 seed 1), "
                        "device-resident input, qdt_gpu_execute_device" % (
                            n, "uniform in [0,1)^2" if args.dist == "uniform" else "Gaussian (0.5, 0.1)"),
            "points": n,
            "distinct": m,
            "triangles": T,
            "step_ms_median": round(statistics.median(step_ms), 3),
            "stage_ms": {k: round(v, 3) for k, v in stats["stage_ms"].items()},
            "filter_failure_rate": {
                "orient2d": stats["orient_exact"] / max(1, stats["orient_calls"]),
                "incircle": stats["incircle_exact"] / max(1, stats["incircle_calls"]),
            },
        },
        "roofline": {
            "kernel": "k_cell (LDS cells, one launch per step)",
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_launch(),
            "avg_launch_ms": round(cell_avg_ms, 4),
            "bytes_per_launch": cell_bytes,
            "cells": cells,
            "cell_level": int(stats["cell_level"]),
            "k_merge": {"launches_per_step": launches, "avg_launch_ms": round(avg_launch_ms, 4),
                        "edges_per_step": int(merge_edges[-1]), "bytes_per_launch": bytes_per_launch,
                        "achieved_gbs": round(merge_achieved, 3)},
            "seam_levels_ms": round(stats["seam_level_ms_sum"], 3),
            "pipeline_device_ms": round(stats["device_ms"], 3),
            "pipeline_alg_bytes": 36 * n + 40 * m + 48 * E + 12 * T,
        },
        "verify": {"check": "qdt_gpu_verify on the last step's result (exact orientation, manifold edges, "
                            "local Delaunay, convex boundary, Euler, every distinct point used)",
                   "ok": ver["ok"], "device_ms": round(ver["device_ms"], 3),
                   "non_delaunay": ver["non_delaunay"], "boundary_edges": ver["boundary_edges"]},
    }

    if not args.no_mesh:
        d_tris, Tm = ctx.device_triangles()
        he = torch.empty((3 * Tm, 4), dtype=torch.int32, device="cuda")
        vert = torch.empty((min(n, 3 * Tm), 2), dtype=torch.int32, device="cuda")
        o2l = torch.empty(n, dtype=torch.int32, device="cuda")
        infos = [Q.build_mesh(d_tris, Tm, n, he.data_ptr(), vert.data_ptr(), o2l.data_ptr()) for _ in range(4)]
        mesh_ms = statistics.median(i["device_ms"] for i in infos[1:])
        nv = int(infos[-1]["vertices"])
        mesh_bytes = 12.0 * Tm + 16.0 * 3 * Tm + 8.0 * nv + 4.0 * n
        line["mesh"] = {"call": "qdt_gpu_build_mesh on the last step's triangles (device arrays)",
                        "ms": round(mesh_ms, 3), "halfedges": 3 * Tm, "vertices": nv,
                        "boundary": int(infos[-1]["boundary"]), "unmatched": int(infos[-1]["unmatched"]),
                        "alg_bytes": mesh_bytes, "achieved_gbs": round(mesh_bytes / (mesh_ms * 1e-3) / 1e9, 2)}
        del he, vert, o2l
        if not args.no_cpu_baseline:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            try:
                import pyoracle

                if pyoracle.ref_available():
                    dt, he_ref = pyoracle.ref_mesh_time(pts, parallel=True)
                    line["mesh"]["cpu_baseline"] = {
                        "ms": round(dt * 1e3, 1), "kind": "reference", "cores": min(os.cpu_count() or 1, 64),
                        "sample": "reference ensure_mesh_parallel_ on its own result of the same input (after an "
                                  "untimed qdt_execute_parallel), one run; half-edges %d" % he_ref}
            except Exception as e:
                line["mesh"]["cpu_baseline"] = {"error": str(e)}

    if not args.no_dropin:
        st = ctx.execute(pts)  # warm the host path (pinned buffers)
        assert st == Q.QDT_OK
        ts = time.perf_counter()
        st = ctx.execute(pts)
        dt = time.perf_counter() - ts
        assert st == Q.QDT_OK
        line["dropin"] = {"call": "qdt_execute (host input, host results)", "ms": round(dt * 1e3, 3),
                          "value": round(n / dt / 1e6, 3), "unit": "Mpts/s"}
    ctx.close()

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        try:
            import pyoracle

            if pyoracle.ref_available():
                times = []
                tri_ref = None
                for _ in range(args.cpu_reps):
                    dt, tri_ref = pyoracle.ref_time(pts, parallel=True)
                    times.append(dt)
                med = statistics.median(times)
                hw = os.cpu_count() or 1
                line["cpu_baseline"] = {
                    "value": round(n / med / 1e6, 3),
                    "unit": "Mpts/s",
                    "cores": min(hw, 64),
                    "kind": "reference",
                    "sample": "the same %d-point input, reference qdt_execute_parallel (libqdt built with "
                              "g++ -O2 from the reference sources), median of %d runs, %.2f s each; "
                              "triangles %d" % (n, args.cpu_reps, med, tri_ref),
                    "nproc": len(os.sched_getaffinity(0)),
                }
            else:
                line["cpu_baseline"] = None
        except Exception as e:  # the baseline is a report, never the measurement
            line["cpu_baseline"] = {"error": str(e)}

    if not args.no_batched:
        print("batched form ...", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(args, Q, W, np, torch)
    if not args.no_serve:
        print("serving form ...", file=sys.stderr, flush=True)
        line["serve"] = bench_serve(args, Q, W, np)

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
