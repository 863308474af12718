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

  c2_gaussian   the second 10M input of the headline configuration (Gaussian,
                mean 0.5, sigma 0.1): GPU ms per call (median of the steps),
                filter-failure rates, and the reference's threaded time.
  c3            the degenerate-heavy 5.2M set (2000 x 2000 lattice step 1e-3,
                200,000 copies, 1,000,000 jittered circle points): GPU ms,
                filter-failure rates, cells rebuilt exactly, reference time.
  c1            100,000 uniform points (seed 0): the reference's serial
                qdt_execute and the GPU drop-in (host input), medians of 5.
  readme_sizes  the reference's own benchmark program (test/qdt_bench.cpp,
                scenario batch: 100 .. 2,624,409 uniform points in
                [0,1000)^2, median of 5, 3 above 2M) built twice, against the
                reference's library and against this one (oracle/_ref/
                qdt_bench_ref, qdt_bench_amd): serial and parallel call
                latencies side by side.
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
                queueing included in its latency; the plain path is one FIFO
                server, so its latencies follow from each request's measured
                service time and the arrival times).  CPU baseline: the
                reference, one context per host thread, on the first
                requests of the trace.

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


def host_info():
    """CPU model, usable cores and the reference build's compiler and flags."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        import subprocess

        gxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except Exception:
        gxx = None
    return {"cpu": model, "cores_usable": len(os.sched_getaffinity(0)), "compiler": gxx,
            "flags": "-O2 -std=c++14 -DNDEBUG -pthread -ffp-contract=off -fno-fast-math (oracle/Makefile)"}


def readme_sizes(args):
    """The reference's own benchmark program, scenario batch, against both
    libraries; returns rows per size."""
    import csv
    import subprocess

    out = {}
    for tag, exe in (("reference", "qdt_bench_ref"), ("gpu_dropin", "qdt_bench_amd")):
        path = os.path.join(ROOT, "oracle", "_ref", exe)
        if not os.path.exists(path):
            return {"error": "%s not built" % exe}
        csvp = os.path.join("/tmp", "qdt_bench_%s_%d.csv" % (tag, os.getpid()))
        r = subprocess.run([path, "--scenario", "batch", "--repeat", "5", "--csv", csvp, "--output", os.devnull],
                           capture_output=True, text=True, timeout=900)
        print("readme sizes (%s): rc %d" % (tag, r.returncode), file=sys.stderr, flush=True)
        if r.returncode != 0:
            return {"error": "%s exited %d: %s" % (exe, r.returncode, r.stderr[-500:])}
        for row in csv.DictReader(open(csvp)):
            n = int(row["points"])
            d = out.setdefault(n, {"points": n})
            m = row["metric"]
            if m == "triangles":
                d.setdefault("triangles", {})[tag] = int(row["value"])
            elif m in ("serial_ms_median", "parallel_ms_median"):
                d["%s_%s" % (tag, m.replace("_median", ""))] = round(float(row["value"]), 4)
        os.unlink(csvp)
    return {"program": "test/qdt_bench.cpp --scenario batch --repeat 5 (medians; 3 above 2M points), uniform "
                       "points in [0,1000)^2, built against each library",
            "rows": [out[k] for k in sorted(out)]}


def gpu_section(Q, torch, pts, steps, stats_keys=True):
    """Median GPU ms of qdt_gpu_execute_device on device-resident pts, and
    the last run's counters."""
    dev = torch.from_numpy(pts).to("cuda:0")
    torch.cuda.synchronize()
    ctx = Q.Triangulation()
    assert ctx.execute_device(dev.data_ptr(), len(pts)) == Q.QDT_OK
    ms = []
    for _ in range(steps):
        t0 = time.perf_counter()
        assert ctx.execute_device(dev.data_ptr(), len(pts)) == Q.QDT_OK
        ms.append((time.perf_counter() - t0) * 1e3)
    st = ctx.stats()
    ver = ctx.verify()
    ctx.close()
    del dev
    return {
        "points": len(pts), "distinct": int(st["unique_points"]), "triangles": int(st["triangles"]),
        "ms_median": round(statistics.median(ms), 3), "mpts_per_s": round(len(pts) / statistics.median(ms) / 1e3, 2),
        "filter_failure_rate": {"orient2d": st["orient_exact"] / max(1, st["orient_calls"]),
                                "incircle": st["incircle_exact"] / max(1, st["incircle_calls"]),
                                "extract_exact": int(st["extract_exact"])},
        "cells_exact": int(st["cells_exact"]), "cells": int(st["cells"]),
        "verify_ok": bool(ver["ok"]),
    }


def ref_threaded(pts, reps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    if not pyoracle.ref_available():
        return None
    times, tri = [], None
    for _ in range(reps):
        dt, tri = pyoracle.ref_time(pts, parallel=True)
        times.append(dt)
    return {"ms_median": round(1e3 * statistics.median(times), 1), "runs": reps, "triangles": int(tri),
            "kind": "reference", "call": "qdt_execute_parallel", "cores": min(os.cpu_count() or 1, 64)}


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
    # the same sets without graphs (one eager pass per call), tuned alike
    b0 = Q.Batch()
    b0.set_graphs(False)
    assert b0.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
    te = []
    for _ in range(max(args.steps // 2, 2)):
        t0 = time.perf_counter()
        assert b0.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
        te.append(time.perf_counter() - t0)
    t_eager = statistics.median(te)
    b0.close()
    # the batched form as shipped: autotuned, the first call captures the
    # shape's graphs, the timed calls replay them
    b = Q.Batch()
    for _ in range(2):
        assert b.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
    ts = []
    replays = 0
    for _ in range(max(args.steps // 2, 2)):
        t0 = time.perf_counter()
        assert b.execute_device(dev.data_ptr(), offs) == Q.QDT_OK
        ts.append(time.perf_counter() - t0)
        replays += int(b.stats()["graph_replay"])
    t_batch = statistics.median(ts)
    tri_total = int(b.all_triangles().shape[0])
    bst = b.stats()
    b.close()
    m_b = int(bst["unique_points"])
    # per set b_i = 2 m_i - 2 - T_i summed: E = 3 m - 3 k - b, written via T
    E_b = 3 * m_b - 3 * len(sets) - (2 * m_b - 2 * len(sets) - tri_total)
    alg_b = 36 * n + 40 * m_b + 48 * E_b + 12 * tri_total
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
        "graph": {"timed_calls_replayed": replays, "timed_calls": len(ts),
                  "launch": "two HIP graph launches per call (the second after the first's guards are read)",
                  "ms_per_batch_without_graphs": round(t_eager * 1e3, 3)},
        "tuning": {"cell_points": int(bst["cell_points"]), "merge_lanes": int(bst["merge_lanes"]),
                   "source": ["default", "timed", "cache", "environment"][int(bst["tune_source"])]},
        "device_ms": round(bst["device_ms"], 3),
        "pipeline_alg_bytes": alg_b,
        "pipeline_frac": alg_b / (bst["device_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS if bst["device_ms"] else None,
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
    nuniq = min(nreq, 250)
    sizes = W.log_uniform_sizes(nuniq, 1000, 2_000_000, 5)
    uniq = [W.mixed_set(int(sz), 5_000_011 + i)[0] for i, sz in enumerate(sizes)]
    rng = np.random.default_rng(17)
    # the trace: nreq arrivals, each one of the distinct point sets
    pick = np.concatenate([np.arange(nuniq), rng.integers(0, nuniq, size=max(0, nreq - nuniq))])
    reqs = [uniq[i] for i in pick]
    npts = int(sum(len(r) for r in reqs))
    # plain path: each distinct request alone through qdt_execute (host input,
    # host results), its service time measured once
    t = Q.Triangulation()
    t.execute(uniq[0])
    dur_u = []
    for r in uniq:
        t0 = time.perf_counter()
        assert t.execute(r) == Q.QDT_OK
        dur_u.append(time.perf_counter() - t0)
    t.close()
    dur = [dur_u[i] for i in pick]
    gaps = rng.exponential(1.0, size=nreq)
    rows = []
    with Q.Server(hbm_budget_bytes=200 << 30, workers=args.serve_workers) as srv:
        # warm the workers' workspaces
        for rid in [srv.submit(reqs[i]) for i in range(min(8, nreq))]:
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
            span = (last - first) * 1e-3
            rows.append({
                "qps": qps,
                "requests": nreq,
                "span_s": round(span, 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2),
                "mpts_per_s": round(npts / span / 1e6, 2),
                "batches": b1[0] - b0[0],
                "plain_p50_ms": round(float(np.percentile(slat, 50)), 2),
                "plain_p99_ms": round(float(np.percentile(slat, 99)), 2),
                "plain_mpts_per_s": round(float(npts / (end - arr[0]) / 1e6), 2),
                "plain_saturated": bool(sum(dur) > arr[-1]),
            })
            rows[-1]["vs_plain"] = round(rows[-1]["mpts_per_s"] / rows[-1]["plain_mpts_per_s"], 2)
            print("serve qps %s: %s" % (qps, rows[-1]), file=sys.stderr, flush=True)
        held, peak = srv.memory()
    out = {
        "workload": "%d requests (%d distinct point sets, sizes log-uniform [1000, 2000000], seed 5), uniform / "
                    "Gaussian / 32-cluster mix, %d points in all (mean %d), Poisson arrivals, host input and "
                    "results, 200 GB budget, %d workers" % (nreq, nuniq, npts, npts // nreq, args.serve_workers),
        "plain_path_ms_per_request_mean": round(1e3 * statistics.mean(dur), 2),
        "device_bytes_peak": int(peak),
        "by_qps": rows,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        try:
            import pyoracle

            if pyoracle.ref_available():
                k = min(args.serve_cpu_requests, nreq)
                sample = reqs[:k]
                pts_s, offs = concat(sample)
                thr = min(len(os.sched_getaffinity(0)), 64)
                dt, _ = pyoracle.ref_batch_time(pts_s, offs, thr)
                out["cpu_baseline"] = {"mpts_per_s": round(len(pts_s) / dt / 1e6, 3), "cores": thr,
                                       "kind": "reference",
                                       "sample": "the first %d requests of the trace (%d points), reference "
                                                 "qdt_execute per request, one context per host thread, %.2f s, "
                                                 "no arrival gaps (its best case)" % (k, len(pts_s), dt)}
        except Exception as e:
            out["cpu_baseline"] = {"error": str(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--dist", default="uniform", choices=["uniform", "gaussian"])
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-sections", action="store_true", help="skip c1 / c2_gaussian / c3")
    ap.add_argument("--no-readme-sizes", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-mesh", action="store_true")
    ap.add_argument("--batch-sets", type=int, default=4096)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-cpu-requests", type=int, default=100)
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
            "pipeline_frac": (36 * n + 40 * m + 48 * E + 12 * T) / (stats["device_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
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
        ds = []
        for _ in range(3):
            ts = time.perf_counter()
            st = ctx.execute(pts)
            ds.append(time.perf_counter() - ts)
            assert st == Q.QDT_OK
        dt = statistics.median(ds)
        line["dropin"] = {"call": "qdt_execute (host input, host results), median of 3", "ms": round(dt * 1e3, 3),
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
                    "host": host_info(),
                }
                # the serial path on the same input: one run (about 10 s)
                dt_s, tri_s = pyoracle.ref_time(pts, parallel=False)
                line["cpu_baseline"]["serial"] = {"value": round(n / dt_s / 1e6, 3), "unit": "Mpts/s",
                                                  "ms": round(dt_s * 1e3, 1), "runs": 1, "call": "qdt_execute",
                                                  "cores": 1, "triangles": int(tri_s)}
            else:
                line["cpu_baseline"] = None
        except Exception as e:  # the baseline is a report, never the measurement
            line["cpu_baseline"] = {"error": str(e)}

    if not args.no_sections:
        print("c2 gaussian / c3 / c1 ...", file=sys.stderr, flush=True)
        g = W.gaussian(n, 1)
        line["c2_gaussian"] = gpu_section(Q, torch, g, max(args.steps, 3))
        if not args.no_cpu_baseline:
            line["c2_gaussian"]["cpu_baseline"] = ref_threaded(g, 1)
        del g
        c3 = W.lattice_dups_circle()
        line["c3"] = gpu_section(Q, torch, c3, max(args.steps, 3))
        line["c3"]["input"] = "2000 x 2000 lattice (step 1e-3) + 200,000 copies + 1,000,000 circle points (jitter 1e-12)"
        if not args.no_cpu_baseline:
            line["c3"]["cpu_baseline"] = ref_threaded(c3, 1)
        del c3
        c1 = W.uniform(100_000, 0)
        t1 = Q.Triangulation()
        t1.execute(c1)
        g1 = []
        for _ in range(5):
            ts = time.perf_counter()
            assert t1.execute(c1) == Q.QDT_OK
            g1.append((time.perf_counter() - ts) * 1e3)
        line["c1"] = {"points": 100_000, "triangles": t1.triangles_size(),
                      "gpu_dropin_ms_median": round(statistics.median(g1), 3)}
        t1.close()
        if not args.no_cpu_baseline:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle

            if pyoracle.ref_available():
                pyoracle.ref_time(c1, parallel=False)
                r1 = [pyoracle.ref_time(c1, parallel=False)[0] * 1e3 for _ in range(5)]
                line["c1"]["reference_serial_ms_median"] = round(statistics.median(r1), 3)
    if not args.no_readme_sizes:
        print("readme sizes (the reference's bench program, both libraries) ...", file=sys.stderr, flush=True)
        try:
            line["readme_sizes"] = readme_sizes(args)
        except Exception as e:  # a report, never the measurement
            line["readme_sizes"] = {"error": str(e)}

    if not args.no_batched:
        print("batched form ...", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(args, Q, W, np, torch)
    if not args.no_serve:
        print("serving form ...", file=sys.stderr, flush=True)
        line["serve"] = bench_serve(args, Q, W, np)

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
