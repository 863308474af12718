"""Benchmark of the GPU batch Delaunay path (one MI355X).

A step is one qdt_gpu_execute_device() call on 10,000,000 device-resident
FP64 points uniform in [0,1)^2 (splitmix64 seed 1): dedupe + median-cut tree +
merges + triangle extraction, results left in HBM.  `value` is points per
second over the K timed steps.

Extra fields:
  roofline      the dominant kernel, k_merge (the Guibas-Stolfi merge of every
                node of one kd level, one thread per node; one launch per level
                whose nodes are merged in one piece): algorithmic bytes per
                launch = 48 B (one 24 B quad-edge record written once and read
                once) per edge the launch creates, counted on the device, over
                the launch's average duration, measured with HIP events on the
                engine's stream around every k_merge launch of the timed steps.
  cpu_baseline  the reference's threaded qdt_execute_parallel, compiled from
                the reference sources into oracle/_ref/, on the same 10M input,
                on this host's cores (median of --cpu-reps runs).
  dropin        qdt_execute() from host memory (H2D, pipeline, D2H, host
                result store): the PCIe-inclusive rate of the drop-in call.

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
    """HBM bytes per k_merge launch from the PMC runs summarised in
    profiles/k_merge_10M_uniform.md (FETCH_SIZE and WRITE_SIZE, each scaled
    by a calibration copy); None when that summary is absent."""
    p = os.path.join(ROOT, "profiles", "k_merge_10M_uniform.md")
    try:
        for line in open(p):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


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

    step_ms, merge_ms, merge_launches, merge_edges, stats = [], [], [], [], None
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
    torch.cuda.synchronize()
    total_s = time.perf_counter() - t0
    ms_per_step = total_s * 1e3 / args.steps
    value = n / (total_s / args.steps) / 1e6

    T = int(stats["triangles"])
    m = int(stats["unique_points"])
    E = 3 * m - 3 - (2 * m - 2 - T)  # Euler: E = 3m - 3 - b with b = 2m - 2 - T
    # k_merge roofline: each edge a launch creates is a 24 B record (four
    # onext + two origins) written once and read once -> 48 B per edge.
    launches = merge_launches[-1] or 1
    avg_launch_ms = statistics.mean(merge_ms) / launches
    bytes_per_launch = 48.0 * statistics.mean(merge_edges) / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9

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
            "kernel": "k_merge (one launch per kd level merged in one piece per node)",
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_launch(),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "launches_per_step": launches,
            "edges_per_step": int(merge_edges[-1]),
            "bytes_per_launch": bytes_per_launch,
            "seam_levels_ms": round(stats["seam_level_ms_sum"], 3),
            "pipeline_device_ms": round(stats["device_ms"], 3),
            "pipeline_alg_bytes": 36 * n + 40 * m + 48 * E + 12 * T,
        },
    }

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

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
