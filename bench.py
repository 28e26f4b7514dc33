"""Benchmark of the SGM hot path on one MI355X.

Workload (one step): one synthetic rectified pair at 1600x1067, census 9x7 (62-bit
descriptors), 256 disparities, P1=8, P2=96, 8 paths, uint16 cost and aggregate
volumes resident in HBM, both matching directions and the left-right check, through
the C-ABI device entry point (inputs already in HBM when the timed region starts).

Prints one JSON line: the metric is SGM cell-updates per second
(W*H*D*paths*2 directions / step time), with the path-sweep kernel's HBM roofline
and the CPU oracle's rate on the host beside it.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0


def workload_params(sgm):
    return sgm.SgmParams(window_w=9, window_h=7, disparities=256, min_disparity=0, cost=0, p1=8,
                         p2=96, edge_scale=32, directions=8, aggregate_bytes=0, census_weight=500,
                         ad_truncation=30, gradient_truncation=60, reserved=0, uniqueness=0.03,
                         lr_threshold=1.0)


def path_bytes_per_step(w, h, D, paths):
    """Algorithmic HBM bytes of the path sweeps for one pair (both directions):
    first sweep reads C and writes S (4 B/cell), middle sweeps read C, read S and
    write S (6 B/cell), the last sweep reads C and S and selects (4 B/cell)."""
    cells = 2 * w * h * D
    return cells * (4 + 6 * (paths - 2) + 4)


def cpu_baseline(w, D, rows=256, trials=1):
    """The C oracle (single thread) on a bounded strip of the same workload."""
    import numpy as np
    from oracle import oracle as O
    from pose_amd import synth
    left, right, _ = synth.stereo_pair(w, rows, D, 0, seed=2)
    op = O.params(window_w=9, window_h=7, disparities=D, p1=8, p2=96, lr_threshold=1.0)
    O.compute(op, left[:64], right[:64], w, 64)  # warm-up
    times = []
    for _ in range(trials):
        t0 = time.perf_counter()
        O.compute(op, left, right, w, rows)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    updates = w * rows * D * 8 * 2
    return dict(value=round(updates / t / 1e9, 4), unit="Gupdates/s", cores=1, kind="port",
                sample="%dx%d strip of the 1600x1067 D=%d workload, both directions + L-R, "
                       "C oracle single-threaded, %.2f s" % (w, rows, D, t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one MI355X; multi-GPU "
              "sharding of this path is design-only (see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from pose_amd import sgm, synth

    W, H, D, P = 1600, 1067, 256, 8
    params = workload_params(sgm)
    left, right, _ = synth.stereo_pair(W, H, D, 0, seed=2)
    n = W * H
    ctx = sgm.Context(0)
    d_left = ctx.alloc(n).upload(left)
    d_right = ctx.alloc(n).upload(right)
    d_out = ctx.alloc(4 * n)
    d_stats = ctx.alloc(128).upload(np.zeros(16, np.int64))

    def step():
        ctx.match_pair_device(params, d_left, d_right, W, H, d_out, d_stats=d_stats)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        ctx.synchronize()
        lat.append(time.perf_counter() - s0)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    updates = W * H * D * P * 2
    value = updates / (elapsed / args.steps) / 1e9
    valid = int(d_stats.download(np.int64, 16)[10])

    # Kernel durations with HIP events on the library's stream (separate pass).
    ctx.set_profiling(True)
    ctx.profile()  # reset
    for _ in range(args.profile_steps):
        step()
    ctx.synchronize()
    prof = ctx.profile()["device_kernel_milliseconds"]
    ctx.set_profiling(False)
    path_ms = prof["path"]["total_ms"] / args.profile_steps
    path_launches = prof["path"]["launches"] / args.profile_steps
    bytes_step = path_bytes_per_step(W, H, D, P)
    achieved = bytes_step / (path_ms / 1e3) / 1e9
    kernels = {k: round(v["total_ms"] / args.profile_steps, 4) for k, v in prof.items()}

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "path_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, D)

    lat_ms = sorted(x * 1e3 for x in lat)
    line = {
        "metric": "SGM Gupdates/s (px*disp*path)",
        "value": round(value, 3),
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": "plain SGM pair 1600x1067, census 9x7, D=256, P1=8 P2=96, "
                               "8 paths, uint16 volumes, left+right + L-R check",
                   "width": W, "height": H, "disparities": D, "paths": P, "pairs_per_step": 1},
        "pairs_per_s": round(1e3 / ms_per_step, 3),
        "latency_ms": {"p50": round(lat_ms[len(lat_ms) // 2], 4),
                       "p99": round(lat_ms[min(len(lat_ms) - 1, int(math.ceil(0.99 * len(lat_ms))) - 1)], 4)},
        "valid_pixels": valid,
        "kernel_ms_per_step": kernels,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": traffic, "kernel": "sgm::path_kernel",
                     "launches_per_step": path_launches,
                     "algorithmic_bytes_per_launch": bytes_step / max(1, path_launches)},
        "cpu_baseline": cpu,
    }
    for b in (d_left, d_right, d_out, d_stats):
        b.free()
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
