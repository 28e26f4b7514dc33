"""Benchmark of the SGM hot path on one MI355X.

Workload (one step): one synthetic rectified pair at 1600x1067, census 9x7 (62-bit
descriptors), 256 disparities, P1=8, P2=96, 8 paths, uint8 cost and uint16 aggregate
volumes resident in HBM, both matching directions and the left-right check, through
the C-ABI device entry point (inputs already in HBM when the timed region starts).

Prints one JSON line: the metric is SGM cell-updates per second
(W*H*D*paths*2 directions / step time), with the path-sweep kernel's HBM roofline
and the CPU oracle's rate on the host beside it.
"""
import argparse
import hashlib
import json
import math
import os
import subprocess
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


def path_bytes_per_step(w, h, D, paths, schedule, cost_bytes=1):
    """HBM bytes this build's path sweeps move for one pair (both directions)
    with a uint8 cost volume C, by sweep schedule (the plan's `sweeps`).
    Path-cost planes (3): the census cost pass writes C (1 B/cell), every
    direction but the last reads C and writes its plane (1 + 1 B/cell each), the
    last sweep reads C and the P-1 planes and selects (1 + 7 B/cell): 23 B per
    cell at 8 paths.  One uint16 aggregate S
    (1, 2): first 1 + 2, middle 1 + 2 + 2 (a paired launch's second member 1 +
    2, and the last sweep then reads both partial sums), last 1 + 2: 36 B per
    cell at 8 paths either way."""
    cells = 2 * w * h * D
    if schedule == 3:
        per_cell = cost_bytes + (cost_bytes + 1) * (paths - 1) + cost_bytes + (paths - 1)
    else:
        per_cell = (cost_bytes + 2) + (cost_bytes + 4) * (paths - 2) + (cost_bytes + 2)
    return cells * per_cell


def algorithmic_bytes_per_step(w, h, D, paths):
    """The problem's algorithmic HBM bytes for one pair (both directions) with
    uint16 cost and aggregate volumes: cost-build write 2 B, P cost reads 2P B,
    P-1 aggregate reads 2(P-1) B, P aggregate writes 2P B, less the final
    aggregate write that a selection fused into the last sweep does not make:
    6P - 2 = 46 B per cell at 8 paths."""
    return 2 * w * h * D * (6 * paths - 2)


# Sources whose change invalidates the committed HBM traffic measurement of the
# path sweeps (profiles/path_traffic.json carries the hash it was measured at).
TRAFFIC_SOURCES = ["pose_amd/csrc/sgm_paths.h", "pose_amd/csrc/sgm_kernels.h",
                   "pose_amd/csrc/sgm_path_launch.h"]


def kernel_source_hash():
    h = hashlib.sha256()
    for f in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def cpu_info():
    """Host CPU model (lscpu) and the processors this process may use."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0))}


def timed_trials(fn, trials=3):
    """MatchingCostBenchmark protocol: one discarded warm-up, then `trials`
    timed runs; the median."""
    fn()
    times = []
    for _ in range(trials):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], times


def cpu_baseline(w, h, D, rows=256, trials=3):
    """The C oracle on the host: (a) one thread in the reference's loop order on
    a `rows`-row strip of the workload, (b) split over every host core on the
    whole pair; one warm-up and `trials` timed runs each, median."""
    from oracle import oracle as O
    from pose_amd import synth
    left, right, _ = synth.stereo_pair(w, h, D, 0, seed=2)
    op = O.params(window_w=9, window_h=7, disparities=D, p1=8, p2=96, lr_threshold=1.0)
    t1, runs1 = timed_trials(lambda: O.compute(op, left[:rows], right[:rows], w, rows), trials)
    single = w * rows * D * 8 * 2 / t1 / 1e9
    threads = len(os.sched_getaffinity(0))
    tn, runsn = timed_trials(lambda: O.compute_parallel(op, left, right, w, h, threads=threads),
                             trials)
    allcores = w * h * D * 8 * 2 / tn / 1e9
    info = cpu_info()
    return dict(value=round(allcores, 4), unit="Gupdates/s", cores=threads, kind="port",
                sample="whole %dx%d D=%d pair, both directions + L-R, C oracle split over %d "
                       "host threads (OpenMP); 1 warm-up + %d trials, median %.2f s"
                       % (w, h, D, threads, trials, tn),
                trials_s=[round(t, 3) for t in runsn], cpu_model=info["model"],
                nproc=info["nproc"],
                single_thread={"value": round(single, 4), "unit": "Gupdates/s", "cores": 1,
                               "sample": "%dx%d strip of the same pair, C oracle in the "
                                         "reference's loop order; 1 warm-up + %d trials, "
                                         "median %.2f s" % (w, rows, trials, t1),
                               "trials_s": [round(t, 3) for t in runs1]})


def cpu_baseline_configs(trials=3):
    """The C oracle over every host core on the other workloads, rates in
    updates of the sample: the 640x480 D=64 pair whole and a quarter-height
    strip of the FP32 ZNCC pair (its FP64 oracle), one warm-up and `trials`
    timed runs, median; the batched distribution on its two smallest pairs of
    the 48 (one warm-up, one timed run)."""
    from oracle import oracle as O
    from pose_amd import synth
    threads = len(os.sched_getaffinity(0))
    out = {}
    left, right, _ = synth.stereo_pair(640, 480, 64, 0, seed=1)
    op = O.params(window_w=9, window_h=7, disparities=64, p1=10, p2=120, lr_threshold=1.0)
    t, runs = timed_trials(lambda: O.compute_parallel(op, left, right, 640, 480, threads=threads),
                           trials)
    out["640x480_d64"] = {"value": round(640 * 480 * 64 * 16 / t / 1e9, 4), "unit": "Gupdates/s",
                          "cores": threads, "sample": "whole pair, median of %d" % trials,
                          "trials_s": [round(x, 3) for x in runs]}
    left, right, _ = synth.stereo_pair(1600, 1067, 192, 0, seed=3, gain=True)
    rows = 267  # a quarter of the pair's rows bounds the FP64 oracle's time
    op = O.params(window_w=7, window_h=7, disparities=192, cost=4, p1=12, p2=150,
                  lr_threshold=1.0)
    t, runs = timed_trials(lambda: O.compute_parallel(op, left[:rows], right[:rows], 1600, rows,
                                                      threads=threads), trials)
    out["zncc_1600x1067_d192"] = {"value": round(1600 * rows * 192 * 16 / t / 1e9, 4),
                                  "unit": "Gupdates/s", "cores": threads,
                                  "sample": "1600x%d strip of the pair, FP64 oracle, median of %d"
                                            % (rows, trials),
                                  "trials_s": [round(x, 3) for x in runs]}
    spec = sorted(batch_workload(), key=lambda s: s[0] * s[1] * s[2])[:2]
    pairs = []
    for (w, h, D, md, seed) in spec:
        left, right, _ = synth.stereo_pair(w, h, D, md, seed=seed)
        pairs.append((O.params(window_w=9, window_h=7, disparities=D, min_disparity=md, p1=10,
                               p2=120, lr_threshold=1.0), left, right, w, h))

    def run_subset():
        for (op, left, right, w, h) in pairs:
            O.compute_parallel(op, left, right, w, h, threads=threads)
    t, runs = timed_trials(run_subset, 1)
    upd = sum(w * h * op.disparities * 16 for (op, _, _, w, h) in pairs)
    out["batched_distribution"] = {"value": round(upd / t / 1e9, 4), "unit": "Gupdates/s",
                                   "cores": threads,
                                   "sample": "the 2 smallest of the 48 pairs (%s), one warm-up "
                                             "and one timed run, %.2f s" % (
                                                 ", ".join("%dx%d D=%d" % (w, h, D)
                                                           for (w, h, D, _, _) in spec), t)}
    return out


def batch_workload(count=48, seed=4):
    """The batched workload: uneven landscape pairs, long side U[800, 3000], aspect
    4:3 or 3:2, D U[64, 384], min disparity U[-32, 32], per-pair seeds."""
    from pose_amd import synth
    return synth.batch_pairs(count=count, seed=seed)


def run_batched(sgm, params_for, steps=2, spec=None, label=None):
    """Same pairs through (a) the plain path, one pair per launch sequence,
    (b) the phase-major batch and (c) the single persistent work-queue launch,
    both replayed from a hipGraph; inputs in HBM.  Default: the 48 uneven pairs
    of the batched workload."""
    import numpy as np
    from pose_amd import synth
    spec = batch_workload() if spec is None else spec
    # the largest budget the library allows (95% of device memory): the batch
    # keeps its cost volumes resident when they fit, else recomputes census costs
    ctx = sgm.Context(0, budget_bytes=1 << 62)
    bufs, recs = [], []
    updates = 0
    t_gen = time.perf_counter()
    for (w, h, D, md, seed) in spec:
        left, right, _ = synth.stereo_pair(w, h, D, md, seed=seed)
        n = w * h
        b = dict(left=ctx.alloc(n).upload(left), right=ctx.alloc(n).upload(right),
                 checked=ctx.alloc(4 * n), plain=ctx.alloc(4 * n), stats=ctx.alloc(128))
        bufs.append(b)
        recs.append(sgm.SgmPair(width=w, height=h, disparities=D, min_disparity=md,
                                left=b["left"].ptr, right=b["right"].ptr, checked=b["checked"].ptr,
                                stats=b["stats"].ptr))
        updates += w * h * D * 8 * 2
    t_gen = time.perf_counter() - t_gen
    base = params_for(64, 0)
    # plain: one pair per launch sequence, all pairs back to back
    def plain_all():
        for (w, h, D, md, _), b in zip(spec, bufs):
            ctx.match_pair_device(params_for(D, md), b["left"], b["right"], w, h, b["plain"])
        ctx.synchronize()
    plain_all()  # warm-up (allocates the plain path's buffers)
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_all()
    t_plain = (time.perf_counter() - t0) / steps
    ctx.release_buffers()

    def timed_batch(phased):
        for b in bufs:
            b["checked"].upload(np.zeros(b["checked"].nbytes // 4, np.uint32))
        batch = sgm.DeviceBatch(ctx, base, recs, graph=True, phased=phased)
        batch.run()  # warm-up
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run()
        t = (time.perf_counter() - t0) / steps
        same = all(np.array_equal(b["checked"].download(np.uint32, (w * h)),
                                  b["plain"].download(np.uint32, (w * h)))
                   for (w, h, _, _, _), b in zip(spec, bufs))
        out = (t, same, batch.device_bytes(), batch.describe())
        batch.close()
        return out

    # (b) phase-major batch: the work-queue kernel takes the census and L-R
    # items, each direction runs one specialised sweep launch per lane group
    # over every pair's lines; one graph
    t_batch, same, dev_bytes, shape = timed_batch(True)
    # (c) the whole batch as one persistent work-queue launch
    t_queue, same_queue, _, queue_shape = timed_batch(False)
    for b in bufs:
        for v in b.values():
            v.free()
    ctx.close()
    return {
        "workload": label or ("48 uneven pairs, long side U[800,3000], aspect 4:3/3:2, "
                              "D U[64,384], min disparity U[-32,32], census 9x7, P1=10 P2=120, "
                              "8 paths, L+R + L-R"),
        "pairs": len(spec), "cell_updates": updates,
        "plain_ms": round(t_plain * 1e3, 3), "batched_ms": round(t_batch * 1e3, 3),
        "plain_gupdates_per_s": round(updates / t_plain / 1e9, 3),
        "batched_gupdates_per_s": round(updates / t_batch / 1e9, 3),
        "plain_pairs_per_s": round(len(spec) / t_plain, 3),
        "batched_pairs_per_s": round(len(spec) / t_batch, 3),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "bit_exact_vs_plain": bool(same), "batch_device_bytes": dev_bytes,
        "batch": shape,
        "single_kernel_ms": round(t_queue * 1e3, 3),
        "single_kernel_speedup_vs_plain": round(t_plain / t_queue, 3),
        "single_kernel_bit_exact_vs_plain": bool(same_queue),
        "single_kernel_batch": queue_shape,
        "input_generation_s": round(t_gen, 2),
    }


def run_filtered(sgm, left, right, W, H, steps=3):
    """The benchmark pair through the reference's post-L-R filters (default
    options: edge-preserving 3x3 median, 7x7 photometric ZNCC >= 0.55, speckles
    < 50 px): device kernel times of the filter stage."""
    p = workload_params(sgm)
    f = sgm.SgmFilterParams(median=3, median_classic=0, median_edge_threshold=24,
                            correlation_window=7, speckle_size=50, speckle_connectivity=8,
                            subpixel_refinement=0, reserved1=0, median_max_change=0.5, correlation=0.55,
                            speckle_range=1.0, reserved2=0.0)
    ctx = sgm.Context(0)
    r = ctx.match_pair_filtered(p, f, left, right)
    ctx.set_profiling(True)
    ctx.profile()
    for _ in range(steps):
        r = ctx.match_pair_filtered(p, f, left, right)
    prof = ctx.profile()["device_kernel_milliseconds"]
    ctx.set_profiling(False)
    ctx.close()
    st = r["stats"]
    return {"filters_ms": round(prof["filters"]["total_ms"] / steps, 4),
            "valid_pixels": st["valid"], "post_median_lr_rejected": st["post_median_lr_rejected"],
            "photometric_rejected": st["photometric_rejected"],
            "speckle_rejected": st["speckle_rejected"]}


def run_streamed(sgm, W, H, D, P, d_left, d_right, rows=64, steps=2):
    """The benchmark pair under the strip-streamed plan (strips of `rows` rows;
    strip cost + aggregate on the device, the aggregate parked in pinned host
    memory between the top-down and bottom-up passes): inputs in HBM, the
    aggregate's PCIe round trip inside the timed region."""
    params = workload_params(sgm)
    ctx = sgm.Context(0)
    ctx.plan_mode(sgm.PLAN_STREAMED, rows)
    n = W * H
    out = ctx.alloc(4 * n)
    ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.profile()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    prof = {k: round(v["total_ms"] / steps, 3)
            for k, v in ctx.profile()["device_kernel_milliseconds"].items()}
    ctx.set_profiling(False)
    plan = ctx.plan(params, W, H)
    out.free()
    ctx.close()
    upd = W * H * D * P * 2
    return {"workload": "plain pair 1600x1067 D=256 under the streamed plan, strips of %d rows" % rows,
            "ms_per_pair": round(t * 1e3, 3), "gupdates_per_s": round(upd / t / 1e9, 3),
            "kernel_ms": prof, "device_bytes": plan["streamed_bytes"],
            "host_aggregate_bytes": plan["host_bytes"],
            "pcie_bytes_per_pair": 2 * plan["host_bytes"]}


def run_small_pair(sgm, steps=50):
    """One 640x480 pair, D=64, census 9x7, P1=10, P2=120, 8 paths, L+R + L-R,
    inputs in HBM: the volumes (C + S about 79 MB per direction) fit the
    256 MB MALL, so the pair is latency- and launch-bound, not HBM-bound."""
    import numpy as np
    from pose_amd import synth
    W, H, D = 640, 480, 64
    p = sgm.SgmParams(window_w=9, window_h=7, disparities=D, min_disparity=0, cost=0, p1=10,
                      p2=120, edge_scale=32, directions=8, aggregate_bytes=0, census_weight=500,
                      ad_truncation=30, gradient_truncation=60, reserved=0, uniqueness=0.03,
                      lr_threshold=1.0)
    left, right, _ = synth.stereo_pair(W, H, D, 0, seed=1)
    ctx = sgm.Context(0)
    n = W * H
    dl, dr, out = ctx.alloc(n).upload(left), ctx.alloc(n).upload(right), ctx.alloc(4 * n)
    for _ in range(3):
        ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    lat = []  # one call and its completion at a time
    for _ in range(steps):
        t1 = time.perf_counter()
        ctx.match_pair_device(p, dl, dr, W, H, out)
        ctx.synchronize()
        lat.append(time.perf_counter() - t1)
    ctx.set_profiling(True)
    ctx.profile()
    ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    prof = {k: round(v["total_ms"], 4) for k, v in ctx.profile()["device_kernel_milliseconds"].items()}
    ctx.set_profiling(False)
    valid = int(np.isfinite(out.download(np.float32, n)).sum())
    for b in (dl, dr, out):
        b.free()
    ctx.close()
    upd = W * H * D * 8 * 2
    lat.sort()
    return {"workload": "plain pair 640x480, census 9x7, D=64, P1=10 P2=120, 8 paths, L+R + L-R, "
                        "inputs in HBM; ms_per_pair back to back, latency one call at a time",
            "ms_per_pair": round(t * 1e3, 4), "gupdates_per_s": round(upd / t / 1e9, 3),
            "latency_ms_p50": round(lat[len(lat) // 2] * 1e3, 4),
            "kernel_ms": prof, "kernel_ms_total": round(sum(prof.values()), 4),
            "valid_pixels": valid}


def run_aggregate32(sgm, W, H, D, P, d_left, d_right, steps=3):
    """The benchmark pair with 32-bit aggregates (--gpu-aggregate-bits 32 in the
    reference; also what a P2 beyond the 16-bit bound selects): uint16 cost
    volume, int32 aggregate volume, one direction per launch with the generic
    recurrence; inputs in HBM."""
    params = workload_params(sgm)
    params.aggregate_bytes = 4
    ctx = sgm.Context(0)
    n = W * H
    out = ctx.alloc(4 * n)
    ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    ctx.set_profiling(True)
    ctx.profile()
    ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    prof = {k: round(v["total_ms"], 3) for k, v in ctx.profile()["device_kernel_milliseconds"].items()}
    ctx.set_profiling(False)
    plan = ctx.plan(params, W, H)
    out.free()
    ctx.close()
    upd = W * H * D * P * 2
    return {"workload": "plain pair 1600x1067 D=256 with 32-bit aggregates (uint16 cost volume, "
                        "int32 aggregate volume, generic recurrence)",
            "ms_per_pair": round(t * 1e3, 3), "gupdates_per_s": round(upd / t / 1e9, 3),
            "kernel_ms": prof, "device_bytes": plan["pair_bytes"],
            "aggregate_bytes": plan["aggregate_bytes"]}


def run_esgm(sgm, W, H, D, P, d_left, d_right, rows=64, pool=2, steps=2):
    """The benchmark pair under the compact plan (proposal and sum sweeps per
    path over strips of `rows` rows, certification, exact verification of the
    uncertified pixels): inputs in HBM; the device footprint holds no volume."""
    params = workload_params(sgm)
    ctx = sgm.Context(0)
    ctx.plan_mode(sgm.PLAN_ESGM, rows)
    ctx.esgm_candidates(pool)
    n = W * H
    out = ctx.alloc(4 * n)
    ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.profile()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.match_pair_device(params, d_left, d_right, W, H, out)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    prof = {k: round(v["total_ms"] / steps, 3)
            for k, v in ctx.profile()["device_kernel_milliseconds"].items()}
    ctx.set_profiling(False)
    counts = ctx.esgm_counts()
    plan = ctx.plan(params, W, H)
    out.free()
    ctx.close()
    upd = W * H * D * P * 2
    verified = counts["left"][1] + counts["right"][1]
    return {"workload": "plain pair 1600x1067 D=256 under the compact (esgm) plan, %d proposals "
                        "per path, strips of %d rows" % (plan["esgm_candidates"], plan["esgm_rows"]),
            "ms_per_pair": round(t * 1e3, 3), "gupdates_per_s": round(upd / t / 1e9, 3),
            "kernel_ms": prof, "device_bytes": plan["esgm_bytes"],
            "resident_device_bytes": plan["pair_bytes"],
            "verified_fraction": round(verified / (2.0 * n), 4)}


def run_zncc(sgm, steps=5):
    """The FP32 ZNCC workload: 1600x1067, 7x7 windows, D=192, P1=12, P2=150, gain
    field 0.8-1.2 on the right image; inputs in HBM; uint16 cost volume."""
    import numpy as np
    from pose_amd import synth
    W, H, D = 1600, 1067, 192
    p = sgm.SgmParams(window_w=7, window_h=7, disparities=D, min_disparity=0, cost=4, p1=12,
                      p2=150, edge_scale=32, directions=8, aggregate_bytes=0, census_weight=500,
                      ad_truncation=30, gradient_truncation=60, reserved=0, uniqueness=0.03,
                      lr_threshold=1.0)
    left, right, _ = synth.stereo_pair(W, H, D, 0, seed=3, gain=True)
    ctx = sgm.Context(0)
    n = W * H
    dl, dr, out = ctx.alloc(n).upload(left), ctx.alloc(n).upload(right), ctx.alloc(4 * n)
    ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / steps
    ctx.set_profiling(True)
    ctx.profile()
    ctx.match_pair_device(p, dl, dr, W, H, out)
    ctx.synchronize()
    prof = {k: round(v["total_ms"], 3) for k, v in ctx.profile()["device_kernel_milliseconds"].items()}
    ctx.set_profiling(False)
    valid = int(np.isfinite(out.download(np.float32, n)).sum())
    for b in (dl, dr, out):
        b.free()
    ctx.close()
    upd = W * H * D * 8 * 2
    return {"workload": "ZNCC FP32 pair 1600x1067, 7x7, D=192, P1=12 P2=150, gain field, 8 paths, "
                        "uint16 cost and aggregate volumes, L+R + L-R",
            "ms_per_pair": round(t * 1e3, 3), "gupdates_per_s": round(upd / t / 1e9, 3),
            "kernel_ms": prof, "valid_pixels": valid}


def serve_trace(count, rate, seed=7):
    """Served stream: Poisson arrivals at `rate` pairs/s (rate None: all at t=0);
    sizes 70% from the batched distribution, 25% 1600x1067 D=256, 5% 6000x4000
    D=512.  Requests draw from a pool of pre-generated pairs of each class."""
    import numpy as np
    from pose_amd import synth
    rng = np.random.default_rng(seed)
    mixed = synth.batch_pairs(count=12, seed=seed)
    pool = {"mixed": mixed,
            "1600x1067": [(1600, 1067, 256, 0, seed * 1000 + 100 + i) for i in range(3)],
            "6000x4000": [(6000, 4000, 512, 0, seed * 1000 + 200)]}
    t, trace = 0.0, []
    for i in range(count):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        u = rng.random()
        cls = "mixed" if u < 0.70 else "1600x1067" if u < 0.95 else "6000x4000"
        spec = pool[cls][int(rng.integers(len(pool[cls])))]
        trace.append((t if rate else 0.0, cls, spec))
    return trace


def run_served_sweep(sgm, params_for, rates=(2.0, 5.0, 10.0), budget=250e9, streams=3):
    """The served stream at lower Poisson rates: the plain path (FIFO, one pair
    at a time) and the lane server, p50/p99 latency and sustained pairs/s; about
    three seconds of arrivals per rate (at least 12 requests)."""
    out = {}
    for rate in rates:
        count = max(12, int(rate * 3))
        r = run_served(sgm, params_for, rate=rate, count=count, budget=budget, streams=streams,
                       burst=False, batching=False)
        out["%g" % rate] = r["poisson"]
    return out


def run_served(sgm, params_for, rate=20.0, count=60, budget=250e9, max_batch=8, streams=3,
               burst=True, batching=True):
    """The same request trace through (a) the plain path, one pair at a time in
    arrival order, and (b) the continuous-batching server (admission by HBM
    footprint against `budget`); host buffers in and out for both.  Latency =
    finish - arrival; sustained pairs/s = requests / (last finish - first arrival)."""
    import threading
    import numpy as np
    from pose_amd import synth

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, max(0, int(math.ceil(q * len(v))) - 1))]

    images = {}
    t_gen = time.perf_counter()
    traces = {"poisson": serve_trace(count, rate)}
    if burst:
        traces["saturated"] = serve_trace(count, None, seed=8)
    for tr in traces.values():
        for _, _, (w, h, D, md, s) in tr:
            if s not in images:
                images[s] = synth.stereo_pair(w, h, D, md, seed=s)[:2]
    t_gen = time.perf_counter() - t_gen
    out = {"workload": "served stream: 70%% batched-distribution pairs, 25%% 1600x1067 D=256, "
                       "5%% 6000x4000 D=512; census 9x7, P1=10 P2=120, 8 paths; host buffers "
                       "in/out; Poisson %.0f pairs/s (%d requests) and a saturated burst "
                       "(%d requests at t=0); server budget %.0f GB, %d streams; batching server "
                       "(max batch %d) and lane server (max batch 1, plain kernels per lane)"
                       % (rate, count, count, budget / 1e9, streams, max_batch),
           "input_generation_s": round(t_gen, 2), "distinct_pairs": len(images)}
    for name, trace in traces.items():
        updates = sum(w * h * D * 8 * 2 for _, _, (w, h, D, md, s) in trace)
        classes = {}
        for _, c, _ in trace:
            classes[c] = classes.get(c, 0) + 1
        # (a) plain path, FIFO, one pair at a time
        ctx = sgm.Context(0)
        # warm-up on the largest request: the plain path's buffers are allocated
        # before the clock starts, as the server's pool is
        big = max((spec for _, _, spec in trace), key=lambda t: t[0] * t[1] * t[2])
        ctx.match_pair(params_for(big[2], big[3]), *images[big[4]])
        plain_lat, plain_out, service = [], [], {}
        t0 = time.perf_counter()
        for (ta, c, (w, h, D, md, s)) in trace:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            s0 = time.perf_counter()
            r = ctx.match_pair(params_for(D, md), *images[s])
            t_end = time.perf_counter()
            service.setdefault(c, []).append(t_end - s0)
            plain_lat.append(t_end - t0 - ta)
            plain_out.append(r["checked"])
        plain_span = time.perf_counter() - t0 - trace[0][0]
        ctx.close()
        # (b) servers: a submitter thread replays the arrivals; the batching
        # server (persistent batch kernel) and the lane server (max batch 1:
        # plain kernels per pair on per-lane streams)
        def replay(mb, streams_):
            ctx = sgm.Context(0, budget_bytes=int(budget) + (8 << 30))
            srv = sgm.Server(ctx, params_for(64, 0), budget_bytes=int(budget), max_batch=mb,
                             streams=streams_)
            if mb == 1:
                # lanes allocate on first use: warm every lane on the largest request
                warm = [srv.submit(*images[big[4]], big[2], big[3]) for _ in range(streams_)]
                for t in warm:
                    srv.wait(t, timeout=600.0)
            tickets = [None] * len(trace)

            def submit():
                t1 = time.perf_counter()
                for i, (ta, _, (w, h, D, md, s)) in enumerate(trace):
                    now = time.perf_counter() - t1
                    if ta > now:
                        time.sleep(ta - now)
                    tickets[i] = srv.submit(*images[s], D, md)
            th = threading.Thread(target=submit)
            th.start()
            lat, first, last, same, batches = [], None, 0.0, True, set()
            for i in range(len(trace)):
                while tickets[i] is None:
                    time.sleep(0.0005)
                got = srv.wait(tickets[i], timeout=600.0)
                info = got["info"]
                lat.append(info["finished_s"] - info["arrival_s"])
                first = info["arrival_s"] if first is None else min(first, info["arrival_s"])
                last = max(last, info["finished_s"])
                batches.add(info["batch_id"])
                same &= bool(np.array_equal(got["checked"].view(np.uint32),
                                            plain_out[i].view(np.uint32)))
            th.join()
            srv.close()
            ctx.close()
            span = last - first
            return span, {"p50_ms": round(pct(lat, 0.5) * 1e3, 2),
                          "p99_ms": round(pct(lat, 0.99) * 1e3, 2),
                          "pairs_per_s": round(len(trace) / span, 3),
                          "gupdates_per_s": round(updates / span / 1e9, 3),
                          "batches": len(batches), "bit_exact_vs_plain": same}
        span, served = replay(max_batch, streams) if batching else (None, None)
        span_l, lanes = replay(1, streams)
        out[name] = {
            "requests": len(trace), "classes": classes, "cell_updates": updates,
            "served": served,
            "served_lanes": lanes,
            "plain": {"p50_ms": round(pct(plain_lat, 0.5) * 1e3, 2),
                      "p99_ms": round(pct(plain_lat, 0.99) * 1e3, 2),
                      "pairs_per_s": round(len(trace) / plain_span, 3),
                      "gupdates_per_s": round(updates / plain_span / 1e9, 3),
                      "mean_service_ms": {c: round(sum(v) / len(v) * 1e3, 2)
                                          for c, v in service.items()},
                      "max_service_ms": {c: round(max(v) * 1e3, 2) for c, v in service.items()}},
            "speedup_vs_plain": round(plain_span / span, 3) if span else None,
            "lanes_speedup_vs_plain": round(plain_span / span_l, 3),
            "bit_exact_vs_plain": ((served is None or served["bit_exact_vs_plain"])
                                   and lanes["bit_exact_vs_plain"]),
        }
        print("served leg %s done" % name, file=sys.stderr, flush=True)
    return out


def read_counter_csv(directory, counter):
    """Per-dispatch values of one PMC counter from a rocprofv3 --pmc CSV run:
    [(kernel name, value)] in dispatch order."""
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r.get("Dispatch_Id", 0)), r["Kernel_Name"],
                                 float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def write_traffic(fetch_dir, write_dir, copy_bytes=2 << 30, paths=8):
    """profiles/path_traffic.json from two rocprofv3 passes over
    `bench.py --steps 1 --warmup 1 --profile-steps 1 --no-...` (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch).  Each counter is scaled by what it
    reads for the library's copy kernel (copy_bytes read + copy_bytes written
    per launch); path launches are grouped by kind, and the per-launch figure
    weights them as one pair does under the schedule the passes ran (path-cost
    planes: the census cost pass, the P - 1 directions of the planes, the last
    sweep, one launch each; one aggregate: one first, P - 2 middle, one last)."""
    W, H, D = 1600, 1067, 256
    fetch, write = read_counter_csv(fetch_dir, "FETCH_SIZE"), read_counter_csv(write_dir, "WRITE_SIZE")

    def mean_of(rows, pred):
        v = [x for k, x in rows if pred(k)]
        if not v:
            raise SystemExit("no dispatches matched in the counter CSV")
        return sum(v) / len(v)

    is_copy = lambda k: "copy_kernel" in k
    fetch_scale = copy_bytes / 1024.0 / mean_of(fetch, is_copy)
    write_scale = copy_bytes / 1024.0 / mean_of(write, is_copy)
    cells = 2 * W * H * D
    # the workload's instances only (lane group 4, uint16 aggregates): other
    # legs of the same run launch other lane groups and 32-bit aggregates
    sweep = lambda k, phase: "path_kernel<4, unsigned short, %d, " % phase in k
    planes = any("path_multi_kernel<4, " in k for k, _ in fetch)
    if planes:
        kinds = (("cost", lambda k: sweep(k, 3), cells, 1),
                 ("multi", lambda k: "path_multi_kernel<4, " in k, 2 * (paths - 1) * cells, 1),
                 ("last", lambda k: sweep(k, 2), paths * cells, 1))
    else:
        kinds = (("first", lambda k: sweep(k, 0), 3 * cells, 1),
                 ("mid", lambda k: sweep(k, 1), 5 * cells, paths - 2),
                 ("last", lambda k: sweep(k, 2), 3 * cells, 1))
    out = {"workload": "1600x1067, census 9x7, D=256, 8 paths, both matching directions per launch",
           "sweeps": "planes" if planes else "sequential",
           "source_hash": kernel_source_hash(), "fetch_scale": round(fetch_scale, 4),
           "write_scale": round(write_scale, 4),
           "calibration": "copy_kernel: %d B read + %d B written per launch; FETCH_SIZE read %.0f "
                          "KiB, WRITE_SIZE %.0f KiB" % (copy_bytes, copy_bytes,
                                                        mean_of(fetch, is_copy),
                                                        mean_of(write, is_copy)),
           "per_launch": {}}
    total_hbm = total_build = launches = 0
    for name, pred, build, count in kinds:
        names = sorted({k for k, _ in fetch if pred(k)})
        rd = mean_of(fetch, pred) * 1024 * fetch_scale
        wr = mean_of(write, pred) * 1024 * write_scale
        out["per_launch"][name] = {"kernel": names, "read_bytes": int(rd), "write_bytes": int(wr),
                                   "hbm_bytes": int(rd + wr), "build_bytes": build,
                                   "ratio": round((rd + wr) / build, 4)}
        total_hbm += (rd + wr) * count
        total_build += build * count
        launches += count
    out["bytes_per_launch"] = int(total_hbm / launches)
    out["build_bytes_per_launch"] = int(total_build / launches)
    with open(os.path.join(ROOT, "profiles", "path_traffic.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--write-traffic":
        write_traffic(sys.argv[2], sys.argv[3])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--no-served", action="store_true", help="skip the served-vs-plain leg")
    ap.add_argument("--no-zncc", action="store_true", help="skip the FP32 ZNCC leg")
    ap.add_argument("--no-streamed", action="store_true", help="skip the strip-streamed leg")
    ap.add_argument("--no-esgm", action="store_true", help="skip the compact (esgm) leg")
    ap.add_argument("--no-filtered", action="store_true", help="skip the post-L-R filter leg")
    ap.add_argument("--no-agg32", action="store_true", help="skip the 32-bit aggregate leg")
    ap.add_argument("--no-small", action="store_true", help="skip the 640x480 D=64 pair leg")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrivals, pairs/s")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-streams", type=int, default=3, help="server streams / lanes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this path runs on a single MI355X; "
              "sharding pairs across GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from pose_amd import sgm, synth

    W, H, D, P = 1600, 1067, 256, 8
    params = workload_params(sgm)
    left, right, _ = synth.stereo_pair(W, H, D, 0, seed=2)
    n = W * H
    dev = sgm.device_info(0)
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
    schedule = ctx.plan(params, W, H)["sweeps"]
    # every path-sweep launch of the pair (first, middle / paired / multi-direction, last)
    path_keys = [k for k in prof if k.startswith("path")]
    path_ms = sum(prof[k]["total_ms"] for k in path_keys) / args.profile_steps
    path_launches = sum(prof[k]["launches"] for k in path_keys) / args.profile_steps
    launch_s = path_ms / 1e3 / path_launches
    alg_launch = algorithmic_bytes_per_step(W, H, D, P) / path_launches
    build_launch = path_bytes_per_step(W, H, D, P, schedule) / path_launches
    achieved = alg_launch / launch_s / 1e9
    kernels = {k: round(v["total_ms"] / args.profile_steps, 4) for k, v in prof.items()}
    # each path launch of the schedule on its own: duration and the build's bytes
    cells = 2 * W * H * D
    own_bytes = ({"path_cost": 1, "path_multi": 2 * (P - 1), "path_last": 1 + (P - 1)}
                 if schedule == 3 else {"path_first": 3, "path_mid": 5, "path_pair": 8, "path_last": 3})
    path_kernels = {}
    for k in path_keys:
        ms = prof[k]["total_ms"] / prof[k]["launches"]
        b = own_bytes.get(k, 0) * cells
        path_kernels[k] = {"launches_per_step": prof[k]["launches"] / args.profile_steps,
                           "ms": round(ms, 4), "build_bytes": b,
                           "build_achieved": round(b / (ms / 1e3) / 1e9, 1),
                           "build_frac": round(b / (ms / 1e3) / 1e9 / PEAK_HBM_GBPS, 4)}

    # Achievable HBM bandwidth: the library's plain streaming copy kernel over
    # 2 GiB (reads 2 GiB, writes 2 GiB); the same launch calibrates the HBM
    # traffic counters in the rocprof runs.
    copy_bytes = 2 << 30
    c_src, c_dst = ctx.alloc(copy_bytes), ctx.alloc(copy_bytes)
    ctx.device_copy(c_dst, c_src, copy_bytes)
    ctx.synchronize()
    ctx.set_profiling(True)
    ctx.profile()
    for _ in range(3):
        ctx.device_copy(c_dst, c_src, copy_bytes)
    ctx.synchronize()
    copy_ms = ctx.profile()["device_kernel_milliseconds"]["copy"]["total_ms"] / 3
    ctx.set_profiling(False)
    c_src.free()
    c_dst.free()
    hbm_copy = {"bytes_per_launch": 2 * copy_bytes, "ms": round(copy_ms, 4),
                "gbps": round(2 * copy_bytes / (copy_ms / 1e3) / 1e9, 1)}

    # HBM bytes per launch from the calibrated FETCH_SIZE / WRITE_SIZE passes,
    # valid only for the kernel sources they were measured on
    traffic, traffic_hash, source_hash = None, None, kernel_source_hash()
    tfile = os.path.join(ROOT, "profiles", "path_traffic.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        traffic_hash = tj.get("source_hash")
        if traffic_hash == source_hash:
            traffic = tj.get("bytes_per_launch")

    streamed = None if args.no_streamed else run_streamed(sgm, W, H, D, P, d_left, d_right)
    esgm = None if args.no_esgm else run_esgm(sgm, W, H, D, P, d_left, d_right)
    agg32 = None if args.no_agg32 else run_aggregate32(sgm, W, H, D, P, d_left, d_right)
    small_pair = None if args.no_small else run_small_pair(sgm)
    for b in (d_left, d_right, d_out, d_stats):
        b.free()
    ctx.close()

    def params_for(Dv, md):
        return sgm.SgmParams(window_w=9, window_h=7, disparities=Dv, min_disparity=md, cost=0,
                             p1=10, p2=120, edge_scale=32, directions=8, aggregate_bytes=0,
                             census_weight=500, ad_truncation=30, gradient_truncation=60,
                             reserved=0, uniqueness=0.03, lr_threshold=1.0)
    zncc = None if args.no_zncc else run_zncc(sgm)
    filtered = None if args.no_filtered else run_filtered(sgm, left, right, W, H)
    batched = None if args.no_batched else run_batched(sgm, params_for)
    batched_small = None
    if not args.no_batched:
        # many small pairs: the per-launch overhead the persistent launch removes
        batched_small = run_batched(sgm, params_for, steps=5,
                                    spec=[(640, 480, 64, 0, 9000 + i) for i in range(48)],
                                    label="48 pairs 640x480 D=64, census 9x7, P1=10 P2=120, "
                                          "8 paths, L+R + L-R")
    served_sweep = None if args.no_served else run_served_sweep(sgm, params_for,
                                                                streams=args.serve_streams)
    served = None if args.no_served else run_served(sgm, params_for, rate=args.serve_rate,
                                                    count=args.serve_requests,
                                                    streams=args.serve_streams)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, H, D)
        cpu["other_workloads"] = cpu_baseline_configs()

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
                               "8 paths, left+right + L-R check; uint8 cost volume, uint8 "
                               "path-cost planes summed in packed uint16 by the last sweep",
                   "width": W, "height": H, "disparities": D, "paths": P, "pairs_per_step": 1},
        "device": {k: dev[k] for k in ("name", "arch", "memory", "compute_units")},
        "pairs_per_s": round(1e3 / ms_per_step, 3),
        "latency_ms": {"p50": round(lat_ms[len(lat_ms) // 2], 4),
                       "p99": round(lat_ms[min(len(lat_ms) - 1, int(math.ceil(0.99 * len(lat_ms))) - 1)], 4)},
        "valid_pixels": valid,
        "kernel_ms_per_step": kernels,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": traffic,
                     "kernel": "path sweeps of one pair (sgm::path_kernel first and last, "
                               "sgm::path_multi_kernel for the middle directions)",
                     "sweeps": ["auto", "sequential", "paired", "planes"][schedule],
                     "launches_per_step": path_launches,
                     "launch_ms": round(launch_s * 1e3, 4),
                     "algorithmic_bytes_per_launch": alg_launch,
                     "algorithmic_model": "6P-2 B per cell (uint16 C and S, selection fused "
                                          "into the last sweep), 2*W*H*D cells per pair; the "
                                          "plane schedule moves 22 B per cell, so the model's "
                                          "rate can pass the HBM peak (build_frac is the "
                                          "fraction of peak used)",
                     "path_kernels": path_kernels,
                     "build_bytes_per_launch": build_launch,
                     "build_achieved": round(build_launch / launch_s / 1e9, 1),
                     "build_frac": round(build_launch / launch_s / 1e9 / PEAK_HBM_GBPS, 4),
                     "traffic_achieved": None if traffic is None
                     else round(traffic / launch_s / 1e9, 1),
                     "traffic_frac": None if traffic is None
                     else round(traffic / launch_s / 1e9 / PEAK_HBM_GBPS, 4),
                     "traffic_source_hash": traffic_hash, "source_hash": source_hash,
                     "traffic_stale": traffic is None,
                     "copy_kernel": hbm_copy},
        "cpu_baseline": cpu,
        "filtered": filtered,
        "streamed": streamed,
        "esgm": esgm,
        "aggregate32": agg32,
        "small_pair": small_pair,
        "zncc": zncc,
        "batched": batched,
        "batched_small": batched_small,
        "served": served,
        "served_rate_sweep": served_sweep,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
