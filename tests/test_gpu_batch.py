"""GPU tests of the batched (persistent work-queue kernel, hipGraph) and served
(continuous batching) modes: every pair must be bit-exact to the oracle and to
the plain one-pair path, alone or in the mix."""
import json
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth

pytestmark = pytest.mark.gpu


def to_sgm(op):
    return sgm.SgmParams(**{f: getattr(op, f) for f, _ in sgm.SgmParams._fields_})


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def same(got, want, what):
    for k in ("checked", "dl", "dr"):
        assert np.array_equal(bits(got[k]), bits(want[k])), (what, k)
    for k in ("reasons", "boundary"):
        if k in got and k in want:
            assert np.array_equal(got[k], want[k]), (what, k)
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], (what, k, got["stats"][k], want["stats"][k])


def uneven_pairs(seed, count, wrange=(96, 420), drange=(3, 500)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w = int(rng.integers(*wrange))
        h = int(rng.integers(40, 160))
        D = int(rng.integers(drange[0], min(drange[1], w - 20)))
        md = int(rng.integers(-8, 9))
        left, right, _ = synth.stereo_pair(w, h, D, md, seed=seed * 100 + i)
        pr = dict(left=left, right=right, disparities=D, min_disparity=md)
        if i % 3 == 1:
            v = (rng.random((h, w)) > 0.05).astype(np.uint8)
            pr["valid_left"] = v
            pr["valid_right"] = np.roll(v, 3, axis=1)
        out.append(pr)
    return out


def oracle_of(base_kw, pr):
    op = O.params(**dict(base_kw, disparities=pr["disparities"], min_disparity=pr["min_disparity"]))
    h, w = pr["left"].shape
    return O.compute(op, pr["left"], pr["right"], w, h, pr.get("valid_left"), pr.get("valid_right"))


CENSUS97 = dict(window_w=9, window_h=7, p1=10, p2=120, lr_threshold=1.0)


@pytest.mark.parametrize("policy", [sgm.VOLUME_RESIDENT, sgm.VOLUME_RECOMPUTE],
                         ids=["resident", "recompute"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_census_bit_exact(gpu_context, graph, policy):
    pairs = uneven_pairs(1, 7)
    print("batch pairs", [(p["left"].shape, p["disparities"], p["min_disparity"]) for p in pairs],
          flush=True)
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    gpu_context.cost_volume(policy)
    try:
        b = sgm.Batch(gpu_context, params, pairs, graph=graph)
    finally:
        gpu_context.cost_volume(sgm.VOLUME_AUTO)
    try:
        for rep in range(2 if graph else 1):
            res = b.run()
            for i, pr in enumerate(pairs):
                same(res[i], oracle_of(CENSUS97, pr), "pair %d rep %d" % (i, rep))
    finally:
        b.close()


@pytest.mark.parametrize("kw", [
    dict(window_w=5, window_h=5, cost=1, p1=4, p2=24),
    dict(window_w=11, window_h=11, cost=0, p1=20, p2=120, directions=4),
    dict(window_w=7, window_h=7, cost=3, p1=12, p2=100),
    dict(window_w=5, window_h=5, cost=2, p1=40, p2=9000),          # 32-bit aggregates
], ids=["ad-census", "census11x11-4paths", "zncc", "census-gradient-agg32"])
def test_batch_other_costs_bit_exact(gpu_context, kw):
    pairs = uneven_pairs(2, 4, wrange=(96, 260), drange=(3, 200))
    params = to_sgm(O.params(**dict(kw, disparities=64)))
    b = sgm.Batch(gpu_context, params, pairs)
    try:
        res = b.run()
        for i, pr in enumerate(pairs):
            same(res[i], oracle_of(kw, pr), "pair %d" % i)
    finally:
        b.close()


def test_batch_zncc_fp32_equals_plain(gpu_context):
    """FP32 ZNCC (uint16 cost volume, 16-bit aggregates) in a batch: each pair equals
    the plain GPU path bit for bit (same FP32 arithmetic)."""
    kw = dict(window_w=7, window_h=7, cost=4, p1=12, p2=150, lr_threshold=1.0)
    pairs = uneven_pairs(6, 4, wrange=(96, 260), drange=(3, 200))
    params = to_sgm(O.params(**dict(kw, disparities=64)))
    b = sgm.Batch(gpu_context, params, pairs)
    try:
        assert b.describe()["cost_volume"] == "resident-u16"
        res = b.run()
        for i, pr in enumerate(pairs):
            p = to_sgm(O.params(**dict(kw, disparities=pr["disparities"],
                                       min_disparity=pr["min_disparity"])))
            want = gpu_context.match_pair(p, pr["left"], pr["right"], pr.get("valid_left"),
                                          pr.get("valid_right"))
            same(res[i], want, "pair %d" % i)
    finally:
        b.close()


def test_batch_graph_replays_match_eager(gpu_context):
    """Graph replays with eager work between them all equal the eager result."""
    pairs = uneven_pairs(3, 5)
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    eager = sgm.Batch(gpu_context, params, pairs)
    graph = sgm.Batch(gpu_context, params, pairs, graph=True)
    try:
        want = [dict((k, np.copy(v)) if k != "stats" else (k, dict(v)) for k, v in r.items())
                for r in eager.run()]
        for rep in range(3):
            got = graph.run()
            for i in range(len(pairs)):
                same(got[i], want[i], "replay %d pair %d" % (rep, i))
            eager.run()  # eager work between replays
    finally:
        eager.close()
        graph.close()


def test_batch_device_buffers(gpu_context):
    ctx = gpu_context
    pairs = uneven_pairs(4, 4)
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    bufs, recs = [], []
    for pr in pairs:
        h, w = pr["left"].shape
        n = w * h
        b = dict(left=ctx.alloc(n).upload(pr["left"]), right=ctx.alloc(n).upload(pr["right"]),
                 checked=ctx.alloc(4 * n), dl=ctx.alloc(4 * n), dr=ctx.alloc(4 * n),
                 stats=ctx.alloc(128))
        if "valid_left" in pr:
            b["vl"] = ctx.alloc(n).upload(pr["valid_left"])
            b["vr"] = ctx.alloc(n).upload(pr["valid_right"])
        bufs.append(b)
        recs.append(sgm.SgmPair(width=w, height=h, disparities=pr["disparities"],
                                min_disparity=pr["min_disparity"], left=b["left"].ptr,
                                right=b["right"].ptr,
                                valid_left=b["vl"].ptr if "vl" in b else None,
                                valid_right=b["vr"].ptr if "vr" in b else None,
                                checked=b["checked"].ptr, dl=b["dl"].ptr,
                                dr=b["dr"].ptr, stats=b["stats"].ptr))
    db = sgm.DeviceBatch(ctx, params, recs, graph=True)
    try:
        for rep in range(2):
            db.run()
            for i, pr in enumerate(pairs):
                h, w = pr["left"].shape
                want = oracle_of(CENSUS97, pr)
                n = w * h
                for k in ("checked", "dl", "dr"):
                    assert np.array_equal(bits(bufs[i][k].download(np.float32, n)),
                                          bits(want[k])), (rep, i, k)
                st = bufs[i]["stats"].download(np.int64, 16)
                assert st[5] == want["stats"]["selected"] and st[10] == want["stats"]["valid"]
    finally:
        db.close()
        for b in bufs:
            for v in b.values():
                v.free()


@pytest.mark.parametrize("policy", [sgm.VOLUME_AUTO, sgm.VOLUME_RECOMPUTE],
                         ids=["auto", "recompute"])
def test_server_results_equal_alone(gpu_context, policy):
    """Requests admitted while others run: each result equals its solo plain run."""
    ctx = gpu_context
    pairs = uneven_pairs(5, 9, wrange=(120, 600), drange=(16, 300))
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    alone = []
    for pr in pairs:
        p = to_sgm(O.params(**dict(CENSUS97, disparities=pr["disparities"],
                                   min_disparity=pr["min_disparity"])))
        alone.append(ctx.match_pair(p, pr["left"], pr["right"]))
    ctx.cost_volume(policy)
    srv = sgm.Server(ctx, params, budget_bytes=2 << 30, max_batch=4, streams=2)
    try:
        tickets = []
        for i, pr in enumerate(pairs):
            tickets.append(srv.submit(pr["left"], pr["right"], pr["disparities"], pr["min_disparity"]))
        batches = set()
        for i, t in enumerate(tickets):
            got = srv.wait(t, timeout=60.0)
            batches.add(got["info"]["batch_id"])
            for k in ("checked", "dl", "dr"):
                assert np.array_equal(bits(got[k]), bits(alone[i][k])), (i, k)
            for k in ("selected", "boundary_rejected", "ambiguity_rejected", "lr_rejected", "valid"):
                assert got["stats"][k] == alone[i]["stats"][k], (i, k)
            info = got["info"]
            assert info["arrival_s"] <= info["admitted_s"] <= info["started_s"] <= info["finished_s"]
        assert len(batches) >= 2  # more than one batch was formed while others ran
    finally:
        srv.close()
        ctx.cost_volume(sgm.VOLUME_AUTO)


def test_server_large_request_batched_alone(gpu_context):
    """A request above an eighth of the serving pool runs in a batch of its own;
    the small requests around it still share batches; all equal their solo runs."""
    ctx = gpu_context
    small = uneven_pairs(7, 6, wrange=(96, 200), drange=(16, 64))
    big_l, big_r, _ = synth.stereo_pair(900, 700, 256, 0, seed=77)
    pairs = small[:3] + [dict(left=big_l, right=big_r, disparities=256, min_disparity=0)] + small[3:]
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    srv = sgm.Server(ctx, params, budget_bytes=2 << 30, max_batch=4, streams=2)
    try:
        tickets = [srv.submit(p["left"], p["right"], p["disparities"], p["min_disparity"])
                   for p in pairs]
        got = [srv.wait(t, timeout=60.0) for t in tickets]
    finally:
        srv.close()
    ids = [g["info"]["batch_id"] for g in got]
    assert ids.count(ids[3]) == 1, ids
    assert len(set(ids[:3] + ids[4:])) < 6, ids  # small requests were batched together
    for p, g in zip(pairs, got):
        op = to_sgm(O.params(**dict(CENSUS97, disparities=p["disparities"],
                                    min_disparity=p["min_disparity"])))
        want = ctx.match_pair(op, p["left"], p["right"])
        for k in ("checked", "dl", "dr"):
            assert np.array_equal(bits(g[k]), bits(want[k])), k


def test_server_rejects_what_never_fits(gpu_context):
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    srv = sgm.Server(gpu_context, params, budget_bytes=64 << 20, max_batch=2, streams=1)
    try:
        left = np.zeros((600, 800), np.uint8)
        with pytest.raises(sgm.SgmError) as e:
            srv.submit(left, left, 256)
        assert e.value.code == -6
    finally:
        srv.close()


def test_autotune_writes_and_loads(gpu_context):
    params = to_sgm(O.params(**dict(CENSUS97, disparities=128)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tune.json")
        ctx = sgm.Context(0)
        try:
            ctx.autotune(params, path)
            t = json.load(open(path))
            assert t["lane_group"] == 2 and t["path_waves"] in (1, 2, 4)
            assert t["batch_lines_per_item"] in (1, 2, 4)
            ctx.autotune_load(path)
            left, right, _ = synth.stereo_pair(200, 100, 128, 0, seed=3)
            r = ctx.match_pair(params, left, right)
            want = O.compute(O.params(**dict(CENSUS97, disparities=128)), left, right, 200, 100)
            assert np.array_equal(bits(r["checked"]), bits(want["checked"]))
        finally:
            ctx.close()


def test_server_lanes_results_equal_alone(gpu_context):
    """max_batch 1: each slot is a lane running the plain kernels on its own
    stream; requests overlap across lanes and each result equals its solo run."""
    ctx = gpu_context
    pairs = uneven_pairs(8, 21, wrange=(120, 600), drange=(16, 300))
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    alone = []
    for pr in pairs:
        p = to_sgm(O.params(**dict(CENSUS97, disparities=pr["disparities"],
                                   min_disparity=pr["min_disparity"])))
        alone.append(ctx.match_pair(p, pr["left"], pr["right"]))
    srv = sgm.Server(ctx, params, budget_bytes=2 << 30, max_batch=1, streams=3)
    try:
        tickets = [srv.submit(pr["left"], pr["right"], pr["disparities"], pr["min_disparity"])
                   for pr in pairs]
        spans = []
        for i, t in enumerate(tickets):
            got = srv.wait(t, timeout=60.0)
            for k in ("checked", "dl", "dr"):
                assert np.array_equal(bits(got[k]), bits(alone[i][k])), (i, k)
            for k in ("selected", "boundary_rejected", "ambiguity_rejected", "lr_rejected", "valid",
                      "lower_boundary_rejected", "full_interval_pixels"):
                assert got["stats"][k] == alone[i]["stats"][k], (i, k)
            info = got["info"]
            assert info["arrival_s"] <= info["admitted_s"] <= info["started_s"] <= info["finished_s"]
            spans.append((info["started_s"], info["finished_s"]))
        # some requests ran at the same time on different lanes
        assert any(a[0] < b[1] and b[0] < a[1] for i, a in enumerate(spans) for b in spans[i + 1:])
    finally:
        srv.close()


def test_server_lanes_large_request_after_small_ones(gpu_context):
    """Lanes hold their buffers between requests; a large request behind small
    ones on a tight budget (idle lanes give their buffers back when the sum
    would not fit) is admitted and served exactly."""
    ctx = gpu_context
    small = uneven_pairs(4, 31, wrange=(300, 500), drange=(100, 200))
    big_l, big_r, _ = synth.stereo_pair(900, 700, 256, 0, seed=78)
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    pbig = to_sgm(O.params(**dict(CENSUS97, disparities=256)))
    want = ctx.match_pair(pbig, big_l, big_r)
    budget = int(ctx.plan(pbig, 900, 700)["pair_bytes"] * 1.2)
    srv = sgm.Server(ctx, params, budget_bytes=budget, max_batch=1, streams=2)
    try:
        ts = [srv.submit(pr["left"], pr["right"], pr["disparities"], pr["min_disparity"])
              for pr in small]
        for t in ts:
            srv.wait(t, timeout=60.0)
        got = srv.wait(srv.submit(big_l, big_r, 256, 0), timeout=60.0)
        assert np.array_equal(bits(got["checked"]), bits(want["checked"]))
    finally:
        srv.close()


def test_server_lanes_stay_within_budget_mixed_sizes(gpu_context):
    """Lanes admit on the footprint the plain path will allocate, capped at
    what the other lanes leave: with a budget below two resident 900x700 D=256
    pairs, mixed sizes never make the lanes together hold more than the budget
    (a lane recomputes census costs or waits), and every result equals its solo
    run."""
    ctx = gpu_context
    small = uneven_pairs(9, 6, wrange=(300, 500), drange=(100, 200))
    bigs = [synth.stereo_pair(900, 700, 256, 0, seed=90 + i)[:2] for i in range(3)]
    pairs = []
    for i, pr in enumerate(small):
        pairs.append(pr)
        if i % 2 == 0:
            pairs.append(dict(left=bigs[i // 2][0], right=bigs[i // 2][1], disparities=256,
                              min_disparity=0))
    pbig = to_sgm(O.params(**dict(CENSUS97, disparities=256)))
    plan = ctx.plan(pbig, 900, 700)
    budget = int(plan["pair_bytes"] * 1.6)
    alone = []
    for pr in pairs:
        p = to_sgm(O.params(**dict(CENSUS97, disparities=pr["disparities"],
                                   min_disparity=pr["min_disparity"])))
        alone.append(ctx.match_pair(p, pr["left"], pr["right"]))
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    srv = sgm.Server(ctx, params, budget_bytes=budget, max_batch=1, streams=3)
    try:
        tickets = [srv.submit(pr["left"], pr["right"], pr["disparities"], pr["min_disparity"])
                   for pr in pairs]
        peak = 0
        for i, t in enumerate(tickets):
            got = srv.wait(t, timeout=120.0)
            peak = max(peak, got["info"]["pool_bytes"])
            for k in ("checked", "dl", "dr"):
                assert np.array_equal(bits(got[k]), bits(alone[i][k])), (i, k)
        assert 0 < peak <= budget, (peak, budget)
    finally:
        srv.close()


@pytest.mark.parametrize("policy", [sgm.VOLUME_RESIDENT, sgm.VOLUME_RECOMPUTE],
                         ids=["resident", "recompute"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_phased_bit_exact(gpu_context, graph, policy):
    """Phase-major run (queue kernel for census and L-R, one launch of the
    specialised sweep kernel per direction and lane group): every pair equals
    the oracle, over lane groups 2, 4 and 8, on graph replays too."""
    pairs = uneven_pairs(11, 7, wrange=(96, 420), drange=(3, 500))
    # disparities per lane of each pair's sweeps (16-bit aggregates: at least 2)
    groups = {max(2, 1 << max(0, (p["disparities"] - 1) // 64).bit_length()) for p in pairs}
    assert len(groups) >= 2, groups
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    gpu_context.cost_volume(policy)
    try:
        b = sgm.Batch(gpu_context, params, pairs, graph=graph, phased=True)
    finally:
        gpu_context.cost_volume(sgm.VOLUME_AUTO)
    try:
        d = b.describe()
        assert d["phased"] and d["lane_groups"] >= 2
        for rep in range(2 if graph else 1):
            res = b.run()
            for i, pr in enumerate(pairs):
                same(res[i], oracle_of(CENSUS97, pr), "phased pair %d rep %d" % (i, rep))
    finally:
        b.close()


def plane_scratch_bytes(pr):
    """Scratch of one pair under path-cost planes: uint8 cost volume and the 7
    plane rows per pixel, both matching directions, 256-byte aligned."""
    h, w = pr["left"].shape
    D = pr["disparities"]
    Dp = -(-D // 64) * 64  # padded disparities: whole 64-byte rows
    a = lambda v: -(-v // 256) * 256
    return 2 * (a(w * h * Dp) + a(w * h * Dp * 7))


def test_batch_planes_in_sub_batches():
    """Phase-major batches run the path-cost plane schedule; when the planes of
    every pair do not fit the budget together, the pairs run in sub-batches that
    reuse one scratch region.  Both equal the oracle; the sequential schedule
    (the context's sweep setting) too."""
    pairs = uneven_pairs(13, 6, wrange=(200, 420), drange=(64, 300))
    params = to_sgm(O.params(**dict(CENSUS97, disparities=64)))
    want = [oracle_of(CENSUS97, pr) for pr in pairs]
    ctx = sgm.Context(0)
    try:
        b = sgm.Batch(ctx, params, pairs, phased=True)
        d = b.describe()
        full = d["device_bytes"]
        assert d["sweeps"] == "planes" and d["sub_batches"] == 1, d
        for i, r in enumerate(b.run()):
            same(r, want[i], "planes pair %d" % i)
        b.close()
        ctx.sweeps("sequential")
        b = sgm.Batch(ctx, params, pairs, phased=True)
        assert b.describe()["sweeps"] == "sequential"
        for i, r in enumerate(b.run()):
            same(r, want[i], "sequential pair %d" % i)
        b.close()
    finally:
        ctx.close()
    sizes = [plane_scratch_bytes(pr) for pr in pairs]
    total = sum(sizes)
    assert max(sizes) < total // 2
    budget = full - total // 2 + (8 << 20)
    ctx = sgm.Context(0, budget_bytes=budget)
    try:
        b = sgm.Batch(ctx, params, pairs, phased=True, graph=True)
        d = b.describe()
        print("sub-batches", d, flush=True)
        assert d["sweeps"] == "planes" and d["sub_batches"] >= 2, d
        assert d["device_bytes"] < full
        for rep in range(2):
            for i, r in enumerate(b.run()):
                same(r, want[i], "sub-batched pair %d rep %d" % (i, rep))
        b.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("kw", [
    dict(window_w=7, window_h=7, cost=3, p1=12, p2=100),
    dict(window_w=5, window_h=5, cost=2, p1=40, p2=9000),          # 32-bit aggregates
    dict(window_w=11, window_h=11, cost=0, p1=20, p2=120, directions=4),
], ids=["zncc", "census-gradient-agg32", "census11x11-4paths"])
def test_batch_phased_other_costs_bit_exact(gpu_context, kw):
    pairs = uneven_pairs(12, 4, wrange=(96, 260), drange=(3, 200))
    params = to_sgm(O.params(**dict(kw, disparities=64)))
    b = sgm.Batch(gpu_context, params, pairs, phased=True)
    try:
        res = b.run()
        for i, pr in enumerate(pairs):
            same(res[i], oracle_of(kw, pr), "phased pair %d" % i)
    finally:
        b.close()
