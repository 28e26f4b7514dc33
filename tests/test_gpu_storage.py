"""Streamed plan storages (--gpu-aggregate ram | disk | recompute), the
allocation-failure fallback, the automatic compact choice and interruption.

The storage matrix is the reference's ScalingRegression (ScalingRegression.java:
42-109): 160x47 random bytes shifted by 11 with a masked box, D = 33 from -5,
four costs x {4, 8} paths x {16, 32}-bit aggregates x three storages, strips of
7 rows, and a host budget small enough that the bounded recomputation runs in
several blocks; every result equals the oracle bit for bit, the host memory
stays within the budget and no temporary aggregate file remains.
"""
import os
import tempfile
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from oracle.java_random import random_bytes
from pose_amd import sgm, synth

pytestmark = pytest.mark.gpu


def to_sgm(op):
    return sgm.SgmParams(**{f: getattr(op, f) for f, _ in sgm.SgmParams._fields_})


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def scaling_fixture():
    """ScalingRegression.java:42-50."""
    w, h = 160, 47
    a = random_bytes(219, w * h).reshape(h, w)
    b = a[:, np.minimum(np.arange(w) + 11, w - 1)].copy()
    ma = np.full((h, w), 255, np.uint8)
    mb = np.full((h, w), 255, np.uint8)
    ma[10:20, 23:32] = 0
    return a, b, ma, mb


def same(got, want, what):
    for k in ("checked", "dl", "dr"):
        assert np.array_equal(bits(got[k]), bits(want[k])), (what, k)
    assert np.array_equal(got["reasons"], want["reasons"]), what
    assert np.array_equal(got["boundary"], want["boundary"]), what
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], (what, k)


def last_plan(ctx):
    ctx.set_profiling(True)
    info = ctx.profile()["last_plan"]
    ctx.set_profiling(False)
    return info


@pytest.mark.parametrize("cost", [0, 1, 2, 3])
@pytest.mark.parametrize("directions", [4, 8])
@pytest.mark.parametrize("aggregate", [2, 4])
def test_scaling_regression_storage_matrix(cost, directions, aggregate):
    a, b, ma, mb = scaling_fixture()
    h, w = a.shape
    op = O.params(disparities=33, min_disparity=-5, cost=cost, directions=directions,
                  aggregate_bytes=aggregate)
    want = O.compute(op, a, b, w, h, ma, mb)
    p = to_sgm(op)
    for storage in ("ram", "disk", "recompute"):
        with tempfile.TemporaryDirectory() as tmp:
            ctx = sgm.Context(0)
            try:
                ctx.plan_mode(sgm.PLAN_STREAMED, 7)
                # recompute: host memory for three 7-row strips of both directions
                stripe = 2 * w * 7 * ctx.plan(p, w, h)["padded_disparities"] * aggregate
                budget = 3 * stripe if storage == "recompute" else 128 << 20
                ctx.aggregate_storage(storage, tmp, budget)
                plan = ctx.plan(p, w, h)
                assert plan["streamed"] and plan["strip_rows"] == 7
                assert plan["aggregate_storage"] == sgm.AGGREGATE_STORAGES[storage]
                assert plan["host_bytes"] <= budget
                if storage == "recompute":
                    assert plan["block_rows"] < h  # several blocks
                if storage == "disk":
                    assert plan["disk_bytes"] > 0
                got = ctx.match_pair(p, a, b, ma, mb)
                same(got, want, "cost %d dirs %d %d-bit %s" % (cost, directions, 8 * aggregate,
                                                              storage))
                info = last_plan(ctx)
                assert info["plan"] == "streamed" and info["storage"] == storage
            finally:
                ctx.close()
            assert os.listdir(tmp) == []  # the temporary aggregate is gone


def test_ram_allocation_failure_falls_back_to_recompute():
    """GpuStereo.java:198-208: under --gpu-aggregate auto a RAM aggregate that
    cannot be allocated is replaced by the bounded recomputation."""
    left, right, _ = synth.stereo_pair(300, 120, 48, 0, seed=61)
    op = O.params(window_w=9, window_h=7, disparities=48, p1=10, p2=120, lr_threshold=1.0)
    want = O.compute(op, left, right, 300, 120)
    ctx = sgm.Context(0)
    try:
        ctx.plan_mode(sgm.PLAN_STREAMED, 16)
        assert ctx.plan(to_sgm(op), 300, 120)["aggregate_storage"] == sgm.AGGREGATE_RAM
        ctx.fail_host_allocations(1)
        got = ctx.match_pair(to_sgm(op), left, right)
        same(got, want, "fallback")
        assert last_plan(ctx)["storage"] == "recompute"
        # an explicit RAM storage reports the failure instead
        ctx.release_buffers()
        ctx.aggregate_storage("ram")
        ctx.fail_host_allocations(1)
        with pytest.raises(sgm.SgmError) as e:
            ctx.match_pair(to_sgm(op), left, right)
        assert e.value.code == -3
        same(ctx.match_pair(to_sgm(op), left, right), want, "after failure")
    finally:
        ctx.close()


def test_auto_runs_the_compact_plan_when_only_recompute_fits():
    """GpuStereo.execute (GpuStereo.java:176-193): auto plan and auto storage,
    the pair too large for the device budget and its aggregate too large for
    the host budget: gray Census runs the compact plan, other costs the
    bounded recomputation; both equal the oracle."""
    w, h = 320, 160
    left, right, _ = synth.stereo_pair(w, h, 64, 0, seed=62)
    for cost, plan_name in ((0, "compact"), (1, "streamed")):
        op = O.params(window_w=9, window_h=7, disparities=64, cost=cost, p1=10, p2=120,
                      lr_threshold=1.0)
        p = to_sgm(op)
        probe = sgm.Context(0)
        probe.esgm_candidates(1)
        probe.plan_mode(sgm.PLAN_ESGM, 16)
        pl = probe.plan(p, w, h)
        probe.close()
        pair = pl["pair_bytes"]
        # a device budget below the resident pair and above the compact plan
        budget = (pair + pl["esgm_bytes"]) // 2 if pl["esgm_bytes"] else pair // 2
        assert pl["esgm_bytes"] < budget < pair
        ctx = sgm.Context(0, budget_bytes=budget)
        try:
            stripe = 2 * w * 16 * 64 * 2
            ctx.plan_mode(sgm.PLAN_AUTO, 16)
            ctx.aggregate_storage("auto", None, 2 * stripe)
            plan = ctx.plan(p, w, h)
            assert not plan["resident"]
            assert plan["compact"] == (plan_name == "compact")
            got = ctx.match_pair(p, left, right)
            same(got, O.compute(op, left, right, w, h), "auto %s" % plan_name)
            info = last_plan(ctx)
            assert info["plan"] == plan_name
            if plan_name == "streamed":
                assert info["storage"] == "recompute" and info["block_rows"] < h
        finally:
            ctx.close()


def test_interrupt_stops_a_call_and_the_context_stays_usable():
    op = O.params(window_w=9, window_h=7, disparities=64, p1=10, p2=120, lr_threshold=1.0)
    left, right, _ = synth.stereo_pair(240, 100, 64, 0, seed=63)
    want = O.compute(op, left, right, 240, 100)
    ctx = sgm.Context(0)
    try:
        # requested before the call: the call stops at its first check
        ctx.interrupt()
        with pytest.raises(sgm.SgmError) as e:
            ctx.match_pair(to_sgm(op), left, right)
        assert e.value.code == -9 and "interrupted" in str(e.value)
        same(ctx.match_pair(to_sgm(op), left, right), want, "after interrupt")
        # from another thread during a long bounded recomputation (one-row
        # strips, one strip per block: many passes over the strips)
        W, H, D = 1600, 400, 128
        big_l, big_r, _ = synth.stereo_pair(W, H, D, 0, seed=64)
        pb = to_sgm(O.params(window_w=9, window_h=7, disparities=D, p1=8, p2=96,
                             lr_threshold=1.0))
        ctx.plan_mode(sgm.PLAN_STREAMED, 1)
        ctx.aggregate_storage("recompute", None, 2 * W * 1 * D * 2)
        assert ctx.plan(pb, W, H)["block_rows"] == 1
        done = {}

        def run():
            t0 = time.perf_counter()
            try:
                ctx.match_pair(pb, big_l, big_r)
                done["rc"] = 0
            except sgm.SgmError as err:
                done["rc"] = err.code
            done["t"] = time.perf_counter() - t0

        th = threading.Thread(target=run)
        th.start()
        time.sleep(0.5)
        t_int = time.perf_counter()
        ctx.interrupt()
        th.join(timeout=60)
        assert not th.is_alive()
        print("interrupted after", done, "stop latency", time.perf_counter() - t_int, flush=True)
        assert done["rc"] == -9
        assert time.perf_counter() - t_int < 5.0
        ctx.plan_mode(sgm.PLAN_AUTO, 64)
        ctx.aggregate_storage("auto")
        same(ctx.match_pair(to_sgm(op), left, right), want, "after a threaded interrupt")
    finally:
        ctx.close()
