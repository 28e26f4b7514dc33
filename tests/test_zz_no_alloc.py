"""sgm_match_pair_device_noalloc: the device entry point that allocates nothing.

Through the C ABI: its results equal sgm_match_pair_device's bit for bit on two
calls in a row (the second with a smaller pair, so stale workspace contents
from the first call would show), with the cost volume resident and with the
census cost recomputed per sweep; the device's free memory does not move during
a call; pairs beyond the fixed workspace are refused with an error.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth

pytestmark = pytest.mark.gpu


def params(D, md=0):
    return sgm.SgmParams(window_w=9, window_h=7, disparities=D, min_disparity=md, cost=0, p1=10,
                         p2=120, edge_scale=32, directions=8, aggregate_bytes=0,
                         census_weight=500, ad_truncation=30, gradient_truncation=60, reserved=0,
                         uniqueness=0.03, lr_threshold=1.0)


def device_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def run(ctx, entry, p, left, right):
    """One pair through `entry` ("plain" or "noalloc") with every output in
    device memory; returns the downloaded outputs."""
    h, w = left.shape
    n = w * h
    b = dict(left=ctx.alloc(n).upload(left), right=ctx.alloc(n).upload(right),
             checked=ctx.alloc(4 * n), dl=ctx.alloc(4 * n), dr=ctx.alloc(4 * n),
             stats=ctx.alloc(128).upload(np.full(16, -1, np.int64)),
             reasons=ctx.alloc(n), boundary=ctx.alloc(n))
    fn = ctx.match_pair_device if entry == "plain" else ctx.match_pair_device_noalloc
    fn(p, b["left"], b["right"], w, h, b["checked"], d_dl=b["dl"], d_dr=b["dr"],
       d_stats=b["stats"], d_reasons=b["reasons"], d_boundary=b["boundary"])
    ctx.synchronize()
    out = {k: b[k].download(np.uint32, n) for k in ("checked", "dl", "dr")}
    out["reasons"] = b["reasons"].download(np.uint8, n)
    out["boundary"] = b["boundary"].download(np.uint8, n)
    out["stats"] = b["stats"].download(np.int64, 16)
    for v in b.values():
        v.free()
    return out


def assert_equal(got, want, what):
    for k in ("checked", "dl", "dr", "reasons", "boundary", "stats"):
        assert np.array_equal(got[k], want[k]), "%s: %s differs" % (what, k)


@pytest.mark.parametrize("policy", [0, 2])  # SGM_VOLUME_AUTO, SGM_VOLUME_RECOMPUTE
def test_noalloc_equals_device_entry_point(policy):
    ref = sgm.Context(0)
    ctx = sgm.Context(0)
    try:
        ref.cost_volume(policy)
        ctx.cost_volume(policy)
        # two calls in a row, the second smaller: 640x480 D=64, then 320x200 D=48
        for (w, h, D, md, seed) in [(640, 480, 64, 0, 21), (320, 200, 48, -4, 22)]:
            print("noalloc case", w, h, D, md, "policy", policy, flush=True)
            left, right, _ = synth.stereo_pair(w, h, D, md, seed=seed)
            p = params(D, md)
            want = run(ref, "plain", p, left, right)
            got = run(ctx, "noalloc", p, left, right)
            assert_equal(got, want, "%dx%d D=%d" % (w, h, D))
            assert got["stats"][10] > 0.5 * w * h
        # the smaller pair against the oracle too
        op = O.params(window_w=9, window_h=7, disparities=48, min_disparity=-4, p1=10, p2=120,
                      lr_threshold=1.0)
        o = O.compute(op, left, right, 320, 200)
        assert np.array_equal(got["checked"], np.ascontiguousarray(o["checked"], np.float32).view(np.uint32))
        assert [int(v) for v in got["stats"][:7]] == [o["stats"][k] for k in O.STAT_FIELDS[:7]]
    finally:
        ref.close()
        ctx.close()


def test_noalloc_recomputes_costs_when_the_volume_does_not_fit():
    """700x480 D=64 needs more than the workspace with a cost volume and fits
    without one: the census cost is recomputed per sweep, results unchanged."""
    ref = sgm.Context(0)
    ctx = sgm.Context(0)
    try:
        left, right, _ = synth.stereo_pair(700, 480, 64, 0, seed=23)
        p = params(64)
        print("noalloc recompute case 700x480 D=64", flush=True)
        want = run(ref, "plain", p, left, right)
        got = run(ctx, "noalloc", p, left, right)
        assert_equal(got, want, "700x480 D=64")
    finally:
        ref.close()
        ctx.close()


def test_noalloc_allocates_nothing_and_refuses_large_pairs():
    ctx = sgm.Context(0)
    try:
        w, h, D = 640, 480, 64
        left, right, _ = synth.stereo_pair(w, h, D, 0, seed=24)
        n = w * h
        dl, dr, out = ctx.alloc(n).upload(left), ctx.alloc(n).upload(right), ctx.alloc(4 * n)
        # a small pair first (kernels loaded), then the large one: an entry point
        # that grows cached buffers would take ~130 MB more for the second
        ctx.match_pair_device_noalloc(params(D), dl, dr, 160, 120, out)
        ctx.synchronize()
        before = device_free_bytes()
        for _ in range(3):
            ctx.match_pair_device_noalloc(params(D), dl, dr, w, h, out)
        ctx.synchronize()
        after = device_free_bytes()
        # allow runtime noise below 2 MiB
        assert before - after < (2 << 20), (before, after)
        # 1600x1067 D=256 needs ~1.8 GB: refused before any device work
        big = ctx.alloc(1600 * 1067)
        with pytest.raises(sgm.SgmError) as e:
            ctx.match_pair_device_noalloc(params(256), big, big, 1600, 1067, out)
        assert e.value.code == -1 and "workspace" in str(e.value)
        ctx.keep_volumes(True)
        with pytest.raises(sgm.SgmError):
            ctx.match_pair_device_noalloc(params(D), dl, dr, w, h, out)
        ctx.keep_volumes(False)
        for b in (dl, dr, out, big):
            b.free()
    finally:
        ctx.close()
