"""GPU parity: libsgm_hip (through the C ABI) against the oracle, bit-exact.

Every comparison is on the same seeded inputs: the committed golden fixtures, the
reference's regenerated known-answer fixtures, randomized parameter sweeps at sizes
the oracle finishes in seconds, and the full-size plain workloads.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth
from tests.cases import shifted, upgrade_fixture

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def to_sgm(op):
    return sgm.SgmParams(**{f: getattr(op, f) for f, _ in sgm.SgmParams._fields_})


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_same(got, want, what=""):
    for k in ("checked", "dl", "dr"):
        g, w = bits(got[k]), bits(want[k])
        if not np.array_equal(g, w):
            bad = np.nonzero(g != w)[0]
            raise AssertionError("%s %s: %d pixels differ, first %d: gpu %r oracle %r" % (
                what, k, len(bad), bad[0], got[k][bad[0]], want[k][bad[0]]))
    assert np.array_equal(got["reasons"], want["reasons"]), what + " reasons"
    assert np.array_equal(got["boundary"], want["boundary"]), what + " boundary"
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], (what, k, got["stats"][k], want["stats"][k])


def run_both(ctx, op, left, right, vl=None, vr=None, what=""):
    h, w = left.shape
    print("case", what, w, h, op.disparities, flush=True)
    want = O.compute(op, left, right, w, h, vl, vr)
    got = ctx.match_pair(to_sgm(op), left, right, vl, vr)
    assert_same(got, want, what)
    return got


def test_golden_fixtures(gpu_context):
    files = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
    assert files
    for f in files:
        z = np.load(f, allow_pickle=False)
        kw = {k: (float(v) if k in ("uniqueness", "lr_threshold") else int(v))
              for k, v in zip(z["param_names"], z["param_values"])}
        if kw["cost"] == 4:
            continue  # FP32 ZNCC is held to a tolerance, not bit-exactness (test_zncc_fp32)
        vl = z["valid_left"] if "valid_left" in z.files else None
        vr = z["valid_right"] if "valid_right" in z.files else None
        print("golden", os.path.basename(f), flush=True)
        got = gpu_context.match_pair(sgm.SgmParams(**kw), z["left"], z["right"], vl, vr)
        want = dict(checked=z["checked"], dl=z["dl"], dr=z["dr"], reasons=z["reasons"],
                    boundary=z["boundary"], stats=dict(zip(O.STAT_FIELDS, map(int, z["stats"]))))
        assert_same(got, want, os.path.basename(f))


@pytest.mark.parametrize("case", range(6))
def test_volumes_bit_exact(gpu_context, case):
    """Cost and aggregate volumes of both matching directions equal the oracle's."""
    rng = np.random.default_rng(case)
    D = [5, 41, 64, 96, 130, 257][case]
    w, h = max(40, D + 30), 24 + case * 3
    ww, wh = [(5, 5), (9, 7), (3, 3), (7, 5), (11, 11), (5, 3)][case]
    left, right, _ = synth.stereo_pair(w, h, D, -2 if case % 2 else 0, seed=100 + case)
    valid = (rng.random((h, w)) > 0.05).astype(np.uint8) if case % 3 == 0 else None
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=-2 if case % 2 else 0,
                  cost=case % 4, p1=8 + case, p2=96 + 9 * case, edge_scale=[32, 0][case % 2],
                  directions=[8, 4][case % 2], aggregate_bytes=[0, 4][case // 3],
                  lr_threshold=1.0)
    ctx = gpu_context
    ctx.keep_volumes(True)
    try:
        got = ctx.match_pair(to_sgm(op), left, right, valid, valid)
        n = w * h
        da, db = O.census(left, w, h, ww, wh), O.census(right, w, h, ww, wh)
        ta = O.texture(left, w, h, ww, wh, valid)
        tb = O.texture(right, w, h, ww, wh, valid)
        for d, (a_, b_, A, B, TA, TB) in enumerate([(da, db, left, right, ta, tb),
                                                    (db, da, right, left, tb, ta)]):
            disp, st, rs, bd, C, S = O.match(op, a_, b_, A.reshape(-1), B.reshape(-1), TA, TB,
                                             w, h, d, volumes=True)
            gc, gs = ctx.read_volumes(d, n, D)
            assert np.array_equal(gc, C), "cost volume direction %d" % d
            assert np.array_equal(gs, S), "aggregate volume direction %d" % d
            assert np.array_equal(bits(got["dl" if d == 0 else "dr"]), bits(disp))
    finally:
        ctx.keep_volumes(False)


PARAM_SWEEP = [
    # (w, h, D, minD, window, cost, dirs, p1, p2, edge, agg)
    (64, 40, 3, 0, (5, 5), 0, 8, 4, 24, 32, 0),
    (90, 33, 17, -5, (3, 3), 1, 4, 2, 9, 0, 0),
    (120, 50, 63, 1, (9, 7), 0, 8, 10, 120, 32, 0),
    (130, 45, 65, -32, (7, 7), 2, 8, 7, 80, 16, 0),
    (260, 30, 129, 3, (11, 11), 3, 4, 20, 120, 32, 0),
    (300, 28, 200, -10, (9, 7), 0, 8, 8, 96, 32, 4),
    (360, 26, 300, 0, (5, 7), 1, 8, 5, 60, 255, 0),
    (600, 20, 513, -20, (9, 7), 0, 4, 8, 96, 32, 0),
    (1100, 12, 1030, 0, (3, 3), 0, 8, 2, 10, 32, 0),
]


@pytest.mark.parametrize("cfg", PARAM_SWEEP, ids=[str(i) for i in range(len(PARAM_SWEEP))])
def test_parameter_sweep(gpu_context, cfg):
    w, h, D, md, (ww, wh), cost, dirs, p1, p2, edge, agg = cfg
    left, right, _ = synth.stereo_pair(w, h, D, md, seed=D)
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=md, cost=cost,
                  directions=dirs, p1=p1, p2=p2, edge_scale=edge, aggregate_bytes=agg)
    run_both(gpu_context, op, left, right, what="sweep%r" % (cfg,))


RECOMPUTE_CASES = [
    # (w, h, D, minD, window, dirs, p1, p2, edge, masks)
    (64, 40, 3, 0, (5, 5), 8, 4, 24, 32, False),
    (200, 90, 64, -7, (9, 7), 8, 10, 120, 32, True),
    (256, 70, 100, 5, (7, 5), 4, 6, 80, 0, False),
    (300, 60, 250, -20, (9, 7), 8, 8, 96, 16, True),
    (640, 40, 500, 3, (9, 7), 8, 8, 96, 32, False),
]


@pytest.mark.parametrize("cfg", RECOMPUTE_CASES, ids=[str(i) for i in range(len(RECOMPUTE_CASES))])
def test_recomputed_census_cost_bit_exact(gpu_context, cfg):
    """Census cost recomputed inside every sweep (no cost volume) against the oracle."""
    w, h, D, md, (ww, wh), dirs, p1, p2, edge, masks = cfg
    left, right, _ = synth.stereo_pair(w, h, D, md, seed=w + D)
    vl = vr = None
    if masks:
        rng = np.random.default_rng(D)
        vl = (rng.random((h, w)) > 0.04).astype(np.uint8)
        vr = np.roll(vl, 2, axis=1)
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=md, directions=dirs,
                  p1=p1, p2=p2, edge_scale=edge)
    gpu_context.cost_volume(sgm.VOLUME_RECOMPUTE)
    try:
        run_both(gpu_context, op, left, right, vl, vr, what="recompute%r" % (cfg,))
    finally:
        gpu_context.cost_volume(sgm.VOLUME_AUTO)


@pytest.mark.parametrize("cost", [0, 1, 2, 3])
@pytest.mark.parametrize("directions", [4, 8])
def test_reference_upgrade_fixture(gpu_context, cost, directions):
    # UpgradeStereoRegression.java:36-79: CPU/GPU Arrays.equals on the same inputs.
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3, cost=cost, directions=directions)
    run_both(gpu_context, op, a, b, ma, mb, what="upgrade")


def test_reference_wide_range_693(gpu_context):
    # UpgradeStereoRegression.java:95-113: 800x64, Random(93), shift 130, D=693.
    a, b = shifted(93, 800, 64, 130)
    run_both(gpu_context, O.params(disparities=693), a, b, what="wide693")


def test_reference_selftest_and_noise(gpu_context):
    a, b = shifted(4, 96, 48, 6)
    run_both(gpu_context, O.params(disparities=16, uniqueness=.02), a, b, what="selftest")
    run_both(gpu_context, O.params(disparities=16, min_disparity=-10, uniqueness=.02), b, a,
             what="selftest-signed")
    run_both(gpu_context, O.params(window_w=11, window_h=11, p1=20, p2=120, directions=4,
                                   disparities=16, uniqueness=.02), a, b, what="selftest-11x11")
    c = np.zeros_like(a)
    got = run_both(gpu_context, O.params(disparities=16), c, c, what="constant")
    assert np.isnan(got["checked"]).all()


def test_single_direction_match_merges_like_reference(gpu_context):
    """sgm_match = GpuSgm.match: reasons/boundary updated in place, stats added."""
    w, h = 160, 96
    a, b = shifted(497, w, h, 8)
    op = O.params(min_disparity=8, disparities=17)
    p = to_sgm(op)
    ta, tb = O.texture(a, w, h, 5, 5), O.texture(b, w, h, 5, 5)
    da, db = O.census(a, w, h, 5, 5), O.census(b, w, h, 5, 5)
    for reverse in (0, 1):
        A, B, TA, TB, DA, DB = (a, b, ta, tb, da, db) if not reverse else (b, a, tb, ta, db, da)
        reasons = np.full(w * h, 7, np.uint8)  # arbitrary prior content must survive
        boundary = np.full(w * h, 9, np.uint8)
        want_r, want_b = reasons.copy(), boundary.copy()
        disp_o, st_o, _, _ = O.match(op, DA, DB, A.reshape(-1), B.reshape(-1), TA, TB, w, h,
                                     reverse)
        # the oracle writes into fresh arrays; replay its writes onto the prior content
        r1 = np.full(w * h, 255, np.uint8)
        b1 = np.full(w * h, 255, np.uint8)
        O.lib().oracle_match(op, O._ptr(np.ascontiguousarray(DA)), O._ptr(np.ascontiguousarray(DB)),
                             O._ptr(A.reshape(-1)), O._ptr(B.reshape(-1)), O._ptr(TA), O._ptr(TB),
                             w, h, reverse, O._ptr(np.empty(w * h, np.float32)),
                             O.OracleStats(), O._ptr(r1), O._ptr(b1), None, None)
        want_r[r1 != 255] = r1[r1 != 255]
        want_b[b1 != 255] = b1[b1 != 255]
        st = sgm.SgmStats()
        st.selected = 5  # accumulates, like GpuSgm.java:289-294
        disp, st = gpu_context.match(p, A, B, TA, TB, w, h, reverse, st, reasons, boundary)
        assert np.array_equal(bits(disp), bits(disp_o))
        assert np.array_equal(reasons, want_r)
        assert np.array_equal(boundary, want_b)
        assert st.selected == st_o["selected"] + 5
        assert st.lower_boundary_rejected == st_o["lower_boundary_rejected"]
        assert st.full_interval_pixels == st_o["full_interval_pixels"]


SCHEDULE_CASES = [
    # (w, h, D, minD, window, cost, dirs, p2, masks)
    (200, 90, 64, -7, (9, 7), 0, 8, 96, True),     # uint8 costs, lane group 2
    (240, 70, 100, 5, (7, 5), 1, 4, 96, False),    # 4 paths: pairs, no planes
    (260, 60, 250, -20, (9, 7), 2, 8, 96, True),   # lane group 4
    (180, 50, 90, 0, (5, 5), 3, 8, 96, False),     # integer ZNCC
    (640, 40, 500, 3, (9, 7), 0, 8, 96, False),    # lane group 8
    (220, 60, 80, 0, (11, 11), 0, 8, 135, True),   # 2-word census: 120 + 135 = 255, planes
    (220, 60, 80, 0, (11, 11), 0, 8, 136, False),  # 120 + 136 > 255: no planes
    (150, 80, 48, -3, (9, 7), 0, 8, 193, True),    # 62 + 193 = 255, planes
]


@pytest.mark.parametrize("cfg", SCHEDULE_CASES, ids=[str(i) for i in range(len(SCHEDULE_CASES))])
def test_sweep_schedules_equal_the_oracle(cfg):
    """Every sweep schedule (one direction per launch, paired partial
    aggregates, path-cost planes) equals the oracle bit for bit; the plan and
    the profile name the schedule that ran, and the planes run exactly where
    every valid path cost fits a byte (window bits - 1 + P2 <= 255, 8 paths)."""
    w, h, D, md, (ww, wh), cost, dirs, p2, masks = cfg
    left, right, _ = synth.stereo_pair(w, h, D, md, seed=300 + D)
    rng = np.random.default_rng(D)
    vl = (rng.random((h, w)) > 0.05).astype(np.uint8) if masks else None
    vr = (rng.random((h, w)) > 0.05).astype(np.uint8) if masks else None
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=md, cost=cost,
                  directions=dirs, p1=8, p2=p2, lr_threshold=1.0)
    want = O.compute(op, left, right, w, h, vl, vr)
    planes_apply = dirs == 8 and ww * wh - 1 + p2 <= 255
    for name in ("sequential", "paired", "planes", "auto"):
        ctx = sgm.Context(0)
        try:
            ctx.sweeps(name)
            ran = ctx.plan(to_sgm(op), w, h)["sweeps"]
            expect = {"sequential": 1, "paired": 2, "planes": 3 if planes_apply else 2,
                      "auto": 3 if planes_apply else 2}[name]
            assert ran == expect, (name, ran)
            ctx.set_profiling(True)
            print("schedule", name, cfg, flush=True)
            got = ctx.match_pair(to_sgm(op), left, right, vl, vr)
            assert ctx.profile()["last_plan"]["sweeps"] == ["", "sequential", "paired",
                                                           "planes"][expect]
            assert_same(got, want, "%s %r" % (name, cfg))
        finally:
            ctx.close()


def test_schedules_fall_back_when_their_volumes_do_not_fit():
    w, h, D = 300, 120, 64
    left, right, _ = synth.stereo_pair(w, h, D, 0, seed=77)
    op = O.params(window_w=9, window_h=7, disparities=D, p1=10, p2=120, lr_threshold=1.0)
    want = O.compute(op, left, right, w, h)
    probe = sgm.Context(0)
    plan = probe.plan(to_sgm(op), w, h)
    probe.close()
    assert plan["sweeps"] == sgm.SWEEPS_PLANES
    cells = 2 * w * h * plan["padded_disparities"]
    base = plan["pair_bytes"] + cells  # resident with the uint8 cost volume
    # room for the second partial aggregate, not for the planes; then for neither
    for extra, expect in ((2 * cells, sgm.SWEEPS_PAIRED), (cells, sgm.SWEEPS_SEQUENTIAL)):
        ctx = sgm.Context(0, budget_bytes=base + extra + (1 << 20))
        try:
            assert ctx.plan(to_sgm(op), w, h)["sweeps"] == expect
            ctx.set_profiling(True)
            got = ctx.match_pair(to_sgm(op), left, right)
            info = ctx.profile()["last_plan"]
            assert info["plan"] == "resident"
            assert info["sweeps"] == ["", "sequential", "paired", "planes"][expect]
            assert_same(got, want, "schedule %d under a tight budget" % expect)
        finally:
            ctx.close()


def test_plain_workload_640x480(gpu_context):
    # Plain SGM, 640x480, census 9x7, D=64, P1=10, P2=120, 8 paths, L-R tolerance 1.
    left, right, _ = synth.stereo_pair(640, 480, 64, 0, seed=1)
    for edge in (32, 0):
        op = O.params(window_w=9, window_h=7, disparities=64, p1=10, p2=120, edge_scale=edge,
                      lr_threshold=1.0)
        run_both(gpu_context, op, left, right, what="cfg640 edge%d" % edge)


def test_plain_workload_1600x1067_d256(gpu_context):
    # Plain GPU SGM at the default processing size: 1600x1067, D=256, P1=8, P2=96.
    left, right, _ = synth.stereo_pair(1600, 1067, 256, 0, seed=2)
    op = O.params(window_w=9, window_h=7, disparities=256, p1=8, p2=96, lr_threshold=1.0)
    got = run_both(gpu_context, op, left, right, what="cfg1600")
    assert got["stats"]["valid"] > 0.5 * 1600 * 1067


def test_device_entry_point_matches_host(gpu_context):
    left, right, _ = synth.stereo_pair(200, 120, 48, -4, seed=9)
    op = O.params(window_w=9, window_h=7, disparities=48, min_disparity=-4, p1=10, p2=120,
                  lr_threshold=1.0)
    p = to_sgm(op)
    ctx = gpu_context
    host = ctx.match_pair(p, left, right)
    n = left.size
    d_left = ctx.alloc(n).upload(left)
    d_right = ctx.alloc(n).upload(right)
    out = ctx.alloc(4 * n)
    st = ctx.alloc(128).upload(np.zeros(16, np.int64))
    reasons = ctx.alloc(n)
    bnd = ctx.alloc(n)
    for _ in range(2):  # twice: buffers are reused
        ctx.match_pair_device(p, d_left, d_right, 200, 120, out, d_stats=st, d_reasons=reasons,
                              d_boundary=bnd)
        ctx.synchronize()
        assert np.array_equal(bits(out.download(np.float32, n)), bits(host["checked"]))
        assert np.array_equal(reasons.download(np.uint8, n), host["reasons"])
        assert np.array_equal(bnd.download(np.uint8, n), host["boundary"])
        s = st.download(np.int64, 16)
        assert s[5] == host["stats"]["selected"] and s[10] == host["stats"]["valid"]
    for b in (d_left, d_right, out, st, reasons, bnd):
        b.free()


def test_repeatable_and_recovers_from_allocation_failure():
    ctx = sgm.Context(0)
    try:
        left, right, _ = synth.stereo_pair(150, 100, 32, 0, seed=5)
        op = O.params(window_w=9, window_h=7, disparities=32, p1=10, p2=120)
        p = to_sgm(op)
        ctx.fail_allocations(1)
        with pytest.raises(sgm.SgmError) as e:
            ctx.match_pair(p, left, right)
        assert e.value.code == -3
        a = ctx.match_pair(p, left, right)
        b = ctx.match_pair(p, left, right)
        assert np.array_equal(bits(a["checked"]), bits(b["checked"]))
        want = O.compute(op, left, right, 150, 100)
        assert_same(a, want, "after-failure")
    finally:
        ctx.close()


def test_image_limits_are_validated(gpu_context):
    """Sizes the kernels cannot launch are refused before any work: images
    taller than 65535 rows (one grid row per image row in the cost kernels),
    above 2^28 pixels, or smaller than the census window plus a border."""
    p = to_sgm(O.params(window_w=9, window_h=7, disparities=16))
    for (w, h) in ((16, 65536), (1 << 15, (1 << 13) + 1), (10, 200), (200, 8)):
        with pytest.raises(sgm.SgmError) as e:
            gpu_context.plan(p, w, h)
        assert e.value.code == -1, (w, h)
    assert gpu_context.plan(p, 16, 65535)["padded_disparities"] == 64


def test_zncc_fp32_against_fp64_oracle(gpu_context):
    """FP32 ZNCC variant: at most 0.1% of pixels may differ by more than 1 disparity
    from the FP64 oracle (tolerance of the ZNCC workload)."""
    left, right, _ = synth.stereo_pair(400, 300, 96, 0, seed=31, gain=True)
    op = O.params(window_w=7, window_h=7, disparities=96, cost=4, p1=12, p2=150, lr_threshold=1.0)
    want = O.compute(op, left, right, 400, 300)
    got = gpu_context.match_pair(to_sgm(op), left, right)
    for k in ("dl", "checked"):
        a, b = got[k], want[k]
        both = np.isfinite(a) & np.isfinite(b)
        one = np.isfinite(a) ^ np.isfinite(b)
        bad = (np.abs(a[both] - b[both]) > 1).sum() + one.sum()
        assert bad <= 0.001 * a.size, (k, int(bad))


def test_zncc_fp32_workload_strip_1600_d192(gpu_context):
    """The ZNCC workload's shape (1600 wide, D=192, 7x7, P1=12, P2=150, gain field
    0.8-1.2) on a 1600x160 strip the FP64 oracle finishes in seconds: at most 0.1%
    of pixels with |dd| > 1 (or valid on one side only)."""
    w, h = 1600, 160
    left, right, _ = synth.stereo_pair(w, h, 192, 0, seed=33, gain=True)
    op = O.params(window_w=7, window_h=7, disparities=192, cost=4, p1=12, p2=150,
                  lr_threshold=1.0)
    print("zncc strip", w, h, flush=True)
    want = O.compute(op, left, right, w, h)
    got = gpu_context.match_pair(to_sgm(op), left, right)
    for k in ("dl", "dr", "checked"):
        a, b = got[k], want[k]
        both = np.isfinite(a) & np.isfinite(b)
        one = np.isfinite(a) ^ np.isfinite(b)
        bad = int((np.abs(a[both] - b[both]) > 1).sum() + one.sum())
        print(k, "differ", bad, "of", a.size, flush=True)
        assert bad <= 0.001 * a.size, (k, bad)
