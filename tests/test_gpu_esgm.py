"""Compact plan (--gpu-sgm esgm, GpuCandidateStereo.java): its results equal the
oracle and the resident plan bit for bit, for every proposal count and strip
height, with certified and verified pixels both present; verification in
groups when the budget is tight; unsupported parameters are refused
(GpuStereo.java:185-197)."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth
from tests.cases import shifted, upgrade_fixture
from tests.test_gpu_parity import assert_same, to_sgm

pytestmark = pytest.mark.gpu


def esgm(ctx, p, left, right, vl=None, vr=None, rows=64, pool=2, what=""):
    print("esgm", what, "rows", rows, "pool", pool, flush=True)
    ctx.plan_mode(sgm.PLAN_ESGM, rows)
    ctx.esgm_candidates(pool)
    try:
        return ctx.match_pair(p, left, right, vl, vr), ctx.esgm_counts()
    finally:
        ctx.plan_mode(sgm.PLAN_AUTO)
        ctx.esgm_candidates(2)


@pytest.mark.parametrize("directions", [4, 8])
def test_upgrade_fixture_esgm(gpu_context, directions):
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3, cost=0, directions=directions)
    h, w = a.shape
    want = O.compute(op, a, b, w, h, ma, mb)
    verified = certified = 0
    for rows, pool in ((1, 1), (7, 2), (32, 4), (h, 3)):
        got, counts = esgm(gpu_context, to_sgm(op), a, b, ma, mb, rows, pool,
                           "upgrade p%d" % directions)
        assert_same(got, want, "esgm dirs %d rows %d pool %d" % (directions, rows, pool))
        for side in ("left", "right"):
            c, v = counts[side]
            assert c + v == w * h
            certified += c
            verified += v
    assert certified > 0 and verified > 0


@pytest.mark.parametrize("case", range(6))
def test_esgm_equals_resident(gpu_context, case):
    """Lane groups R = 2..16, signed ranges, rectangular windows, edge-adaptive
    P2 on and off; horizontal paths with on-the-fly census costs (1-word
    windows, R <= 8) and from the strip cost volume (R = 16, 2-word 11x11)."""
    D = [3, 64, 130, 300, 700, 96][case]
    w, h = [60, 150, 170, 220, 760, 140][case], [30, 61, 68, 47, 40, 52][case]
    left, right, _ = synth.stereo_pair(w, h, D, -4 if case % 2 else 0, seed=400 + case)
    p = sgm.SgmParams(window_w=[3, 9, 5, 7, 9, 11][case], window_h=[3, 7, 5, 5, 7, 11][case],
                      disparities=D, min_disparity=-4 if case % 2 else 0, cost=0,
                      p1=[4, 8, 10, 8, 10, 20][case], p2=[30, 96, 120, 96, 120, 120][case],
                      edge_scale=[0, 32, 32, 0, 32, 32][case], directions=[4, 8, 8, 4, 8, 8][case],
                      aggregate_bytes=0, census_weight=500, ad_truncation=30,
                      gradient_truncation=60, reserved=0, uniqueness=0.03, lr_threshold=1.0)
    want = gpu_context.match_pair(p, left, right)
    # strips (and their carries) only where the costs are not computed on the fly
    for rows, pool in ((3, 2), (16, 4), (h, 1)) if case >= 4 else ((16, 2), (h, 1)):
        got, counts = esgm(gpu_context, p, left, right, rows=rows, pool=pool, what="case %d" % case)
        assert_same(got, want, "case %d rows %d pool %d" % (case, rows, pool))


def test_verification_in_small_groups():
    """A budget a few pixels above the compact footprint: verification runs
    in many groups and the result is unchanged."""
    left, right, _ = synth.stereo_pair(120, 50, 48, 0, seed=77)
    p = sgm.SgmParams(window_w=9, window_h=7, disparities=48, min_disparity=0, cost=0, p1=8,
                      p2=96, edge_scale=32, directions=8, aggregate_bytes=0, census_weight=500,
                      ad_truncation=30, gradient_truncation=60, reserved=0, uniqueness=0.03,
                      lr_threshold=1.0)
    ref = sgm.Context(0)
    try:
        want = ref.match_pair(p, left, right)
        plan = ref.plan(p, 120, 50)
    finally:
        ref.close()
    assert plan["esgm_candidates"] == 2 and plan["esgm_rows"] == 50
    budget = plan["esgm_bytes"] + 2 * plan["padded_disparities"] * 2 * 5
    ctx = sgm.Context(0, budget)
    try:
        got, counts = esgm(ctx, p, left, right, rows=64, pool=2, what="small groups")
        assert counts["left"][1] > 6 and counts["right"][1] > 6
        assert_same(got, want, "small verification groups")
        # groups of 6 pixels (the budget's room), reported by the profile
        most = max(counts["left"][1], counts["right"][1])
        assert ctx.profile()["last_plan"]["verify_groups"] == -(-most // 6)
        # GpuStereo.limit under the compact plan: the largest range it serves
        ctx.plan_mode(sgm.PLAN_ESGM)
        try:
            lim = ctx.disparity_limit(p, 120, 50)
            assert lim >= 48
            q = sgm.SgmParams(**{f: getattr(p, f) for f, _ in sgm.SgmParams._fields_})
            q.disparities = lim
            assert ctx.plan(q, 120, 50)["esgm_candidates"] >= 1
            if lim < 4096:
                q.disparities = lim + 1
                assert ctx.plan(q, 120, 50)["esgm_candidates"] == 0
        finally:
            ctx.plan_mode(sgm.PLAN_AUTO)
        # a tighter budget: the plan drops to one proposal per path
        # (GpuCandidateStereo.plan); costs are computed on the fly, no strips
        cut = 120 * 50 * 2 * 8 * 8
        small = sgm.Context(0, plan["esgm_bytes"] - cut)
        try:
            pl = small.plan(p, 120, 50)
            assert pl["esgm_candidates"] == 1 and pl["esgm_rows"] == 50
            assert pl["esgm_bytes"] <= plan["esgm_bytes"] - cut
        finally:
            small.close()
    finally:
        ctx.close()


def test_single_direction_and_filtered_entry_points(gpu_context):
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3)
    p = to_sgm(op)
    h, w = a.shape
    ta = O.texture(a, w, h, op.window_w, op.window_h, ma)
    tb = O.texture(b, w, h, op.window_w, op.window_h, mb)
    for rev in (0, 1):
        ga, gb, TA, TB = (a, b, ta, tb) if not rev else (b, a, tb, ta)
        ga, gb = np.ascontiguousarray(ga), np.ascontiguousarray(gb)
        want, st_want = gpu_context.match(p, ga, gb, TA, TB, w, h, rev)
        print("esgm single direction", rev, flush=True)
        gpu_context.plan_mode(sgm.PLAN_ESGM, 9)
        try:
            got, st_got = gpu_context.match(p, ga, gb, TA, TB, w, h, rev)
            counts = gpu_context.esgm_counts()
        finally:
            gpu_context.plan_mode(sgm.PLAN_AUTO)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert sum(counts["right" if rev else "left"]) == w * h
        assert sum(counts["left" if rev else "right"]) == 0
        for k in O.STAT_FIELDS:
            assert getattr(st_got, k) == getattr(st_want, k), k
    f = O.filters()
    want = O.compute_filtered(op, f, a, b, w, h, ma, mb)
    gpu_context.plan_mode(sgm.PLAN_ESGM, 16)
    try:
        got = gpu_context.match_pair_filtered(p, sgm.SgmFilterParams(
            **{k: getattr(f, k) for k, _ in sgm.SgmFilterParams._fields_}), a, b, ma, mb)
    finally:
        gpu_context.plan_mode(sgm.PLAN_AUTO)
    assert np.array_equal(got["disparity"].view(np.uint32),
                          np.ascontiguousarray(want["disparity"], np.float32).view(np.uint32))
    assert np.array_equal(got["reasons"], want["reasons"])


def test_unsupported_parameters_are_refused(gpu_context):
    a, b, ma, mb = upgrade_fixture()
    bad = [O.params(disparities=41, cost=1), O.params(disparities=41, aggregate_bytes=4),
           O.params(disparities=41, cost=3), O.params(disparities=41, cost=4)]
    gpu_context.plan_mode(sgm.PLAN_ESGM, 32)
    try:
        for op in bad:
            with pytest.raises(sgm.SgmError) as e:
                gpu_context.match_pair(to_sgm(op), a, b, ma, mb)
            assert e.value.code == -4 and "compact" in str(e.value)
        left_bgr = np.stack([a, a, a], axis=2)
        right_bgr = np.stack([b, b, b], axis=2)
        with pytest.raises(sgm.SgmError) as e:
            gpu_context.match_pair_color(to_sgm(O.params(disparities=41)), a, b, left_bgr,
                                         right_bgr, None, ma, mb)
        assert e.value.code == -4
    finally:
        gpu_context.plan_mode(sgm.PLAN_AUTO)
    with pytest.raises(sgm.SgmError) as e:
        gpu_context.esgm_candidates(5)
    assert e.value.code == -1


def test_census_sgm_mirror_with_esgm_options():
    """CensusSgm.compute with --gpu-sgm esgm / --esgm-candidates equals the
    default plan through the reference option names; a bad candidate count is
    refused as Options.java:781-782 does."""
    left, right, _ = synth.stereo_pair(200, 120, 48, 0, seed=5)
    base = dict(census=5, disparities=48, p1=4, p2=24, compute="gpu", lrThreshold=1.0,
                speckleSize=30)
    try:
        want = sgm.CensusSgm(sgm.Options(**base)).compute(left, right, 200, 120)
        for pool in (1, 4):
            print("census sgm esgm", pool, flush=True)
            got = sgm.CensusSgm(sgm.Options(gpuSgm="esgm", esgmCandidates=pool, **base)).compute(
                left, right, 200, 120)
            assert np.array_equal(got.disparity.view(np.uint32), want.disparity.view(np.uint32))
            assert got.stats.csv() == want.stats.csv()
        with pytest.raises(ValueError):
            sgm.GpuSgm.configured(sgm.Options(gpuSgm="esgm", esgmCandidates=5, **base))
    finally:
        sgm.GpuSgm.closeSession()


def test_gpu_stereo_report_keys_and_storage():
    """GpuStereo.report (GpuStereo.java:95-168): the reference's keys, the
    storage each plan mode selects, footprints from the plan."""
    o = sgm.Options(census=9, censusWidth=9, censusHeight=7, disparities=96, p1=8, p2=96,
                    rawPenalties=True, compute="gpu")
    keys = {"storage", "rows", "bank_bytes", "estimated_device_bytes", "estimated_host_bytes",
            "host_reserve_bytes", "aggregate_bits", "aggregate_proved_bound",
            "cost_generation_passes", "cost_generation_passes_are_upper_bound",
            "aggregate_volume_reads", "aggregate_volume_writes", "aggregate_disk_bytes",
            "aggregate_ram_bytes", "verification_block_rows", "recomputation_blocks",
            "candidates_per_path", "candidate_device_bytes"}
    try:
        seen = {}
        for mode in ("auto", "streamed", "esgm"):
            o.gpuSgm = mode
            r = sgm.GpuStereo.report(640, 480, o)
            assert set(r) == keys, set(r) ^ keys
            seen[mode] = r
        assert seen["auto"]["storage"] == "resident"
        assert seen["streamed"]["storage"] == "streamed_ram"
        # both directions' uint16 aggregates, rows of 96 disparities padded to 128
        assert seen["streamed"]["aggregate_ram_bytes"] == 2 * 640 * 480 * 128 * 2
        assert seen["esgm"]["storage"] == "compact_with_sparse_exact_verification"
        assert seen["esgm"]["candidates_per_path"] == 2
        assert seen["esgm"]["estimated_device_bytes"] < seen["auto"]["estimated_device_bytes"]
        assert seen["auto"]["aggregate_bits"] == 16
        assert sgm.GpuStereo.limit(640, 480, o) >= 96
        o.gpuSgm, o.matchingCost = "esgm", "ad-census"
        assert "unavailable" in sgm.GpuStereo.report(640, 480, o)
        o.compute = "cpu"
        assert sgm.GpuStereo.report(640, 480, o) is None
    finally:
        sgm.GpuSgm.closeSession()


@pytest.mark.parametrize("case", range(3))
def test_esgm_edge_cases_against_oracle(gpu_context, case):
    """Ranges wider than the image (most targets outside it), large negative
    minimum disparities, heavy validity masks and an untextured image: the
    compact plan equals the oracle bit for bit."""
    rng = np.random.default_rng(900 + case)
    w, h = [60, 90, 70][case], [40, 33, 30][case]
    D, md = [130, 48, 40][case], [0, -60, -5][case]
    a, b = shifted(910 + case, w, h, 5)
    if case == 2:
        a = np.full((h, w), 77, np.uint8)  # constant image: every cost invalid
    ma = (rng.random((h, w)) > [0.1, 0.5, 0.0][case]).astype(np.uint8) * 255
    mb = (rng.random((h, w)) > [0.1, 0.5, 0.0][case]).astype(np.uint8) * 255
    op = O.params(window_w=5, window_h=5, disparities=D, min_disparity=md,
                  directions=[8, 4, 8][case])
    want = O.compute(op, a, b, w, h, ma, mb)
    for pool in (1, 3):
        got, counts = esgm(gpu_context, to_sgm(op), a, b, ma, mb, 64, pool, "edge %d" % case)
        assert_same(got, want, "edge case %d pool %d" % (case, pool))
        assert sum(counts["left"]) == w * h
