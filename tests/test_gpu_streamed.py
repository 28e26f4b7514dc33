"""Strip-streamed plan: every strip height gives the resident plan's results,
which are bit-exact to the oracle (UpgradeStereoRegression.java:36-79 runs
"streamed" with gpuStripRows {1, 7, 32} against the CPU path; :95-113 the
800x64, D=693 carry case)."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth
from tests.cases import shifted, upgrade_fixture
from tests.test_gpu_parity import assert_same, to_sgm

pytestmark = pytest.mark.gpu


def streamed(ctx, op, left, right, vl, vr, rows, what):
    print("streamed", what, "rows", rows, flush=True)
    ctx.plan_mode(sgm.PLAN_STREAMED, rows)
    try:
        return ctx.match_pair(to_sgm(op), left, right, vl, vr)
    finally:
        ctx.plan_mode(sgm.PLAN_AUTO)


@pytest.mark.parametrize("cost", [0, 1, 2, 3])
@pytest.mark.parametrize("directions", [4, 8])
def test_upgrade_fixture_streamed_rows(gpu_context, cost, directions):
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3, cost=cost, directions=directions)
    h, w = a.shape
    want = O.compute(op, a, b, w, h, ma, mb)
    for rows in (1, 7, 32):
        got = streamed(gpu_context, op, a, b, ma, mb, rows, "upgrade c%d p%d" % (cost, directions))
        assert_same(got, want, "upgrade cost %d dirs %d rows %d" % (cost, directions, rows))


def test_wide_range_693_streamed(gpu_context):
    a, b = shifted(93, 800, 64, 130)
    op = O.params(disparities=693)
    want = O.compute(op, a, b, 800, 64)
    for rows in (7, 64):
        assert_same(streamed(gpu_context, op, a, b, None, None, rows, "wide693"), want,
                    "wide693 rows %d" % rows)


@pytest.mark.parametrize("case", range(4))
def test_streamed_equals_resident(gpu_context, case):
    """Rectangular windows, signed ranges, 32-bit aggregates, edge-adaptive P2
    on and off, FP32 ZNCC: streamed output equals the resident plan's."""
    D = [64, 96, 130, 48][case]
    w, h = 150 + 20 * case, 61 + 7 * case
    left, right, _ = synth.stereo_pair(w, h, D, -4 if case % 2 else 0, seed=300 + case,
                                       gain=case == 3)
    p = sgm.SgmParams(window_w=[9, 7, 5, 7][case], window_h=[7, 5, 5, 7][case], disparities=D,
                      min_disparity=-4 if case % 2 else 0, cost=[0, 2, 1, 4][case],
                      p1=[10, 8, 12, 12][case], p2=[120, 96, 150, 150][case],
                      edge_scale=[32, 0, 32, 32][case], directions=[8, 4, 8, 8][case],
                      aggregate_bytes=[0, 4, 0, 0][case], census_weight=500, ad_truncation=30,
                      gradient_truncation=60, reserved=0, uniqueness=0.03, lr_threshold=1.0)
    want = gpu_context.match_pair(p, left, right)
    for rows in (1, 5, 16, h):
        print("streamed-vs-resident", case, "rows", rows, flush=True)
        gpu_context.plan_mode(sgm.PLAN_STREAMED, rows)
        try:
            got = gpu_context.match_pair(p, left, right)
        finally:
            gpu_context.plan_mode(sgm.PLAN_AUTO)
        assert_same(got, want, "case %d rows %d" % (case, rows))


def test_auto_plan_streams_when_resident_does_not_fit():
    """GpuStereo.plan: a budget below the resident footprint selects the
    streamed plan (strip rows halved until the strip fits); RESIDENT then fails."""
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3)
    p = to_sgm(op)
    h, w = a.shape
    probe = sgm.Context(0)
    plan = probe.plan(p, w, h)
    probe.close()
    budget = plan["pair_bytes"] - 1
    ctx = sgm.Context(0, budget)
    try:
        plan = ctx.plan(p, w, h)
        assert not plan["resident"] and plan["streamed"] and plan["strip_rows"] >= 1
        assert plan["streamed_bytes"] <= budget
        got = ctx.match_pair(p, a, b, ma, mb)
        assert_same(got, O.compute(op, a, b, w, h, ma, mb), "auto-streamed")
        ctx.plan_mode(sgm.PLAN_RESIDENT)
        with pytest.raises(sgm.SgmError):
            ctx.match_pair(p, a, b, ma, mb)
        assert ctx.disparity_limit(p, w, h) < 41
        ctx.plan_mode(sgm.PLAN_AUTO)
        assert ctx.disparity_limit(p, w, h) >= 41
    finally:
        ctx.close()


def test_filtered_entry_point_streamed(gpu_context):
    """The whole CensusSgm.compute chain under the streamed plan equals the oracle."""
    a, b, ma, mb = upgrade_fixture()
    op = O.params(disparities=41, min_disparity=-3)
    f = O.filters()
    h, w = a.shape
    want = O.compute_filtered(op, f, a, b, w, h, ma, mb)
    gpu_context.plan_mode(sgm.PLAN_STREAMED, 7)
    try:
        got = gpu_context.match_pair_filtered(to_sgm(op), sgm.SgmFilterParams(
            **{k: getattr(f, k) for k, _ in sgm.SgmFilterParams._fields_}), a, b, ma, mb)
    finally:
        gpu_context.plan_mode(sgm.PLAN_AUTO)
    assert np.array_equal(got["disparity"].view(np.uint32),
                          np.ascontiguousarray(want["disparity"], np.float32).view(np.uint32))
    assert np.array_equal(got["reasons"], want["reasons"])
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], k
