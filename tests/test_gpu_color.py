"""GPU parity of --census-mode color-distance (sgm_match_pair_color): the
colour-distance census (CensusSgm.colorCensus, CensusSgm.java:302-332) on the
device, then the same matching, L-R check and filters; bit-exact to the
oracle's restatement, with and without the post-L-R filters."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm
from tests.cases import bgr_gray, color_pair
from tests.test_gpu_parity import assert_same, bits, to_sgm

pytestmark = pytest.mark.gpu

CASES = [
    # (w, h, shift, D, minD, window, cost, masks)
    (120, 70, 9, 24, 0, (5, 5), 0, False),
    (150, 80, 13, 41, -3, (9, 7), 0, True),
    (140, 64, 7, 32, 0, (11, 11), 0, False),
    (130, 60, 11, 28, 2, (7, 5), 1, True),
    (128, 72, 5, 20, -2, (5, 5), 2, False),
    (136, 66, 8, 24, 0, (7, 7), 3, False),
]


def inputs(case):
    w, h, shift, D, md, (ww, wh), cost, masks = case
    lb, rb = color_pair(w, h, shift, seed=w * h + D)
    vl = vr = None
    if masks:
        rng = np.random.default_rng(D)
        vl = (rng.random((h, w)) > 0.04).astype(np.uint8)
        vr = np.roll(vl, 2, axis=1)
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=md, cost=cost,
                  lr_threshold=1.0)
    return lb, rb, bgr_gray(lb), bgr_gray(rb), vl, vr, op


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_color_census_pair_bit_exact(gpu_context, case):
    lb, rb, left, right, vl, vr, op = inputs(case)
    h, w = left.shape
    print("colour case", case, flush=True)
    want = O.compute_color(op, left, right, lb, rb, w, h, vl, vr)
    got = gpu_context.match_pair_color(to_sgm(op), left, right, lb, rb, None, vl, vr)
    assert_same(got, want, "colour %r" % (case,))
    assert np.isfinite(want["checked"]).sum() > w * h // 3


@pytest.mark.parametrize("case", CASES[:4], ids=[str(i) for i in range(4)])
def test_color_census_filtered_bit_exact(gpu_context, case):
    lb, rb, left, right, vl, vr, op = inputs(case)
    h, w = left.shape
    f = O.filters(speckle_size=20)
    print("colour filtered case", case, flush=True)
    want = O.compute_filtered(op, f, left, right, w, h, vl, vr, left_bgr=lb, right_bgr=rb)
    got = gpu_context.match_pair_color(
        to_sgm(op), left, right, lb, rb,
        sgm.SgmFilterParams(**{k: getattr(f, k) for k, _ in sgm.SgmFilterParams._fields_}), vl, vr)
    assert np.array_equal(bits(got["disparity"]), bits(want["disparity"]))
    assert np.array_equal(got["reasons"], want["reasons"])
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], k


def test_color_census_streamed(gpu_context):
    lb, rb, left, right, vl, vr, op = inputs(CASES[1])
    h, w = left.shape
    want = O.compute_color(op, left, right, lb, rb, w, h, vl, vr)
    gpu_context.plan_mode(sgm.PLAN_STREAMED, 7)
    try:
        got = gpu_context.match_pair_color(to_sgm(op), left, right, lb, rb, None, vl, vr)
    finally:
        gpu_context.plan_mode(sgm.PLAN_AUTO)
    assert_same(got, want, "colour streamed")


def test_census_sgm_mirror_color_mode():
    lb, rb = color_pair(160, 90, 10, seed=4)
    left, right = bgr_gray(lb), bgr_gray(rb)
    o = sgm.Options(census=5, disparities=32, compute="gpu", censusMode="color-distance")
    r = sgm.CensusSgm(o).compute(left, right, 160, 90, leftBgr=lb, rightBgr=rb)
    op = O.params(window_w=5, window_h=5, disparities=32, p1=sgm.StereoCost.p1(o),
                  p2=sgm.StereoCost.p2(o), lr_threshold=o.lrThreshold)
    want = O.compute_filtered(op, O.filters(), left, right, 160, 90, left_bgr=lb, right_bgr=rb)
    assert np.array_equal(bits(r.disparity), bits(want["disparity"]))
    with pytest.raises(ValueError):
        sgm.CensusSgm(o).compute(left, right, 160, 90)
    sgm.GpuSgm.closeSession()
