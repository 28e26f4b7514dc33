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
    f = O.filters(speckle_size=20, subpixel_refinement=CASES.index(case) % 2)
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


@pytest.mark.parametrize("window", [(5, 5), (9, 7), (11, 11)])
def test_match_with_caller_descriptors(gpu_context, window):
    """sgm_match_desc = GpuSgm.match with the caller's long[][] descriptors: the
    colour-distance planes and textures computed on the host give the oracle's
    single-direction result in both matching directions."""
    ww, wh = window
    w, h = 150, 70
    lb, rb = color_pair(w, h, 12, seed=ww * wh)
    a, b = bgr_gray(lb), bgr_gray(rb)
    op = O.params(window_w=ww, window_h=wh, disparities=33, min_disparity=-2)
    p = to_sgm(op)
    da, ta = O.color_census(lb, w, h, ww, wh)
    db, tb = O.color_census(rb, w, h, ww, wh)
    for reverse in (0, 1):
        A, B, TA, TB, DA, DB = (a, b, ta, tb, da, db) if not reverse else (b, a, tb, ta, db, da)
        print("descriptors", window, "reverse", reverse, flush=True)
        disp_o, st_o, r_o, b_o = O.match(op, DA, DB, A.reshape(-1), B.reshape(-1), TA, TB, w, h,
                                         reverse)
        reasons = np.ones(w * h, np.uint8)
        boundary = np.zeros(w * h, np.uint8)
        disp, st = gpu_context.match_desc(p, DA, DB, A, B, TA, TB, w, h, reverse, None, reasons,
                                          boundary)
        assert np.array_equal(bits(disp), bits(disp_o))
        assert np.array_equal(reasons, r_o) and np.array_equal(boundary, b_o)
        for k in ("lower_boundary_rejected", "full_interval_pixels", "boundary_rejected",
                  "upper_boundary_rejected", "ambiguity_rejected", "selected"):
            assert getattr(st, k) == st_o[k], k
    # gray census planes through the descriptor entry equal the gray entry
    dga, dgb = O.census(a, w, h, ww, wh), O.census(b, w, h, ww, wh)
    tga, tgb = O.texture(a, w, h, ww, wh), O.texture(b, w, h, ww, wh)
    d1, _ = gpu_context.match_desc(p, dga, dgb, a, b, tga, tgb, w, h, 0)
    d2, _ = gpu_context.match(p, a, b, tga, tgb, w, h, 0)
    assert np.array_equal(bits(d1), bits(d2))


def test_gpusgm_match_color_mode_uses_caller_descriptors():
    w, h = 160, 80
    lb, rb = color_pair(w, h, 9, seed=77)
    a, b = bgr_gray(lb), bgr_gray(rb)
    o = sgm.Options(census=5, disparities=24, compute="gpu", censusMode="color-distance")
    da, ta = O.color_census(lb, w, h, 5, 5)
    db, tb = O.color_census(rb, w, h, 5, 5)
    st = sgm.Stats()
    disp = sgm.GpuSgm.match(o, da, db, a, b, ta, tb, w, h, False, st, None, None)
    op = O.params(window_w=5, window_h=5, disparities=24, p1=sgm.StereoCost.p1(o),
                  p2=sgm.StereoCost.p2(o))
    want, st_o, _, _ = O.match(op, da, db, a.reshape(-1), b.reshape(-1), ta, tb, w, h, 0)
    assert np.array_equal(bits(disp), bits(want))
    assert st.selected == st_o["selected"]
    sgm.GpuSgm.closeSession()
