"""GPU parity of the post-L-R filters (sgm_match_pair_filtered) against the
oracle's CensusSgm.computeRecorded through 05_final_stereo: final disparity
bit-exact, reason codes and every counter equal."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth

pytestmark = pytest.mark.gpu


def to_sgm(op):
    return sgm.SgmParams(**{f: getattr(op, f) for f, _ in sgm.SgmParams._fields_})


def to_filters(f):
    return sgm.SgmFilterParams(**{k: getattr(f, k) for k, _ in sgm.SgmFilterParams._fields_})


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


CASES = [
    # (w, h, D, minD, window, cost, masks, filter overrides)
    (160, 100, 48, 0, (5, 5), 0, False, {}),
    (200, 120, 64, -4, (9, 7), 0, True, {}),
    (180, 90, 40, 2, (7, 7), 1, False, dict(median=5, median_classic=1)),
    (220, 110, 56, 0, (9, 7), 0, False, dict(speckle_connectivity=4, speckle_size=20)),
    (150, 80, 32, 0, (5, 5), 3, False, dict(median=1, correlation=0.3)),
    (170, 95, 48, -2, (5, 3), 2, True, dict(median=7, median_edge_threshold=0,
                                             median_max_change=1.5, speckle_size=0)),
    (240, 130, 64, 0, (9, 7), 0, False, dict(correlation=-2.0, speckle_range=0.5,
                                             correlation_window=5)),
    # --subpixel-refinement zncc (PhotometricRefinement before the L-R check)
    (160, 100, 48, 0, (5, 5), 0, False, dict(subpixel_refinement=1)),
    (200, 120, 64, -4, (9, 7), 1, True, dict(subpixel_refinement=1, correlation=0.3)),
    (180, 90, 40, 2, (7, 7), 0, False, dict(subpixel_refinement=1, correlation_window=5,
                                             median=1, speckle_size=0)),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_filtered_output_bit_exact(gpu_context, case):
    w, h, D, md, (ww, wh), cost, masks, fo = case
    left, right, _ = synth.stereo_pair(w, h, D, md, seed=w + h)
    vl = vr = None
    if masks:
        rng = np.random.default_rng(D)
        vl = (rng.random((h, w)) > 0.03).astype(np.uint8)
        vr = np.roll(vl, 1, axis=0)
    op = O.params(window_w=ww, window_h=wh, disparities=D, min_disparity=md, cost=cost,
                  lr_threshold=1.0)
    f = O.filters(**fo)
    print("filtered case", case, flush=True)
    want = O.compute_filtered(op, f, left, right, w, h, vl, vr)
    got = gpu_context.match_pair_filtered(to_sgm(op), to_filters(f), left, right, vl, vr)
    g, e = bits(got["disparity"]), bits(want["disparity"])
    if not np.array_equal(g, e):
        bad = np.nonzero(g != e)[0]
        raise AssertionError("%d pixels differ, first %d: gpu %r oracle %r" % (
            len(bad), bad[0], got["disparity"][bad[0]], want["disparity"][bad[0]]))
    assert np.array_equal(got["reasons"], want["reasons"])
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], (k, got["stats"][k], want["stats"][k])


def test_filtered_workload_640x480(gpu_context):
    """The plain 640x480 workload (census 9x7, D=64, P1=10, P2=120) with the
    reference's default filters."""
    left, right, _ = synth.stereo_pair(640, 480, 64, 0, seed=1)
    op = O.params(window_w=9, window_h=7, disparities=64, p1=10, p2=120, lr_threshold=1.0)
    f = O.filters()
    want = O.compute_filtered(op, f, left, right, 640, 480)
    got = gpu_context.match_pair_filtered(to_sgm(op), to_filters(f), left, right)
    assert np.array_equal(bits(got["disparity"]), bits(want["disparity"]))
    assert np.array_equal(got["reasons"], want["reasons"])
    for k in O.STAT_FIELDS:
        assert got["stats"][k] == want["stats"][k], k
    assert got["stats"]["speckle_rejected"] > 0 and got["stats"]["photometric_rejected"] > 0


def test_census_sgm_mirror_runs_the_filters(gpu_context):
    """CensusSgm.compute with reference option names equals the oracle chain."""
    left, right, _ = synth.stereo_pair(200, 120, 48, 0, seed=4)
    o = sgm.Options(census=5, disparities=48, p1=4, p2=24, compute="gpu", lrThreshold=1.0,
                    speckleSize=30, medianMode="classic")
    r = sgm.CensusSgm(o).compute(left, right, 200, 120)
    op = O.params(window_w=5, window_h=5, disparities=48, lr_threshold=1.0)
    want = O.compute_filtered(op, O.filters(speckle_size=30, median_classic=1), left, right, 200, 120)
    assert np.array_equal(bits(r.disparity), bits(want["disparity"]))
    assert r.stats.valid == want["stats"]["valid"]


def test_subpixel_refinement_on_the_analytic_plane(gpu_context):
    """SubpixelRegression (SubpixelRegression.java:11-37) on the device: the
    refined map equals the oracle bit for bit, and against the known plane
    disparity 13.37 refinement lowers the RMS error below 0.8x the parabola
    fit's with coverage above 0.8x."""
    from tests.cases import subpixel_texture
    from tests.test_oracle_filters import _plane_metrics
    w, h, d = 384, 192, 13.37
    a, b = subpixel_texture(w, h, d)
    op = O.params(disparities=32)
    metrics = []
    for refine in (0, 1):
        f = O.filters(median=1, speckle_size=0, correlation=0.5, subpixel_refinement=refine)
        got = gpu_context.match_pair_filtered(to_sgm(op), to_filters(f), a, b)
        want = O.compute_filtered(op, f, a, b, w, h)
        assert np.array_equal(bits(got["disparity"]), bits(want["disparity"])), refine
        metrics.append(_plane_metrics(got["disparity"], w, h, d))
    old, new = metrics
    assert new[0] >= 0.8 * old[0] and new[1] < 0.8 * old[1], (old, new)
