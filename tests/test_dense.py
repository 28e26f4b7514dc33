"""The caller-side disparity search rules (DisparityRange.bounds / expansion,
CensusSgm.Stats pressure, DenseReconstruction stop reasons) and the PFM output,
checked against values worked out by hand from the reference's formulas."""
import os
import tempfile

import numpy as np
import pytest

from pose_amd import dense, sgm


def opts(**kw):
    return sgm.Options(**kw)


def stats(lower=0, upper=0, full=1000, valid=0):
    return dict(lower_boundary_rejected=lower, upper_boundary_rejected=upper,
                full_interval_pixels=full, valid=valid)


def test_bounds():
    o = opts()
    assert dense.bounds(5, 100, 200, o) == (5, 96)
    assert dense.bounds(10, 11, 200, o) == (10, 16)          # span 3 -> at least 16
    assert dense.bounds(-5, 300, 200, o) == (0, 200)          # clipped to [minDisparity, w-1]
    assert dense.bounds(190, 195, 200, o) == (190, 10)        # count capped at width - lo
    assert dense.bounds(5, 100, 200, opts(minDisparity=3)) == (3, 98)  # pinned minimum
    with pytest.raises(ValueError):
        dense.bounds(198, 199, 200, o)
    with pytest.raises(ValueError):
        dense.bounds(0, 5000, 6000, o)


def test_pressure_and_clipping():
    assert not dense.upper_clipped(stats(upper=23, full=100))     # needs >= 24 rejections
    assert dense.upper_clipped(stats(upper=24, full=100))
    assert not dense.upper_clipped(stats(upper=30, full=2000))    # 1.5% <= 2%
    assert dense.lower_pressure(stats(lower=5, full=0)) == 5.0    # max(1, full)


def test_expansion():
    o = opts()
    cur = opts(minDisparity=0, disparities=32)
    # upper pressure: grow = max(16, ceil16(32 / 2)) = 16 -> [0, 47]
    assert dense.expansion(o, cur, stats(upper=100), 200, 0) == (0, 48)
    cur = opts(minDisparity=0, disparities=100)
    # grow = ceil16(50) = 64 -> [0, 163]
    assert dense.expansion(o, cur, stats(upper=100), 200, 0) == (0, 164)
    # lower pressure from a raised minimum (not pinned by the original options)
    cur = opts(minDisparity=40, disparities=32)
    assert dense.expansion(o, cur, stats(lower=100), 200, 1) == (24, 48)
    # the fourth attempt opens the whole physical range
    assert dense.expansion(o, cur, stats(upper=100), 200, 3) == (0, 200)
    # no pressure, explicit ranges, an interval already at the image width
    assert dense.expansion(o, cur, stats(), 200, 0) is None
    assert dense.expansion(opts(disparitiesExplicit=True), cur, stats(upper=100), 200, 0) is None
    assert dense.expansion(o, opts(minDisparity=0, disparities=200), stats(upper=100), 200, 0) is None
    # a pinned minimum is never lowered
    pinned = opts(minDisparity=40, minDisparityExplicit=True)
    assert dense.expansion(pinned, cur, stats(lower=100), 200, 0) is None


def test_stop_reasons():
    o = opts()
    assert dense.stop_reason(opts(disparitiesExplicit=True), stats(upper=100), None, 64) == "explicit"
    assert dense.stop_reason(o, stats(), None, 64) == "no_boundary_pressure"
    assert dense.stop_reason(o, stats(upper=100), None, 64) == "physical_or_user_boundary"
    assert dense.stop_reason(o, stats(upper=100), dense.Bounds(0, 96), 64) == "memory_limit"
    assert dense.stop_reason(o, stats(upper=100), dense.Bounds(0, 48), 64) == "retry_limit"


def test_pfm_round_trip_and_layout():
    rng = np.random.default_rng(1)
    d = rng.random((5, 7)).astype(np.float32) * 40
    d[1, 2] = np.nan
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "pair.pfm")
        dense.write_pfm(path, 7, 5, d.reshape(-1))
        raw = open(path, "rb").read()
        assert raw.startswith(b"Pf\n7 5\n-1.0\n")
        body = np.frombuffer(raw[len(b"Pf\n7 5\n-1.0\n"):], "<f4")
        assert body.size == 35
        assert np.array_equal(body[:7], d[4])                 # first stored row = bottom row
        w, h, back = dense.read_pfm(path)
    assert (w, h) == (7, 5)
    assert np.array_equal(back.view(np.uint32), d.view(np.uint32))


def test_pair_range_regression_expansion_known_answers():
    """PairRangeRegression.expansion (PairRangeRegression.java:158-208): the
    reference's own assertions on DisparityRange.expansion, same inputs."""
    o = opts()
    current = opts(minDisparity=450, disparities=64)
    st = stats(lower=100, full=1000)
    nxt = dense.expansion(o, current, st, 1000, 0)
    assert nxt is not None and nxt.minimum < 450 and nxt.maximum() >= 513   # lower pressure
    st = stats(lower=0, upper=100, full=1000)
    nxt = dense.expansion(o, current, st, 1000, 0)
    assert nxt.minimum == 450 and nxt.maximum() > 513                      # upper pressure
    st = stats(lower=100, upper=100, full=1000)
    nxt = dense.expansion(o, current, st, 1000, 0)
    assert nxt.minimum < 450 and nxt.maximum() > 513                       # both together
    nxt = dense.expansion(o, current, st, 1000, 3)
    assert nxt.minimum == 0 and nxt.maximum() == 999                       # whole-domain fallback
    pinned = opts(minDisparity=450, minDisparityExplicit=True)
    assert dense.expansion(pinned, current, stats(lower=100, upper=0, full=1000), 1000, 0) is None
    fixed = opts(disparitiesExplicit=True)
    assert dense.expansion(fixed, current, stats(lower=100, upper=100, full=1000), 1000, 0) is None
    # incomplete border intervals (no full-interval pixels) cannot trigger a retry
    border = dict(stats(full=0), boundary_rejected=100000)
    assert dense.expansion(o, current, border, 1000, 0) is None
    # sparse endpoint noise does not cause retries
    assert dense.expansion(o, current, stats(lower=3, upper=3, full=1000), 1000, 0) is None


def test_pair_range_regression_bounds_known_answers():
    """PairRangeRegression.java:138-155: an unusable physical minimum and an
    observed span over 4096 hypotheses fail instead of overflowing or
    truncating."""
    with pytest.raises(ValueError):
        dense.bounds(999, 999, 1000, opts(minDisparity=999))
    with pytest.raises(ValueError):
        dense.bounds(0, 8000, 10000, opts())
