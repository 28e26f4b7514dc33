"""The oracle split over host threads equals the literal single-threaded
restatement bit for bit: every cost, 4 and 8 paths, 16- and 32-bit aggregates,
signed ranges, rectangular windows, masks, edge scale 0, and the FP64
ZNCC-1024 cost.  The parallel oracle is the checker of the full-size GPU tests
and the all-cores CPU baseline, so it has to be the same function."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import synth
from tests.cases import shifted, upgrade_fixture


def same(a, b):
    for k in ("checked", "dl", "dr"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert np.array_equal(a["reasons"], b["reasons"])
    assert np.array_equal(a["boundary"], b["boundary"])
    assert a["stats"] == b["stats"]


CASES = [
    dict(window_w=9, window_h=7, disparities=64, p1=10, p2=120, lr_threshold=1.0),
    dict(window_w=5, window_h=5, disparities=33, min_disparity=-9, cost=1, directions=4),
    dict(window_w=7, window_h=5, disparities=48, cost=2, edge_scale=0),
    dict(window_w=5, window_h=5, disparities=40, cost=3, directions=8),
    dict(window_w=7, window_h=7, disparities=48, cost=4, p1=12, p2=150),
    dict(window_w=9, window_h=7, disparities=70, p1=10, p2=120, aggregate_bytes=4),
    dict(window_w=11, window_h=9, disparities=21, min_disparity=3, p1=20, p2=300),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("threads", [0, 3])
def test_parallel_equals_literal_on_synthetic_pairs(case, threads):
    kw = CASES[case]
    op = O.params(**kw)
    left, right, _ = synth.stereo_pair(150, 90, kw["disparities"], kw.get("min_disparity", 0),
                                       seed=40 + case, gain=kw.get("cost") == 4)
    vl = (np.random.default_rng(case).random((90, 150)) > 0.05).astype(np.uint8)
    vr = (np.random.default_rng(case + 100).random((90, 150)) > 0.05).astype(np.uint8)
    for masks in ((None, None), (vl, vr)):
        same(O.compute_parallel(op, left, right, 150, 90, *masks, threads=threads),
             O.compute(op, left, right, 150, 90, *masks))


@pytest.mark.parametrize("cost", [0, 1, 2, 3])
def test_parallel_equals_literal_on_the_reference_fixture(cost):
    a, b, ma, mb = upgrade_fixture()
    op = O.params(window_w=5, window_h=5, disparities=32, cost=cost)
    same(O.compute_parallel(op, a, b, 192, 79, ma, mb), O.compute(op, a, b, 192, 79, ma, mb))


def test_parallel_ties_and_tiny_images():
    # random bytes with a shift: many equal aggregates (first-argmin ties)
    left, right = shifted(7, 24, 17, 3)
    for D in (3, 5, 12):
        op = O.params(window_w=3, window_h=3, disparities=D, p1=1, p2=2)
        same(O.compute_parallel(op, left, right, 24, 17), O.compute(op, left, right, 24, 17))
    assert O.threads() >= 1
