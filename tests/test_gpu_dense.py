"""Batched disparity search (pose_amd.dense.search): pairs whose initial
interval misses their disparity grow it by the reference's expansion rule
until the boundary pressure is gone, every attempt running all searching pairs
in one batch; each final result equals the plain path on the final interval."""
import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import dense, sgm
from tests.cases import shifted
from tests.test_gpu_parity import bits

pytestmark = pytest.mark.gpu


def test_batched_search_expands_and_matches_plain(gpu_context):
    o = sgm.Options(census=5, lrThreshold=1.0)
    pairs = []
    for i, (w, h, shift, lo, count) in enumerate([(200, 60, 30, 0, 24), (180, 50, 10, 0, 32),
                                                  (220, 70, 60, 0, 32)]):
        a, b = shifted(200 + i, w, h, shift)
        pairs.append(dict(left=a, right=b, minimum=lo, count=count, shift=shift))
    print("dense search", [(p["shift"], p["count"]) for p in pairs], flush=True)
    out = dense.search(gpu_context, o, pairs, graph=True)
    params = sgm.params_from_options(o)
    for p, r in zip(pairs, out):
        print("pair", p["shift"], r["stop_reason"], [(a["min_disparity"], a["disparities"])
                                                     for a in r["attempts"]], flush=True)
        w = p["left"].shape[1]
        # every retry is the reference's expansion of the previous attempt's counters
        for prev, nxt in zip(r["attempts"], r["attempts"][1:]):
            cur = o.copy()
            cur.minDisparity, cur.disparities = prev["min_disparity"], prev["disparities"]
            assert dense.expansion(o, cur, prev, w, prev["attempt"]) == (
                nxt["min_disparity"], nxt["disparities"])
        last = r["attempts"][-1]
        assert (last["min_disparity"], last["disparities"]) == (r["minimum"], r["count"])
        # the final attempt equals the plain path on the final interval
        params.min_disparity, params.disparities = r["minimum"], r["count"]
        plain = gpu_context.match_pair(params, p["left"], p["right"])
        assert np.array_equal(bits(plain["checked"]), bits(r["checked"]))
        assert plain["stats"] == r["stats"]
    # a pair inside its first interval stops at once; one just beyond it grows to cover it
    assert len(out[1]["attempts"]) == 1 and out[1]["stop_reason"] == "no_boundary_pressure"
    assert out[0]["minimum"] <= 30 <= out[0]["minimum"] + out[0]["count"] - 1
    d = out[0]["checked"]
    assert (np.abs(d[np.isfinite(d)] - 30) < .6).mean() > .95


def test_explicit_range_is_not_expanded(gpu_context):
    o = sgm.Options(census=5, lrThreshold=1.0, disparitiesExplicit=True)
    a, b = shifted(300, 160, 40, 30)
    out = dense.search(gpu_context, o, [dict(left=a, right=b, minimum=0, count=16)])
    assert out[0]["stop_reason"] == "explicit" and len(out[0]["attempts"]) == 1


def oracle_params(o, minimum, count):
    p = sgm.params_from_options(o)
    p.min_disparity, p.disparities = minimum, count
    return O.params(**{f: getattr(p, f) for f, _ in O.OracleParams._fields_})


def test_search_attempts_equal_the_oracle(gpu_context):
    """Every attempt of the batched search is the oracle's match on that
    attempt's interval: same boundary counters (which drive the expansion) and,
    for the final attempt, the same L-R-validated map bit for bit."""
    o = sgm.Options(census=5, lrThreshold=1.0)
    specs = [(200, 60, 30, 0, 24), (170, 64, 45, 0, 16), (190, 56, 12, 20, 24)]
    pairs = []
    for i, (w, h, shift, lo, count) in enumerate(specs):
        a, b = shifted(500 + i, w, h, shift)
        pairs.append(dict(left=a, right=b, minimum=lo, count=count))
    out = dense.search(gpu_context, o, pairs, graph=False)
    retried = 0
    for p, r in zip(pairs, out):
        h, w = p["left"].shape
        retried += len(r["attempts"]) > 1
        for att in r["attempts"]:
            print("attempt", w, h, att["min_disparity"], att["disparities"], flush=True)
            want = O.compute(oracle_params(o, att["min_disparity"], att["disparities"]),
                             p["left"], p["right"], w, h)
            for k in ("lower_boundary_rejected", "upper_boundary_rejected",
                      "full_interval_pixels"):
                assert att[k] == want["stats"][k], (att, k)
            assert att["final_valid"] == want["stats"]["valid"]
        assert np.array_equal(bits(r["checked"]), bits(want["checked"]))
        assert np.array_equal(r["reasons"], want["reasons"])
        assert np.array_equal(r["boundary"], want["boundary"])
    assert retried >= 2  # the lower and the upper expansion both ran


@pytest.mark.parametrize("low", [True, False])
def test_pair_range_regression_boundary_on_gpu(gpu_context, low):
    """PairRangeRegression.boundary (PairRangeRegression.java:210-249): with a
    clipped range the GPU's lower/upper/eligible counters equal the CPU's, and
    the per-pixel boundary flags equal the counters."""
    w, h = 160, 96
    a, b = shifted(497, w, h, 8)
    o = sgm.Options(minDisparity=8 if low else 3, disparities=17 if low else 6, speckleSize=0)
    p = sgm.params_from_options(o)
    got = gpu_context.match_pair(p, a, b)
    want = O.compute(O.params(**{f: getattr(p, f) for f, _ in O.OracleParams._fields_}),
                     a, b, w, h)
    st = got["stats"]
    for k in ("lower_boundary_rejected", "upper_boundary_rejected", "full_interval_pixels"):
        assert st[k] == want["stats"][k], k
    assert st["full_interval_pixels"] > 5000
    assert dense.lower_clipped(st) if low else dense.upper_clipped(st)
    flag = 1 if low else 2
    hits = int(((got["boundary"] & flag) != 0).sum())
    assert hits == (st["lower_boundary_rejected"] if low else st["upper_boundary_rejected"])
