"""Batched disparity search (pose_amd.dense.search): pairs whose initial
interval misses their disparity grow it by the reference's expansion rule
until the boundary pressure is gone, every attempt running all searching pairs
in one batch; each final result equals the plain path on the final interval."""
import numpy as np
import pytest

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
