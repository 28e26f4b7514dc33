"""Writes the golden fixtures (tests/golden/*.npz) from the C oracle.

The reference holds no SGM golden vectors (its only stored hash covers unshipped
inputs), so these fixtures are produced by the pinned oracle on the reference's
regenerated known-answer inputs and on seeded synthetic pairs.  Run from the repo
root: python tests/golden/make_golden.py

`--large` writes the full-size fixtures (tests/golden/large_*.json): pairs whose
volumes exceed 4 GiB, too large to store as maps, so the fixture holds the
generator arguments, SHA-256 hashes of the input rasters and of every output
map, and the counters, all from the oracle split over host threads (which
equals the single-threaded oracle bit for bit).
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pose_amd import synth  # noqa: E402
from tests.cases import shifted, upgrade_fixture  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    a, b = shifted(4, 96, 48, 6)
    yield "selftest_96x48_census5", a, b, None, None, dict(disparities=16, uniqueness=.02)
    ua, ub, ma, mb = upgrade_fixture()
    for cost, dirs in [(0, 8), (1, 4), (2, 8), (3, 4)]:
        yield ("upgrade_192x79_cost%d_dirs%d" % (cost, dirs), ua, ub, ma, mb,
               dict(disparities=41, min_disparity=-3, cost=cost, directions=dirs))
    l, r, _ = synth.stereo_pair(160, 120, 64, 0, seed=11)
    yield ("synthetic_160x120_census9x7_d64", l, r, None, None,
           dict(window_w=9, window_h=7, disparities=64, p1=10, p2=120, lr_threshold=1.0))
    l, r, _ = synth.stereo_pair(140, 90, 40, -8, seed=12)
    yield ("synthetic_140x90_census7x5_signed_edge0", l, r, None, None,
           dict(window_w=7, window_h=5, disparities=40, min_disparity=-8, p1=6, p2=70,
                edge_scale=0, lr_threshold=1.0))
    l, r, _ = synth.stereo_pair(128, 96, 48, 0, seed=13)
    yield ("synthetic_128x96_agg32", l, r, None, None,
           dict(window_w=5, window_h=5, disparities=48, p1=40, p2=9000, lr_threshold=1.0))
    l, r, _ = synth.stereo_pair(120, 80, 32, 0, seed=14, gain=True)
    yield ("synthetic_120x80_zncc1024_fp64", l, r, None, None,
           dict(window_w=7, window_h=7, disparities=32, cost=4, p1=12, p2=150, lr_threshold=1.0))


def large_cases():
    # 6000x700 with 512 disparities: a uint16 aggregate volume of 4.3 GB per
    # matching direction (pixel offsets past 2^32 bytes), census 9x7, the
    # served stream's parameters
    yield ("large_6000x700_census9x7_d512", dict(width=6000, height=700, disparities=512,
                                                 min_disparity=0, seed=501),
           dict(window_w=9, window_h=7, disparities=512, p1=10, p2=120, lr_threshold=1.0))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main_large():
    for name, gen, kw in large_cases():
        left, right, _ = synth.stereo_pair(gen["width"], gen["height"], gen["disparities"],
                                           gen["min_disparity"], seed=gen["seed"])
        p = O.params(**kw)
        w, h = gen["width"], gen["height"]
        res = O.compute_parallel(p, left, right, w, h)
        out = {"generator": gen, "params": {f: getattr(p, f) for f, _ in O.OracleParams._fields_},
               "input_sha256": {"left": sha(left), "right": sha(right)},
               "output_sha256": {k: sha(res[k].view(np.uint32) if res[k].dtype == np.float32
                                        else res[k]) for k in
                                 ("checked", "dl", "dr", "reasons", "boundary")},
               "stats": res["stats"]}
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(out, fh, indent=1)
        print(name, res["stats"]["valid"], "valid")


def main():
    if "--large" in sys.argv:
        main_large()
        return
    for name, left, right, vl, vr, kw in cases():
        p = O.params(**kw)
        h, w = left.shape
        res = O.compute(p, left, right, w, h, vl, vr)
        names = [f for f, _ in O.OracleParams._fields_]
        values = np.array([getattr(p, f) for f in names], np.float64)
        extra = {}
        if vl is not None:
            extra = dict(valid_left=vl, valid_right=vr)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), left=left, right=right,
                            param_names=np.array(names), param_values=values,
                            checked=res["checked"], dl=res["dl"], dr=res["dr"],
                            reasons=res["reasons"], boundary=res["boundary"],
                            stats=np.array([res["stats"][k] for k in O.STAT_FIELDS], np.int64),
                            **extra)
        print(name, res["stats"]["valid"], "valid")


if __name__ == "__main__":
    main()
