"""The workloads at their stated sizes, on the GPU, against the oracle.

* A 6000x700 pair with 512 disparities (uint16 aggregate volume of 4.3 GB per
  matching direction, pixel offsets past 2^32 bytes) against the oracle's golden
  hashes (tests/golden/large_*.json), under the resident, strip-streamed and
  compact plans: three different addressings of the same volumes.
* The served stream (70% batched-distribution pairs, 25% 1600x1067 D=256, 5%
  6000x4000 D=512, 250 GB budget) through the lane server and the batching
  server: every result equals the pair alone through the plain path; the
  6000x4000 D=512 request also equals the compact plan and meets full-size
  properties (agreement with the generator's truth, counter identities); the
  smaller requests equal the oracle.
* FP32 ZNCC at the full 1600x1067 D=192 workload against the FP64 oracle.
* Pairs drawn from the batched workload's distribution (long side >= 800,
  min disparity in [-32, 32]) against the oracle, plain and batched.
"""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pose_amd import sgm, synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CENSUS97 = dict(window_w=9, window_h=7, p1=10, p2=120, lr_threshold=1.0)
BIG = 1 << 62  # the largest budget the library allows (95% of device memory)


def to_sgm(op):
    return sgm.SgmParams(**{f: getattr(op, f) for f, _ in sgm.SgmParams._fields_})


def census_params(D, md=0):
    return O.params(**dict(CENSUS97, disparities=D, min_disparity=md))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def same(got, want, what, keys=("checked", "dl", "dr")):
    for k in keys:
        g, w = bits(got[k]), bits(want[k])
        if not np.array_equal(g, w):
            bad = np.nonzero(g != w)[0]
            raise AssertionError("%s %s: %d pixels differ, first %d" % (what, k, len(bad), bad[0]))
    for k in ("selected", "boundary_rejected", "ambiguity_rejected", "lr_rejected", "valid",
              "lower_boundary_rejected", "upper_boundary_rejected", "full_interval_pixels"):
        assert got["stats"][k] == want["stats"][k], (what, k, got["stats"][k], want["stats"][k])


def identities(r, what):
    """Counter identities of one L-R-checked result (CensusSgm.java:181-201, 460-515)."""
    st, checked, dl = r["stats"], r["checked"], r["dl"]
    reasons, boundary = r["reasons"], r["boundary"]
    assert np.isfinite(dl).sum() == st["selected"], what
    assert np.isfinite(checked).sum() == st["valid"], what
    assert (reasons == 0).sum() == st["valid"], what
    assert (reasons == 2).sum() == st["boundary_rejected"], what
    assert (reasons == 3).sum() == st["ambiguity_rejected"], what
    # every selected pixel is either valid or rejected by the L-R check (reason 4)
    assert (reasons == 4).sum() == st["selected"] - st["valid"], what
    assert (reasons == 4).sum() >= st["lr_rejected"], what
    assert ((boundary & 4) != 0).sum() == st["full_interval_pixels"], what
    assert ((boundary & 1) != 0).sum() == st["lower_boundary_rejected"], what
    assert ((boundary & 2) != 0).sum() == st["upper_boundary_rejected"], what
    assert st["lower_boundary_rejected"] + st["upper_boundary_rejected"] <= st["boundary_rejected"]


def truth_agreement(checked, disparity, w, h):
    """Fraction of valid left pixels within 1 of the generator's left-view truth."""
    truth = synth.left_view(disparity)
    c = checked.reshape(h, w)
    ok = np.isfinite(c) & np.isfinite(truth)
    return float((np.abs(c[ok] - truth[ok]) <= 1).mean()), int(ok.sum())


def run_plan(mode, op, left, right, w, h, rows=64):
    ctx = sgm.Context(0, budget_bytes=BIG)
    try:
        if mode == "streamed":
            ctx.plan_mode(sgm.PLAN_STREAMED, rows)
        elif mode == "compact":
            ctx.plan_mode(sgm.PLAN_ESGM, rows)
        print("plan", mode, w, h, op.disparities, flush=True)
        return ctx.match_pair(to_sgm(op), left, right)
    finally:
        ctx.close()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "large_*.json"))))
def test_large_golden_every_plan(path):
    g = json.load(open(path))
    gen = g["generator"]
    w, h = gen["width"], gen["height"]
    left, right, disparity = synth.stereo_pair(w, h, gen["disparities"], gen["min_disparity"],
                                               seed=gen["seed"])
    assert sha(left) == g["input_sha256"]["left"] and sha(right) == g["input_sha256"]["right"]
    op = O.params(**g["params"])
    # the aggregate volume of one matching direction passes 2^32 bytes
    ctx = sgm.Context(0, budget_bytes=BIG)
    plan = ctx.plan(to_sgm(op), w, h)
    ctx.close()
    assert w * h * plan["padded_disparities"] * plan["aggregate_bytes"] > (1 << 32)
    for mode in ("resident", "streamed", "compact"):
        r = run_plan(mode, op, left, right, w, h)
        for k, want in g["output_sha256"].items():
            got = sha(bits(r[k]) if r[k].dtype == np.float32 else r[k])
            assert got == want, (os.path.basename(path), mode, k)
        assert r["stats"] == g["stats"], (mode, r["stats"], g["stats"])
        identities(r, mode)
    agree, count = truth_agreement(r["checked"], disparity, w, h)
    print("truth agreement", agree, count, flush=True)
    assert agree >= 0.9


def served_mix():
    """One request of each class of the served stream: two pairs of the batched
    distribution (the smallest of its first dozen), one 1600x1067 D=256 and one
    6000x4000 D=512, in arrival order small, big, 1600, small."""
    spec = sorted(synth.batch_pairs(count=12, seed=7), key=lambda t: t[0] * t[1] * t[2])[:2]
    out = []
    for (w, h, D, md, seed) in spec:
        left, right, disp = synth.stereo_pair(w, h, D, md, seed=seed)
        out.append(dict(cls="mixed", w=w, h=h, D=D, md=md, left=left, right=right, disp=disp))
    for (w, h, D, seed, cls) in ((6000, 4000, 512, 7200, "6000x4000"),
                                 (1600, 1067, 256, 7100, "1600x1067")):
        left, right, disp = synth.stereo_pair(w, h, D, 0, seed=seed)
        out.insert(1, dict(cls=cls, w=w, h=h, D=D, md=0, left=left, right=right, disp=disp))
    return out


def test_served_stream_at_scale():
    mix = served_mix()
    assert [m["cls"] for m in mix] == ["mixed", "1600x1067", "6000x4000", "mixed"]
    # each pair alone through the plain path
    ctx = sgm.Context(0, budget_bytes=BIG)
    alone = []
    try:
        for m in mix:
            print("alone", m["cls"], m["w"], m["h"], m["D"], flush=True)
            alone.append(ctx.match_pair(to_sgm(census_params(m["D"], m["md"])), m["left"],
                                        m["right"]))
            identities(alone[-1], m["cls"])
    finally:
        ctx.close()
    big = mix[2]
    agree, count = truth_agreement(alone[2]["checked"], big["disp"], big["w"], big["h"])
    print("6000x4000 truth agreement", agree, count, flush=True)
    assert agree >= 0.9 and count > 0.7 * big["w"] * big["h"]
    # the large request under the compact plan: no aggregate volume, same result
    compact = run_plan("compact", census_params(512), big["left"], big["right"], big["w"], big["h"])
    same(compact, alone[2], "6000x4000 compact", keys=("checked", "dl", "dr", "reasons", "boundary"))
    # through both servers, all requests in flight together, 250 GB budget
    base = to_sgm(census_params(64))
    for max_batch, streams in ((1, 3), (8, 2)):
        sctx = sgm.Context(0, budget_bytes=BIG)
        srv = sgm.Server(sctx, base, budget_bytes=250 * 10**9, max_batch=max_batch,
                         streams=streams)
        try:
            print("server max_batch", max_batch, flush=True)
            tickets = [srv.submit(m["left"], m["right"], m["D"], m["md"]) for m in mix]
            for m, t, a in zip(mix, tickets, alone):
                got = srv.wait(t, timeout=300.0)
                same(got, a, "served %s max_batch %d" % (m["cls"], max_batch))
                assert got["info"]["pool_bytes"] <= 250 * 10**9
        finally:
            srv.close()
            sctx.close()
    # the smaller requests against the oracle
    for m, a in zip(mix, alone):
        if m["cls"] == "6000x4000":
            continue
        print("oracle", m["cls"], m["w"], m["h"], m["D"], flush=True)
        want = O.compute_parallel(census_params(m["D"], m["md"]), m["left"], m["right"], m["w"],
                                  m["h"])
        same(a, want, "oracle %s" % m["cls"], keys=("checked", "dl", "dr", "reasons", "boundary"))


def test_zncc_fp32_full_workload_against_fp64_oracle():
    """The FP32 ZNCC workload at its size (1600x1067, 7x7, D=192, P1=12, P2=150,
    gain field 0.8-1.2): at most 0.1% of pixels with |dd| > 1 (or valid on one
    side only) against the FP64 oracle, in each matching direction and after
    the L-R check."""
    w, h, D = 1600, 1067, 192
    left, right, _ = synth.stereo_pair(w, h, D, 0, seed=3, gain=True)
    op = O.params(window_w=7, window_h=7, disparities=D, cost=4, p1=12, p2=150,
                  lr_threshold=1.0)
    ctx = sgm.Context(0)
    try:
        got = ctx.match_pair(to_sgm(op), left, right)
    finally:
        ctx.close()
    print("zncc oracle", w, h, D, flush=True)
    want = O.compute_parallel(op, left, right, w, h)
    for k in ("dl", "dr", "checked"):
        a, b = got[k], want[k]
        both = np.isfinite(a) & np.isfinite(b)
        one = np.isfinite(a) ^ np.isfinite(b)
        bad = int((np.abs(a[both] - b[both]) > 1).sum() + one.sum())
        print(k, "differ", bad, "of", a.size, flush=True)
        assert bad <= 0.001 * a.size, (k, bad)


def test_batched_distribution_pairs_against_oracle():
    """Four pairs of the batched workload's distribution (long side >= 800,
    D 64-384, min disparity in [-32, 32]): plain path and one batched
    persistent launch, both bit-exact to the oracle."""
    spec = sorted(synth.batch_pairs(count=48, seed=4), key=lambda t: t[0] * t[1] * t[2])
    picks = [spec[0], spec[1], spec[len(spec) // 3], next(s for s in spec if s[3] < -8)]
    pairs = []
    for (w, h, D, md, seed) in picks:
        left, right, _ = synth.stereo_pair(w, h, D, md, seed=seed)
        pairs.append(dict(left=left, right=right, disparities=D, min_disparity=md))
    assert min(max(p["left"].shape) for p in pairs) >= 800
    assert any(p["min_disparity"] < 0 for p in pairs)
    ctx = sgm.Context(0, budget_bytes=BIG)
    try:
        batch = sgm.Batch(ctx, to_sgm(census_params(64)), pairs, graph=True)
        try:
            outs = batch.run()
        finally:
            batch.close()
        for p, b in zip(pairs, outs):
            h, w = p["left"].shape
            op = census_params(p["disparities"], p["min_disparity"])
            print("pair", w, h, p["disparities"], p["min_disparity"], flush=True)
            plain = ctx.match_pair(to_sgm(op), p["left"], p["right"])
            want = O.compute_parallel(op, p["left"], p["right"], w, h)
            same(plain, want, "plain %dx%d" % (w, h), keys=("checked", "dl", "dr", "reasons",
                                                           "boundary"))
            same(b, want, "batched %dx%d" % (w, h))
    finally:
        ctx.close()
