"""Compact SGM restatement (oracle/restate.py esgm_summary / esgm_match, after
GpuCandidateStereo.java:82-102, 413-478) against ordinary SGM on the CPU.

The certification rule is what makes the compact plan exact: a certified pixel
must select exactly what the full aggregate selects.  So the restated compact
match must equal the restated full match on every case (winner, subpixel
offset, reason, boundary flags, counters), with both certified and verified
pixels present, and a planted mistake in the rule (the off-proposal bound
ignored) must break that equality on some case."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import restate as RS


def _case(seed, D, directions, texture_levels):
    rng = np.random.default_rng(seed)
    w, h = 40, 22
    ga = rng.integers(0, 256, (h, w)).astype(np.uint8)
    if texture_levels:
        ga = (ga // texture_levels * texture_levels).astype(np.uint8)
    gb = np.roll(ga, 3, axis=1)
    gb = np.clip(gb.astype(int) + rng.integers(-4, 5, gb.shape), 0, 255).astype(np.uint8)
    valid = (rng.random((h, w)) > 0.05).astype(np.uint8)
    p = dict(window_w=5, window_h=5, disparities=D, min_disparity=-2, cost=0, p1=4, p2=30,
             edge_scale=32 if seed % 2 else 0, directions=directions, aggregate_bytes=2,
             census_weight=500, ad_truncation=30, gradient_truncation=60, reserved=0,
             uniqueness=0.03, lr_threshold=0.75)
    ta = O.texture(ga, w, h, 5, 5, valid).reshape(h, w).astype(bool)
    tb = O.texture(gb, w, h, 5, 5, valid).reshape(h, w).astype(bool)
    return p, ga, gb, ta, tb


CASES = [(11, 9, 8, 0), (12, 14, 4, 64), (13, 20, 8, 32), (14, 6, 4, 0), (15, 12, 8, 96)]


@pytest.mark.parametrize("seed,D,directions,levels", CASES)
@pytest.mark.parametrize("pool", [1, 2, 4])
def test_compact_match_equals_full_match(seed, D, directions, levels, pool):
    p, ga, gb, ta, tb = _case(seed, D, directions, levels)
    for rev in (0, 1):
        A, B, TA, TB = (ga, gb, ta, tb) if not rev else (gb, ga, tb, ta)
        disp, rs, bd, st, _, _ = RS.match(p, A, B, TA, TB, rev)
        cdisp, crs, cbd, cst, trusted = RS.esgm_match(p, A, B, TA, TB, rev, pool)
        assert np.array_equal(disp.view(np.uint32), cdisp.view(np.uint32))
        assert np.array_equal(rs, crs) and np.array_equal(bd, cbd)
        assert st == cst


def test_both_outcomes_occur():
    certified = verified = 0
    for seed, D, directions, levels in CASES:
        p, ga, gb, ta, tb = _case(seed, D, directions, levels)
        *_, trusted = RS.esgm_match(p, ga, gb, ta, tb, 0, 1)
        certified += int(trusted.sum())
        verified += int((~trusted).sum())
    assert certified > 0 and verified > 0


def test_planted_mistake_is_detected():
    """Certifying without the off-proposal bound (lower) trusts compact winners
    that the full aggregate overturns."""
    broken = 0
    for seed, D, directions, levels in CASES:
        p, ga, gb, ta, tb = _case(seed, D, directions, levels)
        da, db = RS.census(ga, 5, 5), RS.census(gb, 5, 5)
        C = RS.cost_volume(p, ga, gb, da, db, ta, tb, 0)
        vols = RS.path_volumes(p, C, ga, ta)
        S = RS.aggregate(p, vols, 2)
        compact, _ = RS.esgm_summary(p, C, vols, S, 1)
        best, value, second, before, after, flags = compact
        has = best >= 0
        winner = ((before < 0) | (before > value)) & ((after < 0) | (after > value)) & (second > value)
        trusted = ~has | winner
        exact = RS.summary(C, S)
        merged = tuple(np.where(trusted, c, e) for c, e in zip(compact, exact))
        a = RS.select(p, merged, 2, 2)[0]
        b = RS.select(p, exact, 2, 2)[0]
        broken += int((a.view(np.uint32) != b.view(np.uint32)).sum())
    assert broken > 0
