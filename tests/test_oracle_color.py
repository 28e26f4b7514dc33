"""Pinning of the colour-distance census (CensusSgm.colorCensus,
CensusSgm.java:302-332) and the three-channel photometric score
(CensusSgm.zncc, CensusSgm.java:614-644) of the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import bgr_gray


def np_color_census(bgr, ww, wh, valid=None):
    """Independent numpy restatement: distances in double, summed in window order."""
    h, w, _ = bgr.shape
    rx, ry = ww // 2, wh // 2
    words = (ww * wh - 1 + 63) // 64
    desc = np.zeros((words, h, w), np.uint64)
    tex = np.zeros((h, w), np.uint8)
    c = bgr[ry:h - ry, rx:w - rx].astype(np.int64)
    total = np.zeros(c.shape[:2])
    maximum = np.zeros(c.shape[:2])
    dists = []
    for dy in range(-ry, ry + 1):
        for dx in range(-rx, rx + 1):
            if dx == 0 and dy == 0:
                continue
            nb = bgr[ry + dy:h - ry + dy, rx + dx:w - rx + dx].astype(np.int64)
            d = np.sqrt(((nb - c) ** 2).sum(axis=2).astype(np.float64))
            dists.append(d)
            total += d
            maximum = np.maximum(maximum, d)
    mean = total / (ww * wh)
    inner = desc[:, ry:h - ry, rx:w - rx]
    for k, d in enumerate(dists):
        inner[k // 64] |= (d >= mean).astype(np.uint64) << np.uint64(k % 64)
    t = maximum >= 4
    if valid is not None:
        t &= valid[ry:h - ry, rx:w - rx] != 0
    tex[ry:h - ry, rx:w - rx] = t
    return desc.reshape(words, h * w), tex.reshape(-1)


@pytest.mark.parametrize("window", [(3, 3), (5, 5), (9, 7), (7, 9), (11, 11), (5, 11)])
def test_color_census_matches_numpy(window):
    ww, wh = window
    rng = np.random.default_rng(ww * 100 + wh)
    h, w = 23, 37
    bgr = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    bgr[5:12, 8:20] = bgr[5, 8]                           # flat patch: untextured, all bits set
    bgr[15:, 25:] = (bgr[15:, 25:] // 64) * 64 + 7        # coarse levels: ties at the mean
    valid = (rng.random((h, w)) > 0.1).astype(np.uint8)
    want_d, want_t = np_color_census(bgr, ww, wh, valid)
    got_d, got_t = O.color_census(bgr, w, h, ww, wh, valid)
    assert np.array_equal(got_d, want_d)
    assert np.array_equal(got_t, want_t)


def test_color_census_closed_forms():
    h, w = 15, 16
    bgr = np.full((h, w, 3), (10, 200, 77), np.uint8)
    d, t = O.color_census(bgr, w, h, 5, 5)
    d = d.reshape(h, w)
    # every distance 0 >= mean 0: all 24 bits set inside; borders stay 0; untextured
    assert (d[2:h - 2, 2:w - 2] == (1 << 24) - 1).all()
    assert d[:2].sum() == 0 and d[:, :2].sum() == 0
    assert t.sum() == 0
    # one bright centre: every neighbour is far, mean = 24 * dist / 25 < dist
    bgr2 = bgr.copy()
    bgr2[7, 8] = (255, 255, 255)
    d2, t2 = O.color_census(bgr2, w, h, 5, 5)
    assert int(d2.reshape(h, w)[7, 8]) == (1 << 24) - 1 and t2.reshape(h, w)[7, 8] == 1


def np_photo(a, b, w, h, x, y, disparity, r, ch):
    target = x - disparity
    bx = int(np.floor(target))
    f = target - bx
    if x - r < 0 or x + r >= w or y - r < 0 or y + r >= h or bx - r < 0 or bx + r + 1 >= w:
        return -1.0
    n = (2 * r + 1) ** 2
    A = a.reshape(h, w, ch)[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
    B0 = b.reshape(h, w, ch)[y - r:y + r + 1, bx - r:bx + r + 1].astype(np.float64)
    B1 = b.reshape(h, w, ch)[y - r:y + r + 1, bx - r + 1:bx + r + 2].astype(np.float64)
    B = (1 - f) * B0 + f * B1
    cov = sum(((A[..., c] - A[..., c].mean()) * (B[..., c] - B[..., c].mean())).sum() for c in range(ch))
    va = sum(((A[..., c] - A[..., c].mean()) ** 2).sum() for c in range(ch))
    vb = sum(((B[..., c] - B[..., c].mean()) ** 2).sum() for c in range(ch))
    if va < ch * n * 4 or vb < ch * n * 4:
        return -1.0
    return cov / np.sqrt(max(1e-12, va * vb))


def test_photometric_score_three_channels():
    rng = np.random.default_rng(5)
    h, w = 30, 60
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    b = np.roll(a, -4, axis=1)
    b = np.clip(b.astype(int) + rng.integers(-6, 7, b.shape), 0, 255).astype(np.uint8)
    for (x, y, d) in [(20, 10, 4.0), (30, 15, 4.25), (40, 20, 3.6), (10, 5, 7.5), (5, 3, 4.0)]:
        got = O.photo_zncc_ch(a, b, w, h, x, y, d, 3, 3)
        want = np_photo(a, b, w, h, x, y, d, 3, 3)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-12), (x, y, d)
    # one channel: the three-channel path reduces to the single-channel score
    g = a[..., 1].copy()
    gb = b[..., 1].copy()
    assert O.photo_zncc_ch(g, gb, w, h, 20, 10, 4.25, 3, 1) == O.photo_zncc(g, gb, w, h, 20, 10, 4.25, 3)
    # an anti-correlated copy scores -1
    assert O.photo_zncc_ch(a, 255 - a, w, h, 20, 10, 0.0, 3, 3) == pytest.approx(-1.0, abs=1e-12)


def test_color_pair_finds_the_shift():
    """A colour pair with a known shift: the colour-distance census path finds it
    on most pixels (the UpgradeStereoRegression coverage check, CensusSgm.java:109-121)."""
    h, w, shift = 60, 120, 9
    rng = np.random.default_rng(8)
    bgr = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    idx = np.minimum(np.arange(w) + shift, w - 1)
    rbgr = bgr[:, idx].copy()
    gray = bgr_gray(bgr)
    rgray = bgr_gray(rbgr)
    p = O.params(window_w=5, window_h=5, disparities=24, min_disparity=0)
    r = O.compute_color(p, gray, rgray, bgr, rbgr, w, h)
    d = r["checked"]
    valid = np.isfinite(d)
    assert valid.sum() > w * h // 2
    assert (np.abs(d[valid] - shift) < .6).mean() > .98
