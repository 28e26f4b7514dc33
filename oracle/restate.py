"""Second, independent restatement of the SGM hot path in numpy -- test infrastructure only.

Where oracle/sgm_oracle.c follows CensusSgm.match's scalar rolling-row loop
(CensusSgm.java:405-515), this file follows the structure of the reference's
own full-volume test oracle (EsgmInvestigation.java:481-588): a full cost
volume, one full L volume per path, then a per-pixel summary
(best, value, second, before, after, endpoint flags -- the GPU `winners` layout,
GpuStereo.java:758-769) and the host `select` rules (GpuStereo.java:631-687).
The recurrence is evaluated as a vectorised wavefront (a whole row or column of
lines per step) instead of pixel by pixel, so a loop-order or indexing slip in
either restatement shows up as a mismatch between the two.
"""
import math

import numpy as np

DIRECTIONS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)]


def census(gray, ww, wh):
    """Census descriptors as a list of uint64 word planes (h, w)."""
    g = gray.astype(np.int32)
    h, w = g.shape
    rx, ry = ww // 2, wh // 2
    words = (ww * wh - 1 + 63) // 64
    out = [np.zeros((h, w), np.uint64) for _ in range(words)]
    inner = g[ry:h - ry, rx:w - rx]
    bit = 0
    for yy in range(-ry, ry + 1):
        for xx in range(-rx, rx + 1):
            if xx == 0 and yy == 0:
                continue
            nb = g[ry + yy:h - ry + yy, rx + xx:w - rx + xx]
            plane = out[bit // 64]
            plane[ry:h - ry, rx:w - rx] |= (nb < inner).astype(np.uint64) << np.uint64(bit % 64)
            bit += 1
    return out


def texture(gray, ww, wh, valid=None):
    g = gray.astype(np.int32)
    h, w = g.shape
    rx, ry = ww // 2, wh // 2
    lo = np.full((h - 2 * ry, w - 2 * rx), 255, np.int32)
    hi = np.zeros_like(lo)
    for yy in range(-ry, ry + 1):
        for xx in range(-rx, rx + 1):
            nb = g[ry + yy:h - ry + yy, rx + xx:w - rx + xx]
            lo = np.minimum(lo, nb)
            hi = np.maximum(hi, nb)
    m = np.zeros((h, w), bool)
    m[ry:h - ry, rx:w - rx] = (hi - lo) >= 4
    if valid is not None:
        m &= valid.reshape(h, w) != 0
    return m


def _popcount(v):
    return np.bitwise_count(v.astype(np.uint64)).astype(np.int64)


def java_round(v):
    f = math.floor(v)
    return int(f) + (1 if v - f >= 0.5 else 0)


def cost_volume(p, ga, gb, da, db, ta, tb, reverse):
    """Full cost volume (h, w, D) int32 with 512 for invalid cells."""
    h, w = ga.shape
    D, rx, ry = p["disparities"], p["window_w"] // 2, p["window_h"] // 2
    bits = p["window_w"] * p["window_h"] - 1
    C = np.full((h, w, D), 512, np.int32)
    gai, gbi = ga.astype(np.int64), gb.astype(np.int64)
    for k in range(D):
        d = p["min_disparity"] + k
        for x in range(rx, w - rx):
            t = x + d if reverse else x - d
            if t < rx or t >= w - rx:
                continue
            rows = np.arange(ry, h - ry)
            ok = ta[rows, x] & tb[rows, t]
            if not ok.any():
                continue
            ham = np.zeros(len(rows), np.int64)
            for wa, wb in zip(da, db):
                ham += _popcount(wa[rows, x] ^ wb[rows, t])
            mode = p["cost"]
            if mode == 0:
                val = ham
            elif mode in (1, 2):
                if mode == 1:
                    diff = np.abs(gai[rows, x] - gbi[rows, t])
                    cap = p["ad_truncation"]
                else:
                    ax = gai[rows, x + 1] - gai[rows, x - 1]
                    ay = gai[rows + 1, x] - gai[rows - 1, x]
                    bx = gbi[rows, t + 1] - gbi[rows, t - 1]
                    by = gbi[rows + 1, t] - gbi[rows - 1, t]
                    diff = np.abs(ax - bx) + np.abs(ay - by)
                    cap = p["gradient_truncation"]
                wgt, den = p["census_weight"], 1000 * cap
                val = (wgt * ham * cap + (1000 - wgt) * np.minimum(diff, cap) * bits + den // 2) // den
            else:
                val = np.empty(len(rows), np.int64)
                for i, y in enumerate(rows):
                    wa_ = gai[y - ry:y + ry + 1, x - rx:x + rx + 1]
                    wb_ = gbi[y - ry:y + ry + 1, t - rx:t + rx + 1]
                    n = wa_.size
                    sa, sb = int(wa_.sum()), int(wb_.sum())
                    va = n * int((wa_ * wa_).sum()) - sa * sa
                    vb = n * int((wb_ * wb_).sum()) - sb * sb
                    if va <= 0 or vb <= 0:
                        val[i] = 512
                        continue
                    corr = (n * int((wa_ * wb_).sum()) - sa * sb) / math.sqrt(float(va) * float(vb))
                    c = max(-1.0, min(1.0, corr))
                    val[i] = java_round((1 - c) * .5 * bits)
            C[rows[ok], x, k] = val[ok]
    return C


def _transition(prev, C, has, p1, p2l):
    """Recurrence for a batch of lines: prev (m, D), C (m, D), has (m,), p2l (m,)."""
    big = np.int64(1) << 40
    mn = prev.min(axis=1)
    t = np.minimum(prev, (mn + p2l)[:, None])
    left = np.concatenate([np.full((prev.shape[0], 1), big), prev[:, :-1]], axis=1)
    right = np.concatenate([prev[:, 1:], np.full((prev.shape[0], 1), big)], axis=1)
    t = np.minimum(t, left + p1)
    t = np.minimum(t, right + p1)
    L = C + t - mn[:, None]
    return np.where(has[:, None], L, C)


def path_volumes(p, C, ga, ta):
    """One full L volume per direction (list of (h, w, D) int64)."""
    h, w, D = C.shape
    p1 = p["p1"]
    p2 = max(p["p1"] + 1, p["p2"])
    scale = p["edge_scale"]
    g = ga.astype(np.int64)
    C = C.astype(np.int64)
    vols = []
    for dx, dy in DIRECTIONS[: p["directions"]]:
        L = np.zeros((h, w, D), np.int64)
        if dy != 0:
            order = range(h) if dy > 0 else range(h - 1, -1, -1)
            for y in order:
                py = y - dy
                xs = np.arange(w)
                px = xs - dx
                inside = (px >= 0) & (px < w) & (0 <= py < h)
                has = np.zeros(w, bool)
                if 0 <= py < h:
                    has[inside] = ta[py, px[inside]]
                prev = np.zeros((w, D), np.int64)
                p2l = np.full(w, p2, np.int64)
                if 0 <= py < h:
                    prev[inside] = L[py, px[inside]]
                    if scale:
                        diff = np.zeros(w, np.int64)
                        diff[inside] = np.abs(g[y, xs[inside]] - g[py, px[inside]])
                        p2l = np.maximum(p1 + 1, (p2 * scale) // (scale + diff))
                L[y] = _transition(prev, C[y], has, p1, p2l)
        else:
            order = range(w) if dx > 0 else range(w - 1, -1, -1)
            for x in order:
                px = x - dx
                has = np.zeros(h, bool)
                prev = np.zeros((h, D), np.int64)
                p2l = np.full(h, p2, np.int64)
                if 0 <= px < w:
                    has = ta[:, px].copy()
                    prev = L[:, px]
                    if scale:
                        diff = np.abs(g[:, x] - g[:, px])
                        p2l = np.maximum(p1 + 1, (p2 * scale) // (scale + diff))
                L[:, x] = _transition(prev, C[:, x], has, p1, p2l)
        vols.append(L)
    return vols


def aggregate(p, vols, sum_bytes):
    S = np.zeros(vols[0].shape, np.int64)
    for L in vols:
        S += L
    if sum_bytes == 2:
        S &= 0xFFFF
    return S


def summary(C, S, invalid=512):
    """Per-pixel (best, value, second, before, after, flags) like GpuStereo winners."""
    h, w, D = C.shape
    valid = C != invalid
    big = np.iinfo(np.int64).max
    Sv = np.where(valid, S, big)
    best = np.argmin(Sv, axis=2)  # first index on ties
    value = np.take_along_axis(Sv, best[..., None], 2)[..., 0]
    has = valid.any(axis=2)
    k = np.arange(D)[None, None, :]
    far = np.abs(k - best[..., None]) > 1
    second = np.where(far & valid, S, big).min(axis=2)
    bm1 = np.clip(best - 1, 0, D - 1)
    bp1 = np.clip(best + 1, 0, D - 1)
    vb = np.take_along_axis(valid, bm1[..., None], 2)[..., 0] & (best > 0)
    va = np.take_along_axis(valid, bp1[..., None], 2)[..., 0] & (best + 1 < D)
    before = np.where(vb, np.take_along_axis(S, bm1[..., None], 2)[..., 0], -1)
    after = np.where(va, np.take_along_axis(S, bp1[..., None], 2)[..., 0], -1)
    flags = valid[..., 0].astype(np.int64) | (valid[..., D - 1].astype(np.int64) << 1)
    best = np.where(has, best, -1)
    return best, value, second, before, after, flags


def select(p, summ, rx, ry):
    """GpuStereo.select (GpuStereo.java:631-687) for interior pixels."""
    best, value, second, before, after, flags = summ
    h, w = best.shape
    D = p["disparities"]
    out = np.full((h, w), np.nan, np.float32)
    reasons = np.ones((h, w), np.uint8)
    boundary = np.zeros((h, w), np.uint8)
    st = dict(lower_boundary_rejected=0, full_interval_pixels=0, boundary_rejected=0,
              upper_boundary_rejected=0, ambiguity_rejected=0, selected=0)
    big = np.iinfo(np.int64).max
    for y in range(ry, h - ry):
        for x in range(rx, w - rx):
            b = int(best[y, x])
            if b < 0:
                continue
            full = flags[y, x] == 3
            if full:
                st["full_interval_pixels"] += 1
                boundary[y, x] = 4
            if b == 0 or b == D - 1:
                st["boundary_rejected"] += 1
                reasons[y, x] = 2
                if full and b == 0:
                    st["lower_boundary_rejected"] += 1
                    boundary[y, x] |= 1
                if full and b == D - 1:
                    st["upper_boundary_rejected"] += 1
                    boundary[y, x] |= 2
                continue
            v, s = int(value[y, x]), int(second[y, x])
            if s != big and (s - v) <= max(1.0, v * p["uniqueness"]):
                st["ambiguity_rejected"] += 1
                reasons[y, x] = 3
                continue
            off = 0.0
            bf, af = int(before[y, x]), int(after[y, x])
            if bf >= 0 and af >= 0:
                curv = float(bf) + af - 2.0 * v
                if curv > 0:
                    off = max(-.5, min(.5, .5 * (bf - float(af)) / curv))
            out[y, x] = np.float32(float(p["min_disparity"] + b) + off)
            st["selected"] += 1
            reasons[y, x] = 0
    return out, reasons, boundary, st


def match(p, ga, gb, ta, tb, reverse):
    """One direction: returns (disp, reasons, boundary, stats, C, S)."""
    da, db = census(ga, p["window_w"], p["window_h"]), census(gb, p["window_w"], p["window_h"])
    C = cost_volume(p, ga, gb, da, db, ta, tb, reverse)
    bound = p["directions"] * (512 + max(p["p1"] + 1, p["p2"]))
    sum_bytes = p.get("aggregate_bytes") or (2 if bound <= 65535 else 4)
    S = aggregate(p, path_volumes(p, C, ga, ta), sum_bytes)
    disp, reasons, boundary, st = select(p, summary(C, S), p["window_w"] // 2, p["window_h"] // 2)
    return disp, reasons, boundary, st, C, S


def esgm_summary(p, C, vols, S, pool):
    """Compact proposals and certification of GpuCandidateStereo (the COLLECT
    hook, GpuCandidateStereo.java:413-441; esgm_select :456-478; certification
    in execute :82-102) over the full per-path volumes.  Returns the compact
    winner summary and the per-pixel `trusted` mask."""
    h, w, D = C.shape
    valid = C != 512
    big = np.int64(1) << 60
    k = np.arange(D, dtype=np.int64)[None, None, :]
    cands = []
    for L in vols:
        chosen = []
        for _ in range(pool):
            ok = valid.copy()
            for c in chosen:
                ok &= np.abs(k - c[..., None]) > 1
            key = np.where(ok, L * (D + 1) + k, big).min(axis=2)
            chosen.append(np.where(key == big, 65535, key % (D + 1)))
        cands += chosen
    cand = np.stack(cands, axis=2)  # (h, w, Cn)
    Cn = cand.shape[2]
    Sw = S & 0xFFFF
    comp = np.zeros((h, w, Cn, 3), np.int64)
    for j in range(Cn):
        for o in range(3):
            kk = cand[..., j] + o - 1
            inside = (kk >= 0) & (kk < D)
            comp[..., j, o] = np.where(inside, np.take_along_axis(Sw, np.clip(kk, 0, D - 1)[..., None], 2)[..., 0], 0)
    covered = np.zeros((h, w, D), bool)
    for j in range(Cn):
        covered |= np.abs(k - cand[..., j][..., None]) <= 1
    outside = np.zeros((h, w), np.int64)
    for L in vols:
        m = np.where(valid & ~covered, L, big).min(axis=2)
        outside = np.where(m == big, 65535, np.minimum(65535, outside + m))
    # winner over the proposals
    best = np.full((h, w), -1, np.int64)
    value = np.full((h, w), big, np.int64)
    owner = np.full((h, w), -1, np.int64)
    for j in range(Cn):
        kj, vj = cand[..., j], comp[..., j, 1]
        take = (kj < D) & ((vj < value) | ((vj == value) & ((best < 0) | (kj < best))))
        best = np.where(take, kj, best)
        value = np.where(take, vj, value)
        owner = np.where(take, j, owner)
    def valid_at(kk):
        inside = (kk >= 0) & (kk < D)
        return inside & np.take_along_axis(valid, np.clip(kk, 0, D - 1)[..., None], 2)[..., 0]
    has = owner >= 0
    own = np.clip(owner, 0, Cn - 1)
    at = np.take_along_axis(comp, own[..., None, None], 2)[..., 0, :]
    before = np.where(has & valid_at(best - 1), at[..., 0], -1)
    after = np.where(has & valid_at(best + 1), at[..., 2], -1)
    second = np.full((h, w), big, np.int64)
    for j in range(Cn):
        for o in range(3):
            kk = cand[..., j] + o - 1
            use = has & (np.abs(kk - best) > 1) & valid_at(kk)
            second = np.where(use, np.minimum(second, comp[..., j, o]), second)
    flags = valid[..., 0].astype(np.int64) | (valid[..., D - 1].astype(np.int64) << 1)
    lower = outside
    winner = (lower > value) & ((before < 0) | (before > value)) & ((after < 0) | (after > value)) \
        & (second > value)
    margin = np.maximum(1.0, value * p["uniqueness"])
    trusted = ~has | (winner & ((best == 0) | (best == D - 1) | (second - value.astype(np.float64) <= margin)
                                | ((lower - value) > margin)))
    return (best, value, second, before, after, flags), trusted


def esgm_match(p, ga, gb, ta, tb, reverse, pool):
    """GpuCandidateStereo.execute: certified pixels take the compact summary,
    the others the exact one (the verification pass).  Returns (disp, reasons,
    boundary, stats, trusted)."""
    da, db = census(ga, p["window_w"], p["window_h"]), census(gb, p["window_w"], p["window_h"])
    C = cost_volume(p, ga, gb, da, db, ta, tb, reverse)
    vols = path_volumes(p, C, ga, ta)
    S = aggregate(p, vols, 2)
    compact, trusted = esgm_summary(p, C, vols, S, pool)
    exact = summary(C, S)
    merged = tuple(np.where(trusted, c, e) for c, e in zip(compact, exact))
    disp, reasons, boundary, st = select(p, merged, p["window_w"] // 2, p["window_h"] // 2)
    return disp, reasons, boundary, st, trusted
