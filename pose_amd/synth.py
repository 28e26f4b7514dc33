"""Seeded synthetic rectified stereo pairs (the benchmark workloads' inputs).

Scene: a band-limited 4-octave Perlin texture; 3-6 slanted disparity planes over a
Voronoi partition of the image; the right image is the left texture warped by the
plane disparity (bilinear along x), plus Gaussian noise sigma=2, clamped to
[0, 255].  ``gain=True`` multiplies the right image by a smooth per-pixel gain
field in [0.8, 1.2] (the ZNCC workload).  Pure numpy, deterministic per seed.
"""
import numpy as np


def _fade(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


def perlin(h, w, spacing, rng):
    """One octave of 2-D gradient noise with lattice spacing ``spacing`` px, in ~[-1, 1]."""
    gh, gw = h // spacing + 2, w // spacing + 2
    ang = rng.uniform(0, 2 * np.pi, (gh, gw)).astype(np.float32)
    gxs, gys = np.cos(ang), np.sin(ang)
    ys = (np.arange(h, dtype=np.float32) + 0.5) / spacing
    xs = (np.arange(w, dtype=np.float32) + 0.5) / spacing
    yi = ys.astype(np.int32)
    xi = xs.astype(np.int32)
    v = (ys - yi)[:, None]
    u = (xs - xi)[None, :]
    out = np.zeros((h, w), np.float32)
    fu, fv = _fade(u), _fade(v)
    for dy in (0, 1):
        for dx in (0, 1):
            gx = gxs[yi[:, None] + dy, xi[None, :] + dx]
            gy = gys[yi[:, None] + dy, xi[None, :] + dx]
            dot = gx * (u - dx) + gy * (v - dy)
            wx = fu if dx else 1 - fu
            wy = fv if dy else 1 - fv
            out += dot * wx * wy
    return out * np.float32(1.41)


def texture(h, w, rng, base_spacing=16, octaves=4):
    t = np.zeros((h, w), np.float32)
    amp, total = 1.0, 0.0
    s = base_spacing
    for _ in range(octaves):
        t += np.float32(amp) * perlin(h, w, max(1, s), rng)
        total += amp
        amp *= 0.5
        s = max(1, s // 2)
    t /= np.float32(total)
    return t


def disparity_planes(h, w, min_disparity, disparities, rng):
    """3-6 slanted planes on a Voronoi partition; values inside [min+2, min+D-3]."""
    count = int(rng.integers(3, 7))
    lo, hi = min_disparity + 2.0, min_disparity + disparities - 3.0
    if hi < lo:  # tiny ranges: use the whole interval
        lo, hi = float(min_disparity), float(min_disparity + disparities - 1)
    seeds = np.stack([rng.uniform(0, w, count), rng.uniform(0, h, count)], 1).astype(np.float32)
    base = rng.uniform(lo + 0.15 * (hi - lo), hi - 0.15 * (hi - lo), count)
    span = (hi - lo) * 0.15
    sx = rng.uniform(-span, span, count) / max(w, 1)
    sy = rng.uniform(-span, span, count) / max(h, 1)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    best = np.full((h, w), np.inf, np.float32)
    label = np.zeros((h, w), np.int32)
    for i in range(count):
        dist = (xx - seeds[i, 0]) ** 2 + (yy - seeds[i, 1]) ** 2
        m = dist < best
        best[m] = dist[m]
        label[m] = i
    d = (base[label] + sx[label] * (xx - seeds[label, 0]) + sy[label] * (yy - seeds[label, 1]))
    return np.clip(d, lo, hi).astype(np.float32)


def stereo_pair(width, height, disparities, min_disparity=0, seed=0, noise=2.0, gain=False):
    """Returns (left, right, disparity) with uint8 (h, w) images and float32 left-view truth."""
    rng = np.random.default_rng(seed)
    h, w = height, width
    margin = int(max(abs(min_disparity), abs(min_disparity + disparities))) + 2
    tex = texture(h, w + 2 * margin, rng)
    tex = np.clip(128.0 + 100.0 * tex, 0, 255).astype(np.float32)
    disp = disparity_planes(h, w, min_disparity, disparities, rng)
    left = tex[:, margin:margin + w]
    # right(xr) = left texture at xr + d (bilinear along x)
    xs = np.arange(w, dtype=np.float32)[None, :] + disp + margin
    x0 = np.floor(xs).astype(np.int32)
    f = xs - x0
    x0 = np.clip(x0, 0, w + 2 * margin - 2)
    rows = np.arange(h)[:, None]
    right = tex[rows, x0] * (1 - f) + tex[rows, x0 + 1] * f
    if gain:
        g = 1.0 + 0.2 * np.clip(perlin(h, w, max(8, min(h, w) // 4), rng), -1, 1)
        right = right * g
    if noise > 0:
        left = left + rng.normal(0, noise, left.shape).astype(np.float32)
        right = right + rng.normal(0, noise, right.shape).astype(np.float32)
    left = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    right = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    return left, right, disp


def batch_pairs(count=48, seed=0, long_side=(800, 3000), disparities=(64, 384), offset=(-32, 32)):
    """Uneven pairs of the batched workload: (w, h, D, minD, seed) per pair."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        L = int(rng.integers(long_side[0], long_side[1] + 1))
        aspect = (4, 3) if rng.random() < 0.5 else (3, 2)
        s = int(round(L * aspect[1] / aspect[0]))
        w, h = L, s  # landscape: the longest side is the scanline
        D = int(rng.integers(disparities[0], disparities[1] + 1))
        md = int(rng.integers(offset[0], offset[1] + 1))
        out.append((w, h, D, md, seed * 1000 + i))
    return out
