"""Seeded synthetic rectified stereo pairs (the benchmark workloads' inputs).

Scene: a band-limited 4-octave Perlin texture; 3-6 slanted disparity planes over a
Voronoi partition of the image; the right image is the left texture warped by the
plane disparity (linear along x), plus Gaussian noise sigma=2, clamped to
[0, 255].  ``gain=True`` multiplies the right image by a smooth per-pixel gain
field in [0.8, 1.2] (the ZNCC workload).  The generator is C (csrc/synth.c,
OpenMP over rows), deterministic per seed.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgm_synth.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build_synth()
        L = ctypes.CDLL(LIB_PATH)
        L.sgm_synth_pair.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint64, ctypes.c_float, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def stereo_pair(width, height, disparities, min_disparity=0, seed=0, noise=2.0, gain=False):
    """Returns (left, right, disparity): uint8 (h, w) images and float32 plane
    disparities on the right image's pixel grid (right[y, x] shows the left
    texture at x + disparity[y, x]): the truth of the right-reference map;
    left_view() warps it to the left image."""
    left = np.empty((height, width), np.uint8)
    right = np.empty((height, width), np.uint8)
    disp = np.empty((height, width), np.float32)
    rc = _load().sgm_synth_pair(width, height, disparities, min_disparity, seed & ((1 << 64) - 1),
                                noise, int(bool(gain)), left.ctypes.data, right.ctypes.data,
                                disp.ctypes.data)
    if rc != 0:
        raise ValueError("bad synthetic pair arguments")
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


def left_view(disparity):
    """Left-image truth from the right-grid disparity of stereo_pair: each right
    pixel x lands on left pixel round(x + d); left pixels no right pixel lands on
    (occluded) are NaN."""
    h, w = disparity.shape
    out = np.full((h, w), np.nan, np.float32)
    xs = np.arange(w)[None, :] + np.rint(disparity).astype(np.int64)
    ys = np.broadcast_to(np.arange(h)[:, None], (h, w))
    m = (xs >= 0) & (xs < w)
    out[ys[m], xs[m]] = disparity[m]
    return out
