"""Shared inputs: the reference's known-answer fixtures (regenerated with the
java.util.Random clone) and seeded synthetic pairs."""
import numpy as np

from oracle.java_random import random_bytes


def shifted(seed, w, h, shift):
    """``new Random(seed).nextBytes(left)``; right[y][x] = left[y][min(w-1, x+shift)]
    (SelfTest.java:83-87, NoiseRegression.java:80-84, UpgradeStereoRegression.java:22-32)."""
    left = random_bytes(seed, w * h).reshape(h, w)
    idx = np.minimum(np.arange(w) + shift, w - 1)
    right = left[:, idx].copy()
    return left, right


def upgrade_fixture():
    """UpgradeStereoRegression.java:22-32: 192x79, shift 13, two masked boxes."""
    w, h = 192, 79
    a, b = shifted(120, w, h, 13)
    ma = np.full((h, w), 255, np.uint8)
    mb = np.full((h, w), 255, np.uint8)
    ma[21:55, 86:97] = 0
    mb[6:18, 53:59] = 0
    return a, b, ma, mb


def subpixel_texture(w, h, d):
    """SubpixelRegression.java:51-59 analytic plane texture."""
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    def tex(x, y):
        v = (128 + 27 * np.sin(.71 * x + .23 * y) + 23 * np.cos(.43 * x - .69 * y)
             + 21 * np.sin(.29 * x + .91 * y) + 14 * np.cos(1.11 * x + .17 * y))
        v = np.clip(v, 0, 255)
        return (np.floor(v) + ((v - np.floor(v)) >= 0.5)).astype(np.uint8)
    return tex(x, y), tex(x + d, y)


def bgr_gray(bgr):
    """Integer luma of interleaved BGR (0.114 B + 0.587 G + 0.299 R, rounded):
    the gray rasters that accompany colour inputs."""
    b = bgr.astype(np.int32)
    return ((29 * b[..., 0] + 150 * b[..., 1] + 77 * b[..., 2] + 128) >> 8).astype(np.uint8)


def color_pair(w, h, shift, seed):
    """Random colour texture, smoothed a little, and its copy shifted by `shift`
    pixels (right[y][x] = left[y][min(w-1, x+shift)]) with mild noise."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (h + 2, w + 2, 3)).astype(np.int32)
    left = ((base[1:-1, 1:-1] * 2 + base[:-2, 1:-1] + base[2:, 1:-1]) // 4).astype(np.uint8)
    idx = np.minimum(np.arange(w) + shift, w - 1)
    right = left[:, idx].astype(np.int32) + rng.integers(-2, 3, (h, w, 3))
    right = np.clip(right, 0, 255).astype(np.uint8)
    return left, right
