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
