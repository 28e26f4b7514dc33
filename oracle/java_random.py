"""java.util.Random restated from its published specification (test infrastructure).

The reference's known-answer tests build their fixtures with
``new Random(seed).nextBytes(buf)`` (e.g. SelfTest.java:84, UpgradeStereoRegression.java:24,
NoiseRegression.java:81-82).  The generator is the 48-bit LCG documented in the
Java SE API (Random.next/nextInt/nextBytes), so those fixtures are regenerated
bit-exactly here without a JVM.
"""
import numpy as np

_MULT = 0x5DEECE66D
_ADD = 0xB
_MASK = (1 << 48) - 1


class JavaRandom:
    def __init__(self, seed):
        self.seed = (seed ^ _MULT) & _MASK

    def next(self, bits):
        self.seed = (self.seed * _MULT + _ADD) & _MASK
        r = self.seed >> (48 - bits)
        if r >= 1 << (bits - 1):  # (int) cast of the top bits
            r -= 1 << bits
        return r

    def next_int(self):
        return self.next(32)

    def next_bytes(self, n):
        out = bytearray(n)
        i = 0
        while i < n:
            rnd = self.next_int() & 0xFFFFFFFF
            for _ in range(min(n - i, 4)):
                out[i] = rnd & 0xFF
                rnd >>= 8
                i += 1
        return np.frombuffer(bytes(out), dtype=np.uint8).copy()


def random_bytes(seed, n):
    """``new Random(seed).nextBytes(new byte[n])`` as uint8 (Java bytes reinterpreted)."""
    return JavaRandom(seed).next_bytes(n)
