"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
entry point (literal and threaded matching, post-L-R filters with subpixel
refinement, colour census) on edge-case inputs -- the smallest accepted
images, ranges wider than the image, signed ranges, every cost, 4 and 8
paths, 16- and 32-bit aggregates, rectangular windows, masks -- with the
threaded oracle equal to the literal one (oracle/sanitize_check.c, built by
`make -C oracle sanitize`)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("gcc") is None,
                    reason="needs make and gcc")
def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 mismatches" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
