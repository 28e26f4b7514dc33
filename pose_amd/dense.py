"""Caller side of the SGM path: the disparity search of the reference's
DenseReconstruction pair loop (DenseReconstruction.java:295-434) run over
many pairs at once, and its PFM disparity output (DenseReconstruction.java:943-956).

The reference matches pair after pair; for each it starts from the pair's
estimated interval, reads the boundary counters of the result and widens the
interval (DisparityRange.expansion, DisparityRange.java:318-338) until no
boundary pressure is left, up to four retries or the memory limit.  Here every
pair still searching is matched in one batched launch per attempt
(``sgm.Batch``), then the same rule decides per pair whether it runs again.
"""
from collections import namedtuple

import numpy as np

from pose_amd import sgm


class Bounds(namedtuple("Bounds", "minimum count")):
    """DisparityRange.Bounds (DisparityRange.java:23-27)."""

    def maximum(self):
        return self.minimum + self.count - 1


def pinned_minimum(o):
    """DisparityRange.pinnedMinimum (DisparityRange.java:164-166)."""
    return o.minDisparityExplicit or o.minDisparity != 0


def bounds(minimum, maximum, width, o):
    """DisparityRange.bounds (DisparityRange.java:300-316)."""
    lo, hi = max(o.minDisparity, minimum), min(width - 1, maximum)
    if pinned_minimum(o):
        lo = o.minDisparity
    if lo > width - 3:
        raise ValueError("Minimum disparity leaves fewer than three usable hypotheses")
    hi = max(lo + 2, hi)
    span = hi - lo + 1
    if span > 4096:
        raise ValueError("Observed disparity interval exceeds 4096 hypotheses; restrict the "
                         "requested range or lower dense resolution")
    return Bounds(lo, int(min(width - lo, max(16, span))))


# CensusSgm.Stats boundary pressure (CensusSgm.java:696-716); this path has no
# coarse pass, so the coarse counters are zero.
def lower_pressure(st):
    return st["lower_boundary_rejected"] / max(1, st["full_interval_pixels"])


def upper_pressure(st):
    return st["upper_boundary_rejected"] / max(1, st["full_interval_pixels"])


def lower_clipped(st):
    return st["lower_boundary_rejected"] >= 24 and lower_pressure(st) > .02


def upper_clipped(st):
    return st["upper_boundary_rejected"] >= 24 and upper_pressure(st) > .02


def expansion(original, current, stats, width, attempt):
    """DisparityRange.expansion (DisparityRange.java:318-338): the next interval
    of a pair whose result shows boundary pressure, or None."""
    if original.disparitiesExplicit:
        return None
    lower = (lower_clipped(stats) and not pinned_minimum(original)
             and current.minDisparity > original.minDisparity)
    upper = upper_clipped(stats) and current.minDisparity + current.disparities - 1 < width - 1
    if not lower and not upper:
        return None
    grow = max(16, ((current.disparities // 2 + 15) // 16) * 16)
    lo = current.minDisparity
    hi = lo + current.disparities - 1
    if lower:
        lo = max(original.minDisparity, lo - grow)
    if upper:
        hi = min(width - 1, hi + grow)
    if attempt >= 3:
        lo, hi = original.minDisparity, width - 1
    nxt = bounds(lo, hi, width, original)
    if nxt.minimum == current.minDisparity and nxt.count == current.disparities:
        return None
    return nxt


def stop_reason(original, stats, nxt, memory_limit):
    """The stop reason the reference records (DenseReconstruction.java:404-416)."""
    clipped = lower_clipped(stats) or upper_clipped(stats)
    if original.disparitiesExplicit:
        return "explicit"
    if not clipped:
        return "no_boundary_pressure"
    if nxt is None:
        return "physical_or_user_boundary"
    return "memory_limit" if nxt.count > memory_limit else "retry_limit"


def search(ctx, options, pairs, graph=False, max_retries=4):
    """Disparity search over many pairs.

    ``pairs``: dicts with left, right (uint8 (h, w)), the initial interval
    ``minimum`` / ``count`` (the estimated DisparityRange) and optional
    valid_left / valid_right.  Each attempt matches every pair still searching
    in one batch.  Returns per pair: the L-R-validated result of its last
    attempt (checked, dl, dr, reasons, boundary, stats), the final
    ``minimum`` / ``count``, the ``attempts`` list and the ``stop_reason``."""
    params = sgm.params_from_options(options)
    state = []
    for pr in pairs:
        h, w = np.asarray(pr["left"]).shape
        m = options.copy()
        m.minDisparity, m.disparities = int(pr["minimum"]), int(pr["count"])
        limit = ctx.disparity_limit(params, w, h)
        if not options.disparitiesExplicit and m.disparities > limit:
            raise ValueError("Automatic disparity range needs %d disparities, but memory permits "
                             "%d" % (m.disparities, limit))
        state.append(dict(pair=pr, width=w, matching=m, attempt=0, attempts=[], limit=limit,
                          result=None, stop_reason=None))
    pending = list(range(len(state)))
    while pending:
        batch = sgm.Batch(ctx, params, [
            dict(left=state[i]["pair"]["left"], right=state[i]["pair"]["right"],
                 disparities=state[i]["matching"].disparities,
                 min_disparity=state[i]["matching"].minDisparity,
                 valid_left=state[i]["pair"].get("valid_left"),
                 valid_right=state[i]["pair"].get("valid_right")) for i in pending], graph=graph)
        try:
            results = batch.run()
        finally:
            batch.close()
        still = []
        for i, r in zip(pending, results):
            s = state[i]
            m, st = s["matching"], r["stats"]
            s["result"] = r
            s["attempts"].append(dict(
                attempt=s["attempt"], min_disparity=m.minDisparity, disparities=m.disparities,
                max_disparity=m.minDisparity + m.disparities - 1,
                lower_boundary_rejected=st["lower_boundary_rejected"],
                upper_boundary_rejected=st["upper_boundary_rejected"],
                full_interval_pixels=st["full_interval_pixels"],
                lower_pressure=lower_pressure(st), upper_pressure=upper_pressure(st),
                final_valid=st["valid"]))
            nxt = expansion(options, m, st, s["width"], s["attempt"])
            if (options.disparitiesExplicit or nxt is None or s["attempt"] >= max_retries
                    or nxt.count > s["limit"]):
                s["stop_reason"] = stop_reason(options, st, nxt, s["limit"])
                continue
            m.minDisparity, m.disparities = nxt.minimum, nxt.count
            s["attempt"] += 1
            still.append(i)
        pending = still
    return [dict(s["result"], minimum=s["matching"].minDisparity,
                 count=s["matching"].disparities, attempts=s["attempts"],
                 stop_reason=s["stop_reason"]) for s in state]


def write_pfm(path, width, height, disparity):
    """DenseReconstruction.writeDisparity (DenseReconstruction.java:943-956):
    "Pf", width height, scale -1.0 (little-endian), float32 rows bottom-up."""
    d = np.asarray(disparity, np.float32).reshape(height, width)
    with open(path, "wb") as f:
        f.write(("Pf\n%d %d\n-1.0\n" % (width, height)).encode("ascii"))
        f.write(np.ascontiguousarray(d[::-1]).astype("<f4").tobytes())


def read_pfm(path):
    """(width, height, float32 (height, width) top-down) of a single-channel PFM."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"Pf":
            raise ValueError("not a single-channel PFM")
        width, height = (int(v) for v in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(4 * width * height), "<f4" if scale < 0 else ">f4")
    return width, height, data.reshape(height, width)[::-1].astype(np.float32)
