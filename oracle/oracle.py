"""ctypes wrapper of the C oracle (liboracle.so) -- test infrastructure only.

Every function mirrors one function of oracle/sgm_oracle.c, which restates the
reference's CensusSgm / StereoCost CPU path (file:line citations live there).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CENSUS, AD_CENSUS, CENSUS_GRADIENT, ZNCC, ZNCC1024 = 0, 1, 2, 3, 4


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("window_w", ctypes.c_int32), ("window_h", ctypes.c_int32),
        ("disparities", ctypes.c_int32), ("min_disparity", ctypes.c_int32),
        ("cost", ctypes.c_int32), ("p1", ctypes.c_int32), ("p2", ctypes.c_int32),
        ("edge_scale", ctypes.c_int32), ("directions", ctypes.c_int32),
        ("aggregate_bytes", ctypes.c_int32), ("census_weight", ctypes.c_int32),
        ("ad_truncation", ctypes.c_int32), ("gradient_truncation", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("uniqueness", ctypes.c_double), ("lr_threshold", ctypes.c_double),
    ]


class OracleFilter(ctypes.Structure):
    _fields_ = [
        ("median", ctypes.c_int32), ("median_classic", ctypes.c_int32),
        ("median_edge_threshold", ctypes.c_int32), ("correlation_window", ctypes.c_int32),
        ("speckle_size", ctypes.c_int32), ("speckle_connectivity", ctypes.c_int32),
        ("subpixel_refinement", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("median_max_change", ctypes.c_double), ("correlation", ctypes.c_double),
        ("speckle_range", ctypes.c_double), ("reserved2", ctypes.c_double),
    ]


def filters(**kw):
    """Post-L-R filter options; defaults are the reference's Options defaults
    (Options.java:19-23,67,76): edge-preserving 3x3 median, max change 0.5, edge
    threshold 24, 7x7 photometric ZNCC >= 0.55, speckles < 50 px, range 1, 8-connected."""
    d = dict(median=3, median_classic=0, median_edge_threshold=24, correlation_window=7,
             speckle_size=50, speckle_connectivity=8, subpixel_refinement=0, reserved1=0,
             median_max_change=0.5, correlation=0.55, speckle_range=1.0, reserved2=0.0)
    d.update(kw)
    return OracleFilter(**d)


STAT_FIELDS = [
    "lower_boundary_rejected", "full_interval_pixels", "boundary_rejected",
    "upper_boundary_rejected", "ambiguity_rejected", "selected", "lr_rejected",
    "post_median_lr_rejected", "photometric_rejected", "speckle_rejected", "valid",
]


class OracleStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in STAT_FIELDS] + [("reserved", ctypes.c_int64 * 5)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f in STAT_FIELDS}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.oracle_census.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.oracle_texture.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.oracle_match.argtypes = [ctypes.POINTER(OracleParams), P, P, P, P, P, P, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(OracleStats),
                                   P, P, P, P]
        L.oracle_compute.argtypes = [ctypes.POINTER(OracleParams), P, P, P, P, ctypes.c_int,
                                     ctypes.c_int, P, P, P, ctypes.POINTER(OracleStats), P, P]
        L.oracle_postfilter.argtypes = [ctypes.POINTER(OracleParams), ctypes.POINTER(OracleFilter),
                                        P, P, P, ctypes.c_int, ctypes.c_int, P,
                                        ctypes.POINTER(OracleStats), P]
        L.oracle_photo_zncc.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_int]
        L.oracle_photo_zncc.restype = ctypes.c_double
        L.oracle_photo_zncc_ch.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.oracle_photo_zncc_ch.restype = ctypes.c_double
        L.oracle_color_census.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, P, P]
        L.oracle_compute_color.argtypes = [ctypes.POINTER(OracleParams), P, P, P, P, P, P,
                                           ctypes.c_int, ctypes.c_int, P, P, P,
                                           ctypes.POINTER(OracleStats), P, P]
        L.oracle_compute_ex.argtypes = [ctypes.POINTER(OracleParams), ctypes.POINTER(OracleFilter),
                                        P, P, P, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P,
                                        ctypes.POINTER(OracleStats), P, P]
        L.oracle_postfilter_ch.argtypes = [ctypes.POINTER(OracleParams),
                                           ctypes.POINTER(OracleFilter), P, P, P, ctypes.c_int,
                                           P, ctypes.c_int, ctypes.c_int, P,
                                           ctypes.POINTER(OracleStats), P]
        L.oracle_match_parallel.argtypes = [ctypes.POINTER(OracleParams), P, P, P, P, P, P,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                            ctypes.POINTER(OracleStats), P, P, ctypes.c_int]
        L.oracle_compute_parallel.argtypes = [ctypes.POINTER(OracleParams), P, P, P, P,
                                              ctypes.c_int, ctypes.c_int, P, P, P,
                                              ctypes.POINTER(OracleStats), P, P, ctypes.c_int]
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_round_float.argtypes = [ctypes.c_float]
        L.oracle_round_float.restype = ctypes.c_int32
        L.oracle_round_double.argtypes = [ctypes.c_double]
        L.oracle_round_double.restype = ctypes.c_int64
        L.oracle_effective_p2.argtypes = [ctypes.POINTER(OracleParams)]
        L.oracle_aggregate_bytes.argtypes = [ctypes.POINTER(OracleParams)]
        L.oracle_aggregate_bound.argtypes = [ctypes.POINTER(OracleParams)]
        L.oracle_aggregate_bound.restype = ctypes.c_int64
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def params(**kw):
    """Oracle parameters; defaults are the reference's Options defaults for a 5x5 census
    (Options.java:17-40,67) with penalties already scaled (4/24)."""
    d = dict(window_w=5, window_h=5, disparities=128, min_disparity=0, cost=CENSUS, p1=4, p2=24,
             edge_scale=32, directions=8, aggregate_bytes=0, census_weight=500,
             ad_truncation=30, gradient_truncation=60, reserved=0, uniqueness=0.03,
             lr_threshold=0.75)
    d.update(kw)
    return OracleParams(**d)


def census(gray, w, h, ww, wh):
    words = (ww * wh - 1 + 63) // 64
    out = np.zeros(words * w * h, dtype=np.uint64)
    lib().oracle_census(_ptr(np.ascontiguousarray(gray, np.uint8)), w, h, ww, wh, _ptr(out))
    return out.reshape(words, h * w)


def texture(gray, w, h, ww, wh, valid=None):
    out = np.zeros(w * h, dtype=np.uint8)
    v = None if valid is None else np.ascontiguousarray(valid, np.uint8)
    lib().oracle_texture(_ptr(np.ascontiguousarray(gray, np.uint8)), _ptr(v), w, h, ww, wh, _ptr(out))
    return out


def color_census(bgr, w, h, ww, wh, valid=None):
    """CensusSgm.colorCensus on interleaved BGR (h, w, 3): (descriptor words, textured)."""
    words = (ww * wh - 1 + 63) // 64
    out = np.zeros(words * w * h, dtype=np.uint64)
    tex = np.zeros(w * h, dtype=np.uint8)
    v = None if valid is None else np.ascontiguousarray(valid, np.uint8)
    lib().oracle_color_census(_ptr(np.ascontiguousarray(bgr, np.uint8)), _ptr(v), w, h, ww, wh,
                              _ptr(out), _ptr(tex))
    return out.reshape(words, h * w), tex


def match(p, da, db, ga, gb, ta, tb, w, h, reverse, volumes=False):
    """One matching direction. Returns (disp, stats, reasons, boundary[, cost, agg])."""
    n = w * h
    disp = np.empty(n, np.float32)
    reasons = np.ones(n, np.uint8)
    boundary = np.zeros(n, np.uint8)
    st = OracleStats()
    cost = np.empty(n * p.disparities, np.uint16) if volumes else None
    agg = np.empty(n * p.disparities, np.uint32) if volumes else None
    rc = lib().oracle_match(ctypes.byref(p), _ptr(np.ascontiguousarray(da)), _ptr(np.ascontiguousarray(db)),
                            _ptr(ga), _ptr(gb), _ptr(ta), _ptr(tb), w, h, int(reverse), _ptr(disp),
                            ctypes.byref(st), _ptr(reasons), _ptr(boundary), _ptr(cost), _ptr(agg))
    if rc != 0:
        raise ValueError("oracle_match failed: %d" % rc)
    out = (disp, st.as_dict(), reasons, boundary)
    return out + (cost, agg) if volumes else out


def compute(p, left, right, w, h, valid_left=None, valid_right=None):
    """Both directions + L-R check. Returns dict of checked, dl, dr, stats, reasons, boundary."""
    n = w * h
    left = np.ascontiguousarray(left, np.uint8).reshape(-1)
    right = np.ascontiguousarray(right, np.uint8).reshape(-1)
    vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8).reshape(-1)
    vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8).reshape(-1)
    checked = np.empty(n, np.float32)
    dl = np.empty(n, np.float32)
    dr = np.empty(n, np.float32)
    reasons = np.empty(n, np.uint8)
    boundary = np.empty(n, np.uint8)
    st = OracleStats()
    rc = lib().oracle_compute(ctypes.byref(p), _ptr(left), _ptr(right), _ptr(vl), _ptr(vr), w, h,
                              _ptr(checked), _ptr(dl), _ptr(dr), ctypes.byref(st), _ptr(reasons),
                              _ptr(boundary))
    if rc != 0:
        raise ValueError("oracle_compute failed: %d" % rc)
    return dict(checked=checked, dl=dl, dr=dr, stats=st.as_dict(), reasons=reasons,
                boundary=boundary)


def compute_parallel(p, left, right, w, h, valid_left=None, valid_right=None, threads=0):
    """compute() split over host threads (OpenMP, threads <= 0: all): the same
    results bit for bit, for the full-size workloads."""
    n = w * h
    left = np.ascontiguousarray(left, np.uint8).reshape(-1)
    right = np.ascontiguousarray(right, np.uint8).reshape(-1)
    vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8).reshape(-1)
    vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8).reshape(-1)
    checked = np.empty(n, np.float32)
    dl = np.empty(n, np.float32)
    dr = np.empty(n, np.float32)
    reasons = np.empty(n, np.uint8)
    boundary = np.empty(n, np.uint8)
    st = OracleStats()
    rc = lib().oracle_compute_parallel(ctypes.byref(p), _ptr(left), _ptr(right), _ptr(vl),
                                       _ptr(vr), w, h, _ptr(checked), _ptr(dl), _ptr(dr),
                                       ctypes.byref(st), _ptr(reasons), _ptr(boundary),
                                       int(threads))
    if rc != 0:
        raise ValueError("oracle_compute_parallel failed: %d" % rc)
    return dict(checked=checked, dl=dl, dr=dr, stats=st.as_dict(), reasons=reasons,
                boundary=boundary)


def threads():
    """Host threads a parallel oracle call uses."""
    return int(lib().oracle_threads())


def compute_color(p, left, right, left_bgr, right_bgr, w, h, valid_left=None, valid_right=None):
    """compute() with colour-distance census descriptors and texture from BGR."""
    n = w * h
    arrs = [np.ascontiguousarray(a, np.uint8).reshape(-1) for a in (left, right, left_bgr, right_bgr)]
    vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8).reshape(-1)
    vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8).reshape(-1)
    checked, dl, dr = (np.empty(n, np.float32) for _ in range(3))
    reasons = np.empty(n, np.uint8)
    boundary = np.empty(n, np.uint8)
    st = OracleStats()
    rc = lib().oracle_compute_color(ctypes.byref(p), *(_ptr(a) for a in arrs), _ptr(vl), _ptr(vr),
                                    w, h, _ptr(checked), _ptr(dl), _ptr(dr), ctypes.byref(st),
                                    _ptr(reasons), _ptr(boundary))
    if rc != 0:
        raise ValueError("oracle_compute_color failed: %d" % rc)
    return dict(checked=checked, dl=dl, dr=dr, stats=st.as_dict(), reasons=reasons,
                boundary=boundary)


def compute_ex(p, f, left, right, w, h, valid_left=None, valid_right=None, left_bgr=None,
               right_bgr=None):
    """compute() / compute_color() with the filter options' subpixel refinement
    (PhotometricRefinement before the L-R check when f.subpixel_refinement)."""
    n = w * h
    arrs = [None if a is None else np.ascontiguousarray(a, np.uint8).reshape(-1)
            for a in (left, right, left_bgr, right_bgr, valid_left, valid_right)]
    checked, dl, dr = (np.empty(n, np.float32) for _ in range(3))
    reasons = np.empty(n, np.uint8)
    boundary = np.empty(n, np.uint8)
    st = OracleStats()
    rc = lib().oracle_compute_ex(ctypes.byref(p), ctypes.byref(f), *(_ptr(a) for a in arrs), w, h,
                                 _ptr(checked), _ptr(dl), _ptr(dr), ctypes.byref(st),
                                 _ptr(reasons), _ptr(boundary))
    if rc != 0:
        raise ValueError("oracle_compute_ex failed: %d" % rc)
    return dict(checked=checked, dl=dl, dr=dr, stats=st.as_dict(), reasons=reasons,
                boundary=boundary)


def postfilter(p, f, left, right, dr, checked, w, h, stats=None, reasons=None, photo=None):
    """Median, reciprocity + photometric ZNCC and speckles on a copy of `checked`.
    ``photo`` = (left_bgr, right_bgr) scores the photometric test over three
    channels (colour-distance mode).  Returns (filtered, stats dict, reasons)."""
    out = np.array(checked, np.float32, copy=True).reshape(-1)
    st = OracleStats()
    if stats:
        for k, v in stats.items():
            setattr(st, k, v)
    rs = None if reasons is None else np.array(reasons, np.uint8, copy=True).reshape(-1)
    left = np.ascontiguousarray(left, np.uint8).reshape(-1)
    right = np.ascontiguousarray(right, np.uint8).reshape(-1)
    dr = np.ascontiguousarray(dr, np.float32).reshape(-1)
    if photo is None:
        pl, pr, ch = left, right, 1
    else:
        pl, pr = (np.ascontiguousarray(a, np.uint8).reshape(-1) for a in photo)
        ch = 3
    rc = lib().oracle_postfilter_ch(ctypes.byref(p), ctypes.byref(f), _ptr(left), _ptr(pl),
                                    _ptr(pr), ch, _ptr(dr), w, h, _ptr(out), ctypes.byref(st),
                                    _ptr(rs))
    if rc != 0:
        raise ValueError("oracle_postfilter failed: %d" % rc)
    return out, st.as_dict(), rs


def photo_zncc(a, b, w, h, x, y, disparity, r):
    a = np.ascontiguousarray(a, np.uint8).reshape(-1)
    b = np.ascontiguousarray(b, np.uint8).reshape(-1)
    return lib().oracle_photo_zncc(_ptr(a), _ptr(b), w, h, x, y, float(disparity), r)


def photo_zncc_ch(a, b, w, h, x, y, disparity, r, channels):
    a = np.ascontiguousarray(a, np.uint8).reshape(-1)
    b = np.ascontiguousarray(b, np.uint8).reshape(-1)
    return lib().oracle_photo_zncc_ch(_ptr(a), _ptr(b), w, h, x, y, float(disparity), r, channels)


def compute_filtered(p, f, left, right, w, h, valid_left=None, valid_right=None, left_bgr=None,
                     right_bgr=None):
    """CensusSgm.computeRecorded through 05_final_stereo (colour-distance census
    and three-channel photometric test when the BGR rasters are given)."""
    r = compute_ex(p, f, left, right, w, h, valid_left, valid_right, left_bgr, right_bgr)
    photo = None if left_bgr is None else (left_bgr, right_bgr)
    out, st, rs = postfilter(p, f, left, right, r["dr"], r["checked"], w, h, r["stats"],
                             r["reasons"], photo=photo)
    return dict(r, disparity=out, stats=st, reasons=rs)


def round_float(v):
    return lib().oracle_round_float(v)


def round_double(v):
    return lib().oracle_round_double(v)
