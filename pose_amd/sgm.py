"""ctypes binding of libsgm_hip and a Python mirror of the reference's GPU-SGM backend.

The reference (Java) reaches its GPU through ``GpuSgm`` (src/main/java/sgm/GpuSgm.java)
and ``GpuStereo`` (src/main/java/sgm/GpuStereo.java), called from
``CensusSgm.match`` (src/main/java/sgm/CensusSgm.java:359-374).  This module keeps
that interface -- ``Options`` field names, ``StereoCost`` unit conversions,
``GpuSgm.use / disparityLimit / match`` and ``CensusSgm.compute`` -- on top of the
C ABI declared in include/sgm_hip.h.  Every computation runs in the HIP library;
there is no CPU fallback in this package: when the library is missing or fails,
calls raise.
"""
import ctypes
import json
import math
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgm_hip.so")

SGM_OK = 0
ERRORS = {
    -1: "invalid argument", -2: "no device", -3: "out of memory", -4: "no plan",
    -5: "HIP error", -6: "rejected", -7: "timeout", -8: "not found", -9: "interrupted",
}

VOLUME_AUTO, VOLUME_RESIDENT, VOLUME_RECOMPUTE = 0, 1, 2
SWEEPS_AUTO, SWEEPS_SEQUENTIAL, SWEEPS_PAIRED, SWEEPS_PLANES = 0, 1, 2, 3
SWEEPS = {"auto": SWEEPS_AUTO, "sequential": SWEEPS_SEQUENTIAL, "paired": SWEEPS_PAIRED,
          "planes": SWEEPS_PLANES}
# aggregation plan: --gpu-sgm auto|resident|streamed|esgm (Options.java:32,780)
PLAN_AUTO, PLAN_RESIDENT, PLAN_STREAMED, PLAN_ESGM = 0, 1, 2, 3
# fixed per-device workspace of match_pair_device_noalloc (SGM_NOALLOC_WORKSPACE_BYTES)
NOALLOC_WORKSPACE_BYTES = 128 << 20
PLAN_MODES = {"auto": PLAN_AUTO, "resident": PLAN_RESIDENT, "streamed": PLAN_STREAMED,
              "esgm": PLAN_ESGM}
# streamed aggregate storage: --gpu-aggregate auto|ram|disk|recompute (GpuStereo.java:56-83)
AGGREGATE_AUTO, AGGREGATE_RAM, AGGREGATE_DISK, AGGREGATE_RECOMPUTE = 0, 1, 2, 3
AGGREGATE_STORAGES = {"auto": AGGREGATE_AUTO, "ram": AGGREGATE_RAM, "disk": AGGREGATE_DISK,
                      "recompute": AGGREGATE_RECOMPUTE}
COSTS = {"census": 0, "ad-census": 1, "census-gradient": 2, "zncc": 3, "zncc-fp32": 4}


class SgmParams(ctypes.Structure):
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


STAT_FIELDS = [
    "lower_boundary_rejected", "full_interval_pixels", "boundary_rejected",
    "upper_boundary_rejected", "ambiguity_rejected", "selected", "lr_rejected",
    "post_median_lr_rejected", "photometric_rejected", "speckle_rejected", "valid",
]


class SgmFilterParams(ctypes.Structure):
    _fields_ = [
        ("median", ctypes.c_int32), ("median_classic", ctypes.c_int32),
        ("median_edge_threshold", ctypes.c_int32), ("correlation_window", ctypes.c_int32),
        ("speckle_size", ctypes.c_int32), ("speckle_connectivity", ctypes.c_int32),
        ("subpixel_refinement", ctypes.c_int32), ("reserved1", ctypes.c_int32),
        ("median_max_change", ctypes.c_double), ("correlation", ctypes.c_double),
        ("speckle_range", ctypes.c_double), ("reserved2", ctypes.c_double),
    ]


class SgmStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int64) for f in STAT_FIELDS] + [("reserved", ctypes.c_int64 * 5)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f in STAT_FIELDS}


class SgmDeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64),
                ("memory", ctypes.c_int64), ("allocation", ctypes.c_int64),
                ("local_memory", ctypes.c_int64), ("group_size", ctypes.c_int32),
                ("compute_units", ctypes.c_int32)]


class SgmPlan(ctypes.Structure):
    _fields_ = [("resident", ctypes.c_int32), ("aggregate_bytes", ctypes.c_int32),
                ("padded_disparities", ctypes.c_int32),
                ("lanes_per_disparity_group", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("pair_bytes", ctypes.c_int64),
                ("budget_bytes", ctypes.c_int64), ("aggregate_bound", ctypes.c_int64),
                ("streamed", ctypes.c_int32), ("strip_rows", ctypes.c_int32),
                ("streamed_bytes", ctypes.c_int64), ("host_bytes", ctypes.c_int64),
                ("esgm_candidates", ctypes.c_int32), ("esgm_rows", ctypes.c_int32),
                ("esgm_bytes", ctypes.c_int64), ("aggregate_storage", ctypes.c_int32),
                ("block_rows", ctypes.c_int32), ("disk_bytes", ctypes.c_int64),
                ("compact", ctypes.c_int32), ("sweeps", ctypes.c_int32)]


class SgmPair(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("disparities", ctypes.c_int32), ("min_disparity", ctypes.c_int32),
                ("left", ctypes.c_void_p), ("right", ctypes.c_void_p),
                ("valid_left", ctypes.c_void_p), ("valid_right", ctypes.c_void_p),
                ("checked", ctypes.c_void_p), ("dl", ctypes.c_void_p), ("dr", ctypes.c_void_p),
                ("reasons", ctypes.c_void_p), ("boundary", ctypes.c_void_p),
                ("stats", ctypes.c_void_p)]


class SgmRequestInfo(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("admitted_s", ctypes.c_double),
                ("started_s", ctypes.c_double), ("finished_s", ctypes.c_double),
                ("device_bytes", ctypes.c_int64), ("status", ctypes.c_int32),
                ("batch_id", ctypes.c_int32), ("pool_bytes", ctypes.c_int64)]


EXPORTS = [
    "sgm_device_count", "sgm_device_info_get", "sgm_open", "sgm_close", "sgm_plan_get",
    "sgm_disparity_limit", "sgm_match", "sgm_match_desc", "sgm_match_pair", "sgm_match_pair_device",
    "sgm_match_pair_device_noalloc",
    "sgm_match_pair_filtered", "sgm_match_pair_color",
    "sgm_set_keep_volumes", "sgm_read_volumes", "sgm_set_cost_volume", "sgm_set_sweeps",
    "sgm_set_plan_mode",
    "sgm_set_esgm_candidates", "sgm_esgm_counts", "sgm_synchronize",
    "sgm_set_aggregate_storage", "sgm_interrupt", "sgm_debug_fail_host_allocations",
    "sgm_release_buffers", "sgm_set_profiling",
    "sgm_profile_json", "sgm_stream", "sgm_debug_fail_allocations", "sgm_last_error",
    "sgm_device_alloc", "sgm_device_free", "sgm_memcpy", "sgm_device_copy",
    "sgm_batch_create", "sgm_batch_run", "sgm_batch_destroy", "sgm_batch_device_bytes",
    "sgm_batch_describe",
    "sgm_autotune", "sgm_autotune_load",
    "sgm_server_create", "sgm_server_submit", "sgm_server_wait", "sgm_server_now",
    "sgm_server_destroy",
]

_lib = None
_lib_lock = threading.Lock()


class SgmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s (%s, code %d)" % (message, ERRORS.get(code, "error"), code))
        self.code = code


def lib():
    """Load libsgm_hip.so; raises if it has not been built (no silent fallback)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OSError("libsgm_hip.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
            pp = ctypes.POINTER(SgmParams)
            ps = ctypes.POINTER(SgmStats)
            L.sgm_last_error.restype = ctypes.c_char_p
            L.sgm_device_count.argtypes = [ctypes.POINTER(I)]
            L.sgm_device_info_get.argtypes = [I, ctypes.POINTER(SgmDeviceInfo)]
            L.sgm_open.argtypes = [I, I64, ctypes.POINTER(P)]
            L.sgm_close.argtypes = [P]
            L.sgm_plan_get.argtypes = [P, pp, I, I, ctypes.POINTER(SgmPlan)]
            L.sgm_disparity_limit.argtypes = [P, pp, I, I, ctypes.POINTER(I)]
            L.sgm_match.argtypes = [P, pp, P, P, P, P, I, I, I, P, ps, P, P]
            L.sgm_match_desc.argtypes = [P, pp, P, P, P, P, P, P, I, I, I, P, ps, P, P]
            L.sgm_match_pair.argtypes = [P, pp, P, P, P, P, I, I, P, P, P, ps, P, P]
            L.sgm_match_pair_device.argtypes = [P, pp, P, P, P, P, I, I, P, P, P, P, P, P]
            L.sgm_match_pair_device_noalloc.argtypes = [P, pp, P, P, P, P, I, I, P, P, P, P, P, P]
            L.sgm_match_pair_filtered.argtypes = [P, pp, ctypes.POINTER(SgmFilterParams), P, P, P,
                                                  P, I, I, P, ps, P]
            L.sgm_match_pair_color.argtypes = [P, pp, ctypes.POINTER(SgmFilterParams), P, P, P,
                                               P, P, P, I, I, P, P, P, ps, P, P]
            L.sgm_set_keep_volumes.argtypes = [P, I]
            L.sgm_set_cost_volume.argtypes = [P, I]
            L.sgm_set_sweeps.argtypes = [P, I]
            L.sgm_set_plan_mode.argtypes = [P, I, I]
            L.sgm_set_esgm_candidates.argtypes = [P, I]
            L.sgm_set_aggregate_storage.argtypes = [P, I, ctypes.c_char_p, I64]
            L.sgm_interrupt.argtypes = [P]
            L.sgm_debug_fail_host_allocations.argtypes = [P, I]
            L.sgm_esgm_counts.argtypes = [P, ctypes.POINTER(I64)]
            L.sgm_read_volumes.argtypes = [P, I, P, P]
            L.sgm_synchronize.argtypes = [P]
            L.sgm_release_buffers.argtypes = [P]
            L.sgm_set_profiling.argtypes = [P, I]
            L.sgm_profile_json.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
            L.sgm_stream.argtypes = [P]
            L.sgm_stream.restype = P
            L.sgm_debug_fail_allocations.argtypes = [P, I]
            L.sgm_device_alloc.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
            L.sgm_device_free.argtypes = [P, P]
            L.sgm_memcpy.argtypes = [P, P, P, ctypes.c_size_t, I]
            L.sgm_device_copy.argtypes = [P, P, P, ctypes.c_size_t]
            L.sgm_batch_create.argtypes = [P, pp, ctypes.POINTER(SgmPair), I, I, ctypes.POINTER(P)]
            L.sgm_batch_run.argtypes = [P]
            L.sgm_batch_destroy.argtypes = [P]
            L.sgm_batch_device_bytes.argtypes = [P]
            L.sgm_batch_device_bytes.restype = I64
            L.sgm_batch_describe.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
            L.sgm_autotune.argtypes = [P, pp, ctypes.c_char_p]
            L.sgm_autotune_load.argtypes = [P, ctypes.c_char_p]
            L.sgm_server_create.argtypes = [P, pp, I64, I, I, ctypes.POINTER(P)]
            L.sgm_server_submit.argtypes = [P, ctypes.POINTER(SgmPair), ctypes.POINTER(ctypes.c_uint64)]
            L.sgm_server_wait.argtypes = [P, ctypes.c_uint64, ctypes.c_double,
                                          ctypes.POINTER(SgmRequestInfo)]
            L.sgm_server_now.argtypes = [P]
            L.sgm_server_now.restype = ctypes.c_double
            L.sgm_server_destroy.argtypes = [P]
            _lib = L
        return _lib


def check(rc):
    if rc != SGM_OK:
        raise SgmError(rc, lib().sgm_last_error().decode(errors="replace"))
    return rc


def _p(a):
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data


# ---------------------------------------------------------------------------
# Reference-interface mirror: Options subset, StereoCost, Stats, Result.


def java_round(v):
    """Math.round(double): closest integer, ties toward +infinity."""
    f = math.floor(v)
    return int(f) + (1 if v - f >= 0.5 else 0)


class Options:
    """SGM subset of the reference's Options (Options.java:17-40,67), same field names.

    ``censusWidth``/``censusHeight`` generalise ``census`` to rectangular windows
    (0 = use ``census``)."""

    def __init__(self, **kw):
        self.census = 5
        self.censusWidth = 0
        self.censusHeight = 0
        self.disparities = 128
        self.minDisparity = 0
        self.disparitiesExplicit = False      # --disparities given (Options.java)
        self.minDisparityExplicit = False     # --min-disparity given
        self.matchingCost = "census"
        self.p1 = 4.0
        self.p2 = 24.0
        self.sgmEdgeScale = 32
        self.directions = 8
        self.uniqueness = 0.03
        self.lrThreshold = 0.75
        self.sgmAggregate = "auto"
        self.censusWeight = 0.5
        self.adTruncation = 30
        self.gradientTruncation = 60
        self.compute = "auto"
        self.gpuDevice = 0
        self.gpuMemoryMb = 0
        self.gpuSgm = "auto"        # auto | resident | streamed | esgm (Options.java:32)
        self.esgmCandidates = 2     # Options.java:35
        self.censusMode = "gray"    # gray | color-distance (Options.java:31)
        self.gpuStripRows = 64      # Options.java:39
        self.gpuAggregate = "auto"  # auto | ram | disk | recompute (streamed aggregate storage)
        self.memoryMb = 0           # host memory of the streamed plan (0: 65% of RAM)
        self.output = None          # directory of the disk aggregate storage
        self.rawPenalties = False  # p1/p2 already in raw scaled units
        # post-L-R filters (Options.java:19-23,67,76)
        self.median = 3
        self.medianMode = "edge-preserving"
        self.medianMaxChange = 0.5
        self.medianEdgeThreshold = 24
        self.correlationWindow = 7
        self.correlation = 0.55
        self.speckleSize = 50
        self.speckleRange = 1.0
        self.speckleConnectivity = 8
        self.subpixelRefinement = "none"   # none | zncc (Options.java:37)
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError("unknown option " + k)
            setattr(self, k, v)

    def copy(self):
        o = Options()
        o.__dict__.update(self.__dict__)
        return o

    def window(self):
        return (self.censusWidth or self.census, self.censusHeight or self.census)


class StereoCost:
    """StereoCost.java:7-48."""

    @staticmethod
    def code(o):
        if o.matchingCost not in COSTS:
            raise ValueError("Unknown matching cost")
        return COSTS[o.matchingCost]

    @staticmethod
    def penalty(value, o):
        ww, wh = o.window()
        return int(java_round(value * (ww * wh - 1) / 24.0))

    @staticmethod
    def p1(o):
        return int(o.p1) if o.rawPenalties else StereoCost.penalty(o.p1, o)

    @staticmethod
    def p2(o):
        raw = int(o.p2) if o.rawPenalties else StereoCost.penalty(o.p2, o)
        return max(StereoCost.p1(o) + 1, raw)

    @staticmethod
    def aggregateBound(o):
        maxc = 2048 if o.matchingCost == "zncc-fp32" else 512
        return o.directions * (maxc + StereoCost.p2(o))

    @staticmethod
    def aggregateBytes(o):
        fits = StereoCost.aggregateBound(o) <= 65535
        if o.sgmAggregate == "16" and not fits:
            raise ValueError("16-bit SGM aggregate cannot hold proved bound %d"
                             % StereoCost.aggregateBound(o))
        return 2 if o.sgmAggregate != "32" and fits else 4


def filters_from_options(o):
    """sgm_filter_params from the reference's option names."""
    if o.medianMode not in ("edge-preserving", "classic"):
        raise ValueError("--median-mode must be edge-preserving or classic")
    if o.subpixelRefinement not in ("none", "zncc"):
        raise ValueError("--subpixel-refinement must be none or zncc")
    return SgmFilterParams(
        median=o.median, median_classic=1 if o.medianMode == "classic" else 0,
        median_edge_threshold=o.medianEdgeThreshold, correlation_window=o.correlationWindow,
        speckle_size=o.speckleSize, speckle_connectivity=4 if o.speckleConnectivity == 4 else 8,
        subpixel_refinement=1 if o.subpixelRefinement == "zncc" else 0, reserved1=0,
        median_max_change=o.medianMaxChange,
        correlation=o.correlation, speckle_range=o.speckleRange, reserved2=0.0)


def params_from_options(o):
    ww, wh = o.window()
    return SgmParams(
        window_w=ww, window_h=wh, disparities=o.disparities, min_disparity=o.minDisparity,
        cost=StereoCost.code(o), p1=StereoCost.p1(o), p2=StereoCost.p2(o),
        edge_scale=o.sgmEdgeScale, directions=o.directions,
        aggregate_bytes=StereoCost.aggregateBytes(o),
        census_weight=int(java_round(o.censusWeight * 1000)),
        ad_truncation=o.adTruncation, gradient_truncation=o.gradientTruncation, reserved=0,
        uniqueness=o.uniqueness, lr_threshold=o.lrThreshold)


class Stats:
    """CensusSgm.Stats (CensusSgm.java:683-737)."""

    def __init__(self):
        for f in STAT_FIELDS:
            setattr(self, f, 0)

    def add_match(self, d):
        # GpuSgm.java:289-294
        for f in ("lower_boundary_rejected", "full_interval_pixels", "boundary_rejected",
                  "upper_boundary_rejected", "ambiguity_rejected", "selected"):
            setattr(self, f, getattr(self, f) + d[f])

    def set_all(self, d):
        for f in STAT_FIELDS:
            setattr(self, f, d[f])

    def csv(self):
        return ",".join(str(getattr(self, f)) for f in (
            "boundary_rejected", "upper_boundary_rejected", "ambiguity_rejected", "selected",
            "lr_rejected", "post_median_lr_rejected", "photometric_rejected", "speckle_rejected",
            "valid"))

    def as_dict(self):
        return {f: getattr(self, f) for f in STAT_FIELDS}


class Result:
    """CensusSgm.Result (CensusSgm.java:739-767): final disparity, stats, reason codes."""

    def __init__(self, width, height, disparity, stats, reasons=None, boundary=None,
                 left=None, right=None):
        self.width, self.height = width, height
        self.disparity = disparity
        self.stats = stats
        self.reasons = reasons
        self.searchBoundary = boundary
        self.left = left
        self.right = right


# ---------------------------------------------------------------------------
# Context: one device session (GpuSgm session, GpuSgm.java:163-221).


class Context:
    def __init__(self, device=0, budget_bytes=0):
        h = ctypes.c_void_p()
        check(lib().sgm_open(device, int(budget_bytes), ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib().sgm_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, params, w, h):
        out = SgmPlan()
        check(lib().sgm_plan_get(self.h, ctypes.byref(params), w, h, ctypes.byref(out)))
        return {f: getattr(out, f) for f, _ in SgmPlan._fields_}

    def disparity_limit(self, params, w, h):
        v = ctypes.c_int()
        check(lib().sgm_disparity_limit(self.h, ctypes.byref(params), w, h, ctypes.byref(v)))
        return v.value

    def match_pair(self, params, left, right, valid_left=None, valid_right=None):
        """Both directions + L-R check on host arrays (h, w) uint8."""
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        h, w = left.shape
        n = w * h
        vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8)
        vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8)
        checked = np.empty(n, np.float32)
        dl = np.empty(n, np.float32)
        dr = np.empty(n, np.float32)
        reasons = np.empty(n, np.uint8)
        boundary = np.empty(n, np.uint8)
        st = SgmStats()
        check(lib().sgm_match_pair(self.h, ctypes.byref(params), _p(left), _p(right), _p(vl),
                                   _p(vr), w, h, _p(checked), _p(dl), _p(dr), ctypes.byref(st),
                                   _p(reasons), _p(boundary)))
        return dict(checked=checked, dl=dl, dr=dr, stats=st.as_dict(), reasons=reasons,
                    boundary=boundary)

    def match_pair_filtered(self, params, filters, left, right, valid_left=None,
                            valid_right=None):
        """CensusSgm.computeRecorded through 05_final_stereo on host arrays (h, w)."""
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        h, w = left.shape
        n = w * h
        vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8)
        vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8)
        disp = np.empty(n, np.float32)
        reasons = np.empty(n, np.uint8)
        st = SgmStats()
        check(lib().sgm_match_pair_filtered(self.h, ctypes.byref(params), ctypes.byref(filters),
                                            _p(left), _p(right), _p(vl), _p(vr), w, h, _p(disp),
                                            ctypes.byref(st), _p(reasons)))
        return dict(disparity=disp, stats=st.as_dict(), reasons=reasons)

    def match_pair_color(self, params, left, right, left_bgr, right_bgr, filters=None,
                         valid_left=None, valid_right=None):
        """--census-mode color-distance: gray (h, w) and interleaved BGR (h, w, 3)
        rasters.  Without filters the L-R-validated map, else the final map."""
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        h, w = left.shape
        lb = np.ascontiguousarray(left_bgr, np.uint8)
        rb = np.ascontiguousarray(right_bgr, np.uint8)
        if lb.size != 3 * w * h or rb.size != lb.size:
            raise ValueError("Color-distance census requires interleaved BGR inputs")
        n = w * h
        vl = None if valid_left is None else np.ascontiguousarray(valid_left, np.uint8)
        vr = None if valid_right is None else np.ascontiguousarray(valid_right, np.uint8)
        disp, dl, dr = (np.empty(n, np.float32) for _ in range(3))
        reasons = np.empty(n, np.uint8)
        boundary = np.empty(n, np.uint8)
        st = SgmStats()
        fp = None if filters is None else ctypes.byref(filters)
        check(lib().sgm_match_pair_color(self.h, ctypes.byref(params), fp, _p(left), _p(right),
                                         _p(lb), _p(rb), _p(vl), _p(vr), w, h, _p(disp), _p(dl),
                                         _p(dr), ctypes.byref(st), _p(reasons), _p(boundary)))
        return dict(disparity=disp, checked=disp, dl=dl, dr=dr, stats=st.as_dict(),
                    reasons=reasons, boundary=boundary)

    def match_desc(self, params, desc_a, desc_b, gray_a, gray_b, mask_a, mask_b, w, h, reverse,
                   stats=None, reasons=None, boundary=None):
        """sgm_match with the caller's word-major uint64 descriptor planes."""
        words = (params.window_w * params.window_h - 1 + 63) // 64
        da = np.ascontiguousarray(desc_a, np.uint64).reshape(-1)
        db = np.ascontiguousarray(desc_b, np.uint64).reshape(-1)
        if da.size != words * w * h or db.size != da.size:
            raise ValueError("descriptor planes must hold %d words per pixel" % words)
        disp = np.empty(w * h, np.float32)
        st = SgmStats() if stats is None else stats
        check(lib().sgm_match_desc(self.h, ctypes.byref(params), _p(da), _p(db), _p(gray_a),
                                   _p(gray_b), _p(mask_a), _p(mask_b), w, h, int(reverse),
                                   _p(disp), ctypes.byref(st), _p(reasons), _p(boundary)))
        return disp, st

    def match(self, params, gray_a, gray_b, mask_a, mask_b, w, h, reverse, stats=None,
              reasons=None, boundary=None):
        disp = np.empty(w * h, np.float32)
        st = SgmStats() if stats is None else stats
        check(lib().sgm_match(self.h, ctypes.byref(params), _p(gray_a), _p(gray_b), _p(mask_a),
                              _p(mask_b), w, h, int(reverse), _p(disp), ctypes.byref(st),
                              _p(reasons), _p(boundary)))
        return disp, st

    def match_pair_device(self, params, d_left, d_right, w, h, d_checked, d_valid_left=None,
                          d_valid_right=None, d_dl=None, d_dr=None, d_stats=None,
                          d_reasons=None, d_boundary=None):
        check(lib().sgm_match_pair_device(self.h, ctypes.byref(params), _p(d_left), _p(d_right),
                                          _p(d_valid_left), _p(d_valid_right), w, h,
                                          _p(d_checked), _p(d_dl), _p(d_dr), _p(d_stats),
                                          _p(d_reasons), _p(d_boundary)))

    def match_pair_device_noalloc(self, params, d_left, d_right, w, h, d_checked,
                                  d_valid_left=None, d_valid_right=None, d_dl=None, d_dr=None,
                                  d_stats=None, d_reasons=None, d_boundary=None):
        """match_pair_device in the library's fixed per-device workspace: no
        device allocation during the call (pairs up to NOALLOC_WORKSPACE_BYTES)."""
        check(lib().sgm_match_pair_device_noalloc(self.h, ctypes.byref(params), _p(d_left),
                                                  _p(d_right), _p(d_valid_left),
                                                  _p(d_valid_right), w, h, _p(d_checked),
                                                  _p(d_dl), _p(d_dr), _p(d_stats),
                                                  _p(d_reasons), _p(d_boundary)))

    def keep_volumes(self, keep=True):
        check(lib().sgm_set_keep_volumes(self.h, int(keep)))

    def cost_volume(self, policy):
        """Cost-volume policy: VOLUME_AUTO, VOLUME_RESIDENT or VOLUME_RECOMPUTE."""
        check(lib().sgm_set_cost_volume(self.h, int(policy)))

    def sweeps(self, schedule):
        """Sweep schedule of the resident plan: SWEEPS_AUTO (default),
        SWEEPS_SEQUENTIAL, SWEEPS_PAIRED or SWEEPS_PLANES (or their names)."""
        schedule = SWEEPS[schedule] if isinstance(schedule, str) else int(schedule)
        check(lib().sgm_set_sweeps(self.h, schedule))

    def plan_mode(self, mode, strip_rows=64):
        """Aggregation plan: PLAN_AUTO, PLAN_RESIDENT, PLAN_STREAMED or PLAN_ESGM
        (or the reference's names "auto", "resident", "streamed", "esgm") and
        gpuStripRows."""
        mode = PLAN_MODES[mode] if isinstance(mode, str) else int(mode)
        check(lib().sgm_set_plan_mode(self.h, mode, int(strip_rows)))

    def aggregate_storage(self, storage, directory=None, host_budget_bytes=0):
        """Host storage of the streamed aggregate (--gpu-aggregate): AGGREGATE_*
        or its name, the directory of the disk storage, the host budget
        (Options.memoryMb; 0: 65% of physical memory)."""
        storage = AGGREGATE_STORAGES.get(storage, storage)
        check(lib().sgm_set_aggregate_storage(
            self.h, int(storage), None if directory is None else os.fsencode(directory),
            int(host_budget_bytes)))

    def interrupt(self):
        """Stop the running (or next) matching call of this context: it fails
        with SGM_ERR_INTERRUPTED (GpuStereo.java:350-351)."""
        check(lib().sgm_interrupt(self.h))

    def fail_host_allocations(self, count):
        check(lib().sgm_debug_fail_host_allocations(self.h, count))

    def esgm_candidates(self, count):
        """--esgm-candidates: proposals per path of the compact plan (1..4)."""
        check(lib().sgm_set_esgm_candidates(self.h, int(count)))

    def esgm_counts(self):
        """Certified and verified pixels of the last compact run:
        {"left": (certified, verified), "right": (certified, verified)}."""
        out = (ctypes.c_int64 * 4)()
        check(lib().sgm_esgm_counts(self.h, out))
        return {"left": (out[0], out[1]), "right": (out[2], out[3])}

    def read_volumes(self, direction, n, D):
        cost = np.empty(n * D, np.uint16)
        agg = np.empty(n * D, np.uint32)
        check(lib().sgm_read_volumes(self.h, direction, _p(cost), _p(agg)))
        return cost, agg

    def synchronize(self):
        check(lib().sgm_synchronize(self.h))

    def release_buffers(self):
        check(lib().sgm_release_buffers(self.h))

    def set_profiling(self, on=True):
        check(lib().sgm_set_profiling(self.h, int(on)))

    def profile(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().sgm_profile_json(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def stream(self):
        return lib().sgm_stream(self.h)

    # -- device buffers owned by the caller (plain device pointers) --
    def device_copy(self, dst, src, nbytes):
        """Asynchronous device-to-device copy by the library's streaming kernel."""
        check(lib().sgm_device_copy(self.h, dst.ptr, src.ptr, int(nbytes)))

    def alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        check(lib().sgm_device_alloc(self.h, int(nbytes), ctypes.byref(ptr)))
        return DeviceBuffer(self, ptr.value, int(nbytes))

    def fail_allocations(self, count):
        check(lib().sgm_debug_fail_allocations(self.h, count))

    def autotune(self, params, path=None):
        check(lib().sgm_autotune(self.h, ctypes.byref(params),
                                 None if path is None else path.encode()))

    def autotune_load(self, path):
        check(lib().sgm_autotune_load(self.h, path.encode()))


class DeviceBuffer:
    """A device allocation made through the library (freed on close/garbage collection)."""

    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def data_ptr(self):
        return self.ptr

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        check(lib().sgm_memcpy(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes, 0))
        return self

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        check(lib().sgm_memcpy(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().sgm_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count():
    n = ctypes.c_int()
    check(lib().sgm_device_count(ctypes.byref(n)))
    return n.value


def device_info(index):
    d = SgmDeviceInfo()
    check(lib().sgm_device_info_get(index, ctypes.byref(d)))
    return dict(name=d.name.decode(), arch=d.arch.decode(), memory=d.memory,
                allocation=d.allocation, local_memory=d.local_memory, group_size=d.group_size,
                compute_units=d.compute_units)


# ---------------------------------------------------------------------------
# GpuSgm mirror (GpuSgm.java:237-303) -- one session per thread.

_session = threading.local()
_warned = threading.local()


class GpuSgm:
    @staticmethod
    def session(o):
        s = getattr(_session, "ctx", None)
        if s is not None and s.device == o.gpuDevice:
            return s
        GpuSgm.closeSession()
        if o.gpuDevice >= device_count():
            raise RuntimeError("No GPU at index %d" % o.gpuDevice)
        s = Context(o.gpuDevice, o.gpuMemoryMb * 1048576)
        _session.ctx = s
        return s

    @staticmethod
    def configured(o):
        """The session with the options' plan mode (--gpu-sgm, --gpu-strip-rows)."""
        if o.gpuSgm not in PLAN_MODES:
            raise ValueError("--gpu-sgm must be auto, resident, streamed or esgm")
        if not 1 <= o.esgmCandidates <= 4:
            raise ValueError("--esgm-candidates must be 1..4")
        if o.gpuAggregate not in AGGREGATE_STORAGES:
            raise ValueError("--gpu-aggregate must be auto, ram, disk or recompute")
        s = GpuSgm.session(o)
        s.plan_mode(o.gpuSgm, o.gpuStripRows)
        s.esgm_candidates(o.esgmCandidates)
        s.aggregate_storage(o.gpuAggregate, o.output, o.memoryMb * 1048576)
        return s

    @staticmethod
    def closeSession():
        s = getattr(_session, "ctx", None)
        _session.ctx = None
        if s is not None:
            s.close()
        _warned.messages = set()

    @staticmethod
    def failure(o, e):
        # GpuSgm.java:260-265
        if o.compute == "gpu":
            raise ValueError("GPU execution required: %s" % e) from e
        msgs = getattr(_warned, "messages", None)
        if msgs is None:
            msgs = _warned.messages = set()
        if str(e) not in msgs:
            msgs.add(str(e))
            print("GPU unavailable; using CPU: %s" % e)

    @staticmethod
    def disparityLimit(w, h, o):
        if o.compute == "cpu":
            return 0
        try:
            return GpuSgm.configured(o).disparity_limit(params_from_options(o), w, h)
        except (RuntimeError, OSError) as e:
            GpuSgm.failure(o, e)
            return 0

    @staticmethod
    def use(w, h, o):
        # GpuSgm.java:247-258
        if o.compute == "cpu" or (o.compute == "auto" and w * h * o.disparities < 1000000):
            return False
        limit = GpuSgm.disparityLimit(w, h, o)
        if o.disparities <= limit:
            return True
        if o.compute == "gpu":
            raise ValueError("GPU SGM search exceeds VRAM/allocation/local-memory budget; limit=%d"
                             " disparities." % limit)
        return False

    @staticmethod
    def match(o, a, b, grayA, grayB, ta, tb, w, h, reverse, stats, reasons, searchBoundary):
        """GpuSgm.match: returns the disparity map or None (= caller's CPU fallback).

        ``a``/``b`` (descriptors) are accepted for signature parity; the engine
        recomputes census on the device from ``grayA``/``grayB``."""
        if not GpuSgm.use(w, h, o):
            return None
        try:
            ctx = GpuSgm.configured(o)
            st = SgmStats()
            codes = None if reasons is None else np.array(reasons, np.uint8, copy=True)
            bnd = None if searchBoundary is None else np.array(searchBoundary, np.uint8, copy=True)
            g = [np.ascontiguousarray(v, np.uint8) for v in (grayA, grayB, ta, tb)]
            if o.censusMode == "color-distance":
                # the caller's colour-distance descriptors (CensusSgm.java:109-121)
                disp, st = ctx.match_desc(params_from_options(o), np.asarray(a, np.uint64),
                                          np.asarray(b, np.uint64), *g, w, h, reverse, st, codes,
                                          bnd)
            else:
                disp, st = ctx.match(params_from_options(o), *g, w, h, reverse, st, codes, bnd)
            stats.add_match(st.as_dict())
            if codes is not None:
                reasons[:] = codes
            if bnd is not None:
                searchBoundary[:] = bnd
            return disp
        except (RuntimeError, OSError) as e:
            GpuSgm.closeSession()
            GpuSgm.failure(o, e)
            return None


class GpuStereo:
    """Planning and reporting of GpuStereo (GpuStereo.java:33-168) over
    sgm_plan_get / sgm_disparity_limit, under the options' plan mode."""

    @staticmethod
    def plan(w, h, o):
        return GpuSgm.configured(o).plan(params_from_options(o), w, h)

    @staticmethod
    def limit(w, h, o):
        return GpuSgm.configured(o).disparity_limit(params_from_options(o), w, h)

    @staticmethod
    def report(w, h, o):
        """GpuStereo.report (GpuStereo.java:95-168): the storage the pair runs
        with and its footprint, with the reference's keys.  This engine's plans
        differ in form: volumes are flat allocations (no banks), the streamed
        plan moves the aggregate over PCIe once per pair (one read, one write),
        and the compact plan verifies on the device within its own footprint
        (no host aggregate).  None for compute=cpu; {"unavailable": message}
        when planning fails."""
        if o.compute == "cpu":
            return None
        try:
            p = GpuStereo.plan(w, h, o)
            compact = bool(p["compact"])
            if compact and not p["esgm_candidates"]:
                raise SgmError(-4, "No compact Census/16-bit GPU plan fits")
            streamed = not compact and bool(p["streamed"])
            if not compact and not streamed and not p["resident"]:
                raise SgmError(-4, "No GPU SGM memory plan fits the requested range")
            sb = p["aggregate_bytes"]
            computed = compact and p["esgm_rows"] == h
            storage = {v: k for k, v in AGGREGATE_STORAGES.items()}.get(p["aggregate_storage"], "ram")
            disk = streamed and storage == "disk"
            blocks = -(-h // p["block_rows"]) if streamed and p["block_rows"] else 1
            strips = 1 if not (compact or streamed) else -(-h // (p["esgm_rows"] if compact else p["strip_rows"]))
            return {
                "storage": ("compact_with_sparse_exact_verification" if compact
                            else "streamed_" + storage if streamed else "resident"),
                "rows": p["esgm_rows"] if compact else p["strip_rows"] if streamed else h,
                "bank_bytes": 0,
                "estimated_device_bytes": (p["esgm_bytes"] if compact else p["streamed_bytes"]
                                           if streamed else p["pair_bytes"]),
                "estimated_host_bytes": p["host_bytes"] if streamed else 0,
                "host_reserve_bytes": 0,
                "aggregate_bits": 8 * sb,
                "aggregate_proved_bound": p["aggregate_bound"],
                # cost volume builds per matching direction: once (resident),
                # once per strip pass (streamed), none when the compact sweeps
                # compute their costs (3 passes x paths otherwise)
                "cost_generation_passes": (0 if computed else 3 * 2 if compact
                                           else 2 * blocks if streamed else 1),
                "cost_generation_passes_are_upper_bound": streamed and storage == "recompute",
                "aggregate_volume_reads": 1 if streamed else 0,
                "aggregate_volume_writes": 1 if streamed else 0,
                "aggregate_disk_bytes": p["disk_bytes"] if disk else 0,
                "aggregate_ram_bytes": p["host_bytes"] if streamed and not disk else 0,
                "verification_block_rows": h if compact else p["block_rows"] if streamed else None,
                "recomputation_blocks": strips if compact else blocks,
                "candidates_per_path": p["esgm_candidates"] if compact else None,
                "candidate_device_bytes": p["esgm_bytes"] if compact else 0,
            }
        except (RuntimeError, OSError, ValueError) as e:
            return {"unavailable": str(e)}


class CensusSgm:
    """CensusSgm.compute (CensusSgm.java:18-280) through the device path: matching,
    the left-right check and the post-L-R filters (median, reciprocity and
    photometric ZNCC, speckles), all on the GPU."""

    def __init__(self, options):
        self.options = options

    def compute(self, left, right, width, height, validLeft=None, validRight=None,
                leftBgr=None, rightBgr=None):
        o = self.options
        if o.censusMode not in ("gray", "color-distance"):
            raise ValueError("--census-mode must be gray or color-distance")
        ww, wh = o.window()
        left = np.ascontiguousarray(left, np.uint8).reshape(height, width)
        right = np.ascontiguousarray(right, np.uint8).reshape(height, width)
        if width < ww + 2 or height < wh + 2 or left.size != width * height or right.size != left.size:
            raise ValueError("Invalid stereo image dimensions")
        if (validLeft is None) != (validRight is None):
            raise ValueError("Invalid stereo mask dimensions")
        if o.compute == "cpu":
            raise RuntimeError("this engine has no CPU SGM path (compute=cpu)")
        try:
            ctx = GpuSgm.configured(o)
            if o.censusMode == "color-distance":
                # CensusSgm.java:109-116
                if leftBgr is None or rightBgr is None:
                    raise ValueError("Color-distance census requires interleaved BGR inputs")
                r = ctx.match_pair_color(params_from_options(o), left, right, leftBgr, rightBgr,
                                         filters_from_options(o), validLeft, validRight)
            else:
                r = ctx.match_pair_filtered(params_from_options(o), filters_from_options(o),
                                            left, right, validLeft, validRight)
        except (RuntimeError, OSError) as e:
            GpuSgm.closeSession()
            GpuSgm.failure(o, e)
            raise RuntimeError("GPU SGM failed and this engine has no CPU path: %s" % e) from e
        st = Stats()
        st.set_all(r["stats"])
        return Result(width, height, r["disparity"], st, r["reasons"])


# ---------------------------------------------------------------------------
# Batched and served modes.

BATCH_DEVICE, BATCH_GRAPH, BATCH_NO_SYNC, BATCH_PHASED = 1, 2, 4, 8


class Batch:
    """Many uneven pairs in one persistent launch (host numpy buffers).

    ``pairs``: list of dicts with left, right (uint8 (h, w)), disparities,
    min_disparity and optional valid_left / valid_right.  Outputs (checked, dl, dr,
    reasons, boundary, stats) are allocated here and refreshed by ``run()``."""

    def __init__(self, ctx, params, pairs, graph=False, phased=False):
        self.ctx = ctx
        self.params = params
        self.keep = []
        arr = (SgmPair * len(pairs))()
        self.outputs = []
        for i, pr in enumerate(pairs):
            left = np.ascontiguousarray(pr["left"], np.uint8)
            right = np.ascontiguousarray(pr["right"], np.uint8)
            h, w = left.shape
            n = w * h
            o = dict(checked=np.empty(n, np.float32), dl=np.empty(n, np.float32),
                     dr=np.empty(n, np.float32), reasons=np.empty(n, np.uint8),
                     boundary=np.empty(n, np.uint8), stats=SgmStats())
            vl = pr.get("valid_left")
            vr = pr.get("valid_right")
            vl = None if vl is None else np.ascontiguousarray(vl, np.uint8)
            vr = None if vr is None else np.ascontiguousarray(vr, np.uint8)
            self.keep += [left, right, vl, vr]
            arr[i] = SgmPair(width=w, height=h, disparities=pr["disparities"],
                             min_disparity=pr.get("min_disparity", 0), left=_p(left),
                             right=_p(right), valid_left=_p(vl), valid_right=_p(vr),
                             checked=_p(o["checked"]), dl=_p(o["dl"]), dr=_p(o["dr"]),
                             reasons=_p(o["reasons"]), boundary=_p(o["boundary"]),
                             stats=ctypes.addressof(o["stats"]))
            self.outputs.append(o)
        self.arr = arr
        h = ctypes.c_void_p()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_PHASED if phased else 0)
        check(lib().sgm_batch_create(ctx.h, ctypes.byref(params), arr, len(pairs), flags,
                                     ctypes.byref(h)))
        self.h = h

    def run(self):
        check(lib().sgm_batch_run(self.h))
        return [dict(checked=o["checked"], dl=o["dl"], dr=o["dr"], reasons=o["reasons"],
                     boundary=o["boundary"], stats=o["stats"].as_dict()) for o in self.outputs]

    def device_bytes(self):
        return lib().sgm_batch_device_bytes(self.h)

    def describe(self):
        buf = ctypes.create_string_buffer(1024)
        check(lib().sgm_batch_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if self.h:
            lib().sgm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    """A batch over caller device buffers (inputs resident in HBM): build it from
    sgm_pair records whose pointers are device pointers."""

    def __init__(self, ctx, params, records, graph=True, no_sync=False, phased=False):
        self.ctx = ctx
        arr = (SgmPair * len(records))(*records)
        self.arr = arr
        flags = (BATCH_DEVICE | (BATCH_GRAPH if graph else 0) | (BATCH_NO_SYNC if no_sync else 0)
                 | (BATCH_PHASED if phased else 0))
        h = ctypes.c_void_p()
        check(lib().sgm_batch_create(ctx.h, ctypes.byref(params), arr, len(records), flags,
                                     ctypes.byref(h)))
        self.h = h

    def run(self):
        check(lib().sgm_batch_run(self.h))

    def device_bytes(self):
        return lib().sgm_batch_device_bytes(self.h)

    def describe(self):
        buf = ctypes.create_string_buffer(1024)
        check(lib().sgm_batch_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if self.h:
            lib().sgm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server over host buffers."""

    def __init__(self, ctx, params, budget_bytes=0, max_batch=16, streams=2):
        self.ctx = ctx
        self.params = params
        h = ctypes.c_void_p()
        check(lib().sgm_server_create(ctx.h, ctypes.byref(params), int(budget_bytes), max_batch,
                                      streams, ctypes.byref(h)))
        self.h = h
        self.live = {}

    def submit(self, left, right, disparities, min_disparity=0):
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        h, w = left.shape
        n = w * h
        out = dict(checked=np.empty(n, np.float32), dl=np.empty(n, np.float32),
                   dr=np.empty(n, np.float32), stats=SgmStats(), left=left, right=right)
        rec = SgmPair(width=w, height=h, disparities=disparities, min_disparity=min_disparity,
                      left=_p(left), right=_p(right), checked=_p(out["checked"]),
                      dl=_p(out["dl"]), dr=_p(out["dr"]), stats=ctypes.addressof(out["stats"]))
        t = ctypes.c_uint64()
        check(lib().sgm_server_submit(self.h, ctypes.byref(rec), ctypes.byref(t)))
        self.live[t.value] = out
        return t.value

    def wait(self, ticket, timeout=-1.0):
        info = SgmRequestInfo()
        rc = lib().sgm_server_wait(self.h, ticket, float(timeout), ctypes.byref(info))
        out = self.live.pop(ticket)
        check(rc)
        out["info"] = {f: getattr(info, f) for f, _ in SgmRequestInfo._fields_}
        out["stats"] = out["stats"].as_dict()
        return out

    def now(self):
        return lib().sgm_server_now(self.h)

    def close(self):
        if self.h:
            lib().sgm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
