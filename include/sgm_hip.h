/*
 * libsgm_hip -- MI355X (gfx950) semi-global-matching dense-stereo engine, C ABI.
 *
 * Drop-in compute library behind the reference's GPU-SGM backend
 * (`--gpu-sgm resident`), replacing the OpenCL runtime the reference reaches
 * through the JDK FFM API (src/main/java/sgm/GpuSgm.java:30-54).  Plain
 * pointers and sizes only; every entry point returns 0 on success or a
 * negative SGM_ERR_* code and records a message for sgm_last_error().
 *
 * Ownership: the library owns all device memory it allocates; the caller owns
 * every host pointer and every device pointer it passes in.  A context is bound
 * to one device and one HIP stream; calls on one context are serialised by the
 * library (the reference serialises all OpenCL calls, GpuSgm.java:30).
 */
#ifndef SGM_HIP_H
#define SGM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGM_OK 0
#define SGM_ERR_INVALID_ARGUMENT (-1)
#define SGM_ERR_NO_DEVICE (-2)
#define SGM_ERR_OUT_OF_MEMORY (-3)
#define SGM_ERR_NO_PLAN (-4)        /* no device memory plan fits (GpuStereo.java:231-232) */
#define SGM_ERR_HIP (-5)            /* a HIP runtime call failed */
#define SGM_ERR_REJECTED (-6)       /* serving admission refused the request */
#define SGM_ERR_TIMEOUT (-7)
#define SGM_ERR_NOT_FOUND (-8)
#define SGM_ERR_INTERRUPTED (-9)    /* sgm_interrupt stopped the call (GpuStereo.java:350-351) */

/* Matching costs: 0..3 are StereoCost.code (StereoCost.java:7-15); 3 is the
 * reference's exact integer/FP64 ZNCC.  4 is the FP32 ZNCC variant with cost
 * round((1 - ncc) * 1024) from FP32 window sums; invalid cells are stored as
 * 0x8800 and enter the recurrence as 2048 (the flag bit is masked off). */
#define SGM_COST_CENSUS 0
#define SGM_COST_AD_CENSUS 1
#define SGM_COST_CENSUS_GRADIENT 2
#define SGM_COST_ZNCC 3
#define SGM_COST_ZNCC_FP32 4

/* SGM options subset (Options.java:17-40,67).  Penalties are raw scaled ints,
 * exactly as the reference's kernels receive them (GpuStereo.java:394-396);
 * the library applies P2 = max(P1 + 1, P2) (StereoCost.java:21-23). */
typedef struct sgm_params {
  int32_t window_w, window_h;   /* census window; the reference uses census x census */
  int32_t disparities;          /* D, 3..4096 (GpuStereo.java:43-44) */
  int32_t min_disparity;        /* signed */
  int32_t cost;                 /* SGM_COST_* */
  int32_t p1, p2;               /* raw penalties */
  int32_t edge_scale;           /* sgmEdgeScale (StereoCost.java:44-48); 0 disables */
  int32_t directions;           /* 4 or 8 (Options.java:932) */
  int32_t aggregate_bytes;      /* 0 auto, 2 or 4 (StereoCost.java:35-41) */
  int32_t census_weight;        /* round(censusWeight * 1000) (GpuStereo.java:612) */
  int32_t ad_truncation;
  int32_t gradient_truncation;
  int32_t reserved;
  double uniqueness;            /* Options.uniqueness */
  double lr_threshold;          /* Options.lrThreshold */
} sgm_params;

/* CensusSgm.Stats counters (CensusSgm.java:683-694). */
typedef struct sgm_stats {
  int64_t lower_boundary_rejected;
  int64_t full_interval_pixels;
  int64_t boundary_rejected;
  int64_t upper_boundary_rejected;
  int64_t ambiguity_rejected;
  int64_t selected;
  int64_t lr_rejected;
  int64_t post_median_lr_rejected;
  int64_t photometric_rejected;
  int64_t speckle_rejected;
  int64_t valid;                /* finite pixels of the left-right-validated map */
  int64_t reserved[5];
} sgm_stats;

/* GpuSgm.Device (GpuSgm.java:11-18). */
typedef struct sgm_device_info {
  char name[128];
  char arch[64];
  int64_t memory;               /* global memory bytes */
  int64_t allocation;           /* largest single allocation */
  int64_t local_memory;         /* LDS bytes per workgroup */
  int32_t group_size;           /* max threads per workgroup */
  int32_t compute_units;
} sgm_device_info;

/* GpuStereo.Plan (GpuStereo.java:12-20) for the resident mode. */
typedef struct sgm_plan {
  int32_t resident;             /* 1 if the resident plan fits the budget */
  int32_t aggregate_bytes;      /* 2 or 4 */
  int32_t padded_disparities;   /* per-pixel stride of the volumes */
  int32_t lanes_per_disparity_group; /* disparities held per lane in the path sweeps */
  int64_t device_bytes;         /* footprint of one matching direction */
  int64_t pair_bytes;           /* footprint of a left+right pair (both volumes resident) */
  int64_t budget_bytes;
  int64_t aggregate_bound;      /* StereoCost.aggregateBound */
  /* Streamed branch of GpuStereo.plan (GpuStereo.java:56-83), "ram" aggregate storage. */
  int32_t streamed;             /* 1 if a pair runs strip-streamed under the plan mode */
  int32_t strip_rows;           /* rows per strip of the streamed plan (0 if none fits) */
  int64_t streamed_bytes;       /* device footprint of a pair under the streamed plan */
  int64_t host_bytes;           /* pinned host memory of the streamed aggregate */
  /* Compact branch (GpuCandidateStereo.plan, GpuCandidateStereo.java:19-47). */
  int32_t esgm_candidates;      /* proposals per path that fit (0: no compact plan) */
  int32_t esgm_rows;            /* strip rows of the compact plan */
  int64_t esgm_bytes;           /* device footprint of a pair under the compact plan */
  /* Host storage of the streamed aggregate (GpuStereo.Plan.aggregateStorage,
   * blockRows, aggregateBytes; GpuStereo.java:61-83). */
  int32_t aggregate_storage;    /* SGM_AGGREGATE_RAM / _DISK / _RECOMPUTE of the streamed plan */
  int32_t block_rows;           /* rows finished per recomputation block (h unless recompute) */
  int64_t disk_bytes;           /* temporary file of the disk storage */
  int32_t compact;              /* 1 if a pair runs the compact plan under the plan mode */
  int32_t sweeps;               /* SGM_SWEEPS_* schedule a resident pair runs (sgm_set_sweeps) */
} sgm_plan;

typedef struct sgm_context sgm_context;

/* GpuSgm.devices (GpuSgm.java:79-141). */
int sgm_device_count(int *count);
int sgm_device_info_get(int index, sgm_device_info *out);

/* GpuSgm session (GpuSgm.java:163-221); budget_bytes 0 = 70% of device
 * memory, as GpuSgm.budget (GpuSgm.java:629-633). */
int sgm_open(int device, int64_t budget_bytes, sgm_context **out);
int sgm_close(sgm_context *ctx);

/* GpuStereo.plan (GpuStereo.java:33-84): resident and streamed footprints. */
int sgm_plan_get(sgm_context *ctx, const sgm_params *params, int width, int height,
                 sgm_plan *out);
/* GpuSgm.disparityLimit / GpuStereo.limit (GpuSgm.java:237-245, GpuStereo.java:86-94):
 * the largest D for which a plan of the context's plan mode fits, 0 if none. */
int sgm_disparity_limit(sgm_context *ctx, const sgm_params *params, int width, int height,
                        int *limit);

/* GpuSgm.match (GpuSgm.java:267-303): one matching direction.  gray_a/gray_b
 * are the reference/target rasters, mask_a/mask_b the texture-and-valid masks
 * (0/1), reverse selects target x + d.  disp receives the selected disparity
 * (NaN = invalid); stats are ADDED to *stats; reasons/boundary (may be NULL)
 * are updated in place exactly where the reference's select writes them. */
int sgm_match(sgm_context *ctx, const sgm_params *params, const uint8_t *gray_a,
              const uint8_t *gray_b, const uint8_t *mask_a, const uint8_t *mask_b, int width,
              int height, int reverse, float *disp, sgm_stats *stats, uint8_t *reasons,
              uint8_t *boundary);

/* GpuSgm.match with the caller's census descriptors `a` / `b` (GpuSgm.java:267,
 * the long[][] arguments) instead of a census of gray_a / gray_b: word-major
 * uint64 planes desc[word * width*height + p], words = (ww*wh - 1 + 63) / 64,
 * e.g. CensusSgm.colorCensus descriptors (--census-mode color-distance,
 * CensusSgm.java:109-121).  Everything else as sgm_match. */
int sgm_match_desc(sgm_context *ctx, const sgm_params *params, const uint64_t *desc_a,
                   const uint64_t *desc_b, const uint8_t *gray_a, const uint8_t *gray_b,
                   const uint8_t *mask_a, const uint8_t *mask_b, int width, int height,
                   int reverse, float *disp, sgm_stats *stats, uint8_t *reasons,
                   uint8_t *boundary);

/* CensusSgm.computeRecorded up to 02_left_right_validated (CensusSgm.java:45-202):
 * census + texture masks on device, both matching directions, device L-R check.
 * valid_left/valid_right, dl, dr, reasons, boundary may be NULL. *stats is
 * overwritten. */
int sgm_match_pair(sgm_context *ctx, const sgm_params *params, const uint8_t *left,
                   const uint8_t *right, const uint8_t *valid_left, const uint8_t *valid_right,
                   int width, int height, float *checked, float *dl, float *dr, sgm_stats *stats,
                   uint8_t *reasons, uint8_t *boundary);

/* Post-L-R filters of CensusSgm.computeRecorded (CensusSgm.java:203-271):
 * median (Options.median, medianMode, medianMaxChange, medianEdgeThreshold),
 * post-median reciprocity and photometric ZNCC (Options.correlationWindow,
 * correlation), speckle removal (speckleSize, speckleRange,
 * speckleConnectivity).  Reference defaults: 3, edge-preserving (0), 24, 7, 50,
 * 8, 0.5, 0.55, 1.0.  median <= 1 or speckle_size <= 1 disables that step;
 * median windows up to 7. */
typedef struct sgm_filter_params {
  int32_t median;
  int32_t median_classic;       /* 1: medianValid, 0: medianEdgePreserving */
  int32_t median_edge_threshold;
  int32_t correlation_window;
  int32_t speckle_size;
  int32_t speckle_connectivity; /* 4 or 8 */
  int32_t subpixel_refinement;  /* 1: PhotometricRefinement before the L-R check */
  int32_t reserved1;
  double median_max_change;
  double correlation;
  double speckle_range;
  double reserved2;
} sgm_filter_params;

/* CensusSgm.computeRecorded through 05_final_stereo: sgm_match_pair, then the
 * filters above on the device.  disparity receives Result.disparity; stats
 * include post_median_lr_rejected, photometric_rejected, speckle_rejected and
 * the final valid count; reasons (may be NULL) the codes 0..7. */
int sgm_match_pair_filtered(sgm_context *ctx, const sgm_params *params,
                            const sgm_filter_params *filters, const uint8_t *left,
                            const uint8_t *right, const uint8_t *valid_left,
                            const uint8_t *valid_right, int width, int height, float *disparity,
                            sgm_stats *stats, uint8_t *reasons);

/* --census-mode color-distance (CensusSgm.java:109-121): as sgm_match_pair /
 * sgm_match_pair_filtered, with the census descriptors and texture masks of
 * CensusSgm.colorCensus (CensusSgm.java:302-332) computed on the device from
 * the interleaved BGR rasters (3*width*height bytes each); costs and
 * edge-adaptive penalties still read the gray rasters, and the photometric
 * filter scores the BGR rasters over three channels (CensusSgm.java:223-225).
 * filters NULL: disparity receives the L-R-validated map (02 checkpoint);
 * otherwise the final map.  dl, dr, stats, reasons, boundary may be NULL. */
int sgm_match_pair_color(sgm_context *ctx, const sgm_params *params,
                         const sgm_filter_params *filters, const uint8_t *left,
                         const uint8_t *right, const uint8_t *left_bgr, const uint8_t *right_bgr,
                         const uint8_t *valid_left, const uint8_t *valid_right, int width,
                         int height, float *disparity, float *dl, float *dr, sgm_stats *stats,
                         uint8_t *reasons, uint8_t *boundary);

/* Same as sgm_match_pair with every buffer in device memory (no transfers, no
 * host synchronisation).  d_stats (device, sgm_stats) may be NULL.  Work is
 * enqueued on the context's stream; call sgm_synchronize before reading.
 * Under SGM_PLAN_ESGM the call waits once for its own stream mid-way: the
 * number of pixels to verify sizes the verification groups. */
int sgm_match_pair_device(sgm_context *ctx, const sgm_params *params, const void *d_left,
                          const void *d_right, const void *d_valid_left,
                          const void *d_valid_right, int width, int height, void *d_checked,
                          void *d_dl, void *d_dr, void *d_stats, void *d_reasons,
                          void *d_boundary);

/* sgm_match_pair_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as sgm_match_pair_device under
 * the resident plan (the plan mode is ignored), with every intermediate buffer
 * (volumes, descriptors, masks, maps, counters) carved from a fixed workspace
 * of SGM_NOALLOC_WORKSPACE_BYTES that the library holds on each device for the
 * life of the process; the census cost is recomputed per sweep when the cost
 * volume would not fit.  A pair that needs more workspace (about
 * 640x480 with 64 disparities at most with a cost volume) fails with
 * SGM_ERR_INVALID_ARGUMENT; so does a context with sgm_set_keep_volumes on.
 * Every call on one device shares that device's workspace: calls that
 * overlap on one device (from two contexts or threads) race and must be
 * ordered by the caller. */
#define SGM_NOALLOC_WORKSPACE_BYTES ((size_t)128 << 20)
int sgm_match_pair_device_noalloc(sgm_context *ctx, const sgm_params *params,
                                  const void *d_left, const void *d_right,
                                  const void *d_valid_left, const void *d_valid_right, int width,
                                  int height, void *d_checked, void *d_dl, void *d_dr,
                                  void *d_stats, void *d_reasons, void *d_boundary);

/* Device memory for the *_device entry points (callers without their own HIP
 * allocator, e.g. a JVM through FFM).  sgm_memcpy is ordered on the context's
 * stream and returns after the copy completes; kind: 0 host->device,
 * 1 device->host, 2 device->device. */
int sgm_device_alloc(sgm_context *ctx, size_t bytes, void **out);
int sgm_device_free(sgm_context *ctx, void *ptr);
int sgm_memcpy(sgm_context *ctx, void *dst, const void *src, size_t bytes, int kind);
/* Device-to-device copy by a plain streaming kernel on the context's stream
 * (asynchronous; 16-byte aligned pointers and size).  No counterpart in the
 * reference: it measures the achievable HBM bandwidth and calibrates HBM
 * traffic counters against a kernel of known size. */
int sgm_device_copy(sgm_context *ctx, void *dst, const void *src, size_t bytes);

/* Copy the last call's cost (uint16, n*D, 512 = invalid) and aggregate
 * (widened to uint32, n*D) volumes of one matching direction (0 = left
 * reference, 1 = right reference) for parity checks.  Volumes are only kept
 * when keep_volumes was set. */
int sgm_set_keep_volumes(sgm_context *ctx, int keep);
int sgm_read_volumes(sgm_context *ctx, int direction, uint16_t *cost, uint32_t *aggregate);

/* Aggregation plan (the reference's --gpu-sgm auto|resident|streamed and
 * --gpu-strip-rows, Options.java:32,39; GpuStereo.plan, GpuStereo.java:33-84).
 * AUTO runs resident when the pair fits the budget and strip-streamed
 * otherwise; RESIDENT fails with SGM_ERR_NO_PLAN when it does not fit;
 * STREAMED always streams.  Streamed runs keep the cost and aggregate volumes
 * of one strip of `strip_rows` rows on the device (halving the rows until the
 * strip fits the budget, as the reference does), carry vertical and diagonal
 * paths across strip seams, and park the aggregate of finished strips in
 * pinned host memory between the top-down and the bottom-up pass.  Results
 * are identical to the resident plan for every strip height. */
#define SGM_PLAN_AUTO 0
#define SGM_PLAN_RESIDENT 1
#define SGM_PLAN_STREAMED 2
/* ESGM (--gpu-sgm esgm; GpuCandidateStereo.execute, GpuCandidateStereo.java:
 * 49-171): no aggregate volume.  Two strip sweeps per path keep a few
 * separated proposals per path and the exact aggregate around them plus a
 * lower bound elsewhere; pixels whose outcome these settle are certified, the
 * rest are verified exactly over the full range.  Results equal the other
 * plans.  Gray Census with 16-bit aggregates, 3 <= D <= 4096, 4 or 8 paths;
 * other parameters or the colour / descriptor entry points fail with
 * SGM_ERR_NO_PLAN. */
#define SGM_PLAN_ESGM 3
int sgm_set_plan_mode(sgm_context *ctx, int mode, int strip_rows);

/* Host storage of the streamed plan's aggregate between its top-down and
 * bottom-up pass (--gpu-aggregate, GpuStereo.plan, GpuStereo.java:56-83;
 * GpuAggregate.java).  RAM: pinned host memory for the whole aggregate of both
 * matching directions.  DISK: a temporary file in `directory` (removed as soon
 * as it is open; free space is checked first), staged through two pinned strip
 * slots.  RECOMPUTE: host memory for a block of rows only; blocks are finished
 * one after another and each block's passes re-sweep the strips above and
 * below it for the path state.  AUTO: RAM when the aggregate fits the host
 * budget, else RECOMPUTE, and under SGM_PLAN_AUTO a pair that only fits with
 * RECOMPUTE runs the compact plan when it supports the pair
 * (GpuStereo.java:176-193); a RAM allocation that fails falls back to
 * RECOMPUTE (GpuStereo.java:198-208).  host_budget_bytes bounds the host
 * memory of the streamed plan (0: 65% of physical memory; Options.memoryMb).
 * Results are identical under every storage. */
#define SGM_AGGREGATE_AUTO 0
#define SGM_AGGREGATE_RAM 1
#define SGM_AGGREGATE_DISK 2
#define SGM_AGGREGATE_RECOMPUTE 3
int sgm_set_aggregate_storage(sgm_context *ctx, int storage, const char *directory,
                              int64_t host_budget_bytes);

/* Interrupt the context's running (or next) matching call from any thread:
 * the call stops at its next strip, block or path launch, waits for the work
 * already queued and fails with SGM_ERR_INTERRUPTED ("GPU SGM interrupted",
 * the reference's per-strip Thread.isInterrupted check, GpuStereo.java:
 * 350-351).  The request is consumed by the call that sees it; the context
 * stays usable. */
int sgm_interrupt(sgm_context *ctx);
/* --esgm-candidates (Options.java:35,781-782): proposals per path, 1..4; the
 * plan lowers it until the pair fits the budget. */
int sgm_set_esgm_candidates(sgm_context *ctx, int candidates);
/* Outcome of the last compact run ("Compact SGM ... certified pixels ...
 * require exact verification", GpuCandidateStereo.java:103-111): counts[0..1]
 * certified / verified pixels of the left-reference direction, counts[2..3]
 * of the right-reference direction (zeros for a direction not run). */
int sgm_esgm_counts(sgm_context *ctx, int64_t counts[4]);

/* Cost-volume policy for 1-word Census with uint16 aggregates (no counterpart
 * in the reference, whose GpuStereo.plan drops to the streamed plan when the
 * cost volume does not fit).  AUTO keeps the uint16 cost volume resident when
 * the pair (or batch) fits the budget with it and otherwise recomputes the
 * census cost inside every path sweep (no cost volume, one third less HBM);
 * RESIDENT and RECOMPUTE force one of the two.  Results are identical. */
#define SGM_VOLUME_AUTO 0
#define SGM_VOLUME_RESIDENT 1
#define SGM_VOLUME_RECOMPUTE 2
int sgm_set_cost_volume(sgm_context *ctx, int policy);

/* Sweep schedule of the resident plan (no counterpart in the reference, whose
 * `paths` kernel runs one direction per launch into one aggregate volume,
 * GpuStereo.java:732-757).  Results are identical under every schedule.
 *   SEQUENTIAL: one direction per launch, each adding into the aggregate.
 *   PAIRED: the middle directions two per launch, each pair member adding into
 *     its own partial aggregate volume; the last sweep adds both (uint16 sums
 *     wrap modulo 2^16 in any order).  One more aggregate volume per matching
 *     direction.
 *   PLANES: every direction but the last stores its path costs as bytes into a
 *     plane of its own (the middle directions all in one launch, as they only
 *     read the cost volume); the last sweep adds the 7 planes to its own path
 *     costs before the winner search.  A valid cell's path cost is at most its
 *     cost plus P2, so this needs window bits - 1 + P2 <= 255, uint8 costs, 8
 *     paths and no kept volumes; 7 bytes per cell in place of the 2-byte
 *     aggregate, and 22 instead of 36 bytes of HBM traffic per cell.
 *   AUTO (default): PLANES where it applies and fits the budget, else PAIRED
 *     where that fits, else SEQUENTIAL.  A requested schedule that does not
 *     apply falls back the same way (PLANES -> PAIRED -> SEQUENTIAL). */
#define SGM_SWEEPS_AUTO 0
#define SGM_SWEEPS_SEQUENTIAL 1
#define SGM_SWEEPS_PAIRED 2
#define SGM_SWEEPS_PLANES 3
int sgm_set_sweeps(sgm_context *ctx, int schedule);

int sgm_synchronize(sgm_context *ctx);
/* Free the per-pair buffers the plain entry points cache in the context. */
int sgm_release_buffers(sgm_context *ctx);
/* Kernel timing of the last calls (GpuSgm.Compute.profile, GpuSgm.java:592-604). */
int sgm_set_profiling(sgm_context *ctx, int enable);
int sgm_profile_json(sgm_context *ctx, char *buffer, size_t length);
/* HIP stream handle of the context (hipStream_t as void*). */
void *sgm_stream(sgm_context *ctx);

/* Host-forceable failure hook for tests of the fallback path: the next
 * `count` device allocations fail as if out of memory. */
int sgm_debug_fail_allocations(sgm_context *ctx, int count);
/* The same for pinned host allocations (the reference's AllocationFailure of
 * a planned RAM aggregate, GpuAggregate.java:11-18). */
int sgm_debug_fail_host_allocations(sgm_context *ctx, int count);

/* Message of the last failing call on this thread. */
const char *sgm_last_error(void);

/* ---- batched mode: many uneven pairs in one persistent launch ---------------- */

/* One stereo pair of a batch.  Sizes and range are per pair; the window,
 * cost and penalties come from the batch's sgm_params.  With SGM_BATCH_DEVICE
 * all pointers are device pointers, otherwise host pointers. */
typedef struct sgm_pair {
  int32_t width, height;
  int32_t disparities, min_disparity;
  const uint8_t *left, *right;
  const uint8_t *valid_left, *valid_right;   /* may be NULL */
  float *checked;                            /* required */
  float *dl, *dr;                            /* may be NULL */
  uint8_t *reasons, *boundary;               /* may be NULL */
  sgm_stats *stats;                          /* host (or device with SGM_BATCH_DEVICE); may be NULL */
} sgm_pair;

#define SGM_BATCH_DEVICE 1     /* sgm_pair pointers are device pointers */
#define SGM_BATCH_GRAPH 2      /* capture the launch into a hipGraph and replay it */
#define SGM_BATCH_NO_SYNC 4    /* do not wait for completion (device mode only) */
/* Phase-major run: one census launch over every image (then the work-queue
 * kernel for the cost items of non-census costs), the path sweeps as launches
 * per lane group of the specialised sweep kernels over a table of all pairs'
 * lines (with path-cost planes: per sub-batch of pairs, lane groups on two
 * streams), then one left-right launch over every pair; all in one graph
 * with SGM_BATCH_GRAPH.  Same results. */
#define SGM_BATCH_PHASED 8

typedef struct sgm_batch sgm_batch;

/* Plan a batch: allocate per-pair volumes and the work queue, choose tile
 * shape / waves per path (autotuned table), optionally build a hipGraph. */
int sgm_batch_create(sgm_context *ctx, const sgm_params *params, const sgm_pair *pairs,
                     int count, int flags, sgm_batch **out);
int sgm_batch_run(sgm_batch *batch);
int sgm_batch_destroy(sgm_batch *batch);
int64_t sgm_batch_device_bytes(sgm_batch *batch);
/* JSON summary of a batch: pairs, work items, launch shape, cost-volume mode
 * ("resident-u16", "resident-u8", "resident-u8-fused" or "recomputed"),
 * device bytes, graph. */
int sgm_batch_describe(sgm_batch *batch, char *buffer, size_t length);

/* Autotuner: time the candidate path-sweep configurations on this device for
 * the given params and store the winners in the context (and, if path is not
 * NULL, in a JSON file that sgm_autotune_load reads back). */
int sgm_autotune(sgm_context *ctx, const sgm_params *params, const char *path);
int sgm_autotune_load(sgm_context *ctx, const char *path);

/* ---- serving mode: continuous batching with HBM-footprint admission --------- */

typedef struct sgm_server sgm_server;

typedef struct sgm_request_info {
  double arrival_s, admitted_s, started_s, finished_s;  /* seconds on the server clock */
  int64_t device_bytes;
  int32_t status;
  int32_t batch_id;
  int64_t pool_bytes;           /* device memory the server's requests held once this one started */
} sgm_request_info;

/* budget_bytes bounds the device footprint of admitted requests (cost +
 * aggregate volumes and per-pair buffers); max_batch caps the pairs launched
 * together; streams is the number of batches that may run concurrently.
 * max_batch == 1 selects lanes: each of the `streams` lanes is a plain context
 * of its own, and an admitted pair runs the plain path's kernels on it. */
int sgm_server_create(sgm_context *ctx, const sgm_params *params, int64_t budget_bytes,
                      int max_batch, int streams, sgm_server **out);
/* Submit one pair (host buffers, owned by the caller until the request
 * completes).  Returns a ticket; fails with SGM_ERR_REJECTED if the pair can
 * never fit the budget. */
int sgm_server_submit(sgm_server *srv, const sgm_pair *pair, uint64_t *ticket);
/* Wait for a ticket (timeout_s < 0 waits forever). */
int sgm_server_wait(sgm_server *srv, uint64_t ticket, double timeout_s, sgm_request_info *info);
double sgm_server_now(sgm_server *srv);
int sgm_server_destroy(sgm_server *srv);

#ifdef __cplusplus
}
#endif
#endif
