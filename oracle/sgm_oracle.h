/*
 * CPU oracle for the SGM dense-stereo hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU semantic-global-matching path
 * (src/main/java/sgm/CensusSgm.java, src/main/java/sgm/StereoCost.java).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker; the product library never links or calls it.
 *
 * The parameter and statistics structs have the same field order and types as
 * the product's C-ABI structs (include/sgm_hip.h) so one ctypes layout serves
 * both, but this header is self-contained on purpose.
 */
#ifndef SGM_ORACLE_H
#define SGM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cost modes.  0..3 follow StereoCost.code (StereoCost.java:7-15).  4 is the
 * FP64 oracle of the FP32 ZNCC variant: cost = round((1 - clamp(ncc)) * 1024),
 * invalid cells stored as 0x8800 and entering the recurrence as 2048. */
enum { OR_CENSUS = 0, OR_AD_CENSUS = 1, OR_CENSUS_GRADIENT = 2, OR_ZNCC = 3, OR_ZNCC1024 = 4 };

typedef struct {
  int32_t window_w, window_h;   /* census / similarity window, odd, 3..11 */
  int32_t disparities;          /* D */
  int32_t min_disparity;        /* signed */
  int32_t cost;                 /* OR_* */
  int32_t p1, p2;               /* raw penalties as the kernels receive them */
  int32_t edge_scale;           /* sgmEdgeScale, 0 disables edge-adaptive P2 */
  int32_t directions;           /* 4 or 8 */
  int32_t aggregate_bytes;      /* 2 or 4; 0 = automatic */
  int32_t census_weight;        /* round(censusWeight * 1000) */
  int32_t ad_truncation;
  int32_t gradient_truncation;
  int32_t reserved;
  double uniqueness;
  double lr_threshold;
} oracle_params;

/* Counter order of CensusSgm.Stats (CensusSgm.java:683-694); csv() order is
 * boundary, upper, ambiguity, selected, lr, postMedianLr, photometric,
 * speckle, valid. */
typedef struct {
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
  int64_t valid;
  int64_t reserved[5];
} oracle_stats;

int oracle_words(int window_w, int window_h);
int oracle_effective_p2(const oracle_params *o);
int oracle_aggregate_bytes(const oracle_params *o);
int64_t oracle_aggregate_bound(const oracle_params *o);

/* CensusSgm.census (CensusSgm.java:540-555), rectangular window. desc is
 * word-major: desc[word * w*h + p]. */
void oracle_census(const uint8_t *gray, int w, int h, int ww, int wh, uint64_t *desc);
/* CensusSgm.textureMask (CensusSgm.java:523-538), ANDed with valid (may be
 * NULL) as in CensusSgm.java:127-134. */
void oracle_texture(const uint8_t *gray, const uint8_t *valid, int w, int h, int ww, int wh,
                    uint8_t *mask);

/* CensusSgm.colorCensus (CensusSgm.java:302-332), rectangular window:
 * interleaved BGR in; bit k (row-major, centre skipped) set where the colour
 * distance to the centre >= the window mean (the centre's zero distance
 * counts in the mean); textured = largest distance >= 4, ANDed with valid
 * (may be NULL).  Border pixels keep 0 descriptors and are untextured. */
void oracle_color_census(const uint8_t *bgr, const uint8_t *valid, int w, int h, int ww, int wh,
                         uint64_t *desc, uint8_t *textured);

/* One matching direction, CPU branch of CensusSgm.match (CensusSgm.java:344-517).
 * cost_out (n*D uint16) and agg_out (n*D uint32, the stored aggregate value)
 * are optional.  reasons/boundary may be NULL. Returns 0 or <0 on bad input. */
int oracle_match(const oracle_params *o, const uint64_t *da, const uint64_t *db,
                 const uint8_t *ga, const uint8_t *gb, const uint8_t *ta, const uint8_t *tb,
                 int w, int h, int reverse, float *disp, oracle_stats *stats, uint8_t *reasons,
                 uint8_t *boundary, uint16_t *cost_out, uint32_t *agg_out);

/* Both directions + left-right check: CensusSgm.computeRecorded up to the
 * 02_left_right_validated checkpoint (CensusSgm.java:45-202).  dl/dr/reasons/
 * boundary may be NULL.  stats->valid counts finite checked pixels. */
int oracle_compute(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                   const uint8_t *valid_left, const uint8_t *valid_right, int w, int h,
                   float *checked, float *dl, float *dr, oracle_stats *stats, uint8_t *reasons,
                   uint8_t *boundary);

/* oracle_compute with colour-distance census descriptors and texture masks
 * from interleaved BGR (--census-mode color-distance, CensusSgm.java:109-121);
 * the costs and edge-adaptive penalties still read the gray rasters. */
int oracle_compute_color(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                         const uint8_t *left_bgr, const uint8_t *right_bgr,
                         const uint8_t *valid_left, const uint8_t *valid_right, int w, int h,
                         float *checked, float *dl, float *dr, oracle_stats *stats,
                         uint8_t *reasons, uint8_t *boundary);

/* oracle_match / oracle_compute split over host threads (OpenMP; threads
 * <= 0: all available).  Same results bit for bit; gray census, no filters.
 * oracle_threads() is the number of threads a parallel region gets. */
int oracle_match_parallel(const oracle_params *o, const uint64_t *da, const uint64_t *db,
                          const uint8_t *ga, const uint8_t *gb, const uint8_t *ta,
                          const uint8_t *tb, int w, int h, int reverse, float *disp,
                          oracle_stats *stats, uint8_t *reasons, uint8_t *boundary, int threads);
int oracle_compute_parallel(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                            const uint8_t *valid_left, const uint8_t *valid_right, int w, int h,
                            float *checked, float *dl, float *dr, oracle_stats *stats,
                            uint8_t *reasons, uint8_t *boundary, int threads);
int oracle_threads(void);

/* Post-L-R filters of CensusSgm.computeRecorded (CensusSgm.java:203-271);
 * same layout as sgm_filter_params (include/sgm_hip.h). */
typedef struct {
  int32_t median;               /* window; <= 1 disables (Options.median) */
  int32_t median_classic;       /* 1: medianValid, 0: medianEdgePreserving */
  int32_t median_edge_threshold;
  int32_t correlation_window;   /* photometric ZNCC window (odd) */
  int32_t speckle_size;         /* minimum component size; <= 1 disables */
  int32_t speckle_connectivity; /* 4 or 8 */
  int32_t subpixel_refinement;  /* 1: PhotometricRefinement before the L-R check */
  int32_t reserved1;
  double median_max_change;
  double correlation;           /* photometric threshold: score < correlation rejects */
  double speckle_range;
  double reserved2;
} oracle_filter;

/* Median (03_median), post-median reciprocity and photometric ZNCC
 * (04_photometric), speckle removal (05_final_stereo) on `checked` in place
 * (CensusSgm.java:203-271, 557-681).  dr: right-reference disparities; the
 * reciprocity border is (window_w/2, window_h/2) of o.  Adds
 * post_median_lr_rejected, photometric_rejected, speckle_rejected and sets
 * valid; reasons (may be NULL) get 5, 6, 7. */
int oracle_postfilter(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                      const uint8_t *right, const float *dr, int w, int h, float *checked,
                      oracle_stats *stats, uint8_t *reasons);

/* oracle_postfilter with the photometric score over `channels` interleaved
 * channels of photo_left / photo_right (the BGR rasters in colour-distance
 * mode, CensusSgm.java:223-225); the median still reads the gray left. */
int oracle_postfilter_ch(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                         const uint8_t *photo_left, const uint8_t *photo_right, int channels,
                         const float *dr, int w, int h, float *checked, oracle_stats *stats,
                         uint8_t *reasons);

/* compute_pair with options: f (may be NULL) with subpixel_refinement set runs
 * PhotometricRefinement (PhotometricRefinement.java:5-61) on both matching
 * directions before the L-R check (CensusSgm.java:168-170); left_bgr /
 * right_bgr (both or neither) select the colour-distance census. */
int oracle_compute_ex(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                      const uint8_t *right, const uint8_t *left_bgr, const uint8_t *right_bgr,
                      const uint8_t *valid_left, const uint8_t *valid_right, int w, int h,
                      float *checked, float *dl, float *dr, oracle_stats *stats,
                      uint8_t *reasons, uint8_t *boundary);

/* CensusSgm.zncc (CensusSgm.java:614-644) over `channels` interleaved channels:
 * the photometric score. */
double oracle_photo_zncc(const uint8_t *a, const uint8_t *b, int w, int h, int x, int y,
                         double disparity, int r);
double oracle_photo_zncc_ch(const uint8_t *a, const uint8_t *b, int w, int h, int x, int y,
                            double disparity, int r, int channels);

/* Java Math.round(float) / Math.round(double) (closest, ties toward +inf). */
int32_t oracle_round_float(float v);
int64_t oracle_round_double(double v);

#ifdef __cplusplus
}
#endif
#endif
