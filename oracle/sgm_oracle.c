/*
 * CPU oracle for the SGM hot path -- TEST INFRASTRUCTURE ONLY (see sgm_oracle.h).
 *
 * Literal restatement of the reference's CPU SGM in plain C, compiled with
 * -O2 -ffp-contract=off and without fast-math so every floating-point step is
 * the IEEE operation the Java source performs (Java >= 17 is strict IEEE):
 *   census / texture      CensusSgm.java:519-555
 *   cost value            StereoCost.java:50-90, cost loop CensusSgm.java:405-420
 *   penalties / bounds    StereoCost.java:17-48
 *   path recurrence       CensusSgm.java:421-459
 *   WTA / subpixel        CensusSgm.java:460-515
 *   left-right check      CensusSgm.java:181-201
 *   16-bit storage        PagedVolume.java:160-214 ((short) store, & 65535 load)
 * Generalisation: the reference uses one square census window (radius
 * census/2 on both axes); here the window is ww x wh with rx = ww/2,
 * ry = wh/2, and a square window reduces to the reference exactly.
 */
#include "sgm_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define INVALID_COST 512

int32_t oracle_round_float(float v) {
  /* Math.round(float): closest int, ties toward +inf (exact, no v+0.5f). */
  float f = floorf(v);
  float frac = v - f; /* exact for |v| < 2^24 */
  return (int32_t)f + (frac >= 0.5f ? 1 : 0);
}

int64_t oracle_round_double(double v) {
  double f = floor(v);
  double frac = v - f; /* exact for |v| < 2^52 */
  return (int64_t)f + (frac >= 0.5 ? 1 : 0);
}

int oracle_words(int ww, int wh) { return (ww * wh - 1 + 63) / 64; }

int oracle_effective_p2(const oracle_params *o) {
  /* StereoCost.p2: max(P1 + 1, P2) (StereoCost.java:21-23). */
  return o->p2 > o->p1 + 1 ? o->p2 : o->p1 + 1;
}

static int max_cost_value(const oracle_params *o) {
  return o->cost == OR_ZNCC1024 ? 2048 : INVALID_COST;
}

int64_t oracle_aggregate_bound(const oracle_params *o) {
  /* StereoCost.aggregateBound: directions * (512 + P2) (StereoCost.java:31-33). */
  return (int64_t)o->directions * (max_cost_value(o) + (int64_t)oracle_effective_p2(o));
}

int oracle_aggregate_bytes(const oracle_params *o) {
  /* StereoCost.aggregateBytes (StereoCost.java:35-41). */
  int fits = oracle_aggregate_bound(o) <= 65535;
  if (o->aggregate_bytes == 2) return fits ? 2 : -1;
  if (o->aggregate_bytes == 4) return 4;
  return fits ? 2 : 4;
}

void oracle_census(const uint8_t *in, int w, int h, int ww, int wh, uint64_t *desc) {
  int rx = ww / 2, ry = wh / 2, words = oracle_words(ww, wh);
  size_t n = (size_t)w * h;
  memset(desc, 0, sizeof(uint64_t) * n * words);
  for (int y = ry; y < h - ry; y++)
    for (int x = rx; x < w - rx; x++) {
      size_t p = (size_t)y * w + x;
      int bit = 0, center = in[p];
      for (int yy = -ry; yy <= ry; yy++)
        for (int xx = -rx; xx <= rx; xx++) {
          if (xx == 0 && yy == 0) continue;
          if (in[(size_t)(y + yy) * w + x + xx] < center)
            desc[(size_t)(bit / 64) * n + p] |= 1ULL << (bit % 64);
          bit++;
        }
    }
}

void oracle_texture(const uint8_t *in, const uint8_t *valid, int w, int h, int ww, int wh,
                    uint8_t *mask) {
  int rx = ww / 2, ry = wh / 2;
  memset(mask, 0, (size_t)w * h);
  for (int y = ry; y < h - ry; y++)
    for (int x = rx; x < w - rx; x++) {
      int low = 255, high = 0;
      for (int yy = -ry; yy <= ry; yy++)
        for (int xx = -rx; xx <= rx; xx++) {
          int v = in[(size_t)(y + yy) * w + x + xx];
          if (v < low) low = v;
          if (v > high) high = v;
        }
      size_t p = (size_t)y * w + x;
      mask[p] = (high - low >= 4) && (valid == NULL || valid[p] != 0);
    }
}

/* CensusSgm.colorCensus (CensusSgm.java:302-332): distances and their sum in
 * double, in window order; mean over all ww*wh cells (centre included). */
void oracle_color_census(const uint8_t *bgr, const uint8_t *valid, int w, int h, int ww, int wh,
                         uint64_t *desc, uint8_t *textured) {
  int rx = ww / 2, ry = wh / 2, words = oracle_words(ww, wh), cells = ww * wh - 1;
  size_t n = (size_t)w * h;
  double *distances = (double *)malloc(sizeof(double) * cells);
  memset(desc, 0, sizeof(uint64_t) * n * words);
  memset(textured, 0, n);
  for (int y = ry; y < h - ry; y++)
    for (int x = rx; x < w - rx; x++) {
      size_t p = (size_t)y * w + x, center = 3 * p;
      int bit = 0;
      double total = 0, maximum = 0;
      for (int dy = -ry; dy <= ry; dy++)
        for (int dx = -rx; dx <= rx; dx++) {
          if (dx == 0 && dy == 0) continue;
          size_t q = 3 * ((size_t)(y + dy) * w + x + dx);
          double squared = 0;
          for (int c = 0; c < 3; c++) {
            int difference = (int)bgr[q + c] - (int)bgr[center + c];
            squared += difference * difference;
          }
          double distance = sqrt(squared);
          distances[bit++] = distance;
          total += distance;
          if (distance > maximum) maximum = distance;
        }
      double mean = total / (ww * wh);
      textured[p] = maximum >= 4 && (valid == NULL || valid[p] != 0);
      for (int k = 0; k < cells; k++)
        if (distances[k] >= mean) desc[(size_t)(k / 64) * n + p] |= 1ULL << (k % 64);
    }
  free(distances);
}

static int popcount64(uint64_t v) { return __builtin_popcountll(v); }

/* StereoCost.value (StereoCost.java:50-90) plus the FP64 ZNCC-1024 oracle. */
static int cost_value(const oracle_params *o, const uint64_t *a, const uint64_t *b,
                      const uint8_t *ga, const uint8_t *gb, size_t p, size_t q, int w,
                      size_t n, int words) {
  int bits = o->window_w * o->window_h - 1;
  int mode = o->cost;
  if (mode == OR_CENSUS || mode == OR_AD_CENSUS || mode == OR_CENSUS_GRADIENT) {
    int hamming = 0;
    for (int word = 0; word < words; word++)
      hamming += popcount64(a[(size_t)word * n + p] ^ b[(size_t)word * n + q]);
    if (mode == OR_CENSUS) return hamming;
    int difference, cap;
    if (mode == OR_AD_CENSUS) {
      difference = abs((int)ga[p] - (int)gb[q]);
      cap = o->ad_truncation;
    } else {
      int ax = (int)ga[p + 1] - (int)ga[p - 1], ay = (int)ga[p + w] - (int)ga[p - w];
      int bx = (int)gb[q + 1] - (int)gb[q - 1], by = (int)gb[q + w] - (int)gb[q - w];
      difference = abs(ax - bx) + abs(ay - by);
      cap = o->gradient_truncation;
    }
    int weight = o->census_weight, denominator = 1000 * cap;
    return (weight * hamming * cap + (1000 - weight) * (difference < cap ? difference : cap) * bits +
            denominator / 2) /
           denominator;
  }
  int rx = o->window_w / 2, ry = o->window_h / 2;
  int64_t count = (int64_t)o->window_w * o->window_h;
  int64_t sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
  for (int dy = -ry; dy <= ry; dy++)
    for (int dx = -rx; dx <= rx; dx++) {
      int64_t x = ga[p + (ptrdiff_t)dy * w + dx], y = gb[q + (ptrdiff_t)dy * w + dx];
      sa += x;
      sb += y;
      saa += x * x;
      sbb += y * y;
      sab += x * y;
    }
  int64_t va = count * saa - sa * sa, vb = count * sbb - sb * sb;
  if (mode == OR_ZNCC) {
    if (va <= 0 || vb <= 0) return INVALID_COST;
    double correlation = (double)(count * sab - sa * sb) / sqrt((double)va * (double)vb);
    double c = correlation < 1 ? correlation : 1; /* Math.min(1, c) */
    c = c > -1 ? c : -1;                          /* Math.max(-1, .) */
    return (int)oracle_round_double((1 - c) * .5 * bits);
  }
  /* OR_ZNCC1024: FP64 oracle of the FP32 variant. */
  if (va <= 0 || vb <= 0) return 0x8800;
  double correlation = (double)(count * sab - sa * sb) / sqrt((double)va * (double)vb);
  double c = correlation < 1 ? correlation : 1;
  c = c > -1 ? c : -1;
  return (int)oracle_round_double((1 - c) * 1024.0);
}

static int edge_p2(int p1, int p2, int scale, const uint8_t *gray, size_t current, size_t previous) {
  /* StereoCost.pathP2 (StereoCost.java:44-48). */
  if (scale == 0) return p2;
  int difference = abs((int)gray[current] - (int)gray[previous]);
  int v = (int)((int64_t)p2 * scale / (scale + difference));
  return p1 + 1 > v ? p1 + 1 : v;
}

int oracle_match(const oracle_params *o, const uint64_t *a, const uint64_t *b,
                 const uint8_t *grayA, const uint8_t *grayB, const uint8_t *texturedA,
                 const uint8_t *texturedB, int width, int height, int reverse, float *result,
                 oracle_stats *stats, uint8_t *reasons, uint8_t *searchBoundary,
                 uint16_t *cost_out, uint32_t *agg_out) {
  int nd = o->disparities, rx = o->window_w / 2, ry = o->window_h / 2;
  int words = oracle_words(o->window_w, o->window_h);
  size_t count = (size_t)width * height, cells = count * (size_t)nd;
  int sum_bytes = oracle_aggregate_bytes(o);
  if (sum_bytes < 0 || nd < 1 || o->directions < 1 || o->directions > 8) return -1;
  int invalid = o->cost == OR_ZNCC1024 ? 0x8800 : INVALID_COST;
  int value_mask = o->cost == OR_ZNCC1024 ? 0x7FFF : 0xFFFF;
  uint16_t *costs = cost_out ? cost_out : (uint16_t *)malloc(cells * sizeof(uint16_t));
  /* Aggregate storage: 16-bit wraps like PagedVolume ((short) v, & 65535). */
  uint16_t *agg16 = NULL;
  int32_t *agg32 = NULL;
  if (sum_bytes == 2) agg16 = (uint16_t *)calloc(cells, sizeof(uint16_t));
  else agg32 = (int32_t *)calloc(cells, sizeof(int32_t));
  int *previous = (int *)malloc(sizeof(int) * (size_t)width * nd);
  int *row = (int *)malloc(sizeof(int) * (size_t)width * nd);
  if (!costs || (!agg16 && !agg32) || !previous || !row) return -2;
  for (size_t i = 0; i < cells; i++) costs[i] = (uint16_t)invalid;
  /* Hot loop 1: cost fill (CensusSgm.java:405-420). */
  for (int y = ry; y < height - ry; y++)
    for (int x = rx; x < width - rx; x++) {
      size_t p = (size_t)y * width + x, base = p * nd;
      if (!texturedA[p]) continue;
      for (int k = 0; k < nd; k++) {
        int disparity = o->min_disparity + k;
        int64_t target = (int64_t)x + (reverse ? disparity : -(int64_t)disparity);
        if (target < rx || target >= width - rx) continue;
        size_t q = (size_t)y * width + (size_t)target;
        if (!texturedB[q]) continue;
        costs[base + k] = (uint16_t)cost_value(o, a, b, grayA, grayB, p, q, width, count, words);
      }
    }
  /* Hot loop 2: path recurrence (CensusSgm.java:421-459). */
  static const int paths[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                  {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  int p1 = o->p1, p2 = oracle_effective_p2(o);
  for (int path = 0; path < o->directions; path++) {
    int dx = paths[path][0], dy = paths[path][1];
    int yStart = dy < 0 ? height - 1 : 0, yStep = dy < 0 ? -1 : 1;
    int xStart = dx < 0 ? width - 1 : 0, xStep = dx < 0 ? -1 : 1;
    for (int yi = 0, y = yStart; yi < height; yi++, y += yStep) {
      for (int xi = 0, x = xStart; xi < width; xi++, x += xStep) {
        int px = x - dx, py = y - dy;
        size_t dest = (size_t)x * nd, base = ((size_t)y * width + x) * nd;
        int hasPrevious = px >= 0 && px < width && py >= 0 && py < height &&
                          texturedA[(size_t)py * width + px];
        int *prev = dy == 0 ? row : previous;
        size_t prevBase = hasPrevious ? (size_t)px * nd : 0;
        int minPrev = 0x7FFFFFFF;
        if (hasPrevious)
          for (int k = 0; k < nd; k++)
            if (prev[prevBase + k] < minPrev) minPrev = prev[prevBase + k];
        int localP2 = hasPrevious ? edge_p2(p1, p2, o->edge_scale, grayA, (size_t)y * width + x,
                                            (size_t)py * width + px)
                                  : p2;
        for (int k = 0; k < nd; k++) {
          int value = costs[base + k] & value_mask;
          if (hasPrevious) {
            int transition = prev[prevBase + k];
            if (minPrev + localP2 < transition) transition = minPrev + localP2;
            if (k > 0 && prev[prevBase + k - 1] + p1 < transition)
              transition = prev[prevBase + k - 1] + p1;
            if (k + 1 < nd && prev[prevBase + k + 1] + p1 < transition)
              transition = prev[prevBase + k + 1] + p1;
            value += transition - minPrev;
          }
          row[dest + k] = value;
          if (agg16) agg16[base + k] = (uint16_t)(agg16[base + k] + (uint32_t)value);
          else agg32[base + k] = (int32_t)((uint32_t)agg32[base + k] + (uint32_t)value);
        }
      }
      int *swap = previous;
      previous = row;
      row = swap;
    }
  }
#define AGG(i) (agg16 ? (int)agg16[i] : (int)agg32[i])
  /* Hot loop 3: WTA, boundary, uniqueness, subpixel (CensusSgm.java:460-515). */
  for (size_t i = 0; i < count; i++) result[i] = NAN;
  for (int y = ry; y < height - ry; y++)
    for (int x = rx; x < width - rx; x++) {
      size_t pixel = (size_t)y * width + x, base = pixel * nd;
      int best = -1, bestCost = 0x7FFFFFFF;
      for (int k = 0; k < nd; k++)
        if (costs[base + k] != invalid && AGG(base + k) < bestCost) {
          best = k;
          bestCost = AGG(base + k);
        }
      if (best < 0) continue;
      int full = costs[base] != invalid && costs[base + nd - 1] != invalid;
      if (full) {
        stats->full_interval_pixels++;
        if (searchBoundary) searchBoundary[pixel] = 4;
      }
      if (best == 0 || best == nd - 1) {
        stats->boundary_rejected++;
        if (reasons) reasons[pixel] = 2;
        if (full && best == 0) {
          stats->lower_boundary_rejected++;
          if (searchBoundary) searchBoundary[pixel] |= 1;
        }
        if (full && best == nd - 1) {
          stats->upper_boundary_rejected++;
          if (searchBoundary) searchBoundary[pixel] |= 2;
        }
        continue;
      }
      int second = 0x7FFFFFFF;
      for (int k = 0; k < nd; k++)
        if (abs(k - best) > 1 && costs[base + k] != invalid && AGG(base + k) < second)
          second = AGG(base + k);
      /* second - bestCost <= Math.max(1, bestCost * uniqueness): int vs double. */
      double allowed = (double)bestCost * o->uniqueness;
      if (!(allowed >= 1.0)) allowed = 1.0;
      if (second != 0x7FFFFFFF && (double)(second - bestCost) <= allowed) {
        stats->ambiguity_rejected++;
        if (reasons) reasons[pixel] = 3;
        continue;
      }
      double offset = 0;
      if (best > 0 && best + 1 < nd && costs[base + best - 1] != invalid &&
          costs[base + best + 1] != invalid) {
        double before = AGG(base + best - 1), after = AGG(base + best + 1);
        double curvature = before + after - 2.0 * bestCost;
        if (curvature > 0) {
          double t = .5 * (before - after) / curvature;
          t = t < .5 ? t : .5;
          offset = t > -.5 ? t : -.5;
        }
      }
      result[pixel] = (float)((double)(o->min_disparity + best) + offset);
      stats->selected++;
      if (reasons) reasons[pixel] = 0;
    }
  if (agg_out)
    for (size_t i = 0; i < cells; i++) agg_out[i] = (uint32_t)AGG(i);
#undef AGG
  if (!cost_out) free(costs);
  free(agg16);
  free(agg32);
  free(previous);
  free(row);
  return 0;
}

static int isfinite_f(float v) { return isfinite(v); }

static int64_t refine(float *disparity, const uint8_t *a, const uint8_t *b, int w, int h,
                      int reverse, const oracle_params *o, const oracle_filter *f);

static int compute_pair(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                        const uint8_t *right, const uint8_t *left_bgr, const uint8_t *right_bgr,
                        const uint8_t *valid_left, const uint8_t *valid_right, int width,
                        int height, float *checked, float *dl_out, float *dr_out,
                        oracle_stats *stats, uint8_t *reasons, uint8_t *searchBoundary) {
  /* CensusSgm.computeRecorded (CensusSgm.java:56-60 dimension check). */
  if (width < o->window_w + 2 || height < o->window_h + 2) return -1;
  size_t n = (size_t)width * height;
  int words = oracle_words(o->window_w, o->window_h);
  uint64_t *a = (uint64_t *)malloc(sizeof(uint64_t) * n * words);
  uint64_t *b = (uint64_t *)malloc(sizeof(uint64_t) * n * words);
  uint8_t *ta = (uint8_t *)malloc(n), *tb = (uint8_t *)malloc(n);
  float *dl = dl_out ? dl_out : (float *)malloc(sizeof(float) * n);
  float *dr = dr_out ? dr_out : (float *)malloc(sizeof(float) * n);
  if (!a || !b || !ta || !tb || !dl || !dr) return -2;
  if (reasons) memset(reasons, 1, n);
  if (searchBoundary) memset(searchBoundary, 0, n);
  if (left_bgr) {
    /* CensusSgm.java:109-121 and the validity AND of :127-134 */
    oracle_color_census(left_bgr, valid_left, width, height, o->window_w, o->window_h, a, ta);
    oracle_color_census(right_bgr, valid_right, width, height, o->window_w, o->window_h, b, tb);
  } else {
    oracle_census(left, width, height, o->window_w, o->window_h, a);
    oracle_census(right, width, height, o->window_w, o->window_h, b);
    oracle_texture(left, valid_left, width, height, o->window_w, o->window_h, ta);
    oracle_texture(right, valid_right, width, height, o->window_w, o->window_h, tb);
  }
  oracle_stats discarded;
  memset(&discarded, 0, sizeof discarded);
  int rc = oracle_match(o, a, b, left, right, ta, tb, width, height, 0, dl, stats, reasons,
                        searchBoundary, NULL, NULL);
  if (rc == 0)
    rc = oracle_match(o, b, a, right, left, tb, ta, width, height, 1, dr, &discarded, NULL, NULL,
                      NULL, NULL);
  if (rc == 0 && f && f->subpixel_refinement) {
    /* --subpixel-refinement zncc (CensusSgm.java:168-170), on the gray rasters */
    refine(dl, left, right, width, height, 0, o, f);
    refine(dr, right, left, width, height, 1, o, f);
  }
  if (rc == 0) {
    /* Left-right check (CensusSgm.java:181-201). */
    int rx = o->window_w / 2, ry = o->window_h / 2;
    for (size_t i = 0; i < n; i++) checked[i] = NAN;
    for (int y = ry; y < height - ry; y++)
      for (int x = rx; x < width - rx; x++) {
        size_t p = (size_t)y * width + x;
        if (!isfinite_f(dl[p])) continue;
        int xr = oracle_round_float((float)x - dl[p]);
        if (xr < rx || xr >= width - rx) {
          if (reasons) reasons[p] = 4;
          continue;
        }
        float other = dr[(size_t)y * width + xr];
        if (isfinite_f(other) && (double)fabsf(dl[p] - other) <= o->lr_threshold)
          checked[p] = dl[p];
        else {
          stats->lr_rejected++;
          if (reasons) reasons[p] = 4;
        }
      }
    for (size_t i = 0; i < n; i++)
      if (isfinite_f(checked[i])) stats->valid++;
  }
  free(a);
  free(b);
  free(ta);
  free(tb);
  if (!dl_out) free(dl);
  if (!dr_out) free(dr);
  return rc;
}

int oracle_compute(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                   const uint8_t *valid_left, const uint8_t *valid_right, int width, int height,
                   float *checked, float *dl, float *dr, oracle_stats *stats, uint8_t *reasons,
                   uint8_t *boundary) {
  return compute_pair(o, NULL, left, right, NULL, NULL, valid_left, valid_right, width, height,
                      checked, dl, dr, stats, reasons, boundary);
}

int oracle_compute_ex(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                      const uint8_t *right, const uint8_t *left_bgr, const uint8_t *right_bgr,
                      const uint8_t *valid_left, const uint8_t *valid_right, int width,
                      int height, float *checked, float *dl, float *dr, oracle_stats *stats,
                      uint8_t *reasons, uint8_t *boundary) {
  if ((left_bgr == NULL) != (right_bgr == NULL)) return -1;
  return compute_pair(o, f, left, right, left_bgr, right_bgr, valid_left, valid_right, width,
                      height, checked, dl, dr, stats, reasons, boundary);
}

int oracle_compute_color(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                         const uint8_t *left_bgr, const uint8_t *right_bgr,
                         const uint8_t *valid_left, const uint8_t *valid_right, int width,
                         int height, float *checked, float *dl, float *dr, oracle_stats *stats,
                         uint8_t *reasons, uint8_t *boundary) {
  if (!left_bgr || !right_bgr) return -1;
  return compute_pair(o, NULL, left, right, left_bgr, right_bgr, valid_left, valid_right, width,
                      height, checked, dl, dr, stats, reasons, boundary);
}


/* ------------------------------------------------------------------------ */
/* Post-L-R filters (CensusSgm.java:203-271, 557-681).                       */

/* Float.compare order for finite values (-0.0 < 0.0). */
static int float_cmp(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  if (x < y) return -1;
  if (x > y) return 1;
  int sx = signbit(x) != 0, sy = signbit(y) != 0;
  return sy - sx;
}

/* CensusSgm.medianValid (CensusSgm.java:557-573). */
static void median_valid(const float *a, int w, int h, int window, float *out) {
  int r = window / 2;
  float *values = (float *)malloc(sizeof(float) * window * window);
  memcpy(out, a, sizeof(float) * (size_t)w * h);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      if (!isfinite_f(a[(size_t)y * w + x])) continue;
      int n = 0;
      for (int yy = y - r < 0 ? 0 : y - r; yy <= (y + r < h - 1 ? y + r : h - 1); yy++)
        for (int xx = x - r < 0 ? 0 : x - r; xx <= (x + r < w - 1 ? x + r : w - 1); xx++) {
          float v = a[(size_t)yy * w + xx];
          if (isfinite_f(v)) values[n++] = v;
        }
      qsort(values, n, sizeof(float), float_cmp);
      out[(size_t)y * w + x] = values[n / 2];
    }
  free(values);
}

/* CensusSgm.medianEdgePreserving (CensusSgm.java:576-600). */
static void median_edge(const float *a, int w, int h, int window, double max_change,
                        const uint8_t *gray, int edge, float *out) {
  int r = window / 2;
  float *values = (float *)malloc(sizeof(float) * window * window);
  median_valid(a, w, h, window, out);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float center = a[(size_t)y * w + x];
      if (!isfinite_f(center)) continue;
      if ((double)fabsf(out[(size_t)y * w + x] - center) < max_change) continue;
      int n = 0;
      for (int yy = y - r < 0 ? 0 : y - r; yy <= (y + r < h - 1 ? y + r : h - 1); yy++)
        for (int xx = x - r < 0 ? 0 : x - r; xx <= (x + r < w - 1 ? x + r : w - 1); xx++) {
          float value = a[(size_t)yy * w + xx];
          if (isfinite_f(value) && (double)fabsf(value - center) <= max_change &&
              (edge == 0 || abs((int)gray[(size_t)yy * w + xx] - (int)gray[(size_t)y * w + x]) <= edge))
            values[n++] = value;
        }
      if (n > 0) {
        qsort(values, n, sizeof(float), float_cmp);
        out[(size_t)y * w + x] = values[n / 2];
      }
    }
  free(values);
}

/* CensusSgm.reciprocal (CensusSgm.java:602-612), border per axis. */
static int reciprocal(const float *right, int w, int h, int x, int y, float disparity,
                      double tolerance, int bx, int by) {
  if (!isfinite_f(disparity) || y < by || y >= h - by) return 0;
  int xr = oracle_round_float((float)x - disparity);
  return xr >= bx && xr < w - bx && isfinite_f(right[(size_t)y * w + xr]) &&
         (double)fabsf(disparity - right[(size_t)y * w + xr]) <= tolerance;
}

/* CensusSgm.zncc (CensusSgm.java:614-644) over `channels` interleaved channels. */
static double photo_zncc(const uint8_t *a, const uint8_t *b, int w, int h, int x, int y,
                         double disparity, int r, int channels) {
  double target = x - disparity;
  int bx = (int)floor(target);
  double fraction = target - bx;
  if (x - r < 0 || x + r >= w || y - r < 0 || y + r >= h || bx - r < 0 || bx + r + 1 >= w)
    return -1;
  int n = (2 * r + 1) * (2 * r + 1);
  double covariance = 0, var_a = 0, var_b = 0;
  for (int c = 0; c < channels; c++) {
    double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
    for (int dy = -r; dy <= r; dy++)
      for (int dx = -r; dx <= r; dx++) {
        size_t ai = ((size_t)(y + dy) * w + x + dx) * channels + c;
        size_t bi = ((size_t)(y + dy) * w + bx + dx) * channels + c;
        double av = a[ai];
        double bv = (1 - fraction) * b[bi] + fraction * b[bi + channels];
        sa += av;
        sb += bv;
        saa += av * av;
        sbb += bv * bv;
        sab += av * bv;
      }
    covariance += sab - sa * sb / n;
    var_a += saa - sa * sa / n;
    var_b += sbb - sb * sb / n;
  }
  if (var_a < channels * n * 4 || var_b < channels * n * 4) return -1;
  double prod = var_a * var_b;
  return covariance / sqrt(prod > 1e-12 ? prod : 1e-12);
}

/* PhotometricRefinement.refine (PhotometricRefinement.java:5-61): quarter-step
 * ZNCC search around each disparity away from depth edges, a parabolic
 * eighth-step refinement, and acceptance of clear improvements only. */
static int64_t refine(float *disparity, const uint8_t *a, const uint8_t *b, int w, int h,
                      int reverse, const oracle_params *o, const oracle_filter *f) {
  size_t n = (size_t)w * h;
  float *original = (float *)malloc(sizeof(float) * n);
  memcpy(original, disparity, sizeof(float) * n);
  int64_t changed = 0;
  int r = f->correlation_window / 2;
  for (int y = r; y < h - r; y++)
    for (int x = r; x < w - r; x++) {
      size_t p = (size_t)y * w + x;
      double d = original[p];
      if (!isfinite(d)) continue;
      int edge = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          float q = original[p + (ptrdiff_t)dy * w + dx];
          if (isfinite_f(q) && fabs((double)q - d) > 1.5) edge = 1;
        }
      if (edge) continue;
      double initial = photo_zncc(a, b, w, h, x, y, reverse ? -d : d, r, 1), best = initial,
             chosen = d;
      int slot = -1;
      double scores[7];
      for (int k = 0; k < 7; k++) {
        double candidate = d + (k - 3) * .25;
        if (candidate <= o->min_disparity || candidate >= o->min_disparity + o->disparities - 1) {
          scores[k] = -1;
          continue;
        }
        scores[k] = photo_zncc(a, b, w, h, x, y, reverse ? -candidate : candidate, r, 1);
        if (scores[k] > best) {
          best = scores[k];
          chosen = candidate;
          slot = k;
        }
      }
      if (slot > 0 && slot < 6 && scores[slot - 1] > -1 && scores[slot + 1] > -1) {
        double curvature = 2 * best - scores[slot - 1] - scores[slot + 1];
        if (curvature > 1e-8) {
          double offset = .125 * (scores[slot + 1] - scores[slot - 1]) / curvature;
          offset = fmax(-.125, fmin(.125, offset));
          double candidate = chosen + offset,
                 score = photo_zncc(a, b, w, h, x, y, reverse ? -candidate : candidate, r, 1);
          if (score > best) {
            best = score;
            chosen = candidate;
          }
        }
      }
      if (best >= fmax(.5, f->correlation) && best > initial + .002 && fabs(chosen - d) <= .75) {
        disparity[p] = (float)chosen;
        changed++;
      }
    }
  free(original);
  return changed;
}

/* CensusSgm.removeSpeckles (CensusSgm.java:651-681). */
static int64_t remove_speckles(float *a, int w, int h, int minimum, double range, int connectivity) {
  if (minimum <= 1) return 0;
  size_t n = (size_t)w * h;
  uint8_t *seen = (uint8_t *)calloc(n, 1);
  int *queue = (int *)malloc(sizeof(int) * n);
  int64_t rejected = 0;
  for (size_t start = 0; start < n; start++) {
    if (seen[start] || !isfinite_f(a[start])) continue;
    size_t head = 0, tail = 0;
    queue[tail++] = (int)start;
    seen[start] = 1;
    while (head < tail) {
      int p = queue[head++], x = p % w, y = p / w;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if ((dx == 0 && dy == 0) || (connectivity == 4 && dx != 0 && dy != 0)) continue;
          int xx = x + dx, yy = y + dy;
          if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
          int q = yy * w + xx;
          if (seen[q] || !isfinite_f(a[q]) || (double)fabsf(a[q] - a[p]) > range) continue;
          seen[q] = 1;
          queue[tail++] = q;
        }
    }
    if ((int64_t)tail < minimum)
      for (size_t i = 0; i < tail; i++) {
        a[queue[i]] = NAN;
        rejected++;
      }
  }
  free(seen);
  free(queue);
  return rejected;
}

double oracle_photo_zncc(const uint8_t *a, const uint8_t *b, int w, int h, int x, int y,
                         double disparity, int r) {
  return photo_zncc(a, b, w, h, x, y, disparity, r, 1);
}

double oracle_photo_zncc_ch(const uint8_t *a, const uint8_t *b, int w, int h, int x, int y,
                            double disparity, int r, int channels) {
  return photo_zncc(a, b, w, h, x, y, disparity, r, channels);
}

int oracle_postfilter(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                      const uint8_t *right, const float *dr, int w, int h, float *checked,
                      oracle_stats *stats, uint8_t *reasons) {
  return oracle_postfilter_ch(o, f, left, left, right, 1, dr, w, h, checked, stats, reasons);
}

int oracle_postfilter_ch(const oracle_params *o, const oracle_filter *f, const uint8_t *left,
                         const uint8_t *photo_left, const uint8_t *photo_right, int channels,
                         const float *dr, int w, int h, float *checked, oracle_stats *stats,
                         uint8_t *reasons) {
  size_t n = (size_t)w * h;
  int bx = o->window_w / 2, by = o->window_h / 2;
  if (f->median > 1) {
    float *out = (float *)malloc(sizeof(float) * n);
    if (f->median_classic)
      median_valid(checked, w, h, f->median, out);
    else
      median_edge(checked, w, h, f->median, f->median_max_change, left, f->median_edge_threshold,
                  out);
    memcpy(checked, out, sizeof(float) * n);
    free(out);
  }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t p = (size_t)y * w + x;
      float disparity = checked[p];
      if (!isfinite_f(disparity)) continue;
      if (!reciprocal(dr, w, h, x, y, disparity, o->lr_threshold, bx, by)) {
        checked[p] = NAN;
        stats->post_median_lr_rejected++;
        if (reasons) reasons[p] = 5;
        continue;
      }
      double score = photo_zncc(photo_left, photo_right, w, h, x, y, disparity,
                                f->correlation_window / 2, channels);
      if (score < f->correlation) {
        checked[p] = NAN;
        stats->photometric_rejected++;
        if (reasons) reasons[p] = 6;
      }
    }
  float *before = reasons ? (float *)malloc(sizeof(float) * n) : NULL;
  if (before) memcpy(before, checked, sizeof(float) * n);
  stats->speckle_rejected +=
      remove_speckles(checked, w, h, f->speckle_size, f->speckle_range, f->speckle_connectivity);
  stats->valid = 0;
  for (size_t i = 0; i < n; i++)
    if (isfinite_f(checked[i])) stats->valid++;
  if (before) {
    for (size_t i = 0; i < n; i++)
      if (isfinite_f(before[i]) && !isfinite_f(checked[i])) reasons[i] = 7;
    free(before);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The same restatement split over host threads (OpenMP).                    */
/*                                                                            */
/* oracle_match's arithmetic, loop for loop, with the work divided where the  */
/* reference's data dependences allow it:                                     */
/*   cost fill (CensusSgm.java:405-420)   every cell independent: rows split  */
/*   path recurrence (:421-459)           a horizontal path's rows are        */
/*       independent (each keeps its own previous-pixel vector); for the      */
/*       other paths a pixel reads only the previous row, so each row's       */
/*       pixels are split and the rows run in path order                      */
/*   WTA / subpixel (:460-515)            every pixel independent; counters   */
/*       summed                                                               */
/* Each cell's aggregate receives the same integer terms in the same path     */
/* order, so results equal oracle_match bit for bit (checked by the tests).   */

static void set_threads(int threads) {
#ifdef _OPENMP
  omp_set_num_threads(threads > 0 ? threads : omp_get_num_procs());
#else
  (void)threads;
#endif
}

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* One pixel step of the recurrence: L(p, .) from the predecessor's vector
 * `prev` (NULL: no predecessor), added into the aggregate. */
static inline void path_step(const uint16_t *costs, uint16_t *agg16,
                             int32_t *agg32, const int *prev, int *out, size_t base, int nd,
                             int p1, int localP2, int value_mask) {
  int minPrev = 0x7FFFFFFF;
  if (prev)
    for (int k = 0; k < nd; k++)
      if (prev[k] < minPrev) minPrev = prev[k];
  for (int k = 0; k < nd; k++) {
    int value = costs[base + k] & value_mask;
    if (prev) {
      int transition = prev[k];
      if (minPrev + localP2 < transition) transition = minPrev + localP2;
      if (k > 0 && prev[k - 1] + p1 < transition) transition = prev[k - 1] + p1;
      if (k + 1 < nd && prev[k + 1] + p1 < transition) transition = prev[k + 1] + p1;
      value += transition - minPrev;
    }
    out[k] = value;
    if (agg16) agg16[base + k] = (uint16_t)(agg16[base + k] + (uint32_t)value);
    else agg32[base + k] = (int32_t)((uint32_t)agg32[base + k] + (uint32_t)value);
  }
}

int oracle_match_parallel(const oracle_params *o, const uint64_t *a, const uint64_t *b,
                          const uint8_t *grayA, const uint8_t *grayB, const uint8_t *texturedA,
                          const uint8_t *texturedB, int width, int height, int reverse,
                          float *result, oracle_stats *stats, uint8_t *reasons,
                          uint8_t *searchBoundary, int threads) {
  int nd = o->disparities, rx = o->window_w / 2, ry = o->window_h / 2;
  int words = oracle_words(o->window_w, o->window_h);
  size_t count = (size_t)width * height, cells = count * (size_t)nd;
  int sum_bytes = oracle_aggregate_bytes(o);
  if (sum_bytes < 0 || nd < 1 || o->directions < 1 || o->directions > 8) return -1;
  set_threads(threads);
  int invalid = o->cost == OR_ZNCC1024 ? 0x8800 : INVALID_COST;
  int value_mask = o->cost == OR_ZNCC1024 ? 0x7FFF : 0xFFFF;
  uint16_t *costs = (uint16_t *)malloc(cells * sizeof(uint16_t));
  uint16_t *agg16 = NULL;
  int32_t *agg32 = NULL;
  if (sum_bytes == 2) agg16 = (uint16_t *)malloc(cells * sizeof(uint16_t));
  else agg32 = (int32_t *)malloc(cells * sizeof(int32_t));
  int *previous = (int *)malloc(sizeof(int) * (size_t)width * nd);
  int *row = (int *)malloc(sizeof(int) * (size_t)width * nd);
  if (!costs || (!agg16 && !agg32) || !previous || !row) {
    free(costs);
    free(agg16);
    free(agg32);
    free(previous);
    free(row);
    return -2;
  }
  /* cost fill (CensusSgm.java:405-420), zeroed aggregates */
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < height; y++) {
    size_t first = (size_t)y * width * nd, last = first + (size_t)width * nd;
    for (size_t i = first; i < last; i++) costs[i] = (uint16_t)invalid;
    if (agg16) memset(agg16 + first, 0, (last - first) * sizeof(uint16_t));
    else memset(agg32 + first, 0, (last - first) * sizeof(int32_t));
    if (y < ry || y >= height - ry) continue;
    for (int x = rx; x < width - rx; x++) {
      size_t p = (size_t)y * width + x, base = p * nd;
      if (!texturedA[p]) continue;
      for (int k = 0; k < nd; k++) {
        int disparity = o->min_disparity + k;
        int64_t target = (int64_t)x + (reverse ? disparity : -(int64_t)disparity);
        if (target < rx || target >= width - rx) continue;
        size_t q = (size_t)y * width + (size_t)target;
        if (!texturedB[q]) continue;
        costs[base + k] = (uint16_t)cost_value(o, a, b, grayA, grayB, p, q, width, count, words);
      }
    }
  }
  /* path recurrence (CensusSgm.java:421-459), directions in the reference order */
  static const int paths[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                  {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  int p1 = o->p1, p2 = oracle_effective_p2(o);
  for (int path = 0; path < o->directions; path++) {
    int dx = paths[path][0], dy = paths[path][1];
    int yStart = dy < 0 ? height - 1 : 0, yStep = dy < 0 ? -1 : 1;
    int xStart = dx < 0 ? width - 1 : 0, xStep = dx < 0 ? -1 : 1;
    if (dy == 0) {
#pragma omp parallel
      {
        int *line = (int *)malloc(sizeof(int) * 2 * (size_t)nd);
#pragma omp for schedule(dynamic, 4)
        for (int y = 0; y < height; y++)
          for (int xi = 0, x = xStart; xi < width; xi++, x += xStep) {
            int px = x - dx;
            int has = px >= 0 && px < width && texturedA[(size_t)y * width + px];
            int *out = line + (size_t)(xi & 1) * nd, *prev = line + (size_t)((xi + 1) & 1) * nd;
            int localP2 = has ? edge_p2(p1, p2, o->edge_scale, grayA, (size_t)y * width + x,
                                        (size_t)y * width + px)
                              : p2;
            path_step(costs, agg16, agg32, has ? prev : NULL, out,
                      ((size_t)y * width + x) * nd, nd, p1, localP2, value_mask);
          }
        free(line);
      }
      continue;
    }
    for (int yi = 0, y = yStart; yi < height; yi++, y += yStep) {
      int py = y - dy;
#pragma omp parallel for schedule(static)
      for (int x = 0; x < width; x++) {
        int px = x - dx;
        int has = px >= 0 && px < width && py >= 0 && py < height &&
                  texturedA[(size_t)py * width + px];
        int localP2 = has ? edge_p2(p1, p2, o->edge_scale, grayA, (size_t)y * width + x,
                                    (size_t)py * width + px)
                          : p2;
        path_step(costs, agg16, agg32, has ? previous + (size_t)px * nd : NULL,
                  row + (size_t)x * nd, ((size_t)y * width + x) * nd, nd, p1, localP2,
                  value_mask);
      }
      int *swap = previous;
      previous = row;
      row = swap;
    }
  }
#define AGG(i) (agg16 ? (int)agg16[i] : (int)agg32[i])
  /* WTA, boundary, uniqueness, subpixel (CensusSgm.java:460-515) */
  int64_t full_n = 0, boundary_n = 0, lower_n = 0, upper_n = 0, ambiguity_n = 0, selected_n = 0;
#pragma omp parallel for schedule(dynamic, 4) \
    reduction(+ : full_n, boundary_n, lower_n, upper_n, ambiguity_n, selected_n)
  for (int y = 0; y < height; y++) {
    for (int x = 0; x < width; x++) result[(size_t)y * width + x] = NAN;
    if (y < ry || y >= height - ry) continue;
    for (int x = rx; x < width - rx; x++) {
      size_t pixel = (size_t)y * width + x, base = pixel * nd;
      int best = -1, bestCost = 0x7FFFFFFF;
      for (int k = 0; k < nd; k++)
        if (costs[base + k] != invalid && AGG(base + k) < bestCost) {
          best = k;
          bestCost = AGG(base + k);
        }
      if (best < 0) continue;
      int full = costs[base] != invalid && costs[base + nd - 1] != invalid;
      if (full) {
        full_n++;
        if (searchBoundary) searchBoundary[pixel] = 4;
      }
      if (best == 0 || best == nd - 1) {
        boundary_n++;
        if (reasons) reasons[pixel] = 2;
        if (full && best == 0) {
          lower_n++;
          if (searchBoundary) searchBoundary[pixel] |= 1;
        }
        if (full && best == nd - 1) {
          upper_n++;
          if (searchBoundary) searchBoundary[pixel] |= 2;
        }
        continue;
      }
      int second = 0x7FFFFFFF;
      for (int k = 0; k < nd; k++)
        if (abs(k - best) > 1 && costs[base + k] != invalid && AGG(base + k) < second)
          second = AGG(base + k);
      double allowed = (double)bestCost * o->uniqueness;
      if (!(allowed >= 1.0)) allowed = 1.0;
      if (second != 0x7FFFFFFF && (double)(second - bestCost) <= allowed) {
        ambiguity_n++;
        if (reasons) reasons[pixel] = 3;
        continue;
      }
      double offset = 0;
      if (best > 0 && best + 1 < nd && costs[base + best - 1] != invalid &&
          costs[base + best + 1] != invalid) {
        double before = AGG(base + best - 1), after = AGG(base + best + 1);
        double curvature = before + after - 2.0 * bestCost;
        if (curvature > 0) {
          double t = .5 * (before - after) / curvature;
          t = t < .5 ? t : .5;
          offset = t > -.5 ? t : -.5;
        }
      }
      result[pixel] = (float)((double)(o->min_disparity + best) + offset);
      selected_n++;
      if (reasons) reasons[pixel] = 0;
    }
  }
#undef AGG
  stats->full_interval_pixels += full_n;
  stats->boundary_rejected += boundary_n;
  stats->lower_boundary_rejected += lower_n;
  stats->upper_boundary_rejected += upper_n;
  stats->ambiguity_rejected += ambiguity_n;
  stats->selected += selected_n;
  free(costs);
  free(agg16);
  free(agg32);
  free(previous);
  free(row);
  return 0;
}

/* oracle_compute through oracle_match_parallel (gray census, no filters). */
int oracle_compute_parallel(const oracle_params *o, const uint8_t *left, const uint8_t *right,
                            const uint8_t *valid_left, const uint8_t *valid_right, int width,
                            int height, float *checked, float *dl_out, float *dr_out,
                            oracle_stats *stats, uint8_t *reasons, uint8_t *searchBoundary,
                            int threads) {
  if (width < o->window_w + 2 || height < o->window_h + 2) return -1;
  size_t n = (size_t)width * height;
  int words = oracle_words(o->window_w, o->window_h);
  uint64_t *a = (uint64_t *)malloc(sizeof(uint64_t) * n * words);
  uint64_t *b = (uint64_t *)malloc(sizeof(uint64_t) * n * words);
  uint8_t *ta = (uint8_t *)malloc(n), *tb = (uint8_t *)malloc(n);
  float *dl = dl_out ? dl_out : (float *)malloc(sizeof(float) * n);
  float *dr = dr_out ? dr_out : (float *)malloc(sizeof(float) * n);
  if (!a || !b || !ta || !tb || !dl || !dr) return -2;
  if (reasons) memset(reasons, 1, n);
  if (searchBoundary) memset(searchBoundary, 0, n);
  oracle_census(left, width, height, o->window_w, o->window_h, a);
  oracle_census(right, width, height, o->window_w, o->window_h, b);
  oracle_texture(left, valid_left, width, height, o->window_w, o->window_h, ta);
  oracle_texture(right, valid_right, width, height, o->window_w, o->window_h, tb);
  oracle_stats discarded;
  memset(&discarded, 0, sizeof discarded);
  int rc = oracle_match_parallel(o, a, b, left, right, ta, tb, width, height, 0, dl, stats,
                                 reasons, searchBoundary, threads);
  if (rc == 0)
    rc = oracle_match_parallel(o, b, a, right, left, tb, ta, width, height, 1, dr, &discarded,
                               NULL, NULL, threads);
  if (rc == 0) {
    /* left-right check (CensusSgm.java:181-201) */
    int rx = o->window_w / 2, ry = o->window_h / 2;
    for (size_t i = 0; i < n; i++) checked[i] = NAN;
    for (int y = ry; y < height - ry; y++)
      for (int x = rx; x < width - rx; x++) {
        size_t p = (size_t)y * width + x;
        if (!isfinite_f(dl[p])) continue;
        int xr = oracle_round_float((float)x - dl[p]);
        if (xr < rx || xr >= width - rx) {
          if (reasons) reasons[p] = 4;
          continue;
        }
        float other = dr[(size_t)y * width + xr];
        if (isfinite_f(other) && (double)fabsf(dl[p] - other) <= o->lr_threshold)
          checked[p] = dl[p];
        else {
          stats->lr_rejected++;
          if (reasons) reasons[p] = 4;
        }
      }
    for (size_t i = 0; i < n; i++)
      if (isfinite_f(checked[i])) stats->valid++;
  }
  free(a);
  free(b);
  free(ta);
  free(tb);
  if (!dl_out) free(dl);
  if (!dr_out) free(dr);
  return rc;
}
