/*
 * Sanitizer run of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Built with -fsanitize=address,undefined (oracle/Makefile, target
 * `sanitize`) together with sgm_oracle.c, it drives every oracle entry point
 * over edge-case inputs: the smallest accepted images (window + 2), ranges
 * wider than the image, signed ranges, every cost, 4 and 8 paths, 16- and
 * 32-bit aggregates, rectangular windows, masks, the post-L-R filters with
 * subpixel refinement, the colour census, and the threaded oracle, which must
 * equal the literal one bit for bit.  Any out-of-bounds access, overflow or
 * other undefined behaviour aborts the run; a mismatch exits with status 1.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sgm_oracle.h"

static uint64_t rng = 88172645463325252ULL;
static uint32_t next_u32(void) {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint32_t)(rng >> 11);
}

static void fill_pair(uint8_t *a, uint8_t *b, int w, int h, int shift) {
  for (int i = 0; i < w * h; i++) a[i] = (uint8_t)(next_u32() & 255);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int sx = x + shift < 0 ? 0 : x + shift >= w ? w - 1 : x + shift;
      b[y * w + x] = a[y * w + sx];
    }
}

static int same_maps(const float *a, const float *b, size_t n) { return memcmp(a, b, n * 4) == 0; }

static int run_case(int w, int h, int ww, int wh, int D, int minD, int cost, int dirs, int agg,
                    int masked) {
  const size_t n = (size_t)w * h;
  uint8_t *l = malloc(n), *r = malloc(n), *vl = malloc(n), *vr = malloc(n);
  uint8_t *reasons = malloc(n), *boundary = malloc(n), *reasons2 = malloc(n), *boundary2 = malloc(n);
  float *checked = malloc(n * 4), *dl = malloc(n * 4), *dr = malloc(n * 4);
  float *checked2 = malloc(n * 4), *dl2 = malloc(n * 4), *dr2 = malloc(n * 4);
  fill_pair(l, r, w, h, (int)(next_u32() % 7) - 3);
  for (size_t i = 0; i < n; i++) {
    vl[i] = (next_u32() % 17) != 0;
    vr[i] = (next_u32() % 13) != 0;
  }
  oracle_params o;
  memset(&o, 0, sizeof o);
  o.window_w = ww;
  o.window_h = wh;
  o.disparities = D;
  o.min_disparity = minD;
  o.cost = cost;
  o.p1 = 4 + (int)(next_u32() % 8);
  o.p2 = o.p1 + 10 + (int)(next_u32() % 100);
  o.edge_scale = (next_u32() & 1) ? 32 : 0;
  o.directions = dirs;
  o.aggregate_bytes = agg;
  o.census_weight = 500;
  o.ad_truncation = 30;
  o.gradient_truncation = 60;
  o.uniqueness = 0.03;
  o.lr_threshold = 1.0;
  oracle_stats s1, s2;
  memset(&s1, 0, sizeof s1);
  memset(&s2, 0, sizeof s2);
  const uint8_t *ml = masked ? vl : NULL, *mr = masked ? vr : NULL;
  int rc = oracle_compute(&o, l, r, ml, mr, w, h, checked, dl, dr, &s1, reasons, boundary);
  int rc2 = oracle_compute_parallel(&o, l, r, ml, mr, w, h, checked2, dl2, dr2, &s2, reasons2,
                                    boundary2, 3);
  int bad = rc != rc2;
  if (!bad && rc == 0)
    bad = !same_maps(checked, checked2, n) || !same_maps(dl, dl2, n) || !same_maps(dr, dr2, n) ||
          memcmp(reasons, reasons2, n) || memcmp(boundary, boundary2, n) ||
          memcmp(&s1, &s2, sizeof s1);
  if (rc == 0 && cost <= 2) {
    /* post-L-R filters, with refinement, on the same pair */
    oracle_filter f;
    memset(&f, 0, sizeof f);
    f.median = 3 + 2 * (int)(next_u32() % 2);
    f.median_classic = (int)(next_u32() & 1);
    f.median_edge_threshold = 24;
    f.correlation_window = ww;
    f.speckle_size = 5 + (int)(next_u32() % 20);
    f.speckle_connectivity = (next_u32() & 1) ? 8 : 4;
    f.subpixel_refinement = 1;
    f.median_max_change = 0.5;
    f.correlation = 0.2;
    f.speckle_range = 1.0;
    oracle_stats s3;
    memset(&s3, 0, sizeof s3);
    if (oracle_compute_ex(&o, &f, l, r, NULL, NULL, ml, mr, w, h, checked, dl, dr, &s3, reasons,
                          boundary) == 0)
      oracle_postfilter(&o, &f, l, r, dr, w, h, checked, &s3, reasons);
    /* colour census over BGR copies of the gray rasters */
    uint8_t *bl = malloc(3 * n), *br = malloc(3 * n);
    for (size_t i = 0; i < n; i++)
      for (int c = 0; c < 3; c++) {
        bl[3 * i + c] = (uint8_t)(l[i] + 7 * c);
        br[3 * i + c] = (uint8_t)(r[i] + 7 * c);
      }
    memset(&s3, 0, sizeof s3);
    oracle_compute_color(&o, l, r, bl, br, ml, mr, w, h, checked, dl, dr, &s3, reasons, boundary);
    free(bl);
    free(br);
  }
  free(l), free(r), free(vl), free(vr), free(reasons), free(boundary), free(reasons2);
  free(boundary2), free(checked), free(dl), free(dr), free(checked2), free(dl2), free(dr2);
  if (bad)
    fprintf(stderr, "mismatch: %dx%d win %dx%d D=%d minD=%d cost %d dirs %d agg %d masked %d\n",
            w, h, ww, wh, D, minD, cost, dirs, agg, masked);
  return bad;
}

int main(void) {
  int bad = 0, cases = 0;
  static const int windows[][2] = {{3, 3}, {5, 5}, {7, 5}, {9, 7}, {11, 11}, {3, 11}};
  for (int wi = 0; wi < 6; wi++)
    for (int cost = 0; cost <= 4; cost++)
      for (int dirs = 4; dirs <= 8; dirs += 4)
        for (int agg = 2; agg <= 4; agg += 2) {
          const int ww = windows[wi][0], wh = windows[wi][1];
          /* smallest accepted image, a small one, one with a range wider than it */
          const int sizes[3][2] = {{ww + 2, wh + 2}, {37, 23}, {29, 17}};
          for (int si = 0; si < 3; si++) {
            const int w = sizes[si][0], h = sizes[si][1];
            const int D = si == 2 ? w + 9 : 3 + (int)(next_u32() % 20);
            const int minD = (int)(next_u32() % 9) - 4;
            bad += run_case(w, h, ww, wh, D, minD, cost, dirs, agg, (int)(next_u32() & 1));
            cases++;
          }
        }
  printf("sanitizer run: %d cases, %d mismatches\n", cases, bad);
  return bad ? 1 : 0;
}
