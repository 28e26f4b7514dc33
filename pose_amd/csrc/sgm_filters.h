// Post-left-right filters of the reference's CensusSgm.computeRecorded
// (src/main/java/sgm/CensusSgm.java:203-271, 557-681), on the device:
//   03_median        median_kernel (classic or edge-preserving)
//   04_photometric   photo_kernel  (post-median reciprocity + subpixel ZNCC)
//   05_final_stereo  cc_* kernels  (speckle removal by connected components)
// Every kernel is per pixel and reproduces the reference's float/double
// arithmetic in the same order (the library is built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgm {

constexpr int kMedianMaxWindow = 7;  // up to 7x7 = 49 values per pixel

struct FilterArgs {
  int w, h;
  int median, median_classic, median_edge;
  double median_max_change;
  int bx, by;            // reciprocity border (census radius per axis)
  double lr_threshold;
  int corr_r;            // photometric window radius
  int channels;          // interleaved channels of the photometric rasters (1 or 3)
  double correlation;
  int speckle_size, speckle_conn;
  double speckle_range;
};

// Float.compare order for finite values (-0.0 before 0.0) as a signed key.
__device__ __forceinline__ int float_order_key(float v) {
  const int b = __float_as_int(v);
  return b < 0 ? b ^ 0x7FFFFFFF : b;
}

// Element n / 2 of the sorted values (insertion sort; n <= 49).
__device__ __forceinline__ float sorted_middle(float *v, int n) {
  for (int i = 1; i < n; i++) {
    const float x = v[i];
    const int kx = float_order_key(x);
    int j = i - 1;
    while (j >= 0 && float_order_key(v[j]) > kx) {
      v[j + 1] = v[j];
      j--;
    }
    v[j + 1] = x;
  }
  return v[n / 2];
}

// CensusSgm.medianValid / medianEdgePreserving (CensusSgm.java:557-600).
static __global__ __launch_bounds__(256) void median_kernel(const float *__restrict__ in,
                                                            float *__restrict__ out,
                                                            const uint8_t *__restrict__ gray,
                                                            FilterArgs F) {
  const size_t n = (size_t)F.w * F.h;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int x = (int)(p % F.w), y = (int)(p / F.w), r = F.median / 2;
  const float center = in[p];
  if (!isfinite(center)) {
    out[p] = center;
    return;
  }
  float v[kMedianMaxWindow * kMedianMaxWindow];
  const int y0 = max(0, y - r), y1 = min(F.h - 1, y + r), x0 = max(0, x - r), x1 = min(F.w - 1, x + r);
  int cnt = 0;
  for (int yy = y0; yy <= y1; yy++)
    for (int xx = x0; xx <= x1; xx++) {
      const float a = in[(size_t)yy * F.w + xx];
      if (isfinite(a)) v[cnt++] = a;
    }
  float m = sorted_middle(v, cnt);
  if (!F.median_classic && !((double)fabsf(m - center) < F.median_max_change)) {
    const int g = gray[p];
    cnt = 0;
    for (int yy = y0; yy <= y1; yy++)
      for (int xx = x0; xx <= x1; xx++) {
        const size_t q = (size_t)yy * F.w + xx;
        const float a = in[q];
        if (isfinite(a) && (double)fabsf(a - center) <= F.median_max_change &&
            (F.median_edge == 0 || abs((int)gray[q] - g) <= F.median_edge))
          v[cnt++] = a;
      }
    if (cnt > 0) m = sorted_middle(v, cnt);
  }
  out[p] = m;
}

// Block-wide sums of two counters, one atomic per block and counter.
__device__ __forceinline__ void add_counts(unsigned long long *stats, int f0, unsigned c0, int f1,
                                           unsigned c1) {
  __shared__ unsigned cnt[2];
  if (threadIdx.x == 0) cnt[0] = cnt[1] = 0;
  __syncthreads();
  if (c0) atomicAdd(&cnt[0], c0);
  if (c1) atomicAdd(&cnt[1], c1);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cnt[0]) atomicAdd(stats + f0, (unsigned long long)cnt[0]);
    if (cnt[1]) atomicAdd(stats + f1, (unsigned long long)cnt[1]);
  }
}

// CensusSgm.zncc (CensusSgm.java:614-644): the photometric score of pixel
// (x, y) at a subpixel disparity over `ch` interleaved channels, in double in
// the reference's summation order; -1 outside the image or on flat windows.
__device__ __forceinline__ double photo_score(const uint8_t *__restrict__ left,
                                              const uint8_t *__restrict__ right, int w, int h,
                                              int x, int y, double disparity, int r, int ch) {
  const double target = (double)x - disparity;
  const int bxi = (int)floor(target);
  const double fraction = target - bxi;
  if (x - r < 0 || x + r >= w || y - r < 0 || y + r >= h || bxi - r < 0 || bxi + r + 1 >= w)
    return -1.0;
  const int cnt = (2 * r + 1) * (2 * r + 1);
  double cov = 0, va = 0, vb = 0;
  for (int c = 0; c < ch; c++) {
    double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
    for (int dy = -r; dy <= r; dy++)
      for (int dx = -r; dx <= r; dx++) {
        const size_t ai = ((size_t)(y + dy) * w + x + dx) * ch + c;
        const size_t bi = ((size_t)(y + dy) * w + bxi + dx) * ch + c;
        const double av = left[ai];
        const double bv = (1 - fraction) * right[bi] + fraction * right[bi + ch];
        sa += av;
        sb += bv;
        saa += av * av;
        sbb += bv * bv;
        sab += av * bv;
      }
    cov += sab - sa * sb / cnt;
    va += saa - sa * sa / cnt;
    vb += sbb - sb * sb / cnt;
  }
  if (va < ch * cnt * 4 || vb < ch * cnt * 4) return -1.0;
  const double prod = va * vb;
  return cov / sqrt(prod > 1e-12 ? prod : 1e-12);
}

// CensusSgm.reciprocal + CensusSgm.zncc with the photometric threshold
// (CensusSgm.java:224-257, 602-644); `left` / `right` hold F.channels
// interleaved channels (gray, or BGR in colour-distance mode).
static __global__ __launch_bounds__(256) void photo_kernel(float *__restrict__ disp,
                                                           const float *__restrict__ dr,
                                                           const uint8_t *__restrict__ left,
                                                           const uint8_t *__restrict__ right,
                                                           uint8_t *__restrict__ reasons,
                                                           unsigned long long *stats, FilterArgs F) {
  const size_t n = (size_t)F.w * F.h;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  unsigned lr_rej = 0, photo_rej = 0;
  if (p < n) {
    const int x = (int)(p % F.w), y = (int)(p / F.w);
    const float d = disp[p];
    if (isfinite(d)) {
      bool ok = y >= F.by && y < F.h - F.by;
      if (ok) {
        const int xr = round_half_up_f((float)x - d);
        ok = xr >= F.bx && xr < F.w - F.bx;
        if (ok) {
          const float o = dr[(size_t)y * F.w + xr];
          ok = isfinite(o) && (double)fabsf(d - o) <= F.lr_threshold;
        }
      }
      if (!ok) {
        disp[p] = __builtin_nanf("");
        lr_rej = 1;
        if (reasons) reasons[p] = 5;
      } else {
        const double score =
            photo_score(left, right, F.w, F.h, x, y, (double)d, F.corr_r, F.channels);
        if (score < F.correlation) {
          disp[p] = __builtin_nanf("");
          photo_rej = 1;
          if (reasons) reasons[p] = 6;
        }
      }
    }
  }
  add_counts(stats, 7, lr_rej, 8, photo_rej);
}

// PhotometricRefinement.refine (PhotometricRefinement.java:5-61) of one
// matching direction (--subpixel-refinement zncc, applied before the L-R check,
// CensusSgm.java:168-170): away from depth edges (a finite 3x3 neighbour more
// than 1.5 away), the ZNCC of the gray windows at the disparity and at seven
// quarter steps around it inside the search range, a parabolic eighth-step
// refinement around the best, and the best kept only if it is a clear
// improvement.  Reads the unrefined map `orig`, writes `disp`.
struct RefineArgs {
  int w, h, r, min_d, count, reverse;
  double correlation;
};

static __global__ __launch_bounds__(256) void refine_kernel(const float *__restrict__ orig,
                                                            float *__restrict__ disp,
                                                            const uint8_t *__restrict__ a,
                                                            const uint8_t *__restrict__ b,
                                                            RefineArgs G) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  const int r = G.r;
  if (x < r || x >= G.w - r || y < r || y >= G.h - r) return;
  const size_t p = (size_t)y * G.w + x;
  const double d = orig[p];
  if (!isfinite(d)) return;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const float q = orig[p + (ptrdiff_t)dy * G.w + dx];
      if (isfinite(q) && fabs((double)q - d) > 1.5) return;  // depth edge
    }
  const double sgn = G.reverse ? -1.0 : 1.0;
  const double initial = photo_score(a, b, G.w, G.h, x, y, sgn * d, r, 1);
  double best = initial, chosen = d;
  int slot = -1;
  double scores[7];
  for (int k = 0; k < 7; k++) {
    const double candidate = d + (k - 3) * .25;
    if (candidate <= G.min_d || candidate >= G.min_d + G.count - 1) {
      scores[k] = -1;
      continue;
    }
    scores[k] = photo_score(a, b, G.w, G.h, x, y, sgn * candidate, r, 1);
    if (scores[k] > best) {
      best = scores[k];
      chosen = candidate;
      slot = k;
    }
  }
  if (slot > 0 && slot < 6 && scores[slot - 1] > -1 && scores[slot + 1] > -1) {
    const double curvature = 2 * best - scores[slot - 1] - scores[slot + 1];
    if (curvature > 1e-8) {
      double offset = .125 * (scores[slot + 1] - scores[slot - 1]) / curvature;
      offset = fmax(-.125, fmin(.125, offset));
      const double candidate = chosen + offset;
      const double score = photo_score(a, b, G.w, G.h, x, y, sgn * candidate, r, 1);
      if (score > best) {
        best = score;
        chosen = candidate;
      }
    }
  }
  if (best >= fmax(.5, G.correlation) && best > initial + .002 && fabs(chosen - d) <= .75)
    disp[p] = (float)chosen;
}

// ---- speckles: connected components of 4/8-connected pixels whose values
// differ by at most the range (CensusSgm.removeSpeckles, CensusSgm.java:651-681),
// labelled by union-find in three steps:
//   cc_local_kernel   union-find inside a 32x8 tile in LDS, one thread per pixel;
//                     the global parent of a pixel becomes its tile-local root
//   cc_border_kernel  lock-free global unions only for links that cross tiles
//   cc_flatten_kernel every pixel points straight at its component's root
// Roots only ever point to smaller indices (a tile-local root is its
// component's smallest pixel; global links go from the larger root to the
// smaller), so there are no cycles; every loop is also capped at n steps and
// raises the error flag if it runs out.

constexpr int kCcTx = 32, kCcTy = 8;

__device__ __forceinline__ int cc_parent(int *parent, int x) {
  return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int cc_find(int *parent, int x, int limit, int *error) {
  for (int it = 0; it < limit; it++) {
    const int px = cc_parent(parent, x);
    if (px == x) return x;
    const int gp = cc_parent(parent, px);
    if (gp != px) atomicCAS(parent + x, px, gp);  // path halving
    x = px;
  }
  atomicExch(error, 1);
  return x;
}

__device__ __forceinline__ void cc_unite(int *parent, int a, int b, int limit, int *error) {
  for (int it = 0; it < limit; it++) {
    a = cc_find(parent, a, limit, error);
    b = cc_find(parent, b, limit, error);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(parent + b, b, a) == b) return;
  }
  atomicExch(error, 1);
}

// LDS counterparts over the kCcTx * kCcTy tile-local indices.
__device__ __forceinline__ int cc_find_lds(int *lp, int x, int *bad) {
  for (int it = 0; it < kCcTx * kCcTy; it++) {
    const int px = __hip_atomic_load(lp + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (px == x) return x;
    x = px;
  }
  *bad = 1;
  return x;
}

__device__ __forceinline__ void cc_unite_lds(int *lp, int a, int b, int *bad) {
  for (int it = 0; it < kCcTx * kCcTy; it++) {
    a = cc_find_lds(lp, a, bad);
    b = cc_find_lds(lp, b, bad);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(lp + b, b, a) == b) return;
  }
  *bad = 1;
}

__device__ __forceinline__ bool cc_linked(const float *a, size_t p, size_t q, double range) {
  const float aq = a[q];
  return isfinite(aq) && !((double)fabsf(aq - a[p]) > range);
}

// Links of pixel (x, y) to its forward neighbours: right, down and, for
// 8-connectivity, down-left and down-right (each undirected link once).
template <typename F>
__device__ __forceinline__ void cc_forward_links(int x, int y, const FilterArgs &A, F &&link) {
  if (x + 1 < A.w) link(x + 1, y);
  if (y + 1 < A.h) {
    link(x, y + 1);
    if (A.speckle_conn == 8) {
      if (x > 0) link(x - 1, y + 1);
      if (x + 1 < A.w) link(x + 1, y + 1);
    }
  }
}

static __global__ __launch_bounds__(256) void cc_local_kernel(const float *__restrict__ a,
                                                              int *parent, int *size, int *error,
                                                              FilterArgs F) {
  __shared__ int lp[kCcTx * kCcTy];
  __shared__ int bad;
  const int tx = threadIdx.x, ty = threadIdx.y, l = ty * kCcTx + tx;
  const int x0 = blockIdx.x * kCcTx, y0 = blockIdx.y * kCcTy;
  const int x = x0 + tx, y = y0 + ty;
  const bool inside = x < F.w && y < F.h;
  const size_t p = (size_t)y * F.w + x;
  const bool valid = inside && isfinite(a[p]);
  if (l == 0) bad = 0;
  lp[l] = valid ? l : -1;
  __syncthreads();
  if (valid)
    cc_forward_links(x, y, F, [&](int qx, int qy) {
      if (qx < x0 || qx >= x0 + kCcTx || qy >= y0 + kCcTy) return;  // crosses the tile
      if (cc_linked(a, p, (size_t)qy * F.w + qx, F.speckle_range))
        cc_unite_lds(lp, l, (qy - y0) * kCcTx + (qx - x0), &bad);
    });
  __syncthreads();
  if (inside) {
    int root = -1;
    if (valid) {
      const int r = cc_find_lds(lp, l, &bad);
      root = (int)((size_t)(y0 + r / kCcTx) * F.w + x0 + r % kCcTx);
    }
    parent[p] = root;
    size[p] = 0;
  }
  __syncthreads();
  if (l == 0 && bad) atomicExch(error, 1);
}

static __global__ __launch_bounds__(256) void cc_border_kernel(const float *__restrict__ a,
                                                               int *parent, int *error,
                                                               FilterArgs F) {
  const int x = blockIdx.x * kCcTx + threadIdx.x, y = blockIdx.y * kCcTy + threadIdx.y;
  if (x >= F.w || y >= F.h) return;
  const size_t p = (size_t)y * F.w + x;
  if (!isfinite(a[p])) return;
  const int tx0 = x - (int)threadIdx.x, ty0 = y - (int)threadIdx.y, limit = F.w * F.h + 1;
  cc_forward_links(x, y, F, [&](int qx, int qy) {
    if (qx >= tx0 && qx < tx0 + kCcTx && qy < ty0 + kCcTy) return;  // inside the tile: done
    const size_t q = (size_t)qy * F.w + qx;
    if (cc_linked(a, p, q, F.speckle_range)) cc_unite(parent, (int)p, (int)q, limit, error);
  });
}

static __global__ __launch_bounds__(256) void cc_flatten_kernel(int *parent, int *error, size_t n) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n || parent[p] < 0) return;
  const int root = cc_find(parent, (int)p, (int)n + 1, error);
  __hip_atomic_store(parent + p, root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Component sizes: lanes of a wave with the same root add their count with
// one atomic (a large component otherwise serialises on its root's counter).
static __global__ __launch_bounds__(256) void cc_size_kernel(const int *parent, int *size,
                                                             size_t n) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int root = p < n ? parent[p] : -1;
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long active = __ballot(root >= 0);
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const int r = __shfl(root, leader);
    const unsigned long long same = __ballot(root == r);
    if (lane == leader) atomicAdd(size + r, (int)__popcll(same));
    active &= ~same;
  }
}

static __global__ __launch_bounds__(256) void cc_remove_kernel(float *a, const int *parent,
                                                               const int *size, uint8_t *reasons,
                                                               unsigned long long *stats,
                                                               FilterArgs F) {
  const size_t n = (size_t)F.w * F.h;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  unsigned removed = 0, valid = 0;
  if (p < n && parent[p] >= 0) {
    if (size[parent[p]] < F.speckle_size) {
      a[p] = __builtin_nanf("");
      removed = 1;
      if (reasons) reasons[p] = 7;
    } else {
      valid = 1;
    }
  }
  add_counts(stats, 9, removed, 10, valid);
}

// Valid count when speckle removal is off.
static __global__ __launch_bounds__(256) void count_valid_kernel(const float *a,
                                                                 unsigned long long *stats,
                                                                 size_t n) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  const unsigned v = (p < n && isfinite(a[p])) ? 1u : 0u;
  add_counts(stats, 10, v, 9, 0u);
}

}  // namespace sgm
