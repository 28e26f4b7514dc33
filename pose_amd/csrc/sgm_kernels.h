// gfx950 device kernels of the SGM hot path.
//
// Layout in HBM (one matching direction of one pair, n = W*H pixels):
//   desc   uint64 word planes      desc[word*n + p]         (GpuStereo.java:243-251 layout)
//   mask   uint8  texture&valid    mask[p]
//   cost   uint16 [p][Dp]          k fastest, Dp = D rounded up to the lane group
//   sum    uint16|int32 [p][Dp]    aggregate S, same indexing as cost
// One wave64 owns one scanline of one path direction; lane l holds the R
// consecutive disparities k = l*R .. l*R+R-1, so a pixel step moves one
// contiguous Dp-vector per volume with a single vector load per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace sgm {

constexpr int kInf = 0x3FFFFFFF;
constexpr int kMaxWords = 2;

// ---------------------------------------------------------------------------
// Java rounding: Math.round(float) / Math.round(double) = closest, ties -> +inf.
__device__ __forceinline__ int round_half_up_f(float v) {
  float f = floorf(v);
  return (int)f + ((v - f) >= 0.5f ? 1 : 0);
}
__device__ __forceinline__ long long round_half_up_d(double v) {
  double f = floor(v);
  return (long long)f + ((v - f) >= 0.5 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// Cross-lane primitives (DPP, wave64).  All lanes must be active.
__device__ __forceinline__ int dpp_shr1(int v, int fill) {  // lane l <- lane l-1
  return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_shl1(int v, int fill) {  // lane l <- lane l+1
  return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xF, 0xF, false);
}

__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0xB1, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0x4E, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0x141, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0x140, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0x142, 0xA, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, 0x143, 0xC, 0xF, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Wave minima of a (signed) and b (unsigned) at once: the two reductions are
// interleaved step by step, so each DPP operation has an independent one
// between it and the operation it reads.  Both results are uniform.
__device__ __forceinline__ void wave_min_and_min_u(int &a, unsigned &b) {
  int v = a, u = (int)b, t;
#define SGM_MIN2_STEP(ctrl, rmask)                                                \
  v = min(v, __builtin_amdgcn_update_dpp(0x7FFFFFFF, v, ctrl, rmask, 0xF, false)); \
  t = __builtin_amdgcn_update_dpp(-1, u, ctrl, rmask, 0xF, false);                  \
  u = (int)min((unsigned)u, (unsigned)t);
  SGM_MIN2_STEP(0xB1, 0xF)
  SGM_MIN2_STEP(0x4E, 0xF)
  SGM_MIN2_STEP(0x141, 0xF)
  SGM_MIN2_STEP(0x140, 0xF)
  SGM_MIN2_STEP(0x142, 0xA)
  SGM_MIN2_STEP(0x143, 0xC)
#undef SGM_MIN2_STEP
  a = __builtin_amdgcn_readlane(v, 63);
  b = (unsigned)__builtin_amdgcn_readlane(u, 63);
}

__device__ __forceinline__ unsigned wave_min_u(unsigned u) {
  int v = (int)u;
  // unsigned min via the int builtins: compare as unsigned
  int t;
  t = __builtin_amdgcn_update_dpp(-1, v, 0xB1, 0xF, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  t = __builtin_amdgcn_update_dpp(-1, v, 0x4E, 0xF, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  t = __builtin_amdgcn_update_dpp(-1, v, 0x141, 0xF, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  t = __builtin_amdgcn_update_dpp(-1, v, 0x140, 0xF, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  t = __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xA, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  t = __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xC, 0xF, false);
  v = (int)min((unsigned)v, (unsigned)t);
  return (unsigned)__builtin_amdgcn_readlane(v, 63);
}

// ---------------------------------------------------------------------------
// Census transform + texture mask (CensusSgm.java:523-555), rectangular window.
struct CensusJob {
  const uint8_t *gray;      // input raster
  const uint8_t *valid;     // optional validity mask (0 = invalid)
  const uint8_t *mask_in;   // optional precomputed texture&valid mask (skips texture)
  uint8_t *gray_copy;       // optional copy of gray (device-resident copy for later kernels)
  uint64_t *desc;           // words planes
  uint8_t *mask;            // texture & valid
  uint16_t *gm;             // gray | (texture & valid) << 8, read by the path sweeps
  int32_t *wsum, *wsum2;    // optional window sums of g and g^2 (ZNCC costs)
  int w, h;
};

constexpr int kCensusTx = 32, kCensusTy = 8;

struct CensusJobs {
  CensusJob job[2];
};

template <int WORDS>
__global__ __launch_bounds__(256) void census_kernel(CensusJobs jobs, int ww, int wh) {
  const CensusJob J = jobs.job[blockIdx.z];
  const int rx = ww >> 1, ry = wh >> 1;
  const int x0 = blockIdx.x * kCensusTx, y0 = blockIdx.y * kCensusTy;
  if (x0 >= J.w || y0 >= J.h) return;
  __shared__ uint8_t tile[kCensusTy + 10][kCensusTx + 10 + 2];
  const int tw = kCensusTx + 2 * rx, th = kCensusTy + 2 * ry;
  const int tid = threadIdx.y * kCensusTx + threadIdx.x;
  for (int i = tid; i < tw * th; i += 256) {
    int ty = i / tw, tx = i - ty * tw;
    int gx = x0 - rx + tx, gy = y0 - ry + ty;
    tile[ty][tx] = (gx >= 0 && gx < J.w && gy >= 0 && gy < J.h) ? J.gray[(size_t)gy * J.w + gx] : 0;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
  if (x >= J.w || y >= J.h) return;
  const size_t n = (size_t)J.w * J.h, p = (size_t)y * J.w + x;
  if (J.gray_copy) J.gray_copy[p] = tile[threadIdx.y + ry][threadIdx.x + rx];
  const bool interior = x >= rx && x < J.w - rx && y >= ry && y < J.h - ry;
  uint64_t d0 = 0, d1 = 0;
  int lo = 255, hi = 0, s1 = 0, s2 = 0;
  if (interior) {
    const int c = tile[threadIdx.y + ry][threadIdx.x + rx];
    int bit = 0;
    for (int yy = 0; yy < wh; yy++)
      for (int xx = 0; xx < ww; xx++) {
        const int v = tile[threadIdx.y + yy][threadIdx.x + xx];
        lo = min(lo, v);
        hi = max(hi, v);
        s1 += v;
        s2 += v * v;
        if (yy == ry && xx == rx) continue;
        const uint64_t m = (uint64_t)(v < c) << (bit & 63);
        if (WORDS == 1 || bit < 64) d0 |= m;
        else d1 |= m;
        bit++;
      }
  }
  uint8_t m;
  if (J.mask_in) m = J.mask_in[p] != 0;
  else m = interior && (hi - lo) >= 4 && (J.valid == nullptr || J.valid[p] != 0);
  // One-word descriptors use at most 62 bits; bit 63 carries texture&valid so a
  // sweep can test a target's usability from its descriptor alone.
  J.desc[p] = WORDS == 1 ? (d0 | ((uint64_t)m << 63)) : d0;
  if (WORDS > 1) J.desc[n + p] = d1;
  J.mask[p] = m;
  if (J.gm) J.gm[p] = (uint16_t)(tile[threadIdx.y + ry][threadIdx.x + rx] | (m << 8));
  if (J.wsum) {
    J.wsum[p] = s1;
    J.wsum2[p] = s2;
  }
}

// ---------------------------------------------------------------------------
// Colour-distance census (CensusSgm.colorCensus, CensusSgm.java:302-332) from
// interleaved BGR, rectangular window: bit k (window order, centre skipped) is
// set where the Euclidean colour distance to the centre is >= the mean
// distance over all ww*wh cells (the centre's zero included); a pixel is
// textured when its largest distance is >= 4.  Distances, their sum and the
// mean are double, summed in window order; the second pass recomputes each
// (correctly rounded) square root instead of keeping up to 120 of them.
// Overwrites the descriptors, the texture&valid mask and the mask byte of gm
// that census_kernel wrote from the gray raster.
struct ColorJob {
  const uint8_t *bgr;       // interleaved B, G, R
  const uint8_t *gray;      // gray raster of the same image (gm plane)
  const uint8_t *valid;     // optional validity mask (0 = invalid)
  uint64_t *desc;
  uint8_t *mask;
  uint16_t *gm;
  int w, h;
};

struct ColorJobs {
  ColorJob job[2];
};

template <int WORDS>
__global__ __launch_bounds__(256) void color_census_kernel(ColorJobs jobs, int ww, int wh) {
  const ColorJob J = jobs.job[blockIdx.z];
  const int rx = ww >> 1, ry = wh >> 1;
  const int x0 = blockIdx.x * kCensusTx, y0 = blockIdx.y * kCensusTy;
  if (x0 >= J.w || y0 >= J.h) return;
  __shared__ uint8_t tile[kCensusTy + 10][(kCensusTx + 10) * 3 + 2];
  const int tw = kCensusTx + 2 * rx, th = kCensusTy + 2 * ry;
  const int tid = threadIdx.y * kCensusTx + threadIdx.x;
  for (int i = tid; i < tw * th * 3; i += 256) {
    const int ty = i / (tw * 3), r = i - ty * tw * 3;
    const int gx = x0 - rx + r / 3, gy = y0 - ry + ty;
    tile[ty][r] = (gx >= 0 && gx < J.w && gy >= 0 && gy < J.h)
                      ? J.bgr[((size_t)gy * J.w + gx) * 3 + r % 3]
                      : 0;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
  if (x >= J.w || y >= J.h) return;
  const size_t n = (size_t)J.w * J.h, p = (size_t)y * J.w + x;
  const bool interior = x >= rx && x < J.w - rx && y >= ry && y < J.h - ry;
  uint64_t d0 = 0, d1 = 0;
  bool textured = false;
  if (interior) {
    const uint8_t *c = &tile[threadIdx.y + ry][(threadIdx.x + rx) * 3];
    const int cb = c[0], cg = c[1], cr = c[2];
    auto distance = [&](int yy, int xx) {
      const uint8_t *q = &tile[threadIdx.y + yy][(threadIdx.x + xx) * 3];
      const int db = q[0] - cb, dg = q[1] - cg, dr = q[2] - cr;
      return sqrt((double)(db * db + dg * dg + dr * dr));
    };
    double total = 0, maximum = 0;
    for (int yy = 0; yy < wh; yy++)
      for (int xx = 0; xx < ww; xx++) {
        if (yy == ry && xx == rx) continue;
        const double d = distance(yy, xx);
        total += d;
        maximum = d > maximum ? d : maximum;
      }
    const double mean = total / (double)(ww * wh);
    textured = maximum >= 4;
    int bit = 0;
    for (int yy = 0; yy < wh; yy++)
      for (int xx = 0; xx < ww; xx++) {
        if (yy == ry && xx == rx) continue;
        const uint64_t m = (uint64_t)(distance(yy, xx) >= mean) << (bit & 63);
        if (WORDS == 1 || bit < 64) d0 |= m;
        else d1 |= m;
        bit++;
      }
  }
  const uint8_t m = interior && textured && (J.valid == nullptr || J.valid[p] != 0);
  J.desc[p] = WORDS == 1 ? (d0 | ((uint64_t)m << 63)) : d0;
  if (WORDS > 1) J.desc[n + p] = d1;
  J.mask[p] = m;
  if (J.gm) J.gm[p] = (uint16_t)(J.gray[p] | (m << 8));
}

// Caller-supplied one-word descriptors (census bits <= 62): bit 63 carries the
// texture&valid flag the sweeps test, as census_kernel writes it.
static __global__ __launch_bounds__(256) void desc_flag_kernel(uint64_t *desc, const uint8_t *mask,
                                                               size_t n) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  desc[p] = (desc[p] & 0x7FFFFFFFFFFFFFFFull) | ((uint64_t)(mask[p] != 0) << 63);
}

// ---------------------------------------------------------------------------
// Cost volume (StereoCost.value, StereoCost.java:50-90; validity test of the
// reference's `costs` kernel, GpuStereo.java:706-731).  One thread = 4
// consecutive disparities of one pixel -> one 8-byte coalesced store.
struct CostJob {
  const uint64_t *desc_a, *desc_b;
  const uint8_t *mask_a, *mask_b;
  const uint8_t *gray_a, *gray_b;
  const int32_t *sum_a, *sum2_a, *sum_b, *sum2_b;   // ZNCC window sums
  uint16_t *cost;
  int w, h, D, Dp, minD, reverse;
  int y0;  // first image row held by `cost` (strips): cost index ((y - y0) * w + x) * Dp + k
};

struct CostArgs {
  int mode, ww, wh, words, weight, adcap, gradcap, inv_code;
  int u8;  // uint8 cost volume (census-family costs <= 120, 255 = invalid 512)
};

// Stores the costs of disparities k0 .. k0+3 of pixel p.
__device__ __forceinline__ void store_cost4(uint16_t *cost, size_t p, int Dp, int k0,
                                            const unsigned *v, bool u8) {
  if (u8) {
    const unsigned b = min(v[0], 255u) | (min(v[1], 255u) << 8) | (min(v[2], 255u) << 16) |
                       (min(v[3], 255u) << 24);
    *reinterpret_cast<uint32_t *>((uint8_t *)cost + p * Dp + k0) = b;
  } else {
    uint2 packed;
    packed.x = v[0] | (v[1] << 16);
    packed.y = v[2] | (v[3] << 16);
    *reinterpret_cast<uint2 *>(cost + p * Dp + k0) = packed;
  }
}

__device__ __forceinline__ int zncc_cost(const CostArgs &A, int sab, long long sa, long long saa,
                                         long long sb, long long sbb);

// FP32 ZNCC pieces: the inverse deviation 1 / sqrt(cnt*saa - sa^2) of a
// window (-1 when the window has no variance), and the cost of a cell from
// the exact covariance numerator and both inverse deviations (invalid code
// when either is -1).
__device__ __forceinline__ float zncc_inv_sd(long long v) {
  return v > 0 ? 1.0f / sqrtf((float)v) : -1.0f;
}
__device__ __forceinline__ int zncc1024(int num, float ia, float ib, int inv_code) {
  // evaluated for every cell, then selected: no branch in the cell loop
  float ncc = ((float)num * ia) * ib;
  ncc = fminf(1.0f, fmaxf(-1.0f, ncc));
  const int v = round_half_up_f((1.0f - ncc) * 1024.0f);
  return (ia > 0.0f && ib > 0.0f) ? v : inv_code;
}

// ZNCC = false compiles only the census-family modes 0..2 (register budget of
// the batched kernel instance that never sees ZNCC).
template <bool ZNCC = true>
__device__ __forceinline__ int cost_value(const CostJob &J, const CostArgs &A, size_t p, size_t q,
                                          size_t n) {
  const int bits = A.ww * A.wh - 1;
  if (!ZNCC || A.mode <= 2) {
    int ham = __popcll(J.desc_a[p] ^ J.desc_b[q]);
    if (A.words > 1) ham += __popcll(J.desc_a[n + p] ^ J.desc_b[n + q]);
    if (A.mode == 0) return ham;
    int difference, cap;
    if (A.mode == 1) {
      difference = abs((int)J.gray_a[p] - (int)J.gray_b[q]);
      cap = A.adcap;
    } else {
      const int w = J.w;
      int ax = (int)J.gray_a[p + 1] - (int)J.gray_a[p - 1];
      int ay = (int)J.gray_a[p + w] - (int)J.gray_a[p - w];
      int bx = (int)J.gray_b[q + 1] - (int)J.gray_b[q - 1];
      int by = (int)J.gray_b[q + w] - (int)J.gray_b[q - w];
      difference = abs(ax - bx) + abs(ay - by);
      cap = A.gradcap;
    }
    const int den = 1000 * cap;
    return (A.weight * ham * cap + (1000 - A.weight) * min(difference, cap) * bits + den / 2) / den;
  }
  if constexpr (!ZNCC) return A.inv_code;
  // ZNCC: exact integer window sums.
  const int rx = A.ww >> 1, ry = A.wh >> 1, w = J.w;
  int sab = 0;
  for (int dy = -ry; dy <= ry; dy++) {
    const uint8_t *ra = J.gray_a + (ptrdiff_t)p + (ptrdiff_t)dy * w;
    const uint8_t *rb = J.gray_b + (ptrdiff_t)q + (ptrdiff_t)dy * w;
    for (int dx = -rx; dx <= rx; dx++) sab += (int)ra[dx] * (int)rb[dx];
  }
  return zncc_cost(A, sab, J.sum_a[p], J.sum2_a[p], J.sum_b[q], J.sum2_b[q]);
}

// ZNCC cost from the exact integer window sums of the reference window (sa,
// saa), the target window (sb, sbb) and their cross products (sab):
// StereoCost.value mode 3 (StereoCost.java:75-89) or the FP32 variant.
__device__ __forceinline__ int zncc_cost(const CostArgs &A, int sab, long long sa, long long saa,
                                         long long sb, long long sbb) {
  const int bits = A.ww * A.wh - 1;
  const long long cnt = (long long)A.ww * A.wh;
  const long long va = cnt * saa - sa * sa, vb = cnt * sbb - sb * sb;
  if (A.mode == 3) {
    if (va <= 0 || vb <= 0) return 512;
    double corr = (double)(cnt * sab - sa * sb) / sqrt((double)va * (double)vb);
    double c = corr < 1.0 ? corr : 1.0;
    c = c > -1.0 ? c : -1.0;
    return (int)round_half_up_d((1.0 - c) * .5 * bits);
  }
  // FP32 ZNCC: the covariance numerator cnt*sab - sa*sb is exact (int32),
  // normalised in FP32 by the inverse window deviations; cost =
  // round((1 - ncc) * 1024).
  return zncc1024((int)(cnt * sab - sa * sb), zncc_inv_sd(va), zncc_inv_sd(vb), A.inv_code);
}

struct CostJobs {
  CostJob job[2];
};

// ZNCC cost volume (modes 3 and 4) over bands of image rows.  One wave owns
// 64 consecutive disparities k of 64 output columns x0 .. x0+63 and sweeps a
// band of rows top-down.  For each window column c (image column
// x0 - RX + c) it keeps in registers the wh-row cross sum
//   col[c] = sum over the window rows r of a(r, c) * b(r, target(c, k)),
// exact integers.  A row step adds the row entering the window and removes the
// row leaving it with one v_dot2 per column: (a_new, a_old) . (b_new, -b_old).
// The target of disparity k at column c + 1 is the target of disparity k -/+ 1
// at column c, so the b values of the next column are this column's shifted
// by one lane (DPP) with one new value entering at the end; the targets' window
// sums slide along x the same way.  Each cell sums 2*RX + 1 columns
// incrementally and normalises with the window sums: the same integers as the
// per-cell window loop (StereoCost.value mode 3, StereoCost.java:75-89, or the
// FP32 variant).
constexpr int kZnccTX = 64, kZnccBand = 32;

// Lane l <- lane l-1 (REV = false) or l+1 (REV = true); the vacated end lane
// takes `fill` (uniform).
template <bool REV>
__device__ __forceinline__ int zncc_shift(int v, int fill) {
  return REV ? dpp_shl1(v, fill) : dpp_shr1(v, fill);
}
template <bool REV>
__device__ __forceinline__ float zncc_shift_f(float v, float fill) {
  return __builtin_bit_cast(float, zncc_shift<REV>(__builtin_bit_cast(int, v),
                                                   __builtin_bit_cast(int, fill)));
}
template <bool REV>
__device__ __forceinline__ double zncc_shift_d(double v, double fill) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned long long f = __builtin_bit_cast(unsigned long long, fill);
  const unsigned lo = (unsigned)zncc_shift<REV>((int)(unsigned)u, (int)(unsigned)f);
  const unsigned hi = (unsigned)zncc_shift<REV>((int)(unsigned)(u >> 32), (int)(unsigned)(f >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int zncc_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float zncc_lane_f(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double zncc_lane_d(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

typedef short zncc_s16x2 __attribute__((ext_vector_type(2)));

// Gray value of `row` at column x, 0 outside the image or for row < 0.
__device__ __forceinline__ int zncc_px(const uint8_t *img, int w, int row, int x) {
  return (row >= 0 && x >= 0 && x < w) ? (int)img[(size_t)row * w + x] : 0;
}

template <int RX, int MODE, bool REV, bool OUT8>
__device__ __forceinline__ void zncc_band(const CostJob &J, const CostArgs &A, int rows) {
  constexpr int XW = kZnccTX + 2 * RX;  // window columns of one tile
  const int ntx = (J.w + kZnccTX - 1) / kZnccTX;
  const int x0 = (int)(blockIdx.x % ntx) * kZnccTX;
  const int kw0 = (int)(blockIdx.x / ntx) * 64;
  const int y_end = min(J.h, J.y0 + rows);
  const int yb = J.y0 + (int)blockIdx.y * kZnccBand, ye = min(y_end, yb + kZnccBand);
  if (kw0 >= J.Dp || yb >= ye) return;  // whole wave
  const int lane = (int)threadIdx.x, k = kw0 + lane;
  const int ry = A.wh >> 1, w = J.w;
  const int cnt = A.ww * A.wh, bits = cnt - 1;
  // target of window column c and lane l: tb + j, j = c + (REV ? l : 63 - l)
  const int tb = REV ? x0 - RX + J.minD + kw0 : x0 - RX - J.minD - kw0 - 63;
  const int j0 = REV ? lane : 63 - lane;
  int col[XW];
#pragma unroll
  for (int c = 0; c < XW; c++) col[c] = 0;
  bool have = false;
  for (int y = yb; y < ye; y++) {
    const bool rows_ok = y >= ry && y < J.h - ry;
    if (rows_ok) {
      // rows entering / leaving the window: a full window after a gap, else one slide
      const int first = have ? y + ry : y - ry, count = have ? 1 : 2 * ry + 1;
      if (!have) {
#pragma unroll
        for (int c = 0; c < XW; c++) col[c] = 0;
      }
      for (int r = 0; r < count; r += 2) {
        // rows P (added) and Q (added when filling, removed when sliding; -1: none)
        const int P = first + r;
        const int Q = have ? y - ry - 1 : (r + 1 < count ? P + 1 : -1);
        const int sq = have ? -1 : 1;
        const int ax0 = x0 - RX + lane, ax1 = ax0 + 64;
        const int av0 = zncc_px(J.gray_a, w, P, ax0) | (sq * zncc_px(J.gray_a, w, Q, ax0)) << 16;
        const int av1 = zncc_px(J.gray_a, w, P, ax1) | (sq * zncc_px(J.gray_a, w, Q, ax1)) << 16;
        // b pairs of j = j0 (the first column) and of j = 64 + lane, 128 + lane
        // (the values entering at the lane end, column after column)
        auto bpair = [&](int j) {
          return (zncc_px(J.gray_b, w, P, tb + j) & 0xFFFF) |
                 (zncc_px(J.gray_b, w, Q, tb + j) << 16);
        };
        int B = bpair(j0);
        const int bv1 = bpair(64 + lane), bv2 = bpair(128 + lane);
#pragma unroll
        for (int c = 0; c < XW; c++) {
          const int apk = c < 64 ? zncc_lane(av0, c) : zncc_lane(av1, c - 64);
          col[c] = __builtin_amdgcn_sdot2(__builtin_bit_cast(zncc_s16x2, apk),
                                          __builtin_bit_cast(zncc_s16x2, B), col[c], false);
          if (c + 1 < XW) {
            const int ins = c < 64 ? zncc_lane(bv1, c) : zncc_lane(bv2, c - 64);
            B = zncc_shift<REV>(B, ins);
          }
        }
      }
      have = true;
    } else {
      have = false;
    }
    // window sums of the targets of columns c = RX + xo (x = x0 + xo): the lane's
    // first target, and the targets entering at the lane end
    auto target_ok = [&](int t) {
      return rows_ok && t >= (A.ww >> 1) && t < w - (A.ww >> 1) && J.mask_b[(size_t)y * w + t];
    };
    // first targets (column RX): j = RX + j0; entering one per column:
    // j = RX + 64 + xo, held by lane xo
    const int jt0 = RX + j0, jt1 = RX + 64 + lane;
    int tsb, sv1;
    float tib = -1.0f, iv1 = -1.0f;
    double tvb = -1.0, vv1 = -1.0;
    auto tdata = [&](int j, int &sb, float &ib, double &vb) {
      const int t = tb + j;
      sb = 0;
      if (target_ok(t)) {
        const size_t q = (size_t)y * w + t;
        sb = J.sum_b[q];
        const long long v = (long long)cnt * J.sum2_b[q] - (long long)sb * sb;
        if (MODE == 4) ib = zncc_inv_sd(v);
        else vb = v > 0 ? (double)v : -1.0;
      }
    };
    tdata(jt0, tsb, tib, tvb);
    tdata(jt1, sv1, iv1, vv1);
    // reference pixel x0 + lane: window sum and deviation, -1 when not valid
    int psa = 0;
    float pia = -1.0f;
    double pva = -1.0;
    {
      const int x = x0 + lane;
      const int rxw = A.ww >> 1;
      if (rows_ok && x >= rxw && x < w - rxw && J.mask_a[(size_t)y * w + x]) {
        const size_t p = (size_t)y * w + x;
        psa = J.sum_a[p];
        const long long v = (long long)cnt * J.sum2_a[p] - (long long)psa * psa;
        if (MODE == 4) pia = zncc_inv_sd(v);
        else pva = v > 0 ? (double)v : -1.0;
      }
    }
    const bool kin = k < J.D, kst = k < J.Dp;
    uint8_t *out = (uint8_t *)J.cost + (((size_t)(y - J.y0) * w + x0) * J.Dp + k) * (OUT8 ? 1 : 2);
    const size_t ostep = (size_t)J.Dp * (OUT8 ? 1 : 2);
    int sab = 0;
#pragma unroll
    for (int c = 0; c <= 2 * RX; c++) sab += col[c];
#pragma unroll
    for (int xo = 0; xo < kZnccTX; xo++) {
      if (xo > 0) sab += col[xo + 2 * RX] - col[xo - 1];
      if (x0 + xo >= w) break;  // uniform
      const int sa = zncc_lane(psa, xo);
      // exact: every operand fits 24 bits (sab < 121 * 255^2 < 2^23)
      const int num = __mul24(cnt, sab) - __mul24(sa, tsb);
      int value;
      if (MODE == 4) {
        value = zncc1024(num, zncc_lane_f(pia, xo), tib, A.inv_code);
      } else {
        // the reference's FP64 normalisation, evaluated for every cell and
        // selected (no branch): StereoCost.java:82-88
        const double va = zncc_lane_d(pva, xo);
        const bool ok = va > 0.0 && tvb > 0.0;
        const double corr = (double)num / sqrt(ok ? va * tvb : 1.0);
        double c = corr < 1.0 ? corr : 1.0;
        c = c > -1.0 ? c : -1.0;
        value = ok ? (int)round_half_up_d((1.0 - c) * .5 * bits) : 512;
      }
      value = kin ? value : (MODE == 4 ? A.inv_code : 512);
      if (kst) {
        if (OUT8) out[(size_t)xo * ostep] = (uint8_t)min(value, 255);
        else *(uint16_t *)(out + (size_t)xo * ostep) = (uint16_t)value;
      }
      if (xo + 1 < kZnccTX) {
        // slide the targets by one column: j = RX + 64 + xo enters at the lane end
        tsb = zncc_shift<REV>(tsb, zncc_lane(sv1, xo));
        if (MODE == 4) tib = zncc_shift_f<REV>(tib, zncc_lane_f(iv1, xo));
        else tvb = zncc_shift_d<REV>(tvb, zncc_lane_d(vv1, xo));
      }
    }
  }
}

// Both matching directions in one launch (blockIdx.z picks the job); the
// direction of the target side is fixed per job, so each block takes one
// branch at entry and the cell loop stays branch-free.
template <int RX, int MODE, bool OUT8>
static __global__ __launch_bounds__(64) void zncc_band_kernel(CostJobs jobs, CostArgs A, int rows) {
  const CostJob &J = jobs.job[blockIdx.z];
  if (J.reverse) zncc_band<RX, MODE, true, OUT8>(J, A, rows);
  else zncc_band<RX, MODE, false, OUT8>(J, A, rows);
}

static __global__ __launch_bounds__(256) void cost_kernel(CostJobs jobs, CostArgs A) {
  const CostJob J = jobs.job[blockIdx.z];
  const int y = blockIdx.y + J.y0;
  if (y >= J.h) return;
  const int quads = J.Dp >> 2;
  const int qi = blockIdx.x * 256 + threadIdx.x;
  if (qi >= J.w * quads) return;
  const int x = qi / quads, k0 = (qi - x * quads) * 4;
  const int rx = A.ww >> 1, ry = A.wh >> 1;
  const size_t n = (size_t)J.w * J.h, p = (size_t)y * J.w + x;
  const bool pix_ok = x >= rx && x < J.w - rx && y >= ry && y < J.h - ry && J.mask_a[p];
  unsigned v[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int k = k0 + j;
    int value = A.inv_code;
    if (pix_ok && k < J.D) {
      const int disp = J.minD + k;
      const long long t = (long long)x + (J.reverse ? (long long)disp : -(long long)disp);
      if (t >= rx && t < J.w - rx) {
        const size_t q = (size_t)y * J.w + (size_t)t;
        if (J.mask_b[q]) value = cost_value(J, A, p, q, n);
      }
    }
    v[j] = (unsigned)value & 0xFFFFu;
  }
  store_cost4(J.cost, p - (size_t)J.y0 * J.w, J.Dp, k0, v, A.u8 != 0);
}

}  // namespace sgm

#include "sgm_paths.h"

namespace sgm {

// ---------------------------------------------------------------------------
// Left-right consistency check (CensusSgm.java:181-201) and the selection
// counters of CensusSgm.Stats (CensusSgm.java:683-694), counted from the
// per-pixel selection codes of the last sweep.  Grid-stride; one global atomic
// per counter per block.
struct LrArgs {
  const float *dl, *dr;            // dr == nullptr: count codes only (single direction)
  const uint8_t *code;             // selection codes of the reference direction
  const uint8_t *reasons_in, *boundary_in;
  float *checked;
  uint8_t *reasons_out, *boundary_out;
  unsigned long long *stats;       // sgm_stats counters
  int w, h, rx, ry;
  double threshold;
};

static __global__ __launch_bounds__(256) void lr_kernel(LrArgs A) {
  __shared__ unsigned int cnt[8];
  if (threadIdx.x < 8) cnt[threadIdx.x] = 0;
  __syncthreads();
  const size_t n = (size_t)A.w * A.h;
  // local counters: lower, full, boundary, upper, ambiguous, selected, lr, valid
  unsigned c_lower = 0, c_full = 0, c_bnd = 0, c_upper = 0, c_amb = 0, c_sel = 0, c_lr = 0,
           c_valid = 0;
  for (size_t p = (size_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (size_t)gridDim.x * 256) {
    if (A.code) {
      const int code = A.code[p];
      c_lower += (code & CODE_LOWER) ? 1 : 0;
      c_full += (code & CODE_FULL) ? 1 : 0;
      c_upper += (code & CODE_UPPER) ? 1 : 0;
      c_bnd += (code & 3) == CODE_BOUNDARY ? 1 : 0;
      c_amb += (code & 3) == CODE_AMBIGUOUS ? 1 : 0;
      c_sel += (code & 3) == CODE_SELECTED ? 1 : 0;
    }
    if (!A.dr) continue;
    const int x = (int)(p % A.w), y = (int)(p / A.w);
    float out = __builtin_nanf("");
    int reason = A.reasons_in ? A.reasons_in[p] : 1;
    const float d = A.dl[p];
    if (x >= A.rx && x < A.w - A.rx && y >= A.ry && y < A.h - A.ry && isfinite(d)) {
      const int xr = round_half_up_f((float)x - d);
      if (xr < A.rx || xr >= A.w - A.rx) {
        reason = 4;
      } else {
        const float other = A.dr[(size_t)y * A.w + xr];
        if (isfinite(other) && (double)fabsf(d - other) <= A.threshold) {
          out = d;
        } else {
          c_lr++;
          reason = 4;
        }
      }
    }
    c_valid += isfinite(out) ? 1 : 0;
    A.checked[p] = out;
    if (A.reasons_out) A.reasons_out[p] = (uint8_t)reason;
    if (A.boundary_out) A.boundary_out[p] = A.boundary_in ? A.boundary_in[p] : 0;
  }
  if (A.stats) {
    const unsigned vals[8] = {c_lower, c_full, c_bnd, c_upper, c_amb, c_sel, c_lr, c_valid};
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (vals[i]) atomicAdd(&cnt[i], vals[i]);
    __syncthreads();
    // sgm_stats field index of each counter
    const int field[8] = {0, 1, 2, 3, 4, 5, 6, 10};
    if (threadIdx.x < 8 && cnt[threadIdx.x])
      atomicAdd(A.stats + field[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
  }
}

// Fill helpers (used instead of hipMemsetAsync so they are safe in graphs).
static __global__ void fill_u8_kernel(uint8_t *p, uint8_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
static __global__ void fill_u64_kernel(unsigned long long *p, unsigned long long v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Plain 16-byte-vector device copy (n vectors), grid-stride: the HBM
// calibration kernel (reads n*16 B, writes n*16 B).
static __global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                          uint4 *__restrict__ dst, size_t n) {
  // four independent 16-byte loads in flight per thread and iteration
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace sgm
