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
  // FP32 ZNCC: per-window mean / variance and covariance in FP32 from the
  // exact window sums; cost = round((1 - ncc) * 1024).
  if (va <= 0 || vb <= 0) return A.inv_code;
  const float inv_n = 1.0f / (float)cnt;
  const float ma = (float)sa * inv_n, mb = (float)sb * inv_n;
  const float vara = (float)saa * inv_n - ma * ma;
  const float varb = (float)sbb * inv_n - mb * mb;
  const float cov = (float)sab * inv_n - ma * mb;
  const float den = sqrtf(vara * varb);
  float ncc = den > 0.0f ? cov / den : 0.0f;
  ncc = fminf(1.0f, fmaxf(-1.0f, ncc));
  return round_half_up_f((1.0f - ncc) * 1024.0f);
}

struct CostJobs {
  CostJob job[2];
};

// ZNCC cost volume rows (modes 3 and 4): one image row, 64 columns x 64
// disparities per workgroup.  The gray rows under the window are staged in
// LDS; the cross sums are separable: for every window column x' and disparity
// d the 7-row (wh-row) column sum of a(x') b(x' -/+ d) is formed once in LDS,
// then each cell adds ww of them.  All sums are exact integers, so the result
// equals the per-cell window loop.
constexpr int kZnccTX = 64, kZnccTD = 64, kZnccMaxR = 5;

static __global__ __launch_bounds__(256) void zncc_cost_kernel(CostJobs jobs, CostArgs A) {
  __shared__ uint8_t arow[2 * kZnccMaxR + 1][kZnccTX + 2 * kZnccMaxR];
  __shared__ uint8_t brow[2 * kZnccMaxR + 1][kZnccTX + 2 * kZnccMaxR + kZnccTD];
  __shared__ int col[kZnccTX + 2 * kZnccMaxR][kZnccTD + 1];
  const CostJob J = jobs.job[blockIdx.z];
  const int y = blockIdx.y + J.y0;
  const int ntx = (J.w + kZnccTX - 1) / kZnccTX;
  const int x0 = (int)(blockIdx.x % ntx) * kZnccTX, d0 = (int)(blockIdx.x / ntx) * kZnccTD;
  if (y >= J.h || d0 >= J.Dp) return;  // whole workgroup
  const int rx = A.ww >> 1, ry = A.wh >> 1;
  const int XW = kZnccTX + 2 * rx, BW = XW + kZnccTD - 1;
  // b column j holds target t = tb0 + j
  const int tb0 = J.reverse ? x0 - rx + J.minD + d0 : x0 - rx - J.minD - d0 - (kZnccTD - 1);
  const bool rows_ok = y >= ry && y < J.h - ry;
  for (int i = threadIdx.x; i < A.wh * XW; i += 256) {
    const int r = i / XW, c = i - r * XW, xx = x0 - rx + c;
    arow[r][c] = (rows_ok && xx >= 0 && xx < J.w) ? J.gray_a[(size_t)(y - ry + r) * J.w + xx] : 0;
  }
  for (int i = threadIdx.x; i < A.wh * BW; i += 256) {
    const int r = i / BW, c = i - r * BW, t = tb0 + c;
    brow[r][c] = (rows_ok && t >= 0 && t < J.w) ? J.gray_b[(size_t)(y - ry + r) * J.w + t] : 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < XW * kZnccTD; i += 256) {
    const int c = i / kZnccTD, dd = i - c * kZnccTD;
    const int j = J.reverse ? c + dd : c - dd + kZnccTD - 1;
    int s = 0;
    for (int r = 0; r < A.wh; r++) s += (int)arow[r][c] * (int)brow[r][j];
    col[c][dd] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kZnccTX * (kZnccTD / 4); i += 256) {
    const int xo = i / (kZnccTD / 4), q4 = i - xo * (kZnccTD / 4);
    const int x = x0 + xo, k0 = d0 + 4 * q4;
    if (x >= J.w || k0 >= J.Dp) continue;
    const size_t p = (size_t)y * J.w + x;
    const bool pix_ok = rows_ok && x >= rx && x < J.w - rx && J.mask_a[p];
    unsigned v[4];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int k = k0 + jj, dd = 4 * q4 + jj;
      int value = A.inv_code;
      if (pix_ok && k < J.D) {
        const int disp = J.minD + k;
        const long long t = (long long)x + (J.reverse ? (long long)disp : -(long long)disp);
        if (t >= rx && t < J.w - rx) {
          const size_t q = (size_t)y * J.w + (size_t)t;
          if (J.mask_b[q]) {
            int sab = 0;
            for (int c = xo; c <= xo + 2 * rx; c++) sab += col[c][dd];
            value = zncc_cost(A, sab, J.sum_a[p], J.sum2_a[p], J.sum_b[q], J.sum2_b[q]);
          }
        }
      }
      v[jj] = (unsigned)value & 0xFFFFu;
    }
    store_cost4(J.cost, p - (size_t)J.y0 * J.w, J.Dp, k0, v, A.u8 != 0);
  }
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
