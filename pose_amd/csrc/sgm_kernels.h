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
  J.desc[p] = d0;
  if (WORDS > 1) J.desc[n + p] = d1;
  uint8_t m;
  if (J.mask_in) m = J.mask_in[p] != 0;
  else m = interior && (hi - lo) >= 4 && (J.valid == nullptr || J.valid[p] != 0);
  J.mask[p] = m;
  if (J.wsum) {
    J.wsum[p] = s1;
    J.wsum2[p] = s2;
  }
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
};

struct CostArgs {
  int mode, ww, wh, words, weight, adcap, gradcap, inv_code;
};

__device__ __forceinline__ int cost_value(const CostJob &J, const CostArgs &A, size_t p, size_t q,
                                          size_t n) {
  const int bits = A.ww * A.wh - 1;
  if (A.mode <= 2) {
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
  // ZNCC: exact integer window sums.
  const int rx = A.ww >> 1, ry = A.wh >> 1, w = J.w;
  int sab = 0;
  for (int dy = -ry; dy <= ry; dy++) {
    const uint8_t *ra = J.gray_a + (ptrdiff_t)p + (ptrdiff_t)dy * w;
    const uint8_t *rb = J.gray_b + (ptrdiff_t)q + (ptrdiff_t)dy * w;
    for (int dx = -rx; dx <= rx; dx++) sab += (int)ra[dx] * (int)rb[dx];
  }
  const long long cnt = (long long)A.ww * A.wh;
  const long long sa = J.sum_a[p], sb = J.sum_b[q], saa = J.sum2_a[p], sbb = J.sum2_b[q];
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

static __global__ __launch_bounds__(256) void cost_kernel(CostJobs jobs, CostArgs A) {
  const CostJob J = jobs.job[blockIdx.z];
  const int y = blockIdx.y;
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
  uint2 packed;
  packed.x = v[0] | (v[1] << 16);
  packed.y = v[2] | (v[3] << 16);
  *reinterpret_cast<uint2 *>(J.cost + p * J.Dp + k0) = packed;
}

// ---------------------------------------------------------------------------
// Path sweeps (Hirschmuller recurrence with edge-adaptive P2):
// CPU CensusSgm.java:421-459, GPU `paths` GpuStereo.java:732-757.
struct PathJob {
  const uint16_t *cost;
  void *sum;
  const uint8_t *mask;     // texture & valid of the reference image
  const uint8_t *gray;     // reference image
  float *disp;             // last phase: selected disparity per pixel
  uint8_t *reasons;        // last phase, optional
  uint8_t *boundary;       // last phase, optional
  unsigned long long *stats;  // last phase, optional: sgm_stats counters
  int w, h, D, Dp, minD, rx, ry;
};

struct PathArgs {
  int dx, dy, cls;         // cls: 0 horizontal, 1 vertical, 2 diagonal
  int p1, p2, edge;
  int inv_code, val_mask;
  int keep_sum;            // last phase: also store the final aggregate
  int njobs, total_lines;
  int split;               // two by-value jobs: first global line of job 1
  double uniq;
};

enum { PHASE_FIRST = 0, PHASE_MID = 1, PHASE_LAST = 2 };

template <int R>
struct CostVec {
  static constexpr int kWords = (R + 1) / 2;
  uint32_t w[kWords];
  __device__ __forceinline__ void load(const uint16_t *ptr) {
    if constexpr (R == 1) {
      w[0] = *ptr;
    } else if constexpr (R == 2) {
      w[0] = *reinterpret_cast<const uint32_t *>(ptr);
    } else if constexpr (R == 4) {
      uint2 t = *reinterpret_cast<const uint2 *>(ptr);
      w[0] = t.x;
      w[1] = t.y;
    } else {
#pragma unroll
      for (int i = 0; i < R / 8; i++) {
        uint4 t = reinterpret_cast<const uint4 *>(ptr)[i];
        w[4 * i] = t.x;
        w[4 * i + 1] = t.y;
        w[4 * i + 2] = t.z;
        w[4 * i + 3] = t.w;
      }
    }
  }
  __device__ __forceinline__ int get(int j) const {
    return (int)((w[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu);
  }
};

template <int R, typename SumT>
struct SumVec;

template <int R>
struct SumVec<R, uint16_t> {
  CostVec<R> v;
  __device__ __forceinline__ void load(const void *ptr) { v.load((const uint16_t *)ptr); }
  __device__ __forceinline__ int get(int j) const { return v.get(j); }
  __device__ __forceinline__ static void store(void *ptr, const int *vals) {
    uint16_t *p = (uint16_t *)ptr;
    if constexpr (R == 1) {
      *p = (uint16_t)vals[0];
    } else if constexpr (R == 2) {
      *reinterpret_cast<uint32_t *>(p) = ((uint32_t)vals[0] & 0xFFFFu) | ((uint32_t)vals[1] << 16);
    } else if constexpr (R == 4) {
      uint2 t;
      t.x = ((uint32_t)vals[0] & 0xFFFFu) | ((uint32_t)vals[1] << 16);
      t.y = ((uint32_t)vals[2] & 0xFFFFu) | ((uint32_t)vals[3] << 16);
      *reinterpret_cast<uint2 *>(p) = t;
    } else {
#pragma unroll
      for (int i = 0; i < R / 8; i++) {
        uint4 t;
        t.x = ((uint32_t)vals[8 * i] & 0xFFFFu) | ((uint32_t)vals[8 * i + 1] << 16);
        t.y = ((uint32_t)vals[8 * i + 2] & 0xFFFFu) | ((uint32_t)vals[8 * i + 3] << 16);
        t.z = ((uint32_t)vals[8 * i + 4] & 0xFFFFu) | ((uint32_t)vals[8 * i + 5] << 16);
        t.w = ((uint32_t)vals[8 * i + 6] & 0xFFFFu) | ((uint32_t)vals[8 * i + 7] << 16);
        reinterpret_cast<uint4 *>(p)[i] = t;
      }
    }
  }
  __device__ __forceinline__ static int wrap(int v) { return v & 0xFFFF; }
};

template <int R>
struct SumVec<R, int32_t> {
  int w[R];
  __device__ __forceinline__ void load(const void *ptr) {
    const int32_t *p = (const int32_t *)ptr;
    if constexpr (R == 1) {
      w[0] = p[0];
    } else if constexpr (R == 2) {
      int2 t = *reinterpret_cast<const int2 *>(p);
      w[0] = t.x;
      w[1] = t.y;
    } else {
#pragma unroll
      for (int i = 0; i < R / 4; i++) {
        int4 t = reinterpret_cast<const int4 *>(p)[i];
        w[4 * i] = t.x;
        w[4 * i + 1] = t.y;
        w[4 * i + 2] = t.z;
        w[4 * i + 3] = t.w;
      }
    }
  }
  __device__ __forceinline__ int get(int j) const { return w[j]; }
  __device__ __forceinline__ static void store(void *ptr, const int *vals) {
    int32_t *p = (int32_t *)ptr;
    if constexpr (R == 1) {
      p[0] = vals[0];
    } else if constexpr (R == 2) {
      *reinterpret_cast<int2 *>(p) = make_int2(vals[0], vals[1]);
    } else {
#pragma unroll
      for (int i = 0; i < R / 4; i++)
        reinterpret_cast<int4 *>(p)[i] =
            make_int4(vals[4 * i], vals[4 * i + 1], vals[4 * i + 2], vals[4 * i + 3]);
    }
  }
  __device__ __forceinline__ static int wrap(int v) { return v; }
};

// Value of element k (uniform) of a lane-distributed vector: the owning lane
// selects its element by comparing the absolute index, then one readlane.
// (No register-array indexing, which would spill the array to scratch.)
template <int R>
__device__ __forceinline__ int lane_pick(const int *a, int kbase, int k) {
  int v = 0;
#pragma unroll
  for (int j = 0; j < R; j++) v |= (kbase + j == k) ? a[j] : 0;
  return __builtin_amdgcn_readlane(v, k / R);
}

// Geometry of line `line` of direction (dx, dy) in a w x h image
// (line enumeration of GpuStereo.java:734-738, image-relative).
__device__ __forceinline__ void line_start(int line, int dx, int dy, int w, int h, int &x, int &y,
                                           int &len) {
  if (dy == 0) {
    y = line;
    x = dx > 0 ? 0 : w - 1;
    len = w;
  } else if (dx == 0) {
    x = line;
    y = dy > 0 ? 0 : h - 1;
    len = h;
  } else {
    if (line < w) {
      x = line;
      y = dy > 0 ? 0 : h - 1;
    } else {
      x = dx > 0 ? 0 : w - 1;
      y = dy > 0 ? line - w + 1 : h - 2 - (line - w);
    }
    const int sx = dx > 0 ? w - x : x + 1;
    const int sy = dy > 0 ? h - y : y + 1;
    len = min(sx, sy);
  }
}

template <int R>
constexpr int prefetch_depth() {
  return R <= 1 ? 8 : R <= 4 ? 4 : R <= 16 ? 2 : 1;
}

// Sweep one line. All 64 lanes of the wave execute this in lockstep.
template <int R, typename SumT, int PHASE>
__device__ __forceinline__ void sweep_line(const PathJob &J, const PathArgs &A, int line,
                                           const int *p2_table) {
  constexpr int PF = prefetch_depth<R>();
  const int lane = threadIdx.x & 63;
  const int kbase = lane * R;
  const bool lane_mem = kbase < J.D;
  int x, y, len;
  line_start(line, A.dx, A.dy, J.w, J.h, x, y, len);
  const ptrdiff_t step_off = (ptrdiff_t)A.dy * J.w + A.dx;
  const ptrdiff_t p0 = (ptrdiff_t)y * J.w + x;
  const uint16_t *cost = J.cost;
  SumT *sum = (SumT *)J.sum;
  const int Dp = J.Dp;

  CostVec<R> cb[PF];
  SumVec<R, SumT> sb[PF];
#pragma unroll
  for (int j = 0; j < PF; j++) {
    cb[j] = CostVec<R>{};
    sb[j] = SumVec<R, SumT>{};
  }
#pragma unroll
  for (int j = 0; j < PF; j++) {
    if (j < len && lane_mem) {
      const ptrdiff_t pp = p0 + (ptrdiff_t)j * step_off;
      cb[j].load(cost + pp * Dp + kbase);
      if (PHASE != PHASE_FIRST) sb[j].load(sum + pp * Dp + kbase);
    }
  }

  int prev[R];
#pragma unroll
  for (int j = 0; j < R; j++) prev[j] = kInf;
  int minp = kInf;
  int g_prev = 0, m_prev = 0;
  // Reference gray value and mask along the line, 64 steps per register,
  // loaded one chunk ahead of use.
  int gchunk = 0, mchunk = 0, gnext = 0, mnext = 0;
  if (lane < len) {
    const ptrdiff_t pl = p0 + (ptrdiff_t)lane * step_off;
    gnext = J.gray[pl];
    mnext = J.mask[pl];
  }
  // stats (last phase)
  int st_full = 0, st_bnd = 0, st_upper = 0, st_lower = 0, st_amb = 0, st_sel = 0;

  for (int s0 = 0; s0 < len; s0 += PF) {
#pragma unroll
    for (int jj = 0; jj < PF; jj++) {
      const int s = s0 + jj;
      if (s >= len) break;
      const ptrdiff_t pp = p0 + (ptrdiff_t)s * step_off;
      if ((s & 63) == 0) {
        gchunk = gnext;
        mchunk = mnext;
        const int si = s + 64 + lane;
        if (si < len) {
          const ptrdiff_t pl = p0 + (ptrdiff_t)si * step_off;
          gnext = J.gray[pl];
          mnext = J.mask[pl];
        }
      }
      const int g_cur = __builtin_amdgcn_readlane(gchunk, s & 63);
      const int m_cur = __builtin_amdgcn_readlane(mchunk, s & 63);
      const bool has = s > 0 && m_prev != 0;
      const int diff = abs(g_cur - g_prev);
      const int P2l = has ? p2_table[diff] : A.p2;
      g_prev = g_cur;
      m_prev = m_cur;

      int c[R], L[R];
#pragma unroll
      for (int j = 0; j < R; j++) c[j] = cb[jj].get(j);
      const int left_nb = dpp_shr1(prev[R - 1], kInf);
      const int right_nb = dpp_shl1(prev[0], kInf);
      int lmin = kInf;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int k = kbase + j;
        const int cv = c[j] & A.val_mask;
        int v = cv;
        if (has) {
          const int lft = j == 0 ? left_nb : prev[j - 1];
          const int rgt = j == R - 1 ? right_nb : prev[j + 1];
          int t = min(prev[j], minp + P2l);
          t = min(t, lft + A.p1);
          t = min(t, rgt + A.p1);
          v = cv + t - minp;
        }
        if (k >= J.D) v = kInf;
        L[j] = v;
        lmin = min(lmin, v);
      }
      int S[R];
      if (PHASE == PHASE_FIRST) {
#pragma unroll
        for (int j = 0; j < R; j++) S[j] = L[j];
      } else {
#pragma unroll
        for (int j = 0; j < R; j++) S[j] = SumVec<R, SumT>::wrap(sb[jj].get(j) + L[j]);
      }
      if (lane_mem && (PHASE != PHASE_LAST || A.keep_sum))
        SumVec<R, SumT>::store(sum + pp * Dp + kbase, S);
      // refill this prefetch slot
      if (s + PF < len && lane_mem) {
        const ptrdiff_t pn = pp + (ptrdiff_t)PF * step_off;
        cb[jj].load(cost + pn * Dp + kbase);
        if (PHASE != PHASE_FIRST) sb[jj].load(sum + pn * Dp + kbase);
      }
      const int newmin = wave_min(lmin);
#pragma unroll
      for (int j = 0; j < R; j++) prev[j] = L[j];
      minp = newmin;

      if constexpr (PHASE == PHASE_LAST) {
        // Winner-take-all on the final aggregate of this pixel
        // (CensusSgm.java:460-515; summary of GpuStereo.java:758-769).
        int valid[R];
#pragma unroll
        for (int j = 0; j < R; j++) valid[j] = (kbase + j < J.D) && (c[j] != A.inv_code);
        int best = -1, value = 0x7FFFFFFF;
        if constexpr (sizeof(SumT) == 2) {
          unsigned key = 0xFFFFFFFFu;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (valid[j]) key = min(key, ((unsigned)S[j] << 12) | (unsigned)(kbase + j));
          key = wave_min_u(key);
          if (key != 0xFFFFFFFFu) {
            best = (int)(key & 0xFFFu);
            value = (int)(key >> 12);
          }
        } else {
          int vm = 0x7FFFFFFF;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (valid[j]) vm = min(vm, S[j]);
          vm = wave_min(vm);
          int km = 0x7FFFFFFF;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (valid[j] && S[j] == vm) km = min(km, kbase + j);
          km = wave_min(km);
          if (km != 0x7FFFFFFF) {
            best = km;
            value = vm;
          }
        }
        const int px = x + s * A.dx, py = y + s * A.dy;
        const bool interior = px >= J.rx && px < J.w - J.rx && py >= J.ry && py < J.h - J.ry;
        float out = __builtin_nanf("");
        int reason = -1, bflag = -1;
        if (interior && best >= 0) {
          int sec = 0x7FFFFFFF;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (valid[j] && abs(kbase + j - best) > 1) sec = min(sec, S[j]);
          const int second = wave_min(sec);
          int before = -1, after = -1;
          if (best > 0) {
            const int sv = lane_pick<R>(S, kbase, best - 1);
            if (lane_pick<R>(valid, kbase, best - 1)) before = sv;
          }
          if (best + 1 < J.D) {
            const int sv = lane_pick<R>(S, kbase, best + 1);
            if (lane_pick<R>(valid, kbase, best + 1)) after = sv;
          }
          const int v0 = __builtin_amdgcn_readlane(valid[0], 0);
          const int vl = lane_pick<R>(valid, kbase, J.D - 1);
          const bool full = v0 && vl;
          if (full) {
            st_full++;
            bflag = 4;
          }
          if (best == 0 || best == J.D - 1) {
            st_bnd++;
            reason = 2;
            if (full && best == 0) {
              st_lower++;
              bflag |= 1;
            }
            if (full && best == J.D - 1) {
              st_upper++;
              bflag |= 2;
            }
          } else {
            double allowed = (double)value * A.uniq;
            if (!(allowed >= 1.0)) allowed = 1.0;
            if (second != 0x7FFFFFFF && (double)(second - value) <= allowed) {
              st_amb++;
              reason = 3;
            } else {
              double offset = 0.0;
              if (before >= 0 && after >= 0) {
                const double curvature = (double)before + (double)after - 2.0 * (double)value;
                if (curvature > 0) {
                  double t = .5 * ((double)before - (double)after) / curvature;
                  t = t < .5 ? t : .5;
                  offset = t > -.5 ? t : -.5;
                }
              }
              out = (float)((double)(J.minD + best) + offset);
              st_sel++;
              reason = 0;
            }
          }
        }
        if (lane == 0) {
          J.disp[pp] = out;
          if (J.reasons && reason >= 0) J.reasons[pp] = (uint8_t)reason;
          if (J.boundary && bflag >= 0) J.boundary[pp] = (uint8_t)bflag;
        }
      }
    }
  }
  if constexpr (PHASE == PHASE_LAST) {
    if (lane == 0 && J.stats) {
      // sgm_stats field order
      if (st_lower) atomicAdd(J.stats + 0, (unsigned long long)st_lower);
      if (st_full) atomicAdd(J.stats + 1, (unsigned long long)st_full);
      if (st_bnd) atomicAdd(J.stats + 2, (unsigned long long)st_bnd);
      if (st_upper) atomicAdd(J.stats + 3, (unsigned long long)st_upper);
      if (st_amb) atomicAdd(J.stats + 4, (unsigned long long)st_amb);
      if (st_sel) atomicAdd(J.stats + 5, (unsigned long long)st_sel);
    }
  }
}

__device__ __forceinline__ void build_p2_table(int *tab, const PathArgs &A) {
  // StereoCost.pathP2 (StereoCost.java:44-48) for every |dI| in 0..255.
  for (int d = threadIdx.x; d < 256; d += blockDim.x) {
    int v = A.p2;
    if (A.edge != 0) {
      v = (int)((long long)A.p2 * A.edge / (A.edge + d));
      v = max(A.p1 + 1, v);
    }
    tab[d] = v;
  }
}

__device__ __forceinline__ int find_job(const int *line_begin, int njobs, int gline) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (line_begin[mid] <= gline) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Job list of one launch: a device table (batches) or up to two jobs passed
// by value in the kernel arguments (the plain left+right pair).
struct PathJobs {
  const PathJob *table;     // device table (batches), or nullptr
  const int *line_begin;    // device: first global line of each table job for this direction
  PathJob local[2];
};

template <int R, typename SumT, int PHASE>
__global__ __launch_bounds__(256) void path_kernel(PathJobs jobs, PathArgs A) {
  __shared__ int p2_table[256];
  build_p2_table(p2_table, A);
  __syncthreads();
  const int waves = blockDim.x >> 6;
  const int gline =
      __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (gline >= A.total_lines) return;
  if (jobs.table) {
    const int ji = __builtin_amdgcn_readfirstlane(find_job(jobs.line_begin, A.njobs, gline));
    const PathJob J = jobs.table[ji];
    sweep_line<R, SumT, PHASE>(J, A, gline - jobs.line_begin[ji], p2_table);
  } else if (A.njobs > 1 && gline >= A.split) {
    const PathJob J = jobs.local[1];
    sweep_line<R, SumT, PHASE>(J, A, gline - A.split, p2_table);
  } else {
    const PathJob J = jobs.local[0];
    sweep_line<R, SumT, PHASE>(J, A, gline, p2_table);
  }
}

// ---------------------------------------------------------------------------
// Left-right consistency check (CensusSgm.java:181-201).
struct LrArgs {
  const float *dl, *dr;
  const uint8_t *reasons_in, *boundary_in;
  float *checked;
  uint8_t *reasons_out, *boundary_out;
  unsigned long long *stats;
  int w, h, rx, ry;
  double threshold;
};

static __global__ __launch_bounds__(256) void lr_kernel(LrArgs A) {
  const size_t n = (size_t)A.w * A.h;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  int rejected = 0, valid = 0;
  if (p < n) {
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
          rejected = 1;
          reason = 4;
        }
      }
    }
    valid = isfinite(out) ? 1 : 0;
    A.checked[p] = out;
    if (A.reasons_out) A.reasons_out[p] = (uint8_t)reason;
    if (A.boundary_out) A.boundary_out[p] = A.boundary_in ? A.boundary_in[p] : 0;
  }
  if (A.stats) {
    const unsigned long long rb = __ballot(rejected), vb = __ballot(valid);
    if ((threadIdx.x & 63) == 0) {
      const int rc = __popcll(rb), vc = __popcll(vb);
      if (rc) atomicAdd(A.stats + 6, (unsigned long long)rc);
      if (vc) atomicAdd(A.stats + 10, (unsigned long long)vc);
    }
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

}  // namespace sgm
