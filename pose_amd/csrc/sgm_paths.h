// Path sweeps of the SGM recurrence (Hirschmuller with edge-adaptive P2):
// CPU CensusSgm.java:421-459, GPU `paths` GpuStereo.java:732-757, and the
// winner-take-all of CensusSgm.java:460-515 / GpuStereo.java:631-687,758-769
// fused into the last sweep.
//
// One wave64 sweeps one scanline.  Lane l holds the R consecutive disparities
// k = l*R .. l*R+R-1 of the current pixel, so each pixel step moves one
// contiguous Dp-vector of cost and aggregate per lane-vector load.  The
// previous pixel's path costs stay in registers; k-1 / k+1 across lane edges
// come from DPP wave shifts, min over k from a DPP reduction.
//
// With 16-bit aggregates (every sum <= 65535, StereoCost.java:25-41) the
// recurrence runs on packed uint16 pairs (v_pk_min_u16 / v_pk_add_u16): all
// intermediate values stay below 2^15 because L <= maxC + P2 and
// minLp + P2 <= 2 * (maxC + P2) <= 65535 / 2.  Adding L to a packed 16-bit S
// wraps exactly like the reference's (short) store.
#pragma once

namespace sgm {

// Lane index within the wave, produced by inline asm so the compiler treats it
// as opaque: values derived from it are recomputed where used instead of being
// hoisted out of the batched kernel's work loop, where they would stay live in
// VGPRs across every inlined sweep variant (2 instructions per use).
__device__ __forceinline__ int lane_id() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}


typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

struct PathJob {
  const uint16_t *cost;     // cost volume (read by non-fused sweeps)
  void *sum;                // aggregate volume
  const uint8_t *mask;      // texture & valid of the reference image
  const uint8_t *gray;      // reference image
  const uint16_t *gm;       // gray | (texture & valid) << 8 of the reference image
  // fused first sweep (census cost computed on the fly, written to cost_out)
  const uint64_t *desc_a, *desc_b;
  const uint8_t *mask_b;
  uint16_t *cost_out;
  // last sweep outputs
  float *disp;
  uint8_t *reasons;         // optional: 0/2/3 where the reference assigns them
  uint8_t *boundary;        // optional: 4|1|2 where the reference assigns them
  uint8_t *code;            // optional: per-pixel selection code for the counters
  int w, h, D, Dp, minD, rx, ry, reverse;
  int R;                    // disparities per lane of this job's sweeps
  // last sweep over paired partial aggregates: the second partial volume,
  // added to `sum` before the winner search (null: a single aggregate)
  const void *sum2;
};

struct PathArgs {
  int dx, dy, cls;         // cls: 0 horizontal, 1 vertical, 2 diagonal
  int p1, p2, edge;
  int inv_code, val_mask;
  int keep_sum;            // last sweep: also store the final aggregate
  int njobs, total_lines;
  int split;               // two by-value jobs: first global line of job 1
  double uniq;
  // dispatch order of the lines of a by-value launch (optional): wave slot i
  // sweeps line order[i]
  const int *order;
};

// Row segment of the census cost pass: one wave per kCostSegment pixels of a
// row, so the pass runs many more waves than rows.
constexpr int kCostSegment = 128;
__host__ __device__ inline int cost_segments(int w) { return (w + kCostSegment - 1) / kCostSegment; }

// PHASE_COST: the fused census cost of a left-to-right line, written to the
// cost volume, with no path recurrence (the cost pass of the plane schedule).
enum { PHASE_FIRST = 0, PHASE_MID = 1, PHASE_LAST = 2, PHASE_COST = 3 };

// Selection code bits (one byte per pixel, counted by the L-R / count kernels).
enum { CODE_SELECTED = 1, CODE_BOUNDARY = 2, CODE_AMBIGUOUS = 3, CODE_FULL = 4, CODE_LOWER = 8,
       CODE_UPPER = 16 };

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

__device__ __forceinline__ int lines_of_dev(int cls, int w, int h) {
  return cls == 0 ? h : cls == 1 ? w : w + h - 1;
}

// Element k (uniform) of a lane-distributed vector, or -1 where invalid.
template <int R>
__device__ __forceinline__ int lane_pick_valid(const int *S, const int *valid, int kbase, int k) {
  int v = -1;
#pragma unroll
  for (int j = 0; j < R; j++) v = (kbase + j == k) ? (valid[j] ? S[j] : -1) : v;
  return __builtin_amdgcn_readlane(v, k / R);
}

// Winner summary of one pixel (best, value, second, before, after, flags):
// the reference GPU `winners` layout (GpuStereo.java:758-769).
struct Winner {
  int best, value, second, before, after, flags;
};

template <int R, bool WIDE>
__device__ __forceinline__ Winner wta_reduce(const int *S, const int *valid, int kbase, int D) {
  Winner r;
  r.best = -1;
  r.value = 0x7FFFFFFF;
  if (!WIDE) {
    unsigned key = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < R; j++)
      if (valid[j]) key = min(key, ((unsigned)S[j] << 12) | (unsigned)(kbase + j));
    key = wave_min_u(key);
    if (key != 0xFFFFFFFFu) {
      r.best = (int)(key & 0xFFFu);
      r.value = (int)(key >> 12);
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
      r.best = km;
      r.value = vm;
    }
  }
  r.second = 0x7FFFFFFF;
  r.before = -1;
  r.after = -1;
  r.flags = 0;
  if (r.best >= 0) {
    int sec = 0x7FFFFFFF;
#pragma unroll
    for (int j = 0; j < R; j++)
      if (valid[j] && abs(kbase + j - r.best) > 1) sec = min(sec, S[j]);
    r.second = wave_min(sec);
    if (r.best > 0) r.before = lane_pick_valid<R>(S, valid, kbase, r.best - 1);
    if (r.best + 1 < D) r.after = lane_pick_valid<R>(S, valid, kbase, r.best + 1);
    const int v0 = __builtin_amdgcn_readlane(valid[0], 0);
    const int vl = lane_pick_valid<R>(valid, valid, kbase, D - 1);  // valid[k] where valid
    r.flags = (v0 ? 1 : 0) | (vl > 0 ? 2 : 0);
  }
  return r;
}

// Winners of up to 64 consecutive pixels of a line, one per lane (lane i holds
// step base+i), are selected together: boundary rejection, uniqueness and the
// double-precision parabolic fit of GpuStereo.select (GpuStereo.java:644-686).
struct Deferred {
  int best, value, second, before, after, flags;
  __device__ __forceinline__ void put(const Winner &w, int slot) {
    const bool mine = (int)(lane_id()) == slot;
    best = mine ? w.best : best;
    value = mine ? w.value : value;
    second = mine ? w.second : second;
    before = mine ? w.before : before;
    after = mine ? w.after : after;
    flags = mine ? w.flags : flags;
  }
};

// Per-lane winner summary of one pixel, decoded (the fields of Winner).
struct WinnerView {
  int best, value, second, before, after, flags;
};

// Winner summaries of up to 64 pixel steps of the packed sweep, lane i holding
// step i, in three registers:
//   f0 = (best + 1) | flags << 13      (best + 1 = 0: no valid disparity)
//   f1 = value | second << 16          (second 0xFFFF: none)
//   f2 = before | after << 16          (0xFFFF: neighbour invalid or absent)
// Aggregates are at most 65532 (dirs * (512 + P2) <= 65535 with dirs 4 or 8),
// so 0xFFFF never collides with a value.
struct DeferredPk {
  uint32_t f0, f1, f2;
  __device__ __forceinline__ void put(bool mine, uint32_t a, uint32_t b, uint32_t c) {
    f0 = mine ? a : f0;
    f1 = mine ? b : f1;
    f2 = mine ? c : f2;
  }
  __device__ __forceinline__ WinnerView view() const {
    WinnerView w;
    w.best = (int)(f0 & 0x1FFFu) - 1;
    w.flags = (int)(f0 >> 13);
    w.value = (int)(f1 & 0xFFFFu);
    w.second = (f1 >> 16) == 0xFFFFu ? 0x7FFFFFFF : (int)(f1 >> 16);
    w.before = (f2 & 0xFFFFu) == 0xFFFFu ? -1 : (int)(f2 & 0xFFFFu);
    w.after = (f2 >> 16) == 0xFFFFu ? -1 : (int)(f2 >> 16);
    return w;
  }
};

// The uniform inputs of select_flush, one per lane of a single VGPR.  The
// flush runs once per 64 steps; holding its ~20 scalars in SGPRs across the
// sweep loop would exhaust the SGPR file (the compiler then spills SGPRs to
// VGPR lanes inside the hot loop), so they are parked here and read back with
// readlane at flush time.
struct FlushArgs {
  int v;
  __device__ __forceinline__ int get(int i) const { return __builtin_amdgcn_readlane(v, i); }
  template <typename T>
  __device__ __forceinline__ T *ptr(int i) const {
    return (T *)(((uint64_t)(uint32_t)get(i + 1) << 32) | (uint32_t)get(i));
  }
};

enum {
  FA_DISP = 0, FA_CODE = 2, FA_REASONS = 4, FA_BOUNDARY = 6, FA_W = 8, FA_H, FA_RX, FA_RY,
  FA_D, FA_MIND, FA_DX, FA_DY, FA_UNIQ = 16, FA_COUNT = 18
};

__device__ __forceinline__ FlushArgs make_flush_args(const PathJob &J, const PathArgs &A) {
  const uint64_t ptrs[4] = {(uint64_t)J.disp, (uint64_t)J.code, (uint64_t)J.reasons,
                            (uint64_t)J.boundary};
  const uint64_t uq = __builtin_bit_cast(uint64_t, A.uniq);
  int vals[FA_COUNT];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    vals[2 * i] = (int)(uint32_t)ptrs[i];
    vals[2 * i + 1] = (int)(uint32_t)(ptrs[i] >> 32);
  }
  vals[FA_W] = J.w;
  vals[FA_H] = J.h;
  vals[FA_RX] = J.rx;
  vals[FA_RY] = J.ry;
  vals[FA_D] = J.D;
  vals[FA_MIND] = J.minD;
  vals[FA_DX] = A.dx;
  vals[FA_DY] = A.dy;
  vals[FA_UNIQ] = (int)(uint32_t)uq;
  vals[FA_UNIQ + 1] = (int)(uint32_t)(uq >> 32);
  const int lane = lane_id();
  int v = 0;
#pragma unroll
  for (int i = 0; i < FA_COUNT; i++) v = lane == i ? vals[i] : v;
  asm volatile("" : "+v"(v));
  return FlushArgs{v};
}

__device__ __forceinline__ void select_flush(const FlushArgs &F, const WinnerView &d, int count,
                                             int base, int x0, int y0) {
  const int lane = lane_id();
  // Every lane of F is read while the whole wave is active: after the lanes
  // >= count have left, their part of F is dead to the compiler, which may
  // restore a spilled F for the remaining lanes only (seen with 64 disparities
  // per lane, where F lives in an AGPR across the sweep).
  const int Jw = F.get(FA_W), Jh = F.get(FA_H), Jrx = F.get(FA_RX), Jry = F.get(FA_RY);
  const int JD = F.get(FA_D), JminD = F.get(FA_MIND), Adx = F.get(FA_DX), Ady = F.get(FA_DY);
  const double uniq = __builtin_bit_cast(
      double, ((uint64_t)(uint32_t)F.get(FA_UNIQ + 1) << 32) | (uint32_t)F.get(FA_UNIQ));
  float *disp = F.ptr<float>(FA_DISP);
  uint8_t *codes = F.ptr<uint8_t>(FA_CODE);
  uint8_t *reasons = F.ptr<uint8_t>(FA_REASONS);
  uint8_t *boundary = F.ptr<uint8_t>(FA_BOUNDARY);
  if (lane >= count) return;
  const int s = base + lane;
  const int px = x0 + s * Adx, py = y0 + s * Ady;
  const size_t pp = (size_t)py * Jw + px;
  const bool interior = px >= Jrx && px < Jw - Jrx && py >= Jry && py < Jh - Jry;
  float out = __builtin_nanf("");
  int reason = -1, bflag = -1, code = 0;
  const int best = d.best, value = d.value;
  if (interior && best >= 0) {
    const bool full = d.flags == 3;
    if (full) {
      bflag = 4;
      code |= CODE_FULL;
    }
    if (best == 0 || best == JD - 1) {
      reason = 2;
      code |= CODE_BOUNDARY;
      if (full && best == 0) {
        bflag |= 1;
        code |= CODE_LOWER;
      }
      if (full && best == JD - 1) {
        bflag |= 2;
        code |= CODE_UPPER;
      }
    } else {
      double allowed = (double)value * uniq;
      if (!(allowed >= 1.0)) allowed = 1.0;
      if (d.second != 0x7FFFFFFF && (double)(d.second - value) <= allowed) {
        reason = 3;
        code |= CODE_AMBIGUOUS;
      } else {
        double offset = 0.0;
        if (d.before >= 0 && d.after >= 0) {
          const double curvature = (double)d.before + (double)d.after - 2.0 * (double)value;
          if (curvature > 0) {
            double t = .5 * ((double)d.before - (double)d.after) / curvature;
            t = t < .5 ? t : .5;
            offset = t > -.5 ? t : -.5;
          }
        }
        out = (float)((double)(JminD + best) + offset);
        reason = 0;
        code |= CODE_SELECTED;
      }
    }
  }
  disp[pp] = out;
  if (codes) codes[pp] = (uint8_t)code;
  if (reasons && reason >= 0) reasons[pp] = (uint8_t)reason;
  if (boundary && bflag >= 0) boundary[pp] = (uint8_t)bflag;
}

// Per-64-step chunk of a per-pixel value along the line, loaded one chunk ahead.
struct LineChunk {
  int cur = 0, next = 0;
};

// ---------------------------------------------------------------------------
// Generic sweep on 32-bit integers: any R, 16- or 32-bit aggregates.
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

template <int R>
constexpr int prefetch_depth() {
  return R <= 1 ? 8 : R <= 4 ? 4 : R <= 16 ? 2 : 1;
}

template <int R, typename SumT, int PHASE>
__device__ __forceinline__ void sweep_line(const PathJob &J, const PathArgs &A, int line,
                                           const int *p2_table) {
  constexpr int PF = prefetch_depth<R>();
  const int lane = lane_id();
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
  LineChunk gch, mch;
  if (lane < len) {
    const ptrdiff_t pl = p0 + (ptrdiff_t)lane * step_off;
    gch.next = J.gray[pl];
    mch.next = J.mask[pl];
  }
  Deferred dfr{};

  for (int s0 = 0; s0 < len; s0 += PF) {
#pragma unroll
    for (int jj = 0; jj < PF; jj++) {
      const int s = s0 + jj;
      if (s >= len) break;
      const ptrdiff_t pp = p0 + (ptrdiff_t)s * step_off;
      if ((s & 63) == 0) {
        gch.cur = gch.next;
        mch.cur = mch.next;
        const int si = s + 64 + lane;
        if (si < len) {
          const ptrdiff_t pl = p0 + (ptrdiff_t)si * step_off;
          gch.next = J.gray[pl];
          mch.next = J.mask[pl];
        }
      }
      const int g_cur = __builtin_amdgcn_readlane(gch.cur, s & 63);
      const int m_cur = __builtin_amdgcn_readlane(mch.cur, s & 63);
      const bool has = s > 0 && m_prev != 0;
      const int P2l = has ? p2_table[abs(g_cur - g_prev)] : A.p2;
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
        if (kbase + j >= J.D) v = kInf;
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
      if (kbase < Dp && (PHASE != PHASE_LAST || A.keep_sum))  // whole rows, pads included
        SumVec<R, SumT>::store(sum + pp * Dp + kbase, S);
      if (s + PF < len && lane_mem) {
        const ptrdiff_t pn = pp + (ptrdiff_t)PF * step_off;
        cb[jj].load(cost + pn * Dp + kbase);
        if (PHASE != PHASE_FIRST) sb[jj].load(sum + pn * Dp + kbase);
      }
      minp = wave_min(lmin);
#pragma unroll
      for (int j = 0; j < R; j++) prev[j] = L[j];

      if constexpr (PHASE == PHASE_LAST) {
        int valid[R];
#pragma unroll
        for (int j = 0; j < R; j++) valid[j] = (kbase + j < J.D) && (c[j] != A.inv_code);
        const Winner win = wta_reduce<R, sizeof(SumT) == 4>(S, valid, kbase, J.D);
        dfr.put(win, s & 63);
        if ((s & 63) == 63 || s == len - 1)
          select_flush(make_flush_args(J, A),
                       WinnerView{dfr.best, dfr.value, dfr.second, dfr.before, dfr.after, dfr.flags},
                       (s & 63) + 1, s & ~63, x, y);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Packed 16-bit sweep (R even), straight-line in the steady state: every load
// and store goes through a buffer resource whose range drops the accesses of
// lanes beyond D, so no branch splits the prefetch pipeline and the compiler
// can keep PF pixel steps of cost/aggregate loads in flight.
// FUSED: the census Hamming cost is computed inside the first (left-to-right)
// sweep from a sliding window of target descriptors and written to the cost
// volume for the remaining sweeps.
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as_pk(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t splat16(uint32_t s) { return (s & 0xFFFFu) | (s << 16); }
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
  return as_u32(__builtin_elementwise_min(as_pk(a), as_pk(b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
  return as_u32(as_pk(a) + as_pk(b));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
  return as_u32(as_pk(a) - as_pk(b));
}

typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

// Raw buffer resource over `bytes` bytes at `base` (gfx9 dword3).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}

// Cache policy of the cost / aggregate volume accesses (the aux field of the
// buffer instructions): POL_NT (nt, the SLC bit position) streams the volumes
// past the caches, POL_DEFAULT keeps the normal policy.  Streaming pays when a
// volume is far larger than the MALL and its rows are whole cache lines; short
// or unaligned rows and MALL-sized volumes run faster with the default policy.
enum { POL_DEFAULT = 0, POL_NT = 2 };

template <int NP, int POL = POL_DEFAULT>
__device__ __forceinline__ void bload_pairs(uint32_t *dst, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (NP == 1) {
    dst[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, POL);
  } else if constexpr (NP == 2) {
    const v2u32 t = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, POL);
    dst[0] = t.x;
    dst[1] = t.y;
  } else {
#pragma unroll
    for (int i = 0; i < NP / 4; i++) {
      const v4u32 t = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * i, 0, POL);
      dst[4 * i] = t.x;
      dst[4 * i + 1] = t.y;
      dst[4 * i + 2] = t.z;
      dst[4 * i + 3] = t.w;
    }
  }
}

template <int NP, int POL = POL_DEFAULT>
__device__ __forceinline__ void bstore_pairs(const uint32_t *src, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (NP == 1) {
    __builtin_amdgcn_raw_buffer_store_b32(src[0], r, off, 0, POL);
  } else if constexpr (NP == 2) {
    v2u32 t;
    t.x = src[0];
    t.y = src[1];
    __builtin_amdgcn_raw_buffer_store_b64(t, r, off, 0, POL);
  } else {
#pragma unroll
    for (int i = 0; i < NP / 4; i++) {
      v4u32 t;
      t.x = src[4 * i];
      t.y = src[4 * i + 1];
      t.z = src[4 * i + 2];
      t.w = src[4 * i + 3];
      __builtin_amdgcn_raw_buffer_store_b128(t, r, off + 16 * i, 0, POL);
    }
  }
}

template <int R>
constexpr int pk_prefetch_depth() {
  return R <= 4 ? 8 : R <= 8 ? 4 : R <= 16 ? 2 : 1;
}

constexpr uint32_t kInfPk = 0x7FFF7FFFu;
constexpr int kOutOfRange = 0x7FFFFF00;  // buffer offset that is always dropped

// Target descriptor with bit 63 set when the target is usable (inside
// [rx, w - rx) and textured): the validity test of the reference's `costs`
// kernel (GpuStereo.java:711-713).  Branch-free: the load index is clamped.
// rd / rm: buffer resources over the target row's descriptors / mask.
__device__ __forceinline__ uint64_t fetch_target(const PathJob &J, __amdgpu_buffer_rsrc_t rd,
                                                 __amdgpu_buffer_rsrc_t rm, int t) {
  const bool ok = t >= J.rx && t < J.w - J.rx;
  const int tc = min(max(t, 0), J.w - 1);
  const v2u32 d = __builtin_amdgcn_raw_buffer_load_b64(rd, tc * 8, 0, 0);
  const uint32_t m = __builtin_amdgcn_raw_buffer_load_b8(rm, tc, 0, 0);
  const uint64_t dd = ((uint64_t)(d.y & 0x7FFFFFFFu) << 32) | d.x;
  return (ok && m) ? (dd | 0x8000000000000000ull) : 0ull;
}

// Cost modes of the packed sweep.
//   CM_LOAD:      read the cost volume.
//   CM_FUSED:     compute the census cost (first, left-to-right sweep) and
//                 write the cost volume for the later sweeps.
//   CM_RECOMPUTE: compute the census cost in every sweep from descriptors
//                 whose bit 63 carries the texture&valid flag; no cost volume.
//                 Horizontal lines slide a DPP-shifted descriptor window,
//                 other lines load each pixel's target row segment (L2-resident).
//   CM_LOAD8:     read a uint8 cost volume (census-family costs are <= 120;
//                 255 marks an invalid cell and is widened back to 512).
// The fused sweep writes the uint8 volume that CM_LOAD8 reads.
enum { CM_LOAD = 0, CM_FUSED = 1, CM_RECOMPUTE = 2, CM_LOAD8 = 3 };

// R bytes per lane through a buffer resource (R = 2 .. 64, byte offset `off`).
template <int R, int POL = POL_DEFAULT>
__device__ __forceinline__ void bload_bytes(uint32_t *dst, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (R == 2) {
    dst[0] = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, POL);
  } else if constexpr (R == 4) {
    dst[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, POL);
  } else {
    bload_pairs<R / 4, POL>(dst, r, off);
  }
}

template <int R, int POL = POL_DEFAULT>
__device__ __forceinline__ void bstore_bytes(const uint32_t *src, __amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (R == 2) {
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)src[0], r, off, 0, POL);
  } else if constexpr (R == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(src[0], r, off, 0, POL);
  } else {
    bstore_pairs<R / 4, POL>(src, r, off);
  }
}

// uint8 costs -> packed uint16 pairs: pair i of the lane from raw bytes 2i, 2i+1;
// 255 (invalid) becomes 512.  Each half is <= 255, so +1 never carries across
// halves; t holds 0x100 in a half that was 255, t + (t >> 8) adds 257 there.
template <int NP>
__device__ __forceinline__ void widen_costs(const uint32_t *raw, uint32_t *c) {
#pragma unroll
  for (int i = 0; i < NP; i++) {
    const uint32_t v = __builtin_amdgcn_perm(0u, raw[i >> 1], (i & 1) ? 0x0C030C02u : 0x0C010C00u);
    const uint32_t t = (v + 0x00010001u) & 0x01000100u;
    c[i] = v + t + (t >> 8);
  }
}

// packed uint16 pairs (values <= 120 or 512) -> uint8 bytes, 512 -> 255.
template <int NP>
__device__ __forceinline__ void narrow_costs(const uint32_t *c, uint32_t *raw) {
  if constexpr (NP == 1) {
    raw[0] = __builtin_amdgcn_perm(0u, pk_min(c[0], 0x00FF00FFu), 0x0C0C0200u);
  } else {
#pragma unroll
    for (int i = 0; i < NP / 2; i++)
      raw[i] = __builtin_amdgcn_perm(pk_min(c[2 * i + 1], 0x00FF00FFu), pk_min(c[2 * i], 0x00FF00FFu),
                                     0x06040200u);
  }
}

// Per-step metadata of a window of PF steps, one step per lane (lane i holds
// step base + (i % PF)): the edge-adaptive P2 of the step, or -1 when the
// predecessor is outside the image or untextured (no transition term).
struct Meta {
  int p2;       // >= 0: P2 to use; -1: L = C
  int ok;       // computed cost: reference pixel interior and textured
  uint32_t da_lo, da_hi;  // computed cost: reference descriptor
  uint32_t e_lo, e_hi;    // sliding window: target descriptor entering after this step
};

// Buffer resources of one line: gray|mask of the reference, the reference
// descriptors and the target row (computed costs).
struct LineRsrc {
  __amdgpu_buffer_rsrc_t gm, da, tdesc, tmask;
};

__device__ __forceinline__ uint64_t ld64(__amdgpu_buffer_rsrc_t r, int off) {
  const v2u32 d = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return ((uint64_t)d.y << 32) | d.x;
}

// Target entering the sliding window after step s of a horizontal line.
__device__ __forceinline__ int slide_target(const PathJob &J, int x0, int dx, int s, int R) {
  const int xn = x0 + (s + 1) * dx;
  const bool up = (dx > 0) != (J.reverse != 0);
  if (up) return J.reverse ? xn + J.minD : xn - J.minD;
  return J.reverse ? xn + J.minD + 64 * R - 1 : xn - J.minD - (64 * R - 1);
}

template <int PF, int CM>
__device__ __forceinline__ void meta_load(const PathJob &J, const PathArgs &A, const LineRsrc &rs,
                                          int base, int len, ptrdiff_t p0, ptrdiff_t step_off,
                                          int x0, int &g_raw, int &gp_raw, uint64_t &da,
                                          uint64_t &e) {
  const int lane = lane_id();
  const int s = min(base + (lane % PF), len - 1);
  const int sp = max(s - 1, 0);
  const ptrdiff_t ps = p0 + (ptrdiff_t)s * step_off;
  g_raw = __builtin_amdgcn_raw_buffer_load_b16(rs.gm, (int)(ps * 2), 0, 0);
  gp_raw = __builtin_amdgcn_raw_buffer_load_b16(rs.gm, (int)((p0 + (ptrdiff_t)sp * step_off) * 2), 0, 0);
  if constexpr (CM == CM_FUSED) {
    da = ld64(rs.da, (int)(ps * 8));
    const int t = J.reverse ? x0 + s + 1 + J.minD + 64 * J.R - 1 : x0 + s + 1 - J.minD;
    e = fetch_target(J, rs.tdesc, rs.tmask, t);
  } else if constexpr (CM == CM_RECOMPUTE) {
    da = ld64(rs.da, (int)(ps * 8));
    if (A.dy == 0) e = ld64(rs.tdesc, slide_target(J, x0, A.dx, s, J.R) * 8);  // dropped if outside
  }
}

template <int PF, int CM>
__device__ __forceinline__ Meta meta_finish(const PathJob &J, const PathArgs &A, int base,
                                            int len, int x0, int y0, int g_raw, int gp_raw,
                                            uint64_t da, uint64_t e, const int *p2_table) {
  const int lane = lane_id();
  const int s = min(base + (lane % PF), len - 1);
  Meta m;
  const int g = g_raw & 0xFF, gp = gp_raw & 0xFF;
  const bool has = s > 0 && (gp_raw >> 8) != 0;
  m.p2 = has ? p2_table[abs(g - gp)] : -1;
  if constexpr (CM == CM_FUSED || CM == CM_RECOMPUTE) {
    const int x = x0 + s * A.dx, y = y0 + s * A.dy;
    m.ok = (y >= J.ry && y < J.h - J.ry && x >= J.rx && x < J.w - J.rx && (g_raw >> 8) != 0);
    m.da_lo = (uint32_t)da;
    m.da_hi = (uint32_t)(da >> 32);
    m.e_lo = (uint32_t)e;
    m.e_hi = (uint32_t)(e >> 32);
  } else {
    m.ok = 0;
    m.da_lo = m.da_hi = m.e_lo = m.e_hi = 0;
  }
  return m;
}

// Census Hamming costs of one lane's R disparities from flagged target
// descriptors (StereoCost.value mode 0 with the validity test of the
// reference `costs` kernel, GpuStereo.java:706-719).
template <int R>
__device__ __forceinline__ void census_costs(const uint64_t *tw, uint32_t alo, uint32_t ahi,
                                             bool ok, int kbase, int D, int inv, uint32_t *c) {
  int cv[R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    const uint32_t wl = (uint32_t)tw[j], wh = (uint32_t)(tw[j] >> 32);
    const int ham = __builtin_popcount(wl ^ alo) + __builtin_popcount((wh ^ ahi) & 0x7FFFFFFFu);
    cv[j] = (ok && (wh >> 31) && kbase + j < D) ? ham : inv;
  }
#pragma unroll
  for (int i = 0; i < R / 2; i++) c[i] = (uint32_t)cv[2 * i] | ((uint32_t)cv[2 * i + 1] << 16);
}

// SLIDE_UP: direction of the sliding target window of a horizontal computed-cost
// line, fixed at compile time (1: towards higher x indices of the window, 0:
// lower) so the window rotates by register renaming in the unrolled steps; -1
// where the line has no sliding window.
// PLANE: path-cost planes instead of a running aggregate.  The planes of a
// matching direction are interleaved per pixel: pixel p holds kPlanes rows of
// Dp bytes at (p * kPlanes + k) * Dp, so the last sweep reads one contiguous
// block per pixel.  A first-phase sweep stores its path costs L as bytes into
// its own plane (`sum` points at plane k of pixel 0); the last sweep adds the
// kPlanes planes of the other directions to its own L (`sum` at plane 0).
constexpr int kPlanes = 7;

template <int R, int PHASE, int CM, int PFCAP = 64, int SLIDE_UP = -1, bool NT = false,
          bool SUM2 = false, bool PLANE = false>
struct PkSweep {
  static constexpr int NP = R / 2;
  static constexpr int POL = NT ? POL_NT : POL_DEFAULT;
  static constexpr int PF = pk_prefetch_depth<R>() < PFCAP ? pk_prefetch_depth<R>() : PFCAP;
  static constexpr bool kComputed = CM == CM_FUSED || CM == CM_RECOMPUTE;
  static constexpr bool kLoaded = CM == CM_LOAD || CM == CM_LOAD8;
  static constexpr int NC = CM == CM_LOAD8 ? (R + 3) / 4 : NP;  // dwords of loaded cost per slot
  static constexpr int kDirect = CM == CM_RECOMPUTE ? PF : 1;  // direct target-row prefetch slots
  static constexpr bool kPlaneIn = PLANE && PHASE == PHASE_LAST;    // reads the planes
  static constexpr bool kPlaneOut = PLANE && PHASE == PHASE_FIRST;  // writes its plane
  static constexpr int NB = (R + 3) / 4;                             // dwords of R bytes
  static_assert(!PLANE || (PHASE != PHASE_MID && PHASE != PHASE_COST),
                "plane sweeps are first-phase or last");
  static_assert(PHASE != PHASE_COST || CM == CM_FUSED, "the cost pass computes the census cost");

  const PathJob &J;
  const PathArgs &A;
  const int *p2_table;
  int lane, kbase, loff, boff, x0, y0, len;
  int sloff, sboff;  // store offsets: every lane inside the padded row
  bool slide;  // computed cost on a horizontal line: sliding window
  ptrdiff_t p0, step_off;
  uint32_t row_bytes, cost_row_bytes;
  const uint8_t *ld_c, *ld_s;  // rows of the next prefetched step
  uint8_t *st_c, *st_s;        // rows of the next stored step
  ptrdiff_t cstep, sstep;      // bytes per pixel step in the cost / aggregate volume
  uint32_t keep[NP], infm[NP];
  uint32_t vmask, p1pk;
  uint32_t cb[kLoaded ? PF : 1][NC], sb[kPlaneIn ? 1 : PF][NP], sb2[SUM2 ? PF : 1][NP];
  uint32_t pl[kPlaneIn ? PF : 1][kPlaneIn ? kPlanes : 1][NB];
  const uint8_t *ld_s2;  // second partial aggregate (SUM2): rows of the next prefetched step
  int poff;              // plane reads: lane offset, dropped for lanes beyond D
  uint64_t win[kComputed ? R : 1];
  uint64_t dsc[kDirect][CM == CM_RECOMPUTE ? R : 1];
  uint32_t prev[NP];
  int minp;
  LineRsrc rs;
  Meta meta;
  DeferredPk dfr;
  FlushArgs flush_args;
  uint32_t kpk[PHASE == PHASE_LAST ? NP : 1], keep01[PHASE == PHASE_LAST ? NP : 1];
  uint32_t invpk;
  int last_lane, last_pair, last_half;
  int sum_store_bytes;

  __device__ __forceinline__ PkSweep(const PathJob &J_, const PathArgs &A_, int line,
                                     const int *tab)
      : J(J_), A(A_), p2_table(tab) {
    lane = lane_id();
    kbase = lane * R;
    loff = kbase < J.D ? kbase * 2 : kOutOfRange;
    boff = kbase < J.D ? kbase : kOutOfRange;  // uint8 cost volume
    // stores cover the pad cells too (saturated values that no load reads):
    // whole rows keep HBM writes to whole granules
    sloff = kbase < J.Dp ? kbase * 2 : kOutOfRange;
    sboff = kbase < J.Dp ? kbase : kOutOfRange;
    line_start(line, A.dx, A.dy, J.w, J.h, x0, y0, len);
    if constexpr (PHASE == PHASE_COST) {
      // a left-to-right row segment: line = row * segments + segment
      const int nseg = cost_segments(J.w);
      y0 = line / nseg;
      x0 = (line - y0 * nseg) * kCostSegment;
      len = min(kCostSegment, J.w - x0);
    }
    step_off = (ptrdiff_t)A.dy * J.w + A.dx;
    p0 = (ptrdiff_t)y0 * J.w + x0;
    row_bytes = (uint32_t)J.Dp * 2u;
    cost_row_bytes = (CM == CM_LOAD8 || CM == CM_FUSED) ? (uint32_t)J.Dp : row_bytes;
    {
      const ptrdiff_t cb_ = (CM == CM_LOAD8 || CM == CM_FUSED) ? 1 : 2;
      cstep = step_off * J.Dp * cb_;
      ld_c = (const uint8_t *)J.cost + p0 * J.Dp * cb_;
      st_c = (uint8_t *)J.cost_out + p0 * J.Dp * cb_;
      const ptrdiff_t sb_ = PLANE ? kPlanes : 2;  // aggregate / plane-block bytes per cell
      sstep = step_off * J.Dp * sb_;
      ld_s = (const uint8_t *)J.sum + p0 * J.Dp * sb_;
      st_s = (uint8_t *)J.sum + p0 * J.Dp * sb_;
      ld_s2 = SUM2 ? (const uint8_t *)J.sum2 + p0 * J.Dp * 2 : nullptr;
    }
    // beyond the range of every plane read (offsets stay below 2^31)
    poff = kbase < J.D ? kbase : 0x40000000;
    sum_store_bytes = kPlaneOut ? J.Dp
                      : kPlaneIn ? 0
                                 : (PHASE != PHASE_LAST || A.keep_sum) ? (int)row_bytes : 0;
    slide = CM == CM_FUSED || (CM == CM_RECOMPUTE && A.dy == 0);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const bool lo = kbase + 2 * i < J.D, hi = kbase + 2 * i + 1 < J.D;
      keep[i] = (lo ? 0xFFFFu : 0u) | (hi ? 0xFFFF0000u : 0u);
      infm[i] = (lo ? 0u : 0x7FFFu) | (hi ? 0u : 0x7FFF0000u);
      prev[i] = kInfPk;
    }
#pragma unroll
    for (int j = 0; j < (kPlaneIn ? 1 : PF); j++)
#pragma unroll
      for (int i = 0; i < NP; i++) sb[j][i] = 0;
    vmask = A.val_mask == 0xFFFF ? 0xFFFFFFFFu : splat16((uint32_t)A.val_mask);
    p1pk = splat16((uint32_t)A.p1);
    minp = 0;
    dfr = DeferredPk{0u, 0u, 0u};
    flush_args = FlushArgs{0};
    invpk = splat16((uint32_t)A.inv_code);
    last_lane = last_pair = last_half = 0;
    if constexpr (PHASE == PHASE_LAST) {
      flush_args = make_flush_args(J, A);
#pragma unroll
      for (int i = 0; i < NP; i++) {
        kpk[i] = (uint32_t)(kbase + 2 * i) | ((uint32_t)(kbase + 2 * i + 1) << 16);
        keep01[i] = keep[i] & 0x00010001u;
      }
      last_lane = (J.D - 1) / R;
      last_pair = ((J.D - 1) % R) >> 1;
      last_half = (J.D - 1) & 1;
    }
    const uint32_t n = (uint32_t)J.w * (uint32_t)J.h;
    rs.gm = make_rsrc(J.gm, n * 2u);
    if constexpr (kComputed) {
      rs.da = make_rsrc(J.desc_a, n * 8u);
      rs.tdesc = make_rsrc(J.desc_b + (size_t)y0 * J.w, (uint32_t)J.w * 8u);
      rs.tmask = make_rsrc(J.mask_b + (size_t)y0 * J.w, (uint32_t)J.w);
    } else {
      rs.da = rs.tdesc = rs.tmask = rs.gm;
    }
  }

  __device__ __forceinline__ ptrdiff_t pix(int s) const { return p0 + (ptrdiff_t)s * step_off; }

  // Target row segment of one pixel step (direct mode): R flagged descriptors.
  __device__ __forceinline__ void load_targets(uint64_t *dst, int s) {
    const int sc = min(s, len - 1);
    const int px = x0 + sc * A.dx, py = y0 + sc * A.dy;
    const __amdgpu_buffer_rsrc_t r = make_rsrc(J.desc_b + (size_t)py * J.w, (uint32_t)J.w * 8u);
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int t = J.reverse ? px + J.minD + kbase + j : px - J.minD - kbase - j;
      dst[j] = ld64(r, t * 8);  // outside the row: dropped (0 = unusable target)
    }
  }

  // Loads of pixel step s (called for s = 0, 1, 2, ... in order): the row
  // pointers advance by one step each call instead of being recomputed with
  // 64-bit multiplies; steps past the line end get an empty range (no access).
  __device__ __forceinline__ void load_step(int slot, int s) {
    const bool in = s < len;
    if constexpr (CM == CM_LOAD)
      bload_pairs<NP, POL>(cb[slot], make_rsrc(ld_c, in ? row_bytes : 0u), loff);
    if constexpr (CM == CM_LOAD8)
      // the plane writers' cost loads keep the default policy: the seven
      // directions of one launch read C seven times, and cached loads catch
      // some of the re-reads (measured 2.57 -> 2.48 ms at 1600x1067 D=256)
      bload_bytes<R, (kPlaneOut ? POL_DEFAULT : POL)>(cb[slot], make_rsrc(ld_c, in ? cost_row_bytes : 0u), boff);
    if constexpr (CM == CM_RECOMPUTE)
      if (!slide) load_targets(dsc[slot], s);
    if constexpr (kPlaneIn) {
      const __amdgpu_buffer_rsrc_t r = make_rsrc(ld_s, in ? (uint32_t)(kPlanes * J.Dp) : 0u);
#pragma unroll
      for (int k = 0; k < kPlanes; k++) bload_bytes<R, POL>(pl[slot][k], r, poff + k * J.Dp);
    } else if constexpr (PHASE == PHASE_MID || PHASE == PHASE_LAST) {
      bload_pairs<NP, POL>(sb[slot], make_rsrc(ld_s, in ? row_bytes : 0u), loff);
    }
    if constexpr (SUM2) {
      bload_pairs<NP, POL>(sb2[slot], make_rsrc(ld_s2, in ? row_bytes : 0u), loff);
      ld_s2 += sstep;
    }
    ld_c += cstep;
    ld_s += sstep;
  }

  __device__ __forceinline__ void init_window() {
    if constexpr (CM == CM_FUSED) {
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int k = kbase + j;
        const int t = J.reverse ? x0 + J.minD + k : x0 - J.minD - k;
        win[j] = fetch_target(J, rs.tdesc, rs.tmask, t);
      }
    } else if constexpr (CM == CM_RECOMPUTE) {
      if (slide) {
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int k = kbase + j;
          const int t = J.reverse ? x0 + J.minD + k : x0 - J.minD - k;
          win[j] = ld64(rs.tdesc, t * 8);
        }
      }
    }
  }

  __device__ __forceinline__ void slide_window(uint32_t elo, uint32_t ehi) {
    const bool up = SLIDE_UP >= 0 ? SLIDE_UP == 1
                    : CM == CM_FUSED ? !J.reverse
                                     : ((A.dx > 0) != (J.reverse != 0));
    if (!up) {
      const uint32_t ilo = (uint32_t)dpp_shl1((int)(uint32_t)win[0], (int)elo);
      const uint32_t ihi = (uint32_t)dpp_shl1((int)(uint32_t)(win[0] >> 32), (int)ehi);
#pragma unroll
      for (int j = 0; j < R - 1; j++) win[j] = win[j + 1];
      win[R - 1] = ((uint64_t)ihi << 32) | ilo;
    } else {
      const uint32_t ilo = (uint32_t)dpp_shr1((int)(uint32_t)win[R - 1], (int)elo);
      const uint32_t ihi = (uint32_t)dpp_shr1((int)(uint32_t)(win[R - 1] >> 32), (int)ehi);
#pragma unroll
      for (int j = R - 1; j > 0; j--) win[j] = win[j - 1];
      win[0] = ((uint64_t)ihi << 32) | ilo;
    }
  }

  // One pixel step with metadata lane `ml` (uniform, compile-time in the
  // unrolled loop) and prefetch slot `slot`.
  __device__ __forceinline__ void step(int s, int slot, int ml) {
    const int p2h = __builtin_amdgcn_readlane(meta.p2, ml);
    uint32_t c[NP];
    if constexpr (kComputed) {
      const uint32_t alo = (uint32_t)__builtin_amdgcn_readlane((int)meta.da_lo, ml);
      const uint32_t ahi = (uint32_t)__builtin_amdgcn_readlane((int)meta.da_hi, ml);
      const bool ok = __builtin_amdgcn_readlane(meta.ok, ml) != 0;
      if (slide) {
        census_costs<R>(win, alo, ahi, ok, kbase, J.D, A.inv_code, c);
        const uint32_t elo = (uint32_t)__builtin_amdgcn_readlane((int)meta.e_lo, ml);
        const uint32_t ehi = (uint32_t)__builtin_amdgcn_readlane((int)meta.e_hi, ml);
        slide_window(elo, ehi);
      } else {
        census_costs<R>(dsc[CM == CM_RECOMPUTE ? slot : 0], alo, ahi, ok, kbase, J.D,
                        A.inv_code, c);
      }
      if constexpr (CM == CM_FUSED) {
        uint32_t raw[(R + 3) / 4];
        narrow_costs<NP>(c, raw);
        bstore_bytes<R, POL>(raw, make_rsrc(st_c, cost_row_bytes), sboff);
        st_c += cstep;
      }
      if constexpr (PHASE == PHASE_COST) return;  // the cost pass: no recurrence
    } else if constexpr (CM == CM_LOAD8) {
      widen_costs<NP>(cb[slot], c);
    } else {
#pragma unroll
      for (int i = 0; i < NP; i++) c[i] = cb[slot][i];
    }
    // Recurrence: with no usable predecessor (p2h < 0) the offsets below are
    // zero and t = min(prev, 0, ...) = 0, so L = C.
    const bool has = p2h >= 0;
    const uint32_t mp2 = splat16(has ? (uint32_t)(minp + p2h) : 0u);
    const uint32_t mpk = splat16(has ? (uint32_t)minp : 0u);
    const uint32_t lft = (uint32_t)dpp_shr1((int)prev[NP - 1], (int)kInfPk);
    const uint32_t rgt = (uint32_t)dpp_shl1((int)prev[0], (int)kInfPk);
    uint32_t nb[NP + 1];
    nb[0] = __builtin_amdgcn_alignbit(prev[0], lft, 16);
#pragma unroll
    for (int i = 1; i < NP; i++) nb[i] = __builtin_amdgcn_alignbit(prev[i], prev[i - 1], 16);
    nb[NP] = __builtin_amdgcn_alignbit(rgt, prev[NP - 1], 16);
    uint32_t L[NP], lm = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      uint32_t t = pk_min(prev[i], mp2);
      t = pk_min(t, pk_add(nb[i], p1pk));
      t = pk_min(t, pk_add(nb[i + 1], p1pk));
      const uint32_t v = pk_sub(pk_add(c[i] & vmask, t), mpk);
      L[i] = (v & keep[i]) | infm[i];
      lm = pk_min(lm, L[i]);
    }
    int lmin = (int)min(lm & 0xFFFFu, lm >> 16);
    uint32_t S[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) {
      if constexpr (SUM2) {
        S[i] = pk_add(pk_add(sb[slot][i], sb2[slot][i]), L[i]);
      } else if constexpr (kPlaneIn) {
        // the other directions' path costs: bytes 2i, 2i+1 of each plane
        uint32_t t = L[i];
#pragma unroll
        for (int k = 0; k < kPlanes; k++)
          t = pk_add(t, __builtin_amdgcn_perm(0u, pl[slot][k][i >> 1],
                                              (i & 1) ? 0x0C030C02u : 0x0C010C00u));
        S[i] = t;
      } else {
        S[i] = PHASE == PHASE_FIRST ? L[i] : pk_add(sb[slot][i], L[i]);
      }
    }
    if constexpr (kPlaneOut) {
      // path costs of valid cells are <= max cost + P2 <= 255 (checked by the
      // host); other cells saturate at 255 and are never selected
      uint32_t raw[NB];
      narrow_costs<NP>(L, raw);
      bstore_bytes<R, POL>(raw, make_rsrc(st_s, (uint32_t)sum_store_bytes), sboff);
    } else if constexpr (!kPlaneIn) {
      bstore_pairs<NP, POL>(S, make_rsrc(st_s, (uint32_t)sum_store_bytes), sloff);
    }
    st_s += sstep;
    if constexpr (PHASE == PHASE_LAST) {
      // the winner key reduces beside the path minimum, interleaved
      uint32_t Sm[NP];
      unsigned key = masked_key(S, c, Sm);
      wave_min_and_min_u(lmin, key);
      minp = lmin;
#pragma unroll
      for (int i = 0; i < NP; i++) prev[i] = L[i];
      select(Sm, key, s);
    } else {
      minp = wave_min(lmin);
#pragma unroll
      for (int i = 0; i < NP; i++) prev[i] = L[i];
    }
  }

  // Pair i (uniform) of the lane group, selected without dynamic register indexing.
  __device__ __forceinline__ static uint32_t pick_pair(const uint32_t *v, int i) {
    uint32_t r = v[0];
#pragma unroll
    for (int j = 1; j < NP; j++) r = i == j ? v[j] : r;
    return r;
  }

  // Pair i (uniform) of lane `src` (uniform): every pair register is read from
  // that lane and the scalar picks one (an indexed pick on the vector side is
  // turned into an LDS round trip by the compiler).
  __device__ __forceinline__ static uint32_t lane_pair(const uint32_t *v, int i, int src) {
    if constexpr (NP <= 4) {
      uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)v[0], src);
#pragma unroll
      for (int j = 1; j < NP; j++) {
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)v[j], src);
        r = i == j ? t : r;
      }
      return r;
    } else {
      return (uint32_t)__builtin_amdgcn_readlane((int)pick_pair(v, i), src);
    }
  }

  // Masked aggregate of disparity k (uniform, 0 <= k < D): 0xFFFF when invalid.
  __device__ __forceinline__ static uint32_t cell(const uint32_t *Sm, int k) {
    const uint32_t w = lane_pair(Sm, (k % R) >> 1, k / R);
    return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
  }

  // Winner-take-all summary of this step (the reference's first argmin over
  // valid disparities, the second best at |k - best| > 1, the aggregates beside
  // the winner and the validity of both range ends; CensusSgm.java:460-515),
  // on packed pairs: invalid cells (cost == invalid code, or k >= D) are masked
  // to 0xFFFF, above every aggregate.
  // Lane-local part of the winner search: masked aggregates and the smallest
  // (aggregate << 12 | k) key of the lane.
  __device__ __forceinline__ unsigned masked_key(const uint32_t *S, const uint32_t *c,
                                                 uint32_t *Sm) const {
    unsigned key = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const uint32_t ne = pk_min(c[i] ^ invpk, 0x00010001u) & keep01[i];  // 1: valid
      Sm[i] = S[i] | pk_sub(ne, 0x00010001u);
      const unsigned k0 = ((Sm[i] & 0xFFFFu) << 12) | (kpk[i] & 0xFFFFu);
      const unsigned k1 = ((Sm[i] >> 16) << 12) | (kpk[i] >> 16);
      key = min(key, min(k0, k1));
    }
    return key;
  }

  // Sm, key: masked aggregates and the wave's smallest key (masked_key reduced).
  __device__ __forceinline__ void select(const uint32_t *Sm, unsigned key, int s) {
    const bool any = key < (0xFFFFu << 12);
    const int best = (int)(key & 0xFFFu);
    // second best: cells with |k - best| <= 1 masked to 0xFFFF
    const u16x2 bm1 = as_pk(splat16((uint32_t)(best - 1) & 0xFFFFu));
    const u16x2 two = as_pk(0x00020002u);
    uint32_t sec = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const u16x2 e = __builtin_elementwise_sub_sat(as_pk(kpk[i]) - bm1, two);  // 0: excluded
      const uint32_t far = pk_min(as_u32(e), 0x00010001u);
      sec = pk_min(sec, Sm[i] | pk_sub(far, 0x00010001u));
    }
    const unsigned second = wave_min_u(min(sec & 0xFFFFu, sec >> 16));
    uint32_t f0 = 0u, f1 = 0u, f2 = 0xFFFFFFFFu;
    if (any) {
      const uint32_t before = best > 0 ? cell(Sm, best - 1) : 0xFFFFu;
      const uint32_t after = best + 1 < J.D ? cell(Sm, best + 1) : 0xFFFFu;
      const uint32_t v0 = (uint32_t)__builtin_amdgcn_readlane((int)Sm[0], 0) & 0xFFFFu;
      const uint32_t wl = lane_pair(Sm, last_pair, last_lane);
      const uint32_t vl = last_half ? (wl >> 16) : (wl & 0xFFFFu);
      const uint32_t flags = (v0 != 0xFFFFu ? 1u : 0u) | (vl != 0xFFFFu ? 2u : 0u);
      f0 = (uint32_t)(best + 1) | (flags << 13);
      f1 = ((key >> 12) & 0xFFFFu) | (second << 16);
      f2 = before | (after << 16);
    }
    dfr.put(lane == (s & 63), f0, f1, f2);
  }

  __device__ __forceinline__ void maybe_flush(int s_done) {
    // after step s_done - 1: flush when a 64-step chunk is complete or the line ends
    if constexpr (PHASE == PHASE_LAST) {
      if ((s_done & 63) == 0 || s_done == len) {
        const int base = (s_done - 1) & ~63;
        select_flush(flush_args, dfr.view(), s_done - base, base, x0, y0);
      }
    }
  }

  __device__ __forceinline__ void run() {
    init_window();
    // prologue: metadata of window 0, loads of steps 0 .. PF-1
    int g_raw, gp_raw;
    uint64_t da = 0, e = 0;
    meta_load<PF, CM>(J, A, rs, 0, len, p0, step_off, x0, g_raw, gp_raw, da, e);
    meta = meta_finish<PF, CM>(J, A, 0, len, x0, y0, g_raw, gp_raw, da, e, p2_table);
#pragma unroll
    for (int j = 0; j < PF; j++) load_step(j, j);
    int s0 = 0;
    auto window = [&](int w0) {
      // metadata of the next window, consumed next iteration
      meta_load<PF, CM>(J, A, rs, w0 + PF, len, p0, step_off, x0, g_raw, gp_raw, da, e);
#pragma unroll
      for (int jj = 0; jj < PF; jj++) {
        step(w0 + jj, jj, jj);
        load_step(jj, w0 + jj + PF);
      }
      meta = meta_finish<PF, CM>(J, A, w0 + PF, len, x0, y0, g_raw, gp_raw, da, e, p2_table);
    };
    if constexpr (PHASE == PHASE_LAST) {
      // 64-step chunks with the selection flush outside the inner loop, so the
      // flush's scalar pressure does not force spills into the step code
      while (s0 + PF <= len) {
        const int lim = min(len, (s0 | 63) + 1);
        do {
          window(s0);
          s0 += PF;
        } while (s0 + PF <= lim);
        if (s0 == lim) maybe_flush(s0);
      }
    } else {
      for (; s0 + PF <= len; s0 += PF) window(s0);
    }
    // tail: fewer than PF steps; their loads are already in the slots
#pragma unroll
    for (int jj = 0; jj < PF - 1; jj++) {
      if (s0 + jj < len) {
        step(s0 + jj, jj, jj);
        maybe_flush(s0 + jj + 1);
      }
    }
  }
};

template <int R, int PHASE, int CM, int PFCAP = 64, bool NT = false, bool SUM2 = false,
          bool PLANE = false>
__device__ __forceinline__ void sweep_line_pk(const PathJob &J, const PathArgs &A, int line,
                                              const int *p2_table) {
  static_assert(R % 2 == 0, "packed sweep needs an even lane group");
  // the fused first sweep runs a copy of its loop per slide direction
  if constexpr (CM == CM_FUSED) {
    if (!J.reverse) {
      PkSweep<R, PHASE, CM, PFCAP, 1, NT, false, PLANE> sw(J, A, line, p2_table);
      sw.run();
    } else {
      PkSweep<R, PHASE, CM, PFCAP, 0, NT, false, PLANE> sw(J, A, line, p2_table);
      sw.run();
    }
    return;
  }
  PkSweep<R, PHASE, CM, PFCAP, -1, NT, SUM2, PLANE> sw(J, A, line, p2_table);
  sw.run();
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

// MODE: 0 generic int sweep; 1 packed u16 reading the uint16 cost volume; 2
// packed u16 first sweep with the census cost fused in (writes the uint8 cost
// volume); 3 packed u16 with the census cost recomputed in the sweep (no cost
// volume); 4 packed u16 reading the uint8 cost volume.  MODE_NT (added to 1, 2
// or 4): the volume accesses stream (nontemporal policy).  MODE_SUM2 (added
// to 1 or 4 of a last sweep): the job's second partial aggregate is added in.
// MODE_PLANE (added to 2 or 4 of a first-phase sweep, or to 4 of a last
// sweep): path-cost planes instead of the running aggregate.
constexpr int MODE_NT = 8;
constexpr int MODE_SUM2 = 16;
constexpr int MODE_PLANE = 32;

template <int R, typename SumT, int PHASE, int MODE>
__device__ __forceinline__ void sweep_dispatch(const PathJob &J, const PathArgs &A, int line,
                                               const int *p2_table) {
  constexpr int M = MODE & 7;
  constexpr bool NT = (MODE & MODE_NT) != 0;
  constexpr bool SUM2 = (MODE & MODE_SUM2) != 0;
  constexpr bool PLANE = (MODE & MODE_PLANE) != 0;
  static_assert(!SUM2 || ((M == 1 || M == 4) && PHASE == PHASE_LAST),
                "a second partial aggregate is read by last sweeps over a cost volume");
  static_assert(!PLANE || (M == 2 && PHASE == PHASE_FIRST) ||
                    (M == 4 && PHASE != PHASE_MID),
                "path-cost planes: first-phase sweeps (fused or uint8 costs) and last sweeps");
  if constexpr (M == 0) sweep_line<R, SumT, PHASE>(J, A, line, p2_table);
  else if constexpr (M == 1) sweep_line_pk<R, PHASE, CM_LOAD, 64, NT, SUM2>(J, A, line, p2_table);
  else if constexpr (M == 2)
    sweep_line_pk<R, PHASE, CM_FUSED, 64, NT, false, PLANE>(J, A, line, p2_table);
  else if constexpr (M == 3) sweep_line_pk<R, PHASE, CM_RECOMPUTE>(J, A, line, p2_table);
  // the last sweep over the planes prefetches 4 steps (7 plane dwords and a
  // cost dword per step): fewer VGPRs, more resident waves (measured faster)
  else sweep_line_pk<R, PHASE, CM_LOAD8, (PLANE && PHASE == PHASE_LAST) ? 4 : 64, NT, SUM2, PLANE>(
      J, A, line, p2_table);
}

// Up to kMultiDirs path directions in one launch, each a first-phase sweep
// storing its path costs into its own plane (PLANE modes).  Lines
// [begin[i], begin[i + 1]) sweep direction i; the first split[i] of them are
// matching direction 0's.  The directions are independent (each reads only
// the cost volume), so one launch carries all their lines.
constexpr int kMultiDirs = 7;
struct MultiDir {
  int ndirs, total_lines;
  int plane0;  // table jobs: plane index of direction 0 of the launch
  int dx[kMultiDirs], dy[kMultiDirs], cls[kMultiDirs], split[kMultiDirs];
  int begin[kMultiDirs];
  void *plane[kMultiDirs][2];  // jobs passed by value: each direction's plane per job
  // jobs in a device table (batches): the first line of each job per line
  // class, and the job count; a job's `sum` points at its plane block, the
  // plane of direction i lies plane0 + i rows of Dp bytes further
  const int *line_begin[3];
  int njobs;
  // dispatch order (optional): wave slot i sweeps line order[i], so the
  // longest lines start first and the launch does not end on a few long
  // lines dispatched last
  const int *order;
};

template <int R, int MODE>
__global__ __launch_bounds__(256) void path_multi_kernel(PathJobs jobs, PathArgs A, MultiDir M) {
  __shared__ int p2_table[256];
  build_p2_table(p2_table, A);  // P1, P2 and the edge scale are shared by every direction
  __syncthreads();
  const int waves = blockDim.x >> 6;
  const int slot = __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (slot >= M.total_lines) return;
  const int gline = M.order ? __builtin_amdgcn_readfirstlane(M.order[slot]) : slot;
  // direction of this line: fields picked with constant indices (uniform)
  PathArgs B = A;
  B.dx = M.dx[0];
  B.dy = M.dy[0];
  B.cls = M.cls[0];
  int dir = 0, begin = M.begin[0], split = M.split[0];
  void *pl0 = M.plane[0][0], *pl1 = M.plane[0][1];
#pragma unroll
  for (int i = 1; i < kMultiDirs; i++) {
    if (i < M.ndirs && gline >= M.begin[i]) {
      B.dx = M.dx[i];
      B.dy = M.dy[i];
      B.cls = M.cls[i];
      dir = i;
      begin = M.begin[i];
      split = M.split[i];
      pl0 = M.plane[i][0];
      pl1 = M.plane[i][1];
    }
  }
  const int local = gline - begin;
  if (jobs.table) {
    const int *lb = B.cls == 0 ? M.line_begin[0] : B.cls == 1 ? M.line_begin[1] : M.line_begin[2];
    const int ji = __builtin_amdgcn_readfirstlane(find_job(lb, M.njobs, local));
    PathJob J = jobs.table[ji];
    J.sum = (uint8_t *)J.sum + (size_t)(M.plane0 + dir) * J.Dp;
    sweep_dispatch<R, uint16_t, PHASE_FIRST, MODE>(J, B, local - lb[ji], p2_table);
  } else if (A.njobs > 1 && local >= split) {
    PathJob J = jobs.local[1];
    J.sum = pl1;
    sweep_dispatch<R, uint16_t, PHASE_FIRST, MODE>(J, B, local - split, p2_table);
  } else {
    PathJob J = jobs.local[0];
    J.sum = pl0;
    sweep_dispatch<R, uint16_t, PHASE_FIRST, MODE>(J, B, local, p2_table);
  }
}

// Two path directions in one launch, each accumulating into its own partial
// aggregate volume (the jobs of `ja` and `jb` point at different volumes):
// lines [0, A.total_lines) sweep direction A in phase PHA, the next
// B.total_lines lines direction B in phase PHB.  Uint16 sums wrap modulo 2^16
// in any order, so the partial volumes add up to the sequential aggregate
// exactly; one launch carries the lines of both directions, twice the waves
// of one, which keeps more volume requests in flight per CU.
template <int R, int PHA, int PHB, int MODE>
__global__ __launch_bounds__(256) void path_pair_kernel(PathJobs ja, PathArgs A, PathJobs jb,
                                                        PathArgs B) {
  __shared__ int p2_table[256];
  build_p2_table(p2_table, A);  // P1, P2 and the edge scale are shared by both directions
  __syncthreads();
  const int waves = blockDim.x >> 6;
  const int slot = __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (slot >= A.total_lines + B.total_lines) return;
  // A.order: dispatch order over both directions' lines (A's, then B's)
  int gline = A.order ? __builtin_amdgcn_readfirstlane(A.order[slot]) : slot;
  if (gline < A.total_lines) {
    if (A.njobs > 1 && gline >= A.split) {
      const PathJob J = ja.local[1];
      sweep_dispatch<R, uint16_t, PHA, MODE>(J, A, gline - A.split, p2_table);
    } else {
      const PathJob J = ja.local[0];
      sweep_dispatch<R, uint16_t, PHA, MODE>(J, A, gline, p2_table);
    }
    return;
  }
  gline -= A.total_lines;
  if (gline >= B.total_lines) return;
  if (B.njobs > 1 && gline >= B.split) {
    const PathJob J = jb.local[1];
    sweep_dispatch<R, uint16_t, PHB, MODE>(J, B, gline - B.split, p2_table);
  } else {
    const PathJob J = jb.local[0];
    sweep_dispatch<R, uint16_t, PHB, MODE>(J, B, gline, p2_table);
  }
}

template <int R, typename SumT, int PHASE, int MODE>
__global__ __launch_bounds__(256) void path_kernel(PathJobs jobs, PathArgs A) {
  __shared__ int p2_table[256];
  build_p2_table(p2_table, A);
  __syncthreads();
  const int waves = blockDim.x >> 6;
  const int slot =
      __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (slot >= A.total_lines) return;
  const int gline = A.order ? __builtin_amdgcn_readfirstlane(A.order[slot]) : slot;
  if (jobs.table) {
    const int ji = __builtin_amdgcn_readfirstlane(find_job(jobs.line_begin, A.njobs, gline));
    const PathJob J = jobs.table[ji];
    sweep_dispatch<R, SumT, PHASE, MODE>(J, A, gline - jobs.line_begin[ji], p2_table);
  } else if (A.njobs > 1 && gline >= A.split) {
    const PathJob J = jobs.local[1];
    sweep_dispatch<R, SumT, PHASE, MODE>(J, A, gline - A.split, p2_table);
  } else {
    const PathJob J = jobs.local[0];
    sweep_dispatch<R, SumT, PHASE, MODE>(J, A, gline, p2_table);
  }
}

}  // namespace sgm
