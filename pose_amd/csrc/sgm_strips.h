// Strip-streamed aggregation: the streamed plan of the reference's GPU SGM
// (GpuStereo.java:55-84 plan, 336-451 strip loop, path carries of the `paths`
// kernel GpuStereo.java:742-753).  The image is cut into horizontal strips of
// `rows` rows; per strip the device holds only the strip's cost and aggregate
// volumes.  Horizontal paths are strip-local.  A vertical or diagonal path
// crosses the seam between two strips: its value at the last row a strip
// reaches is written to a carry row ([x][Dp] int32), which the next strip in
// sweep order reads as the predecessor of its first row.  The recurrence is
// evaluated in the same order on the same values as in one whole-image sweep,
// so strip results are identical to the resident plan for every strip height.
#pragma once

#include "sgm_paths.h"

namespace sgm {

struct StripJob {
  const uint16_t *cost;   // strip cost volume, [(y - y0) * w + x][Dp]
  void *sum;              // strip aggregate volume, same indexing
  const uint8_t *mask;    // texture & valid of the reference image (whole image)
  const uint8_t *gray;    // reference image (whole image)
  const int *carry_in;    // [x][Dp]: path value of the pixel before the strip, or nullptr
  int *carry_out;         // [x][Dp]: path value at the strip's exit row, or nullptr
  float *disp;            // last sweep outputs (whole image), as PathJob
  uint8_t *reasons, *boundary, *code;
  int w, h, D, Dp, minD, rx, ry;
  int y0, rows;           // the strip: image rows [y0, y0 + rows)
};

// Line `line` of direction (dx, dy) clipped to the strip, in the enumeration
// of line_start with the strip in place of the image: horizontal lines are the
// strip's rows; vertical lines its columns; diagonal lines start on the entry
// row (x = line) or on the entry side column (rows after the entry row).
__device__ __forceinline__ void segment_start(int line, int dx, int dy, int w, int y0, int rows,
                                              int &x, int &y, int &len) {
  const int entry = dy > 0 ? y0 : y0 + rows - 1;
  if (dy == 0) {
    y = y0 + line;
    x = dx > 0 ? 0 : w - 1;
    len = w;
  } else if (dx == 0) {
    x = line;
    y = entry;
    len = rows;
  } else {
    if (line < w) {
      x = line;
      y = entry;
    } else {
      x = dx > 0 ? 0 : w - 1;
      y = dy > 0 ? y0 + line - w + 1 : y0 + rows - 2 - (line - w);
    }
    const int sx = dx > 0 ? w - x : x + 1;
    const int sy = dy > 0 ? y0 + rows - y : y - y0 + 1;
    len = min(sx, sy);
  }
}

__host__ __device__ inline int segments_of(int cls, int w, int rows) {
  return cls == 0 ? rows : cls == 1 ? w : w + rows - 1;
}

// One path segment: the generic 32-bit recurrence of sweep_line (the path
// recurrence of CensusSgm.java:421-459), started from the carried predecessor
// and ending with the carry of its exit pixel.
template <int R, typename SumT, int PHASE>
__device__ __forceinline__ void strip_segment(const StripJob &J, const PathArgs &A, int line,
                                              const int *p2_table) {
  constexpr int PF = prefetch_depth<R>();
  const int lane = lane_id();
  const int kbase = lane * R;
  const bool lane_mem = kbase < J.D;
  int x, y, len;
  segment_start(line, A.dx, A.dy, J.w, J.y0, J.rows, x, y, len);
  const ptrdiff_t step_off = (ptrdiff_t)A.dy * J.w + A.dx;
  const ptrdiff_t p0 = (ptrdiff_t)(y - J.y0) * J.w + x;  // strip-local pixel
  const ptrdiff_t img = (ptrdiff_t)J.y0 * J.w;           // strip-local -> image pixel
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

  // predecessor of the first pixel: in the previous strip of the sweep order
  int prev[R];
#pragma unroll
  for (int j = 0; j < R; j++) prev[j] = kInf;
  int minp = kInf;
  int g_prev = 0, m_prev = 0;
  {
    const int px = x - A.dx, py = y - A.dy;
    if (J.carry_in && A.dy != 0 && px >= 0 && px < J.w && py >= 0 && py < J.h) {
      const size_t q = (size_t)py * J.w + px;
      m_prev = J.mask[q];
      g_prev = J.gray[q];
      int lm = kInf;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int k = kbase + j;
        prev[j] = k < J.D ? J.carry_in[(size_t)px * Dp + k] : kInf;
        lm = min(lm, prev[j]);
      }
      minp = wave_min(lm);
    }
  }
  LineChunk gch, mch;
  if (lane < len) {
    const ptrdiff_t pl = img + p0 + (ptrdiff_t)lane * step_off;
    gch.next = J.gray[pl];
    mch.next = J.mask[pl];
  }
  Deferred dfr{};
  FlushArgs fa{0};
  if constexpr (PHASE == PHASE_LAST) {
    PathJob pj{};
    pj.disp = J.disp;
    pj.reasons = J.reasons;
    pj.boundary = J.boundary;
    pj.code = J.code;
    pj.w = J.w;
    pj.h = J.h;
    pj.rx = J.rx;
    pj.ry = J.ry;
    pj.D = J.D;
    pj.minD = J.minD;
    fa = make_flush_args(pj, A);
  }

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
          const ptrdiff_t pl = img + p0 + (ptrdiff_t)si * step_off;
          gch.next = J.gray[pl];
          mch.next = J.mask[pl];
        }
      }
      const int g_cur = __builtin_amdgcn_readlane(gch.cur, s & 63);
      const int m_cur = __builtin_amdgcn_readlane(mch.cur, s & 63);
      const bool has = m_prev != 0;
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
      // whole rows, pads included
      if (kbase < Dp && PHASE != PHASE_LAST) SumVec<R, SumT>::store(sum + pp * Dp + kbase, S);
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
          select_flush(fa,
                       WinnerView{dfr.best, dfr.value, dfr.second, dfr.before, dfr.after, dfr.flags},
                       (s & 63) + 1, s & ~63, x, y);
      }
    }
  }

  // carry of the exit pixel when the path continues in the next strip
  if (J.carry_out && A.dy != 0 && len > 0) {
    const int xe = x + (len - 1) * A.dx, ye = y + (len - 1) * A.dy;
    const int exit_row = A.dy > 0 ? J.y0 + J.rows - 1 : J.y0;
    if (ye == exit_row && lane_mem) {
#pragma unroll
      for (int j = 0; j < R; j++)
        if (kbase + j < J.D) J.carry_out[(size_t)xe * Dp + kbase + j] = prev[j];
    }
  }
}

// The strip of up to two matching directions in one launch: lines [0, split)
// belong to job 0, the rest to job 1.
struct StripJobs {
  StripJob job[2];
  int split;
};

template <int R, typename SumT, int PHASE>
__global__ __launch_bounds__(256) void strip_kernel(StripJobs jobs, PathArgs A) {
  __shared__ int p2_table[256];
  build_p2_table(p2_table, A);
  __syncthreads();
  const int waves = blockDim.x >> 6;
  const int line = __builtin_amdgcn_readfirstlane(blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (line >= A.total_lines) return;
  if (line < jobs.split) strip_segment<R, SumT, PHASE>(jobs.job[0], A, line, p2_table);
  else strip_segment<R, SumT, PHASE>(jobs.job[1], A, line - jobs.split, p2_table);
}

}  // namespace sgm
