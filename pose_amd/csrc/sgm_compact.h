// Compact SGM with certified outcomes ("esgm" plan of the reference:
// GpuCandidateStereo.java:14-47 plan, 49-171 execute, 190-335 run, 373-480
// the collect and select hooks).  No aggregate volume is kept.  Per pixel and
// path, two sweeps over the image keep only:
//   pass CANDIDATES: up to `pool` separated proposals per path (the smallest
//     path costs L_r(p, k), first index on ties, each more than one disparity
//     away from the proposals chosen before it);
//   pass SUMS: for every proposal k of every path, sum_r L_r(p, k-1..k+1)
//     (the exact aggregate at those disparities), and sum_r min L_r(p, k) over
//     the valid disparities that no proposal covers (a lower bound of the
//     aggregate anywhere off the proposals).
// A pixel whose winner, neighbours and uniqueness decision are settled by
// these numbers is certified; every other pixel is verified exactly by a third
// pass that accumulates the full aggregate vector of only those pixels.  The
// selected map, reasons and counters are therefore those of ordinary SGM.
//
// With 1-word census windows and R <= 8 every path sweeps the whole image and
// computes its census costs on the fly from the descriptors (HREC): no cost
// volume at all.  Otherwise sweeps run over horizontal strips with int32 path
// carries at the seams (sgm_strips.h), so the device holds one strip's cost
// volume at a time: a top-down pass over the strips runs the paths with
// dy >= 0, a bottom-up pass the others, each strip's cost computed once per
// pass.  One wave64 per line (segment); lane l holds disparities l*R .. l*R+R-1.
#pragma once

#include "sgm_strips.h"

namespace sgm {

enum { ESGM_CANDIDATES = 0, ESGM_SUMS = 1, ESGM_VERIFY = 2 };

struct CompactJob {
  const uint16_t *cost;      // strip cost volume (uint16, 512 = invalid), [(y - y0) * w + x][Dp]
  const uint8_t *mask;       // texture & valid of the reference image (whole image)
  const uint8_t *gray;       // reference image (whole image)
  const int *carry_in;       // [x][Dp] path value before the strip, or nullptr
  int *carry_out;            // [x][Dp] path value at the strip's exit row, or nullptr
  uint16_t *cand;            // [p][C] proposals, path r at [r * pool, r * pool + pool); 65535 = none
  uint16_t *sums;            // [p][C][3] aggregate at proposal - 1, proposal, proposal + 1
  uint16_t *outside;         // [p] sum over paths of the path minimum off the proposals (65535 = none)
  const int *slot;           // [p] verification slot of the pixel, -1 if certified
  uint16_t *verify;          // [slot - lo][Dp] exact aggregates of the group's pixels
  int lo, hi;                // verification slots of this group
  int w, h, D, Dp, y0, rows;
  // whole-image lines with the census cost computed on the fly (1-word
  // windows): reference / target descriptors (bit 63 = the texture & valid
  // flag of the pixel), the disparity mapping and the window
  const uint64_t *desc_a, *desc_b;
  int minD, reverse, rx, ry;
};

struct CompactJobs {
  CompactJob job[2];
  int split;  // lines [0, split) belong to job 0
};

struct CompactArgs {
  int C;          // proposals per pixel over all paths (directions * pool)
  int pool;       // proposals per path, 1..4
  int direction;  // index of the path in the direction order
};

// The LDS of one wave: the path costs of the current pixel and the coverage
// stamps of the proposals, Dp ints each.
template <int R, int PH, bool HREC>
__device__ __forceinline__ void compact_segment(const CompactJob &J, const PathArgs &A,
                                                const CompactArgs &E, int line,
                                                const int *p2_table, int *lds_l, int *lds_cov) {
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
  const int Dp = J.Dp, D = J.D, C = E.C;

  if (PH == ESGM_SUMS) {
#pragma unroll
    for (int j = 0; j < R; j++)
      if (lane_mem) lds_cov[kbase + j] = -1;
  }

  CostVec<R> cb[PF];
  // sum pass: the proposals, the 3C sums (two lane chunks) and the bound of
  // the pixel PF steps ahead are loaded with its costs
  int pcand[PF], psum0[PF], psum1[PF], pout[PF];
  // HREC (costs computed on the fly): target descriptors of the pixel PF
  // steps ahead (index clamped into its row, in-range bits in pok)
  uint64_t pdesc[HREC ? PF : 1][HREC ? R : 1];
  uint32_t pok[HREC ? PF : 1];
#pragma unroll
  for (int j = 0; j < PF; j++) {
    cb[j] = CostVec<R>{};
    pcand[j] = 65535;
    psum0[j] = psum1[j] = pout[j] = 0;
  }
  auto prefetch = [&](int j, int st) {
    const ptrdiff_t pp = p0 + (ptrdiff_t)st * step_off;
    if constexpr (HREC) {
      const int xs = x + st * A.dx, ys = y + st * A.dy;
      const size_t rowb = (size_t)ys * J.w;
      uint32_t ok = 0;
#pragma unroll
      for (int jr = 0; jr < R; jr++) {
        const int d = J.minD + kbase + jr;
        const int t = J.reverse ? xs + d : xs - d;
        const bool in = kbase + jr < D && t >= J.rx && t < J.w - J.rx;
        ok |= in ? 1u << jr : 0u;
        pdesc[j][jr] = J.desc_b[rowb + (size_t)min(max(t, 0), J.w - 1)];
      }
      pok[j] = ok;
    } else {
      if (lane_mem) cb[j].load(cost + pp * Dp + kbase);
    }
    if constexpr (PH == ESGM_SUMS) {
      const size_t g = (size_t)(img + pp);
      const uint16_t *row = J.sums + g * (size_t)(3 * C);
      pcand[j] = lane < C ? (int)J.cand[g * C + lane] : 65535;
      psum0[j] = lane < 3 * C ? (int)row[lane] : 0;
      psum1[j] = lane + 64 < 3 * C ? (int)row[lane + 64] : 0;
      pout[j] = J.outside[g];
    }
  };
#pragma unroll
  for (int j = 0; j < PF; j++)
    if (j < len) prefetch(j, j);

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
        prev[j] = k < D ? J.carry_in[(size_t)px * Dp + k] : kInf;
        lm = min(lm, prev[j]);
      }
      minp = wave_min(lm);
    }
  }
  LineChunk gch, mch, sch, alo, ahi;
  if (lane < len) {
    const ptrdiff_t pl = img + p0 + (ptrdiff_t)lane * step_off;
    gch.next = J.gray[pl];
    mch.next = J.mask[pl];
    if (PH == ESGM_VERIFY) sch.next = J.slot[pl];
    if (HREC) {
      const uint64_t a = J.desc_a[pl];
      alo.next = (int)(uint32_t)a;
      ahi.next = (int)(uint32_t)(a >> 32);
    }
  }

  for (int s0 = 0; s0 < len; s0 += PF) {
#pragma unroll
    for (int jj = 0; jj < PF; jj++) {
      const int s = s0 + jj;
      if (s >= len) break;
      const ptrdiff_t pp = p0 + (ptrdiff_t)s * step_off;
      const size_t gp = (size_t)(img + pp);
      if ((s & 63) == 0) {
        gch.cur = gch.next;
        mch.cur = mch.next;
        sch.cur = sch.next;
        alo.cur = alo.next;
        ahi.cur = ahi.next;
        const int si = s + 64 + lane;
        if (si < len) {
          const ptrdiff_t pl = img + p0 + (ptrdiff_t)si * step_off;
          gch.next = J.gray[pl];
          mch.next = J.mask[pl];
          if (PH == ESGM_VERIFY) sch.next = J.slot[pl];
          if (HREC) {
            const uint64_t a = J.desc_a[pl];
            alo.next = (int)(uint32_t)a;
            ahi.next = (int)(uint32_t)(a >> 32);
          }
        }
      }
      const int g_cur = __builtin_amdgcn_readlane(gch.cur, s & 63);
      const int m_cur = __builtin_amdgcn_readlane(mch.cur, s & 63);
      const bool has = m_prev != 0;
      const int P2l = has ? p2_table[abs(g_cur - g_prev)] : A.p2;
      g_prev = g_cur;
      m_prev = m_cur;

      int c[R], L[R];
      if constexpr (HREC) {
        // census Hamming cost with the `costs` kernel's validity test
        // (GpuStereo.java:706-719): reference pixel interior and flagged,
        // target inside [rx, w - rx) and flagged; 512 otherwise
        const uint32_t a_lo = (uint32_t)__builtin_amdgcn_readlane(alo.cur, s & 63);
        const uint32_t a_hi = (uint32_t)__builtin_amdgcn_readlane(ahi.cur, s & 63);
        const int xs = x + s * A.dx, ys = y + s * A.dy;
        const bool pix = xs >= J.rx && xs < J.w - J.rx && ys >= J.ry && ys < J.h - J.ry && m_cur;
#pragma unroll
        for (int j = 0; j < R; j++) {
          const uint64_t e = pdesc[jj][j];
          const uint32_t e_lo = (uint32_t)e, e_hi = (uint32_t)(e >> 32);
          const int ham = __builtin_popcount(e_lo ^ a_lo) +
                          __builtin_popcount((e_hi ^ a_hi) & 0x7FFFFFFFu);
          c[j] = pix && ((pok[jj] >> j) & 1u) && (e_hi >> 31) ? ham : A.inv_code;
        }
      } else {
#pragma unroll
        for (int j = 0; j < R; j++) c[j] = cb[jj].get(j);
      }
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
        if (kbase + j >= D) v = kInf;
        L[j] = v;
        lmin = min(lmin, v);
      }
      const int prop = pcand[jj], sum0 = psum0[jj], sum1 = psum1[jj], prior = pout[jj];
      if (s + PF < len) prefetch(jj, s + PF);
      minp = wave_min(lmin);
#pragma unroll
      for (int j = 0; j < R; j++) prev[j] = L[j];

      if constexpr (PH == ESGM_CANDIDATES) {
        // separated proposals: smallest L over valid k, first k on ties
        // (key L * (D + 1) + k), each more than one disparity from the earlier ones
        int chosen[4] = {65535, 65535, 65535, 65535};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          if (q < E.pool) {
            int key = 0x7FFFFFFF;
#pragma unroll
            for (int j = 0; j < R; j++) {
              const int k = kbase + j;
              bool ok = k < D && c[j] != A.inv_code;
#pragma unroll
              for (int u = 0; u < q; u++) ok = ok && abs(k - chosen[u]) > 1;
              if (ok) key = min(key, L[j] * (D + 1) + k);
            }
            key = wave_min(key);
            chosen[q] = key == 0x7FFFFFFF ? 65535 : key % (D + 1);
          }
        }
        if (lane < E.pool) {
          const int mine = lane == 0 ? chosen[0] : lane == 1 ? chosen[1] : lane == 2 ? chosen[2]
                                                                                     : chosen[3];
          J.cand[gp * C + E.direction * E.pool + lane] = (uint16_t)mine;
        }
      } else if constexpr (PH == ESGM_SUMS) {
        // path costs and proposal coverage of this pixel into the wave's LDS
#pragma unroll
        for (int j = 0; j < R; j++)
          if (lane_mem) lds_l[kbase + j] = L[j];
        if (prop < D) {
#pragma unroll
          for (int o = -1; o <= 1; o++) {
            const int kk = prop + o;
            if (kk >= 0 && kk < D) lds_cov[kk] = s;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // entry t of the pixel's 3C sums: proposal t / 3, offset t % 3 - 1
        const int prop0 = __builtin_amdgcn_ds_bpermute((lane / 3) * 4, prop);
        const int prop1 = __builtin_amdgcn_ds_bpermute(((lane + 64) / 3) * 4, prop);
        uint16_t *row = J.sums + gp * (size_t)(3 * C);
        {
          const int t = lane;
          const int k = prop0 + t % 3 - 1;
          if (t < 3 * C && k >= 0 && k < D) row[t] = (uint16_t)(sum0 + lds_l[k]);
        }
        {
          const int t = lane + 64;
          const int k = prop1 + t % 3 - 1;
          if (t < 3 * C && k >= 0 && k < D) row[t] = (uint16_t)(sum1 + lds_l[k]);
        }
        int lower = 0x7FFFFFFF;
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int k = kbase + j;
          if (k < D && c[j] != A.inv_code && lds_cov[k] != s) lower = min(lower, L[j]);
        }
        lower = wave_min(lower);
        if (lane == 0)
          J.outside[gp] = (uint16_t)(lower == 0x7FFFFFFF ? 65535 : min(65535, prior + lower));
        // the next step rewrites the LDS only after every lane has read it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      } else {
        const int sl = __builtin_amdgcn_readlane(sch.cur, s & 63);
        if (sl >= J.lo && sl < J.hi && lane_mem) {
          uint16_t *row = J.verify + (size_t)(sl - J.lo) * Dp + kbase;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (kbase + j < D) row[j] = (uint16_t)(row[j] + L[j]);
        }
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
        if (kbase + j < D) J.carry_out[(size_t)xe * Dp + kbase + j] = prev[j];
    }
  }
}

// Dynamic LDS: 2 * Dp ints per wave.
template <int R, int PH, bool HREC>
__global__ __launch_bounds__(256) void compact_kernel(CompactJobs jobs, PathArgs A, CompactArgs E) {
  __shared__ int p2_table[256];
  extern __shared__ int compact_lds[];
  build_p2_table(p2_table, A);
  __syncthreads();
  const int waves = blockDim.x >> 6, wave = (int)(threadIdx.x >> 6);
  const int line = __builtin_amdgcn_readfirstlane(blockIdx.x * waves + wave);
  if (line >= A.total_lines) return;
  const int Dp = jobs.job[0].Dp;
  int *lds_l = compact_lds + (size_t)wave * 2 * Dp, *lds_cov = lds_l + Dp;
  if (line < jobs.split)
    compact_segment<R, PH, HREC>(jobs.job[0], A, E, line, p2_table, lds_l, lds_cov);
  else
    compact_segment<R, PH, HREC>(jobs.job[1], A, E, line - jobs.split, p2_table, lds_l, lds_cov);
}

// ---------------------------------------------------------------------------
// Winner summaries in the packed form of DeferredPk (f0 = (best + 1) | flags
// << 13, f1 = value | second << 16, f2 = before | after << 16; 0xFFFF = none).

struct SummaryJob {
  const uint16_t *cand, *sums, *outside;
  const uint8_t *mask_a, *mask_b;  // texture & valid of the reference / target image
  uint32_t *summary;               // [p][3]
  int *slot;                       // [p] verification slot or -1
  int *pixels;                     // [slot] pixel
  int *count;                      // verification slots handed out
  const uint16_t *verify;          // [slot - lo][Dp] (verification winners)
  int lo, hi;
  int w, h, D, Dp, minD, rx, ry, reverse;
};

struct SummaryJobs {
  SummaryJob job[2];
};

// Cost validity of (p, k): the cost kernel's test without the cost
// (GpuCandidateStereo.java:452-455 esgm_valid), split into the reference
// pixel's part (interior, texture & valid; once per pixel) and the target's.
__device__ __forceinline__ bool esgm_pixel_ok(const SummaryJob &J, int x, int y) {
  return x >= J.rx && x < J.w - J.rx && y >= J.ry && y < J.h - J.ry &&
         J.mask_a[(size_t)y * J.w + x];
}

__device__ __forceinline__ bool esgm_target_ok(const SummaryJob &J, int x, int y, int k) {
  if (k < 0 || k >= J.D) return false;
  const int d = J.minD + k;
  const long long t = (long long)x + (J.reverse ? (long long)d : -(long long)d);
  return t >= J.rx && t < J.w - J.rx && J.mask_b[(size_t)y * J.w + (size_t)t];
}

__device__ __forceinline__ bool esgm_valid(const SummaryJob &J, int x, int y, int k) {
  return esgm_pixel_ok(J, x, y) && esgm_target_ok(J, x, y, k);
}

__device__ __forceinline__ void put_summary(uint32_t *out, int best, int flags, int value,
                                            int second, int before, int after) {
  out[0] = (uint32_t)(best + 1) | ((uint32_t)flags << 13);
  out[1] = (uint32_t)(value & 0xFFFF) | ((uint32_t)(second == 0x7FFFFFFF ? 0xFFFF : second) << 16);
  out[2] = (uint32_t)(before < 0 ? 0xFFFF : before) | ((uint32_t)(after < 0 ? 0xFFFF : after) << 16);
}

// Winner over the proposals (esgm_select, GpuCandidateStereo.java:456-478) and
// the certification of GpuCandidateStereo.java:82-102; pixels that are not
// certified get a verification slot.  One thread per pixel.
__global__ __launch_bounds__(256) void esgm_summary_kernel(SummaryJobs jobs, int C, double uniq) {
  const SummaryJob J = jobs.job[blockIdx.y];
  const size_t n = (size_t)J.w * J.h;
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  bool unresolved = false;
  if (p < n) {
    const int x = (int)(p % J.w), y = (int)(p / J.w);
    const uint16_t *cand = J.cand + p * C, *sums = J.sums + p * (size_t)(3 * C);
    int best = -1, value = 0x7FFFFFFF, second = 0x7FFFFFFF, before = -1, after = -1, owner = -1;
    for (int j = 0; j < C; j++) {
      const int k = cand[j], v = sums[3 * j + 1];
      if (k < J.D && (v < value || (v == value && (best < 0 || k < best)))) {
        best = k;
        value = v;
        owner = j;
      }
    }
    const bool pix = esgm_pixel_ok(J, x, y);
    if (owner >= 0 && pix) {
      if (esgm_target_ok(J, x, y, best - 1)) before = sums[3 * owner];
      if (esgm_target_ok(J, x, y, best + 1)) after = sums[3 * owner + 2];
      for (int j = 0; j < C; j++) {
        const int kc = (int)cand[j];
#pragma unroll
        for (int o = 0; o < 3; o++) {
          const int k = kc + o - 1;
          if (abs(k - best) > 1 && esgm_target_ok(J, x, y, k))
            second = min(second, (int)sums[3 * j + o]);
        }
      }
    }
    const int flags = pix ? (esgm_target_ok(J, x, y, 0) ? 1 : 0) |
                                (esgm_target_ok(J, x, y, J.D - 1) ? 2 : 0)
                          : 0;
    put_summary(J.summary + 3 * p, best, flags, value, second, before, after);
    if (best >= 0) {
      const long long lower = J.outside[p];
      const bool winner = lower > value && (before < 0 || before > value) &&
                          (after < 0 || after > value) && second > value;
      double margin = (double)value * uniq;
      if (!(margin >= 1.0)) margin = 1.0;
      const bool trusted = winner && (best == 0 || best == J.D - 1 ||
                                      (double)second - (double)value <= margin ||
                                      (double)(lower - value) > margin);
      unresolved = !trusted;
    }
    J.slot[p] = -1;
  }
  // slots handed out per wave: one atomic per wave
  const unsigned long long ballot = __ballot(unresolved);
  if (ballot) {
    const int lane = lane_id();
    int base = 0;
    if (lane == __builtin_ctzll(ballot)) base = atomicAdd(J.count, __popcll(ballot));
    base = __builtin_amdgcn_readlane(base, __builtin_ctzll(ballot));
    if (unresolved) {
      const int mine = base + __popcll(ballot & ((1ull << lane) - 1));
      J.slot[p] = mine;
      J.pixels[mine] = (int)p;
    }
  }
}

// Exact winners of the verified pixels of one group (slots [lo, hi)): one wave
// per pixel over its full aggregate vector.
template <int R>
__global__ __launch_bounds__(256) void esgm_verify_kernel(SummaryJobs jobs) {
  const SummaryJob J = jobs.job[blockIdx.y];
  const int waves = blockDim.x >> 6;
  const int sl = J.lo + __builtin_amdgcn_readfirstlane((int)blockIdx.x * waves + (int)(threadIdx.x >> 6));
  if (sl >= J.hi) return;
  const int p = J.pixels[sl];
  const int x = p % J.w, y = p / J.w;
  const int kbase = lane_id() * R;
  const uint16_t *row = J.verify + (size_t)(sl - J.lo) * J.Dp;
  int S[R], valid[R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    const int k = kbase + j;
    valid[j] = esgm_valid(J, x, y, k) ? 1 : 0;
    S[j] = k < J.D ? (int)row[k] : 0x7FFFFFFF;
  }
  const Winner win = wta_reduce<R, false>(S, valid, kbase, J.D);
  if (lane_id() == 0)
    put_summary(J.summary + 3 * (size_t)p, win.best, win.flags, win.value, win.second, win.before,
                win.after);
}

// Selection of every pixel from its summary (GpuStereo.select, the same
// select_flush as the fused last sweep): one wave per 64 pixels of a row.
struct SelectJob {
  const uint32_t *summary;
  PathJob out;  // disp, reasons, boundary, code, w, h, rx, ry, D, minD
};

struct SelectJobs {
  SelectJob job[2];
};

__global__ __launch_bounds__(256) void esgm_select_kernel(SelectJobs jobs, PathArgs A) {
  const SelectJob &S = jobs.job[blockIdx.z];
  const int waves = blockDim.x >> 6;
  const int chunk = __builtin_amdgcn_readfirstlane((int)blockIdx.x * waves + (int)(threadIdx.x >> 6));
  const int w = S.out.w, y = blockIdx.y;
  const int x0 = chunk * 64;
  if (x0 >= w) return;
  const int count = min(64, w - x0), lane = lane_id();
  DeferredPk d{0, 0, 0};
  if (lane < count) {
    const uint32_t *sm = S.summary + 3 * ((size_t)y * w + x0 + lane);
    d.f0 = sm[0];
    d.f1 = sm[1];
    d.f2 = sm[2];
  }
  select_flush(make_flush_args(S.out, A), d.view(), count, 0, x0, y);
}

}  // namespace sgm
