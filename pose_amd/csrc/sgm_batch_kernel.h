// Persistent work-queue kernel of the batched mode: many uneven stereo pairs
// (per-pair size, disparity count and offset) in one launch.
//
// Work items are scanline tiles of one phase of one job, numbered phase-major:
//   census (per image)  ->  cost (per matching direction, only when the cost is
//   not fused into the first sweep)  ->  path sweep 1 .. P (per direction)  ->
//   left-right check (per pair).
// A wave pops the next item index from a device counter.  An item may only
// start once the items it depends on (previous phase of the same job) have
// all completed; every dependency has a smaller index, so the waves that hold
// them are running and finish -- waits always end.  Each wait is also bounded
// (spin limit): on expiry the wave records an error and stops, so the grid
// always drains.  The reference processes pairs one at a time
// (DenseReconstruction.java:64-72); this mode has no reference counterpart.
#pragma once

#include "sgm_kernels.h"

namespace sgm {

struct BatchImage {
  const uint8_t *gray, *valid;
  uint64_t *desc;
  uint8_t *mask;
  uint16_t *gm;
  int32_t *wsum, *wsum2;
  uint8_t *reasons_init, *boundary_init;  // left image: reason 1 / boundary 0 for every pixel
  int w, h;
};

struct BatchPairDev {
  const float *dl, *dr;
  const uint8_t *code, *reasons_tmp, *boundary_tmp;
  float *checked;
  uint8_t *reasons, *boundary;
  unsigned long long *stats;
  int w, h;
};

enum { BPH_CENSUS = 0, BPH_COST = 1, BPH_PATH0 = 2 };

struct BatchArgs {
  const BatchImage *images;
  const PathJob *jobs;
  const CostJob *cost_jobs;
  const BatchPairDev *pairs;
  const int *phase_begin;  // [nph + 1] first global item of each phase
  const int *job_begin;    // [nph][stride] first item of each job within its phase
  unsigned *head;          // queue head
  unsigned *done;          // [nph][stride] completed items per job
  int *error;
  int nph, stride, npairs, directions, fused;
  int cmode;               // CM_*: the first sweep's cost mode
  int lines_per_item, census_rows, cost_rows, lr_rows;
  PathArgs pa;             // penalties, codes, uniqueness (direction fields set per item)
  CostArgs ca;
  int ww, wh;
  double lr_threshold;
  long long spin_limit;    // wall-clock ticks a dependency wait may take before the error flag
  // the queue range of this launch; a phase-major run launches the census /
  // cost phases and the L-R phase separately (the sweeps run between them),
  // and phases that end before item_lo were completed by earlier launches
  int item_lo, item_hi;
  int phased;
};

__device__ __forceinline__ int batch_phase_of(const BatchArgs &B, int item) {
  int ph = 0;
  while (ph + 1 < B.nph && B.phase_begin[ph + 1] <= item) ph++;
  return ph;
}

__device__ __forceinline__ int batch_job_of(const BatchArgs &B, int ph, int local, int njobs) {
  const int *jb = B.job_begin + (size_t)ph * B.stride;
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jb[mid] <= local) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int batch_items(const BatchArgs &B, int ph, int job) {
  const int *jb = B.job_begin + (size_t)ph * B.stride;
  return jb[job + 1] - jb[job];
}

__device__ __forceinline__ int batch_njobs(const BatchArgs &B, int ph) {
  return ph == B.nph - 1 ? B.npairs : 2 * B.npairs;
}

// Wait until every item of (ph, job) has completed; false when the wall-clock
// limit expires.  The poll is a relaxed device-coherent load (no cache
// invalidation per iteration); the caller issues one acquire fence after the
// wait succeeds.
__device__ __forceinline__ bool batch_wait(const BatchArgs &B, int ph, int job) {
  if (B.phased && B.phase_begin[ph] < B.item_lo) return true;  // an earlier launch ran it
  const unsigned need = (unsigned)batch_items(B, ph, job);
  unsigned *ctr = B.done + (size_t)ph * B.stride + job;
  if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) return true;
  const unsigned long long t0 = wall_clock64();
  while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
    if (wall_clock64() - t0 > (unsigned long long)B.spin_limit) return false;
    __builtin_amdgcn_s_sleep(32);
  }
  return true;
}

// Census + texture mask of rows [r0, r1) of one image; each wave stages its
// own (64 + 2rx) x (rows + 2ry) tile in LDS.
constexpr int kBatchCensusRows = 8;
constexpr int kBatchTileW = 64 + 10, kBatchTileH = kBatchCensusRows + 10;

// One 64-column chunk [x0, x0 + 64) of the row band.
template <int WORDS>
__device__ __forceinline__ void batch_census_chunk(const BatchArgs &B, const BatchImage &I, int x0,
                                                   int r0, int r1, uint8_t *tile) {
  const int lane = lane_id();
  const int rx = B.ww >> 1, ry = B.wh >> 1;
  const size_t n = (size_t)I.w * I.h;
  {
    const int tw = 64 + 2 * rx, th = (r1 - r0) + 2 * ry;
    for (int i = lane; i < tw * th; i += 64) {
      const int ty = i / tw, tx = i - ty * tw;
      const int gx = x0 - rx + tx, gy = r0 - ry + ty;
      tile[ty * kBatchTileW + tx] =
          (gx >= 0 && gx < I.w && gy >= 0 && gy < I.h) ? I.gray[(size_t)gy * I.w + gx] : 0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int x = x0 + lane;
    for (int y = r0; y < r1 && x < I.w; y++) {
      const int ly = y - r0;
      const size_t p = (size_t)y * I.w + x;
      const bool interior = x >= rx && x < I.w - rx && y >= ry && y < I.h - ry;
      uint64_t d0 = 0, d1 = 0;
      int lo = 255, hi = 0, s1 = 0, s2 = 0;
      const int c = tile[(ly + ry) * kBatchTileW + lane + rx];
      if (interior) {
        int bit = 0;
        for (int yy = 0; yy < B.wh; yy++)
          for (int xx = 0; xx < B.ww; xx++) {
            const int v = tile[(ly + yy) * kBatchTileW + lane + xx];
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
      const uint8_t m = interior && (hi - lo) >= 4 && (I.valid == nullptr || I.valid[p] != 0);
      I.desc[p] = WORDS == 1 ? (d0 | ((uint64_t)m << 63)) : d0;  // bit 63: texture & valid
      if (WORDS > 1) I.desc[n + p] = d1;
      I.mask[p] = m;
      I.gm[p] = (uint16_t)(c | (m << 8));
      if (I.wsum) {
        I.wsum[p] = s1;
        I.wsum2[p] = s2;
      }
      if (I.reasons_init) {
        I.reasons_init[p] = 1;
        I.boundary_init[p] = 0;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
}

template <int WORDS>
__device__ __forceinline__ void batch_census(const BatchArgs &B, const BatchImage &I, int r0,
                                             int r1, uint8_t *tile) {
  for (int x0 = 0; x0 < I.w; x0 += 64) batch_census_chunk<WORDS>(B, I, x0, r0, r1, tile);
}

// Cost volume rows [r0, r1) of one matching direction (non-fused costs).
template <bool ZNCC>
__device__ __forceinline__ void batch_cost(const BatchArgs &B, const CostJob &J, int r0, int r1) {
  const int lane = lane_id();
  const int rx = B.ww >> 1, ry = B.wh >> 1;
  const size_t n = (size_t)J.w * J.h;
  const int quads = J.Dp >> 2;
  for (int y = r0; y < r1; y++) {
    for (int qi = lane; qi < J.w * quads; qi += 64) {
      const int x = qi / quads, k0 = (qi - x * quads) * 4;
      const size_t p = (size_t)y * J.w + x;
      const bool pix_ok = x >= rx && x < J.w - rx && y >= ry && y < J.h - ry && J.mask_a[p];
      unsigned v[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int k = k0 + j;
        int value = B.ca.inv_code;
        if (pix_ok && k < J.D) {
          const int disp = J.minD + k;
          const long long t = (long long)x + (J.reverse ? (long long)disp : -(long long)disp);
          if (t >= rx && t < J.w - rx) {
            const size_t q = (size_t)y * J.w + (size_t)t;
            if (J.mask_b[q]) value = cost_value<ZNCC>(J, B.ca, p, q, n);
          }
        }
        v[j] = (unsigned)value & 0xFFFFu;
      }
      store_cost4(J.cost, p, J.Dp, k0, v, B.ca.u8 != 0);
    }
  }
}

// Left-right check + counters (as lr_kernel) of columns x0 + lane,
// x0 + lane + 64, ... below x1 of rows [r0, r1) of one pair: outputs written,
// the lane's eight counts added to c (stats fields 0-6 and 10).
__device__ __forceinline__ void batch_lr_span(const BatchArgs &B, const BatchPairDev &P, int r0,
                                              int r1, int x0, int x1, unsigned *c) {
  const int lane = lane_id();
  const int rx = B.ww >> 1, ry = B.wh >> 1;
  for (int y = r0; y < r1; y++)
    for (int x = x0 + lane; x < x1; x += 64) {
      const size_t p = (size_t)y * P.w + x;
      const int code = P.code[p];
      c[0] += (code & CODE_LOWER) ? 1 : 0;
      c[1] += (code & CODE_FULL) ? 1 : 0;
      c[2] += (code & 3) == CODE_BOUNDARY ? 1 : 0;
      c[3] += (code & CODE_UPPER) ? 1 : 0;
      c[4] += (code & 3) == CODE_AMBIGUOUS ? 1 : 0;
      c[5] += (code & 3) == CODE_SELECTED ? 1 : 0;
      float out = __builtin_nanf("");
      int reason = P.reasons_tmp[p];
      const float d = P.dl[p];
      if (x >= rx && x < P.w - rx && y >= ry && y < P.h - ry && isfinite(d)) {
        const int xr = round_half_up_f((float)x - d);
        if (xr < rx || xr >= P.w - rx) {
          reason = 4;
        } else {
          const float other = P.dr[(size_t)y * P.w + xr];
          if (isfinite(other) && (double)fabsf(d - other) <= B.lr_threshold) {
            out = d;
          } else {
            c[6]++;
            reason = 4;
          }
        }
      }
      c[7] += isfinite(out) ? 1 : 0;
      P.checked[p] = out;
      if (P.reasons) P.reasons[p] = (uint8_t)reason;
      if (P.boundary) P.boundary[p] = P.boundary_tmp[p];
    }
}

// Stats field of count i of batch_lr_span.
__device__ __forceinline__ int lr_field(int i) { return i < 7 ? i : 10; }

// Rows [r0, r1) of one pair, every column.
__device__ __forceinline__ void batch_lr(const BatchArgs &B, const BatchPairDev &P, int r0, int r1) {
  const int lane = lane_id();
  unsigned c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  batch_lr_span(B, P, r0, r1, 0, P.w, c);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int v = (int)c[i];
    // wave sum
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lane == 0 && v) atomicAdd(P.stats + lr_field(i), (unsigned long long)v);
  }
}

// Cost-volume instance of the batch kernel (each its own kernel, so the
// register-heavy recomputing sweeps do not lower the occupancy of the others):
//   VOL_U16:  uint16 cost volume (FP32 ZNCC costs, 32-bit aggregates)
//   VOL_U8:   uint8 cost volume, census cost fused into the first sweep when
//             the batch is 1-word census, else written by the cost phase
//   VOL_NONE: census cost recomputed in every sweep
enum { VOL_U16 = 0, VOL_U8 = 1, VOL_NONE = 2 };

template <int R, typename SumT, int VOL>
__device__ __forceinline__ void batch_sweep_r(const PathJob &J, const PathArgs &A, int line,
                                              int phase, bool fused, const int *p2_table) {
  if constexpr (sizeof(SumT) == 2 && VOL == VOL_NONE) {
    if (phase == PHASE_FIRST) sweep_line_pk<R, PHASE_FIRST, CM_RECOMPUTE>(J, A, line, p2_table);
    else if (phase == PHASE_MID) sweep_line_pk<R, PHASE_MID, CM_RECOMPUTE>(J, A, line, p2_table);
    else sweep_line_pk<R, PHASE_LAST, CM_RECOMPUTE>(J, A, line, p2_table);
  } else if constexpr (sizeof(SumT) == 2 && VOL == VOL_U8) {
    if (phase == PHASE_FIRST) {
      // the fused R = 8 sweep runs without metadata prefetch so it fits the
      // instance's 128-register budget
      if (fused) sweep_line_pk<R, PHASE_FIRST, CM_FUSED, (R >= 8 ? 1 : 64)>(J, A, line, p2_table);
      else sweep_line_pk<R, PHASE_FIRST, CM_LOAD8>(J, A, line, p2_table);
    } else if (phase == PHASE_MID) {
      sweep_line_pk<R, PHASE_MID, CM_LOAD8>(J, A, line, p2_table);
    } else {
      sweep_line_pk<R, PHASE_LAST, CM_LOAD8>(J, A, line, p2_table);
    }
  } else if constexpr (sizeof(SumT) == 2) {
    if (phase == PHASE_FIRST) sweep_line_pk<R, PHASE_FIRST, CM_LOAD>(J, A, line, p2_table);
    else if (phase == PHASE_MID) sweep_line_pk<R, PHASE_MID, CM_LOAD>(J, A, line, p2_table);
    else sweep_line_pk<R, PHASE_LAST, CM_LOAD>(J, A, line, p2_table);
  } else {
    if (phase == PHASE_FIRST) sweep_line<R, SumT, PHASE_FIRST>(J, A, line, p2_table);
    else if (phase == PHASE_MID) sweep_line<R, SumT, PHASE_MID>(J, A, line, p2_table);
    else sweep_line<R, SumT, PHASE_LAST>(J, A, line, p2_table);
  }
}

template <typename SumT, int VOL>
__device__ __forceinline__ void batch_sweep(const PathJob &J, const PathArgs &A, int line, int phase,
                                            bool fused, const int *p2_table) {
  if constexpr (sizeof(SumT) == 2) {
    if (J.R == 2) batch_sweep_r<2, SumT, VOL>(J, A, line, phase, fused, p2_table);
    else if (J.R == 4) batch_sweep_r<4, SumT, VOL>(J, A, line, phase, fused, p2_table);
    else batch_sweep_r<8, SumT, VOL>(J, A, line, phase, fused, p2_table);
  } else {
    if (J.R == 1) batch_sweep_r<1, SumT, VOL_U16>(J, A, line, phase, fused, p2_table);
    else if (J.R == 2) batch_sweep_r<2, SumT, VOL_U16>(J, A, line, phase, fused, p2_table);
    else if (J.R == 4) batch_sweep_r<4, SumT, VOL_U16>(J, A, line, phase, fused, p2_table);
    else batch_sweep_r<8, SumT, VOL_U16>(J, A, line, phase, fused, p2_table);
  }
}

constexpr int kBatchMaxWaves = 8;

// Waves per SIMD the register allocation must allow, the most each instance
// reaches without spilling: cost-volume instances 128 VGPRs (4 waves),
// recomputed census costs 256 (2).
template <typename SumT, int VOL>
constexpr int batch_min_waves() {
  return VOL == VOL_NONE ? 2 : 4;
}

template <typename SumT, int WORDS, int VOL>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(batch_min_waves<SumT, VOL>())))
void batch_kernel(BatchArgs B) {
  __shared__ int p2_table[256];
  __shared__ uint8_t tiles[kBatchMaxWaves][kBatchTileH * kBatchTileW];
  build_p2_table(p2_table, B.pa);
  __syncthreads();
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  uint8_t *tile = tiles[wave];
  const int total = B.item_hi;
  static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  // Completions of the current (phase, job) run are published together, before
  // the wave waits on anything else, so no wave waits while holding them.
  int cur_ph = -1, cur_job = -1;
  unsigned pending = 0;
  const int last_path = BPH_PATH0 + B.directions - 1;
  for (;;) {
    int got = 0;
    if (lane == 0) got = (int)atomicAdd(B.head, 1u);
    const int item = B.item_lo + __builtin_amdgcn_readlane(got, 0);
    if (item >= total) break;
    const int ph = batch_phase_of(B, item);
    const int nj = batch_njobs(B, ph);
    const int local = item - B.phase_begin[ph];
    const int job = __builtin_amdgcn_readfirstlane(batch_job_of(B, ph, local, nj));
    const int tile_idx = local - B.job_begin[(size_t)ph * B.stride + job];
    if (ph != cur_ph || job != cur_job) {
      if (pending) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0)
          __hip_atomic_fetch_add(B.done + (size_t)cur_ph * B.stride + cur_job, pending,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pending = 0;
      }
      bool ok = true;
      if (ph == BPH_COST || (ph == BPH_PATH0 && B.fused)) {
        const int pr = job >> 1;
        ok = batch_wait(B, BPH_CENSUS, 2 * pr) && batch_wait(B, BPH_CENSUS, 2 * pr + 1);
      } else if (ph == BPH_PATH0) {
        ok = batch_wait(B, BPH_COST, job);
      } else if (ph > BPH_PATH0 && ph <= last_path) {
        ok = batch_wait(B, ph - 1, job);
      } else if (ph == last_path + 1) {
        ok = batch_wait(B, last_path, 2 * job) && batch_wait(B, last_path, 2 * job + 1);
      }
      if (!ok) {
        if (lane == 0) atomicExch(B.error, 1);
        break;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      cur_ph = ph;
      cur_job = job;
    }
    if (ph == BPH_CENSUS) {
      const BatchImage I = B.images[job];
      const int r0 = tile_idx * B.census_rows, r1 = min(I.h, r0 + B.census_rows);
      batch_census<WORDS>(B, I, r0, r1, tile);
    } else if (ph == BPH_COST) {
      const CostJob J = B.cost_jobs[job];
      const int r0 = tile_idx * B.cost_rows, r1 = min(J.h, r0 + B.cost_rows);
      batch_cost<VOL == VOL_U16>(B, J, r0, r1);  // ZNCC costs only use uint16 volumes
    } else if (ph <= last_path) {
      const PathJob J = B.jobs[job];
      const int dir = ph - BPH_PATH0;
      PathArgs A = B.pa;
      A.dx = dirs[dir][0];
      A.dy = dirs[dir][1];
      A.cls = A.dy == 0 ? 0 : A.dx == 0 ? 1 : 2;
      const int phase = dir == 0 ? PHASE_FIRST : dir == B.directions - 1 ? PHASE_LAST : PHASE_MID;
      const bool fused = B.cmode == CM_FUSED;
      const int nlines = lines_of_dev(A.cls, J.w, J.h);
      const int l0 = tile_idx * B.lines_per_item, l1 = min(nlines, l0 + B.lines_per_item);
      for (int l = l0; l < l1; l++) batch_sweep<SumT, VOL>(J, A, l, phase, fused, p2_table);
    } else {
      const BatchPairDev P = B.pairs[job];
      const int r0 = tile_idx * B.lr_rows, r1 = min(P.h, r0 + B.lr_rows);
      batch_lr(B, P, r0, r1);
    }
    pending++;
  }
  if (pending) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0)
      __hip_atomic_fetch_add(B.done + (size_t)cur_ph * B.stride + cur_job, pending,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Phase-major runs: the census and L-R phases as regular grids over every
// image / pair of the batch (blockIdx.z), four waves per block, each wave one
// 64-column tile of a row band; the same per-tile work as the queue items.
constexpr int kGridWaves = 4, kGridLrRows = 8;

template <int WORDS>
static __global__ __launch_bounds__(256) void batch_census_grid(BatchArgs B, int z0) {
  __shared__ uint8_t tiles[kGridWaves][kBatchTileH * kBatchTileW];
  const BatchImage I = B.images[z0 + (int)blockIdx.z];
  const int wave = threadIdx.x >> 6;
  const int x0 = (blockIdx.x * kGridWaves + wave) * 64, r0 = blockIdx.y * kBatchCensusRows;
  if (x0 >= I.w || r0 >= I.h) return;  // wave-uniform: the tile lies outside this image
  batch_census_chunk<WORDS>(B, I, x0, r0, min(I.h, r0 + kBatchCensusRows), tiles[wave]);
}

static __global__ __launch_bounds__(256) void batch_lr_grid(BatchArgs B, int z0) {
  __shared__ unsigned sums[kGridWaves][8];
  const BatchPairDev P = B.pairs[z0 + (int)blockIdx.z];
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int x0 = (blockIdx.x * kGridWaves + wave) * 64, r0 = blockIdx.y * kGridLrRows;
  unsigned c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (x0 < P.w && r0 < P.h)
    batch_lr_span(B, P, r0, min(P.h, r0 + kGridLrRows), x0, min(P.w, x0 + 64), c);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int v = (int)c[i];
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (lane == 0) sums[wave][i] = (unsigned)v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    unsigned v = 0;
    for (int w = 0; w < kGridWaves; w++) v += sums[w][threadIdx.x];
    if (v) atomicAdd(P.stats + lr_field((int)threadIdx.x), (unsigned long long)v);
  }
}

static __global__ void fill_u32_kernel(unsigned *p, unsigned v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace sgm
