// Batched mode: many uneven stereo pairs in one persistent-kernel launch,
// optionally captured into a hipGraph, with tile shape and waves per path
// chosen by the autotuner (sgm_tune.hip).  The reference has no batching:
// it matches pairs sequentially (DenseReconstruction.java:64-72).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "sgm_batch_internal.h"

using namespace sgm;

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves one device arena into 256-byte aligned sub-buffers.
struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off = align_up(off + std::max<size_t>(bytes, 8), 256);
    return o;
  }
};

}  // namespace

namespace sgm {

int batch_lines_per_item(sgm_context *ctx) { return ctx->batch_lines_per_item; }

// The persistent kernel instance of a batch.
const void *batch_kernel_of(const sgm_batch *b) {
  if (b->sum_bytes == 2) {
    if (b->recompute) return (const void *)&batch_kernel<uint16_t, 1, VOL_NONE>;
    if (b->cost8) {
      if (b->words == 1) return (const void *)&batch_kernel<uint16_t, 1, VOL_U8>;
      return (const void *)&batch_kernel<uint16_t, 2, VOL_U8>;
    }
    if (b->words == 1) return (const void *)&batch_kernel<uint16_t, 1, VOL_U16>;
    return (const void *)&batch_kernel<uint16_t, 2, VOL_U16>;
  }
  if (b->words == 1) return (const void *)&batch_kernel<int32_t, 1, VOL_U16>;
  return (const void *)&batch_kernel<int32_t, 2, VOL_U16>;
}

}  // namespace sgm

static int validate_pair(const sgm_params *p, const sgm_pair &pr, int flags) {
  sgm_params q = *p;
  q.disparities = pr.disparities;
  q.min_disparity = pr.min_disparity;
  int rc = validate_params(&q, pr.width, pr.height);
  if (rc) return rc;
  if (!pr.left || !pr.right || !pr.checked)
    return fail(SGM_ERR_INVALID_ARGUMENT, "pair needs left, right and checked buffers");
  if ((pr.valid_left == nullptr) != (pr.valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  if (lanes_group(pr.disparities) > 8)
    return fail(SGM_ERR_INVALID_ARGUMENT,
                "batched mode supports up to 512 disparities per pair (got %d)", pr.disparities);
  if ((int64_t)pr.width * pr.height * 8 >= (int64_t)1 << 31)
    return fail(SGM_ERR_INVALID_ARGUMENT, "pair too large for the batched mode");
  (void)flags;
  return SGM_OK;
}

// Upper bound of the phase-major group tables (build_groups): the 2 * count
// path jobs once, four line-prefix arrays per group, 256-byte alignment.  A
// group (one lane group R of one sub-batch) holds at least one pair, so there
// are at most `count` groups.
static size_t group_table_bound(int count) {
  const size_t G = (size_t)count;
  return align_up(2 * (size_t)count * sizeof(PathJob) + 4 * 4 * (2 * (size_t)count + G) +
                      G * 5 * 256,
                  256);
}

// Phase-major tables: the pairs' path jobs grouped by lane group R, and per
// group and line class the first line of each job (prefix sums of the lines),
// in one device allocation.
static int build_groups(sgm_batch *b, const std::vector<PathJob> &jobs, const sgm_pair *pairs,
                        int count) {
  std::vector<int> rs;
  for (const PathJob &J : jobs)
    if (std::find(rs.begin(), rs.end(), J.R) == rs.end()) rs.push_back(J.R);
  std::sort(rs.begin(), rs.end());
  const int sb = b->sum_bytes;
  std::vector<uint8_t> host;
  struct Pending {
    size_t jobs_off, lb_off[4], multi_off = 0, last_off = 0;
  };
  std::vector<Pending> offs;
  b->groups.clear();
  // sub-batch by sub-batch (path-cost planes reuse one scratch region), and
  // within a sub-batch one group per lane group R
  for (int sub = 0; sub < b->nsub; sub++)
  for (int R : rs) {
    sgm_batch::Group g;
    g.R = R;
    g.sub = sub;
    std::vector<PathJob> gj;
    int64_t volume = 0;
    bool aligned = true;
    for (size_t j = 0; j < jobs.size(); j++) {
      const PathJob &J = jobs[j];
      if (J.R == R && b->pair_sub[j / 2] == sub) {
        gj.push_back(J);
        volume += (int64_t)J.w * J.h * J.Dp * (b->cost8 ? 3 : 4);
        aligned &= J.Dp % 64 == 0;
      }
    }
    if (gj.empty()) continue;
    g.njobs = (int)gj.size();
    // stream the volumes past the caches as the plain path does: whole 64-byte
    // lines only, volumes far beyond the 256 MB MALL
    g.nt = sb == 2 && !b->recompute && aligned && R <= 8 && volume > ((int64_t)768 << 20);
    Pending pd;
    pd.jobs_off = align_up(host.size(), 256);
    host.resize(pd.jobs_off + gj.size() * sizeof(PathJob));
    memcpy(host.data() + pd.jobs_off, gj.data(), gj.size() * sizeof(PathJob));
    for (int cls = 0; cls < 4; cls++) {
      // line classes 0..2, and 3: the row segments of the census cost pass
      std::vector<int> lb(gj.size() + 1, 0);
      for (size_t j = 0; j < gj.size(); j++)
        lb[j + 1] = lb[j] + (cls == 3 ? gj[j].h * cost_segments(gj[j].w)
                                      : lines_of(cls, gj[j].w, gj[j].h));
      g.total_lines[cls] = lb.back();
      pd.lb_off[cls] = align_up(host.size(), 256);
      host.resize(pd.lb_off[cls] + lb.size() * sizeof(int));
      memcpy(host.data() + pd.lb_off[cls], lb.data(), lb.size() * sizeof(int));
    }
    if (b->planes && !b->external) {
      // longest-first orders (as the plain path's): the directions but the
      // last in one launch, each over every job's lines of its class, then
      // the last direction alone
      static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
      const int P = b->params.directions;
      int maxlen = 1;
      for (const PathJob &J : gj) maxlen = std::max(maxlen, std::max(J.w, J.h));
      auto lengths = [&](int dir, std::vector<int> &len) {
        const int dx = dirs[dir][0], dy = dirs[dir][1];
        const int cls = dy == 0 ? 0 : dx == 0 ? 1 : 2;
        for (const PathJob &J : gj)
          for (int l = 0; l < lines_of(cls, J.w, J.h); l++) len.push_back(line_length(dx, dy, l, J.w, J.h));
      };
      std::vector<int> len, order;
      for (int d = 0; d < P - 1; d++) lengths(d, len);  // the launch's line numbering
      longest_first(len, maxlen, order);
      pd.multi_off = align_up(host.size(), 256);
      host.resize(pd.multi_off + order.size() * sizeof(int));
      memcpy(host.data() + pd.multi_off, order.data(), order.size() * sizeof(int));
      len.clear();
      lengths(P - 1, len);
      longest_first(len, maxlen, order);
      pd.last_off = align_up(host.size(), 256);
      host.resize(pd.last_off + order.size() * sizeof(int));
      memcpy(host.data() + pd.last_off, order.data(), order.size() * sizeof(int));
    }
    offs.push_back(pd);
    b->groups.push_back(g);
  }
  (void)pairs;
  uint8_t *base;
  if (b->external) {
    // carved batches: the room batch_table_bytes reserves after the job tables
    if (b->groups.size() > (size_t)count || host.size() > group_table_bound(count))
      return fail(SGM_ERR_NO_PLAN, "batch group tables exceed their reserved room");
    base = (uint8_t *)b->tables + align_up(b->tables_bytes, 256);
  } else {
    int rc = context_alloc(b->ctx, &b->group_tables, host.size(), "batch group tables");
    if (rc) return rc;
    b->group_tables_bytes = host.size();
    base = (uint8_t *)b->group_tables;
  }
  hipError_t e = hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch group table upload");
  for (size_t i = 0; i < b->groups.size(); i++) {
    b->groups[i].jobs = (const PathJob *)(base + offs[i].jobs_off);
    for (int cls = 0; cls < 4; cls++) b->groups[i].line_begin[cls] = (const int *)(base + offs[i].lb_off[cls]);
    if (offs[i].multi_off) {
      b->groups[i].order_multi = (const int *)(base + offs[i].multi_off);
      b->groups[i].order_last = (const int *)(base + offs[i].last_off);
    }
  }
  b->phased = true;
  return SGM_OK;
}

// One launch of the work-queue kernel over items [lo, hi) with its own head.
static int queue_launch(sgm_batch *b, int lo, int hi, unsigned *head, const char *what) {
  if (hi <= lo) return SGM_OK;
  BatchArgs A = b->args;
  A.item_lo = lo;
  A.item_hi = hi;
  A.head = head;
  A.phased = 1;
  ProfileScope ps(b->ctx, what);
  void *kargs[] = {(void *)&A};
  hipError_t e = hipLaunchKernel(batch_kernel_of(b), dim3((unsigned)b->blocks),
                                 dim3(b->waves_per_block * 64), kargs, 0, b->stream);
  if (e == hipSuccess) e = hipGetLastError();
  return e == hipSuccess ? SGM_OK : hip_fail(e, "batch kernel launch");
}

// Census (lr = false) or L-R (lr = true) phase of a phase-major run as one
// regular grid over every image or pair: 64-column tiles of row bands, four
// per block (a queue item is a whole row band, too few waves for a few pairs).
static int grid_launch(sgm_batch *b, bool lr) {
  int w = 1, h = 1;
  for (const sgm_pair &pr : b->pairs) {
    w = std::max(w, pr.width);
    h = std::max(h, pr.height);
  }
  const int rows = lr ? kGridLrRows : kBatchCensusRows;
  const int n = (int)b->pairs.size() * (lr ? 1 : 2);
  constexpr int kMaxZ = 65535;  // grid z limit: images / pairs in slices
  ProfileScope ps(b->ctx, lr ? "batch lr" : "batch census");
  for (int z0 = 0; z0 < n; z0 += kMaxZ) {
    dim3 grid((unsigned)((w + 64 * kGridWaves - 1) / (64 * kGridWaves)),
              (unsigned)((h + rows - 1) / rows), (unsigned)std::min(kMaxZ, n - z0));
    if (lr) batch_lr_grid<<<grid, 64 * kGridWaves, 0, b->stream>>>(b->args, z0);
    else if (b->words == 1) batch_census_grid<1><<<grid, 64 * kGridWaves, 0, b->stream>>>(b->args, z0);
    else batch_census_grid<2><<<grid, 64 * kGridWaves, 0, b->stream>>>(b->args, z0);
    int rc = launch_check(lr ? "batch lr grid" : "batch census grid");
    if (rc) return rc;
  }
  return SGM_OK;
}

// Phase-major run: the census grid (and the cost items), the sweeps direction
// by direction and lane group by lane group with the specialised sweep kernel,
// the L-R grid.
static int phased_enqueue(sgm_batch *b) {
  sgm_context *ctx = b->ctx;
  const sgm_params *p = &b->params;
  const std::vector<int> &pb = b->phase_begin;
  int rc;
  if ((rc = grid_launch(b, false))) return rc;
  // cost items (non-fused costs) after the census grid: their waits on the
  // census phase pass at once, an earlier launch having run it
  if ((rc = queue_launch(b, pb[BPH_COST], pb[BPH_PATH0], b->args.head, "batch cost"))) return rc;
  const int sb = b->sum_bytes;
  if (b->planes) {
    // path-cost planes, sub-batch after sub-batch (one scratch region): per
    // lane group the census cost pass, every direction but the last in one
    // launch (planes 0..6), the last sweep over the planes.  With a side
    // stream, the lane groups of a sub-batch alternate between the two
    // streams, so one group's launches fill the tail of the other's.
    int cur_sub = -1, gi = 0;
    auto sync_streams = [&](bool fork) -> int {
      if (!b->side) return SGM_OK;
      hipError_t e = hipSuccess;
      if (cur_sub >= 0) {  // join: the main stream waits for the side stream's groups
        e = hipEventRecord(b->join, b->side);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join, 0);
      }
      if (fork && e == hipSuccess) {  // fork: the side stream starts after the main stream's work
        e = hipEventRecord(b->fork, b->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->fork, 0);
      }
      return e == hipSuccess ? SGM_OK : hip_fail(e, "batch stream fork/join");
    };
    for (const sgm_batch::Group &g : b->groups) {
      if (g.sub != cur_sub) {
        if ((rc = sync_streams(true))) return rc;
        cur_sub = g.sub;
        gi = 0;
      }
      hipStream_t st = (b->side && (gi++ & 1)) ? b->side : b->stream;
      const int nt = g.nt ? MODE_NT : 0;
      auto args = [&](int dir) {
        PathArgs A = path_args(p, dir, g.njobs, 0, 0);
        A.total_lines = g.total_lines[A.cls];
        return A;
      };
      PathJobs pj;
      memset(&pj, 0, sizeof pj);
      pj.table = g.jobs;
      PathArgs Ac = args(0);
      Ac.total_lines = g.total_lines[3];
      pj.line_begin = g.line_begin[3];  // row segments of the cost pass
      if ((rc = launch_path(ctx, g.R, sb, 2 | nt, pj, Ac, PHASE_COST, path_waves(ctx, g.R), st)))
        return rc;
      MultiDir M;
      memset(&M, 0, sizeof M);
      M.ndirs = p->directions - 1;
      for (int k = 0; k < M.ndirs; k++) {
        const PathArgs Ak = args(k);
        M.dx[k] = Ak.dx;
        M.dy[k] = Ak.dy;
        M.cls[k] = Ak.cls;
        M.begin[k] = M.total_lines;
        M.total_lines += Ak.total_lines;
      }
      for (int cls = 0; cls < 3; cls++) M.line_begin[cls] = g.line_begin[cls];
      M.njobs = g.njobs;
      M.plane0 = 0;
      M.order = g.order_multi;
      if ((rc = launch_path_multi(ctx, g.R, 4 | nt | MODE_PLANE, pj, args(0), M,
                                  path_waves(ctx, g.R), st)))
        return rc;
      PathArgs AL = args(p->directions - 1);
      AL.order = g.order_last;
      pj.line_begin = g.line_begin[AL.cls];
      if ((rc = launch_path(ctx, g.R, sb, 4 | nt | MODE_PLANE, pj, AL, PHASE_LAST,
                            path_waves(ctx, g.R), st)))
        return rc;
    }
    if ((rc = sync_streams(false))) return rc;
    return grid_launch(b, true);
  }
  const int base_mode = b->recompute ? 3 : sb == 2 ? (b->cost8 ? 4 : 1) : 0;
  for (int dir = 0; dir < p->directions; dir++) {
    const int phase = dir == 0 ? PHASE_FIRST : dir == p->directions - 1 ? PHASE_LAST : PHASE_MID;
    for (const sgm_batch::Group &g : b->groups) {
      PathArgs A = path_args(p, dir, g.njobs, 0, b->ctx->keep_volumes ? 1 : 0);
      A.total_lines = g.total_lines[A.cls];
      PathJobs pj;
      memset(&pj, 0, sizeof pj);
      pj.table = g.jobs;
      pj.line_begin = g.line_begin[A.cls];
      int mode = (dir == 0 && b->fused) ? 2 : base_mode;
      if (g.nt) mode |= MODE_NT;
      if ((rc = launch_path(ctx, g.R, sb, mode, pj, A, phase, path_waves(ctx, g.R), b->stream)))
        return rc;
    }
  }
  return grid_launch(b, true);
}

static int batch_enqueue(sgm_batch *b) {
  sgm_context *ctx = b->ctx;
  hipStream_t st = b->stream;
  int rc;
  if (!(b->flags & SGM_BATCH_DEVICE)) {
    for (size_t i = 0; i < b->pairs.size(); i++) {
      const sgm_pair &pr = b->pairs[i];
      const size_t n = (size_t)pr.width * pr.height;
      const PairSlots &s = b->slots[i];
      hipError_t e = hipMemcpyAsync(s.gray[0], pr.left, n, hipMemcpyHostToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(s.gray[1], pr.right, n, hipMemcpyHostToDevice, st);
      if (e == hipSuccess && pr.valid_left) {
        e = hipMemcpyAsync(s.valid[0], pr.valid_left, n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(s.valid[1], pr.valid_right, n, hipMemcpyHostToDevice, st);
      }
      if (e != hipSuccess) return hip_fail(e, "batch upload");
    }
  }
  if (b->gate_wait) {
    hipError_t e = hipStreamWaitEvent(st, b->gate_wait, 0);
    if (e != hipSuccess) return hip_fail(e, "batch gate wait");
  }
  {
    const size_t words = b->control_words;
    fill_u32_kernel<<<(unsigned)((words + 255) / 256), 256, 0, st>>>(b->d_control, 0u, words);
    if (hipGetLastError() != hipSuccess) return fail(SGM_ERR_HIP, "batch reset launch failed");
  }
  const int threads = b->waves_per_block * 64;
  const unsigned blocks = (unsigned)b->blocks;
  if (b->phased) {
    if ((rc = phased_enqueue(b))) return rc;
  } else {
    ProfileScope ps(ctx, "batch");
    void *kargs[] = {(void *)&b->args};
    hipError_t e = hipLaunchKernel(batch_kernel_of(b), dim3(blocks), dim3(threads), kargs, 0, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
  }
  if (b->gate_done) {
    hipError_t e = hipEventRecord(b->gate_done, st);
    if (e != hipSuccess) return hip_fail(e, "batch gate record");
  }
  if (!(b->flags & SGM_BATCH_DEVICE)) {
    for (size_t i = 0; i < b->pairs.size(); i++) {
      const sgm_pair &pr = b->pairs[i];
      const size_t n = (size_t)pr.width * pr.height;
      const PairSlots &s = b->slots[i];
      hipError_t e = hipMemcpyAsync(pr.checked, s.checked, n * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess && pr.dl) e = hipMemcpyAsync(pr.dl, s.disp[0], n * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess && pr.dr) e = hipMemcpyAsync(pr.dr, s.disp[1], n * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess && pr.reasons)
        e = hipMemcpyAsync(pr.reasons, s.reasons, n, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess && pr.boundary)
        e = hipMemcpyAsync(pr.boundary, s.boundary, n, hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return hip_fail(e, "batch download");
    }
  }
  (void)rc;
  return SGM_OK;
}

extern "C" int sgm_batch_create(sgm_context *ctx, const sgm_params *params, const sgm_pair *pairs,
                                int count, int flags, sgm_batch **out) {
  if (!ctx || !params || !pairs || count <= 0 || !out)
    return fail(SGM_ERR_INVALID_ARGUMENT, "bad batch arguments");
  *out = nullptr;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  return batch_create(ctx, params, pairs, count, flags, nullptr, 0, ctx->stream, out, -1,
                      (flags & SGM_BATCH_PHASED) ? -1 : 0);
}

namespace sgm {

bool batch_cost8(const sgm_params *params) {
  // uint8 cost volume in a batch: the census-family costs without ZNCC (the
  // ZNCC window arithmetic stays in the uint16-volume kernel instance)
  return cost_u8(params) && params->cost != SGM_COST_ZNCC;
}

bool batch_recompute_possible(const sgm_params *params, bool keep_volumes) {
  // every pair of a batch has R <= 8 (validate_pair): 1-word census with
  // uint16 aggregates can recompute its cost in the sweeps
  return !keep_volumes && params->cost == SGM_COST_CENSUS && words_of(params) == 1 &&
         aggregate_bytes(params) == 2;
}

// Path-cost planes in a phase-major batch: the plain plan's plane conditions
// with the census cost fused into the first sweep (1-word Census, 16-bit sums).
bool batch_planes_possible(const sgm_context *ctx, const sgm_params *params, int flags) {
  return (flags & SGM_BATCH_PHASED) && params->cost == SGM_COST_CENSUS && words_of(params) == 1 &&
         aggregate_bytes(params) == 2 && params->directions == kPlanes + 1 &&
         !ctx->keep_volumes &&
         params->window_w * params->window_h - 1 + effective_p2(params) <= 255 &&
         (ctx->sweeps == SGM_SWEEPS_AUTO || ctx->sweeps == SGM_SWEEPS_PLANES);
}

// Scratch bytes of one pair under path-cost planes (carved as batch_create
// carves them): the uint8 cost volume and the plane block of both matching
// directions.
size_t batch_plane_bytes(const sgm_pair &pr) {
  const size_t n = (size_t)pr.width * pr.height, Dp = padded_disparities(pr.disparities);
  return 2 * (align_up(n * Dp, 256) + align_up(n * Dp * kPlanes, 256));
}

size_t batch_arena_bytes(const sgm_params *params, const sgm_pair *pairs, int count, int flags,
                         bool recompute, bool planes) {
  // must mirror the carving in batch_create; with planes every pair in one
  // sub-batch (scratch for all of them)
  const bool host_io = !(flags & SGM_BATCH_DEVICE);
  const int words = words_of(params), sb = aggregate_bytes(params);
  const bool zncc = params->cost == SGM_COST_ZNCC || params->cost == SGM_COST_ZNCC_FP32;
  Carver cv;
  const int nph = BPH_PATH0 + params->directions + 1, stride = 2 * count + 1;
  cv.take(4);
  cv.take(4);
  cv.take((size_t)nph * stride * 4);
  cv.take((size_t)count * 16 * 8);
  for (int i = 0; i < count; i++) {
    const sgm_pair &pr = pairs[i];
    const size_t n = (size_t)pr.width * pr.height, Dp = padded_disparities(pr.disparities);
    for (int d = 0; d < 2; d++) {
      if (host_io) cv.take(n);
      if (host_io && pr.valid_left) cv.take(n);
      cv.take(n * 8 * words);
      cv.take(n);
      cv.take(n * 2);
      if (zncc) {
        cv.take(n * 4);
        cv.take(n * 4);
      }
      if (!recompute && !planes) cv.take(n * Dp * (batch_cost8(params) ? 1 : 2));
      if (!planes) cv.take(n * Dp * sb);
      const bool own_disp = host_io || (d == 0 ? pr.dl == nullptr : pr.dr == nullptr);
      if (own_disp) cv.take(n * 4);
    }
    cv.take(n);
    cv.take(n);
    cv.take(n);
    if (host_io) cv.take(n * 4);
    if (host_io && pr.reasons) cv.take(n);
    if (host_io && pr.boundary) cv.take(n);
  }
  if (planes)
    for (int i = 0; i < count; i++) cv.off += batch_plane_bytes(pairs[i]);
  return cv.off;
}

static size_t base_table_bytes(int count, const sgm_params *params) {
  const int nph = BPH_PATH0 + params->directions + 1, stride = 2 * count + 1;
  Carver tc;
  tc.take(2 * count * sizeof(BatchImage));
  tc.take(2 * count * sizeof(PathJob));
  tc.take(2 * count * sizeof(CostJob));
  tc.take(count * sizeof(BatchPairDev));
  tc.take((nph + 1) * sizeof(int));
  tc.take((size_t)nph * stride * sizeof(int));
  return tc.off;
}

size_t batch_table_bytes(int count, const sgm_params *params) {
  // job tables, then (carved batches) room for the phase-major group tables
  return align_up(base_table_bytes(count, params), 256) + group_table_bound(count);
}

// Creates a batch.  With `region` non-null the arena and tables are carved from
// that caller-owned device region (serving pool) instead of being allocated.
int batch_create(sgm_context *ctx, const sgm_params *params, const sgm_pair *pairs, int count,
                 int flags, void *region, size_t region_bytes, hipStream_t stream,
                 sgm_batch **out, int recompute, int planes) {
  for (int i = 0; i < count; i++) {
    int rc = validate_pair(params, pairs[i], flags);
    if (rc) return rc;
  }
  const bool host_io = !(flags & SGM_BATCH_DEVICE);
  sgm_batch *b = new sgm_batch();
  b->ctx = ctx;
  b->stream = stream;
  b->params = *params;
  b->flags = flags;
  b->pairs.assign(pairs, pairs + count);
  b->words = words_of(params);
  b->sum_bytes = aggregate_bytes(params);
  const bool zncc = params->cost == SGM_COST_ZNCC || params->cost == SGM_COST_ZNCC_FP32;
  // Census cost: recomputed in the sweeps (no cost volume) when forced by the
  // policy or when the batch with resident cost volumes does not fit the
  // memory at hand; otherwise fused into the first sweep and kept resident.
  const bool possible = batch_recompute_possible(params, ctx->keep_volumes);
  if (recompute < 0) {
    if (!possible || ctx->volume_policy == SGM_VOLUME_RESIDENT) {
      recompute = 0;
    } else if (ctx->volume_policy == SGM_VOLUME_RECOMPUTE) {
      recompute = 1;
    } else {
      const int64_t avail = region ? (int64_t)region_bytes : ctx->budget - ctx->allocated;
      const int64_t need = (int64_t)(batch_arena_bytes(params, pairs, count, flags, false, false) +
                                     batch_table_bytes(count, params));
      recompute = need > avail ? 1 : 0;
    }
  }
  b->recompute = possible && recompute != 0;
  b->cost8 = !b->recompute && batch_cost8(params);
  b->fused = !b->recompute && params->cost == SGM_COST_CENSUS && b->words == 1 &&
             b->sum_bytes == 2;
  // Path-cost planes (phase-major runs of the fused census cost): no resident
  // cost or aggregate volume per pair; the cost volumes and plane blocks of a
  // sub-batch of pairs live in one scratch region that the sub-batches reuse
  // in turn.  planes = 1: all pairs in one sub-batch; -1: sub-batches as large
  // as the memory at hand allows (off when not even the largest pair fits).
  b->nsub = 1;
  b->pair_sub.assign(count, 0);
  if (planes != 0 && b->fused && batch_planes_possible(ctx, params, flags)) {
    size_t total = 0, largest = 0;
    for (int i = 0; i < count; i++) {
      total += batch_plane_bytes(pairs[i]);
      largest = std::max(largest, batch_plane_bytes(pairs[i]));
    }
    size_t cap = total;
    if (planes < 0) {
      const int64_t avail = region ? (int64_t)region_bytes : ctx->budget - ctx->allocated;
      const int64_t rest = avail - (int64_t)(batch_arena_bytes(params, pairs, count, flags, false,
                                                               true) - total) -
                           (int64_t)batch_table_bytes(count, params);
      cap = rest > 0 ? std::min(total, (size_t)rest) : 0;
    }
    if (cap >= largest) {
      b->planes = true;
      size_t used = 0;
      int sub = 0;
      for (int i = 0; i < count; i++) {
        const size_t need = batch_plane_bytes(pairs[i]);
        if (used + need > cap) {
          sub++;
          used = 0;
        }
        b->pair_sub[i] = sub;
        used += need;
        b->scratch_bytes = std::max(b->scratch_bytes, used);
      }
      b->nsub = sub + 1;
    }
  }
  b->waves_per_block = ctx->batch_waves_per_block;
  b->lines_per_item = ctx->batch_lines_per_item;
  {
    // workgroups per CU: as many as stay resident (occupancy of the chosen
    // instance) unless the tuned shape asks for fewer
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, batch_kernel_of(b),
                                                     b->waves_per_block * 64, 0) != hipSuccess ||
        resident < 1) {
      (void)hipGetLastError();
      resident = 1;
    }
    const int per_cu = ctx->batch_blocks_per_cu > 0 ? std::min(ctx->batch_blocks_per_cu, resident)
                                                    : resident;
    b->blocks = ctx->info.compute_units * per_cu;
  }

  // ---- arena layout
  Carver cv;
  struct Off {
    size_t gray[2], valid[2], desc[2], mask[2], gm[2], wsum[2], wsum2[2], cost[2], sum[2], disp[2];
    size_t code, rtmp, btmp, checked, reasons, boundary;
  };
  std::vector<Off> offs(count);
  const int nph = BPH_PATH0 + params->directions + 1;
  const int stride = 2 * count + 1;
  const size_t done_words = (size_t)nph * stride;
  const size_t ctrl_head = cv.take(4), ctrl_error = cv.take(4);
  const size_t ctrl_done = cv.take(done_words * 4);
  const size_t ctrl_stats = cv.take((size_t)count * 16 * 8);
  const size_t ctrl_end = cv.off;
  for (int i = 0; i < count; i++) {
    const sgm_pair &pr = pairs[i];
    const size_t n = (size_t)pr.width * pr.height, Dp = padded_disparities(pr.disparities);
    Off &o = offs[i];
    for (int d = 0; d < 2; d++) {
      o.gray[d] = host_io ? cv.take(n) : 0;
      o.valid[d] = (host_io && pr.valid_left) ? cv.take(n) : 0;
      o.desc[d] = cv.take(n * 8 * b->words);
      o.mask[d] = cv.take(n);
      o.gm[d] = cv.take(n * 2);
      o.wsum[d] = zncc ? cv.take(n * 4) : 0;
      o.wsum2[d] = zncc ? cv.take(n * 4) : 0;
      o.cost[d] = b->recompute || b->planes ? 0 : cv.take(n * Dp * (b->cost8 ? 1 : 2));
      o.sum[d] = b->planes ? 0 : cv.take(n * Dp * b->sum_bytes);
      const bool own_disp = host_io || (d == 0 ? pr.dl == nullptr : pr.dr == nullptr);
      o.disp[d] = own_disp ? cv.take(n * 4) : 0;
    }
    o.code = cv.take(n);
    o.rtmp = cv.take(n);
    o.btmp = cv.take(n);
    o.checked = host_io ? cv.take(n * 4) : 0;
    o.reasons = (host_io && pr.reasons) ? cv.take(n) : 0;
    o.boundary = (host_io && pr.boundary) ? cv.take(n) : 0;
  }
  if (b->planes) {
    // the scratch region, carved again for each sub-batch's pairs
    const size_t scratch = cv.off;
    cv.off += b->scratch_bytes;
    size_t at = scratch;
    for (int i = 0; i < count; i++) {
      if (i > 0 && b->pair_sub[i] != b->pair_sub[i - 1]) at = scratch;
      const size_t n = (size_t)pairs[i].width * pairs[i].height;
      const size_t Dp = padded_disparities(pairs[i].disparities);
      Carver sc;
      sc.off = at;
      for (int d = 0; d < 2; d++) {
        offs[i].cost[d] = sc.take(n * Dp);
        offs[i].sum[d] = sc.take(n * Dp * kPlanes);
      }
      at = sc.off;
    }
  }
  b->arena_bytes = cv.off;
  int rc = SGM_OK;
  const size_t table_bytes = batch_table_bytes(count, params);
  if (region) {
    if (b->arena_bytes + table_bytes > region_bytes) {
      delete b;
      return fail(SGM_ERR_NO_PLAN, "batch region too small");
    }
    b->arena = region;
    b->tables = (uint8_t *)region + b->arena_bytes;
    b->external = true;
  } else {
    if ((int64_t)(b->arena_bytes + table_bytes) > ctx->budget - ctx->allocated) {
      const size_t need = b->arena_bytes;
      delete b;
      return fail(SGM_ERR_NO_PLAN, "batch needs %zu MiB of device memory; budget left %lld MiB",
                  need >> 20, (long long)((ctx->budget - ctx->allocated) >> 20));
    }
    rc = context_alloc(ctx, &b->arena, b->arena_bytes, "batch arena");
    if (rc) {
      delete b;
      return rc;
    }
  }
  uint8_t *base = (uint8_t *)b->arena;
  b->d_control = (unsigned *)(base + ctrl_head);
  b->control_words = (ctrl_end - ctrl_head) / 4;
  b->d_stats = (unsigned long long *)(base + ctrl_stats);

  // ---- job tables
  std::vector<BatchImage> images(2 * count);
  std::vector<PathJob> jobs(2 * count);
  std::vector<CostJob> cjobs(2 * count);
  std::vector<BatchPairDev> pdev(count);
  b->slots.resize(count);
  const int census_rows = kBatchCensusRows, cost_rows = 4, lr_rows = 16;
  std::vector<int> job_begin((size_t)nph * stride, 0), phase_begin(nph + 1, 0);
  auto P8 = [&](size_t o) { return base + o; };
  for (int i = 0; i < count; i++) {
    const sgm_pair &pr = pairs[i];
    const int w = pr.width, h = pr.height, D = pr.disparities, Dp = padded_disparities(D);
    const Off &o = offs[i];
    PairSlots &s = b->slots[i];
    for (int d = 0; d < 2; d++) {
      s.gray[d] = host_io ? P8(o.gray[d]) : (uint8_t *)(d == 0 ? pr.left : pr.right);
      s.valid[d] = pr.valid_left ? (host_io ? P8(o.valid[d])
                                            : (uint8_t *)(d == 0 ? pr.valid_left : pr.valid_right))
                                 : nullptr;
      s.disp[d] = o.disp[d] ? (float *)P8(o.disp[d]) : (d == 0 ? pr.dl : pr.dr);
    }
    s.checked = host_io ? (float *)P8(o.checked) : pr.checked;
    s.reasons = host_io ? (o.reasons ? P8(o.reasons) : nullptr) : pr.reasons;
    s.boundary = host_io ? (o.boundary ? P8(o.boundary) : nullptr) : pr.boundary;
    for (int d = 0; d < 2; d++) {
      BatchImage &I = images[2 * i + d];
      I.gray = s.gray[d];
      I.valid = s.valid[d];
      I.desc = (uint64_t *)P8(o.desc[d]);
      I.mask = P8(o.mask[d]);
      I.gm = (uint16_t *)P8(o.gm[d]);
      I.wsum = zncc ? (int32_t *)P8(o.wsum[d]) : nullptr;
      I.wsum2 = zncc ? (int32_t *)P8(o.wsum2[d]) : nullptr;
      I.reasons_init = d == 0 ? P8(o.rtmp) : nullptr;
      I.boundary_init = d == 0 ? P8(o.btmp) : nullptr;
      I.w = w;
      I.h = h;
    }
    for (int d = 0; d < 2; d++) {
      const int od = 1 - d;
      PathJob &J = jobs[2 * i + d];
      memset(&J, 0, sizeof J);
      J.cost = b->recompute ? nullptr : (const uint16_t *)P8(o.cost[d]);
      J.sum = P8(o.sum[d]);
      J.mask = P8(o.mask[d]);
      J.gray = s.gray[d];
      J.gm = (const uint16_t *)P8(o.gm[d]);
      J.desc_a = (const uint64_t *)P8(o.desc[d]);
      J.desc_b = (const uint64_t *)P8(o.desc[od]);
      J.mask_b = P8(o.mask[od]);
      J.cost_out = b->recompute ? nullptr : (uint16_t *)P8(o.cost[d]);
      J.disp = s.disp[d];
      J.reasons = d == 0 ? P8(o.rtmp) : nullptr;
      J.boundary = d == 0 ? P8(o.btmp) : nullptr;
      J.code = d == 0 ? P8(o.code) : nullptr;
      J.w = w;
      J.h = h;
      J.D = D;
      J.Dp = Dp;
      J.minD = pr.min_disparity;
      J.rx = params->window_w / 2;
      J.ry = params->window_h / 2;
      J.reverse = d;
      J.R = b->sum_bytes == 2 ? std::max(2, lanes_group(D)) : lanes_group(D);
      CostJob &C = cjobs[2 * i + d];
      memset(&C, 0, sizeof C);
      C.desc_a = J.desc_a;
      C.desc_b = J.desc_b;
      C.mask_a = J.mask;
      C.mask_b = J.mask_b;
      C.gray_a = s.gray[d];
      C.gray_b = s.gray[od];
      C.sum_a = images[2 * i + d].wsum;
      C.sum2_a = images[2 * i + d].wsum2;
      C.sum_b = images[2 * i + od].wsum;
      C.sum2_b = images[2 * i + od].wsum2;
      C.cost = b->recompute ? nullptr : (uint16_t *)P8(o.cost[d]);
      C.w = w;
      C.h = h;
      C.D = D;
      C.Dp = Dp;
      C.minD = pr.min_disparity;
      C.reverse = d;
    }
    BatchPairDev &PD = pdev[i];
    PD.dl = s.disp[0];
    PD.dr = s.disp[1];
    PD.code = P8(o.code);
    PD.reasons_tmp = P8(o.rtmp);
    PD.boundary_tmp = P8(o.btmp);
    PD.checked = s.checked;
    PD.reasons = s.reasons;
    PD.boundary = s.boundary;
    PD.stats = b->d_stats + (size_t)i * 16;
    PD.w = w;
    PD.h = h;
  }
  // items per (phase, job): census per image, cost / paths per direction, L-R per pair
  int total = 0;
  for (int ph = 0; ph < nph; ph++) {
    phase_begin[ph] = total;
    const int nj = ph == nph - 1 ? count : 2 * count;
    int acc = 0;
    for (int j = 0; j < nj; j++) {
      job_begin[(size_t)ph * stride + j] = acc;
      const sgm_pair &pr = pairs[ph == nph - 1 ? j : j / 2];
      const int w = pr.width, h = pr.height;
      int items;
      if (ph == BPH_CENSUS) items = (h + census_rows - 1) / census_rows;
      else if (ph == BPH_COST) items = (b->fused || b->recompute) ? 0 : (h + cost_rows - 1) / cost_rows;
      else if (ph == nph - 1) items = (h + lr_rows - 1) / lr_rows;
      else {
        static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
        const int dir = ph - BPH_PATH0;
        const int cls = dirs[dir][1] == 0 ? 0 : dirs[dir][0] == 0 ? 1 : 2;
        items = (lines_of(cls, w, h) + b->lines_per_item - 1) / b->lines_per_item;
      }
      acc += items;
    }
    job_begin[(size_t)ph * stride + nj] = acc;
    total += acc;
  }
  phase_begin[nph] = total;
  b->total_items = total;
  b->phase_begin = phase_begin;

  // ---- upload tables (one small allocation)
  Carver tc;
  const size_t t_img = tc.take(images.size() * sizeof(BatchImage));
  const size_t t_job = tc.take(jobs.size() * sizeof(PathJob));
  const size_t t_cjob = tc.take(cjobs.size() * sizeof(CostJob));
  const size_t t_pair = tc.take(pdev.size() * sizeof(BatchPairDev));
  const size_t t_pb = tc.take(phase_begin.size() * sizeof(int));
  const size_t t_jb = tc.take(job_begin.size() * sizeof(int));
  std::vector<uint8_t> host(tc.off, 0);
  memcpy(host.data() + t_img, images.data(), images.size() * sizeof(BatchImage));
  memcpy(host.data() + t_job, jobs.data(), jobs.size() * sizeof(PathJob));
  memcpy(host.data() + t_cjob, cjobs.data(), cjobs.size() * sizeof(CostJob));
  memcpy(host.data() + t_pair, pdev.data(), pdev.size() * sizeof(BatchPairDev));
  memcpy(host.data() + t_pb, phase_begin.data(), phase_begin.size() * sizeof(int));
  memcpy(host.data() + t_jb, job_begin.data(), job_begin.size() * sizeof(int));
  if (!b->external) rc = context_alloc(ctx, &b->tables, tc.off, "batch tables");
  if (rc == SGM_OK) {
    hipError_t e = hipMemcpyAsync(b->tables, host.data(), tc.off, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) rc = hip_fail(e, "batch table upload");
  }
  if (rc) {
    sgm_batch_destroy(b);
    return rc;
  }
  b->tables_bytes = tc.off;
  uint8_t *tb = (uint8_t *)b->tables;
  BatchArgs &A = b->args;
  memset(&A, 0, sizeof A);
  A.images = (const BatchImage *)(tb + t_img);
  A.jobs = (const PathJob *)(tb + t_job);
  A.cost_jobs = (const CostJob *)(tb + t_cjob);
  A.pairs = (const BatchPairDev *)(tb + t_pair);
  A.phase_begin = (const int *)(tb + t_pb);
  A.job_begin = (const int *)(tb + t_jb);
  A.head = (unsigned *)(base + ctrl_head);
  A.error = (int *)(base + ctrl_error);
  A.done = (unsigned *)(base + ctrl_done);
  A.nph = nph;
  A.stride = stride;
  A.npairs = count;
  A.directions = params->directions;
  A.fused = (b->fused || b->recompute) ? 1 : 0;  // first sweep waits on census, not on a cost phase
  A.cmode = b->recompute ? CM_RECOMPUTE : b->fused ? CM_FUSED : b->cost8 ? CM_LOAD8 : CM_LOAD;
  A.lines_per_item = b->lines_per_item;
  A.census_rows = census_rows;
  A.cost_rows = cost_rows;
  A.lr_rows = lr_rows;
  A.pa = path_args(params, 0, 0, 0, ctx->keep_volumes ? 1 : 0);
  A.ca = cost_args(params);
  A.ca.u8 = b->cost8 ? 1 : 0;
  A.ww = params->window_w;
  A.wh = params->window_h;
  A.lr_threshold = params->lr_threshold;
  A.item_lo = 0;
  A.item_hi = total;
  A.phased = 0;
  if (flags & SGM_BATCH_PHASED) {
    if ((rc = build_groups(b, jobs, pairs, count))) {
      sgm_batch_destroy(b);
      return rc;
    }
  }
  {
    // dependency waits give up after 20 s of wall clock (error flag, grid drains)
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess ||
        khz <= 0)
      khz = 100000;
    A.spin_limit = (long long)khz * 1000LL * 20LL;
  }

  if (b->planes && !b->external) {
    // a second stream where a sub-batch has more than one lane group
    // (measured: 48 uneven pairs 415 -> 401 ms)
    bool several = false;
    for (size_t i = 1; i < b->groups.size(); i++) several |= b->groups[i].sub == b->groups[i - 1].sub;
    if (several) {
      hipError_t e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
      if (e != hipSuccess) {
        sgm_batch_destroy(b);
        return hip_fail(e, "batch side stream");
      }
    }
  }

  if (flags & SGM_BATCH_GRAPH) {
    hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      sgm_batch_destroy(b);
      return hip_fail(e, "hipStreamBeginCapture");
    }
    const bool prof = ctx->profiling;
    ctx->profiling = false;  // no event records inside the graph
    rc = batch_enqueue(b);
    ctx->profiling = prof;
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(stream, &graph);
    if (rc == SGM_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    if (rc == SGM_OK) {
      e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) {
      sgm_batch_destroy(b);
      return rc;
    }
  }
  *out = b;
  return SGM_OK;
}

// Enqueue one run of the batch on its stream (graph replay or direct launches).
int batch_launch(sgm_batch *b) {
  if (b->graph_exec) {
    ProfileScope ps(b->ctx, "batch_graph");
    hipError_t e = hipGraphLaunch(b->graph_exec, b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return SGM_OK;
  }
  return batch_enqueue(b);
}

}  // namespace sgm

extern "C" int sgm_batch_run(sgm_batch *b) {
  if (!b) return fail(SGM_ERR_INVALID_ARGUMENT, "null batch");
  sgm_context *ctx = b->ctx;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  int rc = batch_launch(b);
  if (rc) return rc;
  if ((b->flags & SGM_BATCH_NO_SYNC) && (b->flags & SGM_BATCH_DEVICE)) return SGM_OK;
  return batch_finish(b);
}

namespace sgm {

// Waits for the batch, checks the queue's error flag, delivers the counters.
int batch_finish(sgm_batch *b) {
  sgm_context *ctx = b->ctx;
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch synchronize");
  collect_profile(ctx);
  int err = 0;
  // copies on the batch's own stream: a copy on the null stream can queue
  // behind other streams' kernels on a shared hardware queue
  e = hipMemcpyAsync(&err, b->args.error, sizeof err, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch status read");
  if (err) return fail(SGM_ERR_TIMEOUT, "batch work queue wait expired (dependency never completed)");
  const size_t count = b->pairs.size();
  std::vector<unsigned long long> st(count * 16);
  e = hipMemcpyAsync(st.data(), b->d_stats, st.size() * 8, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "batch counters read");
  for (size_t i = 0; i < count; i++) {
    sgm_stats *out = b->pairs[i].stats;
    if (!out) continue;
    sgm_stats s;
    memset(&s, 0, sizeof s);
    const unsigned long long *c = &st[i * 16];
    s.lower_boundary_rejected = (int64_t)c[0];
    s.full_interval_pixels = (int64_t)c[1];
    s.boundary_rejected = (int64_t)c[2];
    s.upper_boundary_rejected = (int64_t)c[3];
    s.ambiguity_rejected = (int64_t)c[4];
    s.selected = (int64_t)c[5];
    s.lr_rejected = (int64_t)c[6];
    s.valid = (int64_t)c[10];
    if (b->flags & SGM_BATCH_DEVICE) {
      e = hipMemcpyAsync(out, &s, sizeof s, hipMemcpyHostToDevice, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) return hip_fail(e, "batch counters write");
    } else {
      *out = s;
    }
  }
  return SGM_OK;
}

}  // namespace sgm

extern "C" int sgm_batch_destroy(sgm_batch *b) {
  if (!b) return SGM_OK;
  sgm_context *ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(b->stream);
  if (b->side) (void)hipStreamSynchronize(b->side);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->side_owned) {
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
  }
  if (b->external) {
    delete b;
    return SGM_OK;
  }
  if (b->arena) {
    (void)hipFree(b->arena);
    ctx->allocated -= b->arena_bytes;
  }
  if (b->tables) {
    (void)hipFree(b->tables);
    ctx->allocated -= b->tables_bytes;
  }
  if (b->group_tables) {
    (void)hipFree(b->group_tables);
    ctx->allocated -= b->group_tables_bytes;
  }
  delete b;
  return SGM_OK;
}

extern "C" int64_t sgm_batch_device_bytes(sgm_batch *b) {
  return b ? (int64_t)(b->arena_bytes + b->tables_bytes) : 0;
}

extern "C" int sgm_batch_describe(sgm_batch *b, char *buffer, size_t length) {
  if (!b || !buffer || length == 0) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  const int n = snprintf(buffer, length,
                         "{\"pairs\":%zu,\"work_items\":%d,\"workgroups\":%d,"
                         "\"waves_per_workgroup\":%d,\"lines_per_item\":%d,"
                         "\"cost_volume\":\"%s\",\"device_bytes\":%lld,\"graph\":%s,"
                         "\"phased\":%s,\"lane_groups\":%zu,\"sweeps\":\"%s\","
                         "\"sub_batches\":%d}",
                         b->pairs.size(), b->total_items, b->blocks, b->waves_per_block,
                         b->lines_per_item,
                         b->recompute ? "recomputed"
                         : b->fused   ? "resident-u8-fused"
                         : b->cost8   ? "resident-u8"
                                      : "resident-u16",
                         (long long)(b->arena_bytes + b->tables_bytes),
                         b->graph_exec ? "true" : "false", b->phased ? "true" : "false",
                         b->groups.size(), b->planes ? "planes" : "sequential", b->nsub);
  if (n < 0 || (size_t)n >= length) return fail(SGM_ERR_INVALID_ARGUMENT, "buffer too small");
  return SGM_OK;
}
