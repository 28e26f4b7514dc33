// Host-side internals shared by the translation units of libsgm_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sgm_hip.h"
#include "sgm_kernels.h"

struct sgm_context;

namespace sgm {

struct Buffer {
  void *ptr = nullptr;
  size_t cap = 0;
  int ensure(sgm_context *ctx, size_t bytes, const char *what);
  void release(sgm_context *ctx);
};

// Pinned host memory (hipHostMalloc), grown on demand.  `fail_counter`
// (optional) simulates allocation failures while positive.
struct HostBuffer {
  void *ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes, const char *what, int *fail_counter = nullptr);
  void release();
};

// Streamed plan of a pair (GpuStereo.Plan, GpuStereo.java:12-20,56-83).
struct StreamedPlan {
  int rows = 0;        // strip rows (0: no streamed plan fits)
  int storage = 0;     // SGM_AGGREGATE_RAM / _DISK / _RECOMPUTE
  int block_rows = 0;  // rows per recomputation block (h for RAM and DISK)
  int64_t device_bytes = 0, host_bytes = 0, disk_bytes = 0;
};

struct PendingEvent {
  hipEvent_t start, stop;
  const char *name;
};

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int context_alloc(sgm_context *ctx, void **ptr, size_t bytes, const char *what);
int validate_params(const sgm_params *p, int w, int h);
int effective_p2(const sgm_params *p);
int64_t aggregate_bound(const sgm_params *p);
int aggregate_bytes(const sgm_params *p);
int lanes_group(int D);
int padded_disparities(int D);
int words_of(const sgm_params *p);
int invalid_code(const sgm_params *p);
bool cost_u8(const sgm_params *p);
int64_t direction_bytes(const sgm_params *p, int w, int h);
int64_t direction_bytes(const sgm_params *p, int w, int h, bool cost_volume);
bool census_recompute_possible(const sgm_params *p);
bool recompute_cost(sgm_context *ctx, const sgm_params *p, int w, int h);
int64_t pair_bytes(const sgm_params *p, int w, int h);
int64_t pair_bytes(const sgm_params *p, int w, int h, bool cost_volume);
int64_t resident_footprint(sgm_context *ctx, const sgm_params *p, int w, int h, bool cost_volume,
                           int schedule, bool fresh);
bool schedule_applies(const sgm_context *ctx, const sgm_params *p, int schedule, bool cost_volume);
int64_t schedule_bytes(const sgm_params *p, int w, int h, int schedule);
int sweep_schedule(const sgm_context *ctx, const sgm_params *p, int w, int h, bool cost_volume,
                   int64_t budget);
int lines_of(int cls, int w, int h);
// length of line `line` of direction (dx, dy) in a w x h image (line_start)
int line_length(int dx, int dy, int line, int w, int h);
// the indices of `len` by length, longest first, equal lengths in index order
void longest_first(const std::vector<int> &len, int maxlen, std::vector<int> &order);
struct HostCopy {
  void *dst;
  const void *src;
  size_t bytes;
};
// Host-to-host copies split in 4 MiB chunks over up to eight threads: copies
// into fresh caller buffers are bound by first-touch page faults, which
// proceed in parallel.
void parallel_copy(const std::vector<HostCopy> &copies);
void collect_profile(sgm_context *ctx);
// plain pair on the context's stream without waiting (serving lanes)
int lane_enqueue(sgm_context *ctx, const sgm_params *base, const sgm_pair &pr);
int lane_collect(sgm_context *ctx, const sgm_pair &pr, sgm_stats *stats);

struct ProfileScope {
  sgm_context *ctx;
  const char *name;
  PendingEvent ev{};
  ProfileScope(sgm_context *c, const char *n);
  ~ProfileScope();
};

int launch_fill_u8(sgm_context *ctx, uint8_t *p, uint8_t v, size_t n);
int launch_fill_u64(sgm_context *ctx, unsigned long long *p, unsigned long long v, size_t n);
int launch_census(sgm_context *ctx, const CensusJobs &jobs, int njobs, int w, int h,
                  const sgm_params *p);
int launch_cost(sgm_context *ctx, const CostJobs &jobs, int njobs, int w, int h, int Dp,
                const sgm_params *p, int rows = -1, bool force_u16 = false);
int64_t streamed_pair_bytes(const sgm_params *p, int w, int h, int rows);
int64_t streamed_host_bytes(const sgm_params *p, int w, int h);
int streamed_rows(sgm_context *ctx, const sgm_params *p, int w, int h);
StreamedPlan streamed_plan(sgm_context *ctx, const sgm_params *p, int w, int h, int storage);
int64_t host_budget(const sgm_context *ctx);
int64_t esgm_pair_bytes(const sgm_params *p, int w, int h, int rows, int pool);
const char *esgm_unsupported(const sgm_params *p, int w, int h);
int esgm_plan(sgm_context *ctx, const sgm_params *p, int w, int h, int *rows, int *pool);
CostArgs cost_args(const sgm_params *p);
PathArgs path_args(const sgm_params *p, int dir, int njobs, int total_lines, int keep_sum);
// on `stream` (the context's stream when null)
int launch_path(sgm_context *ctx, int R, int sum_bytes, int mode, const PathJobs &jobs,
                const PathArgs &A, int phase, int waves, hipStream_t stream = nullptr);
int launch_check(const char *what);
// two directions in one launch: `ja` (phase MID) and `jb` (phase_b: FIRST or
// MID) accumulate into different partial aggregate volumes
int launch_path_pair(sgm_context *ctx, int R, int mode, int phase_b, const PathJobs &ja,
                     const PathArgs &A, const PathJobs &jb, const PathArgs &B, int waves,
                     hipStream_t stream = nullptr);
// the directions of `M` in one launch, each storing its path-cost plane
int launch_path_multi(sgm_context *ctx, int R, int mode, const PathJobs &jobs, const PathArgs &A,
                      const MultiDir &M, int waves, hipStream_t stream = nullptr);
// one sweep phase's kernel instances each (sgm_path_first/mid/last.hip)
int launch_path_first(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                      const PathArgs &A, int waves);
int launch_path_mid(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                    const PathArgs &A, int waves);
int launch_path_last(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                     const PathArgs &A, int waves);
int launch_path_cost(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                     const PathArgs &A, int waves);
int path_lanes(const sgm_params *p);
bool fused_cost(const sgm_params *p);
int launch_lr(sgm_context *ctx, const LrArgs &A);
int path_waves(sgm_context *ctx, int R);

}  // namespace sgm

struct sgm_context {
  int device = 0;
  sgm_device_info info{};
  int64_t budget = 0;
  int64_t allocated = 0;
  hipStream_t stream = nullptr;
  std::mutex lock;
  std::map<std::string, sgm::Buffer> buffers;
  std::map<std::string, sgm::HostBuffer> host_buffers;
  int plan_mode = 0;       // SGM_PLAN_*
  int strip_rows = 64;     // Options.gpuStripRows default
  bool last_streamed = false;
  int esgm_candidates = 2;            // Options.esgmCandidates default
  int64_t esgm_last[4] = {0, 0, 0, 0};  // certified / verified pixels per matching direction
  int esgm_groups = 0;  // verification groups of the last compact-plan call (0: none needed)
  // streamed plan: copy stream and the events of the two strip aggregate slots
  // (0-1: slot computed, 2-3: slot copied), created on first use
  hipStream_t copy_stream = nullptr;
  hipEvent_t strip_events[4] = {nullptr, nullptr, nullptr, nullptr};
  bool keep_volumes = false;
  int volume_policy = 0;  // SGM_VOLUME_*
  bool profiling = false;
  int fail_allocations = 0;
  int fail_host_allocations = 0;
  // streamed aggregate storage (SGM_AGGREGATE_*), its directory and host budget
  int aggregate_storage = 0;
  std::string aggregate_dir = "/tmp";
  int64_t host_budget = 0;
  // sgm_interrupt: set by any thread, consumed by the matching call that sees it
  std::atomic<int> interrupt{0};
  bool last_compact = false;
  int sweeps = SGM_SWEEPS_AUTO;  // sgm_set_sweeps
  // longest-first line orders of the plane sweeps, by buffer name: the host
  // copy and the (w, h, matching directions, directions, first dx, dy) it is for
  struct LineOrder {
    std::vector<int> host;
    int key[6] = {0, 0, 0, 0, 0, 0};
  };
  std::map<std::string, LineOrder> line_orders;
  int last_sweeps = SGM_SWEEPS_SEQUENTIAL;
  sgm::StreamedPlan splan;  // streamed plan picked for the current call
  std::vector<sgm::PendingEvent> pending_events;
  std::map<std::string, std::pair<double, long long>> kernel_stats;
  std::map<int, int> tuned_waves;   // lane group R -> waves per workgroup in path sweeps
  // batched mode launch shape (autotuned)
  int batch_waves_per_block = 8;
  int batch_lines_per_item = 1;
  int batch_blocks_per_cu = 0;  // 0: as many as stay resident
  int last_w = 0, last_h = 0;
  sgm_params last_params{};
};
