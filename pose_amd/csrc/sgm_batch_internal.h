// Host-side state of the batched and served modes.
#pragma once

#include <vector>

#include "sgm_batch_kernel.h"
#include "sgm_internal.h"

namespace sgm {

// Device buffers of one pair inside a batch arena (or the caller's buffers).
struct PairSlots {
  uint8_t *gray[2] = {nullptr, nullptr};
  uint8_t *valid[2] = {nullptr, nullptr};
  float *disp[2] = {nullptr, nullptr};
  float *checked = nullptr;
  uint8_t *reasons = nullptr, *boundary = nullptr;
};

bool batch_recompute_possible(const sgm_params *params, bool keep_volumes);
size_t batch_arena_bytes(const sgm_params *params, const sgm_pair *pairs, int count, int flags,
                         bool recompute, bool planes);
bool batch_planes_possible(const sgm_context *ctx, const sgm_params *params, int flags);
size_t batch_plane_bytes(const sgm_pair &pr);
size_t batch_table_bytes(int count, const sgm_params *params);
// recompute: 1 census cost recomputed in the sweeps (no cost volume), 0 resident
// cost volume, -1 chosen by the context's policy and the memory at hand.
// planes (phase-major runs): 1 path-cost planes with every pair in one
// sub-batch, 0 one aggregate volume per pair, -1 planes in sub-batches as
// large as the memory at hand allows, where they apply.
int batch_create(sgm_context *ctx, const sgm_params *params, const sgm_pair *pairs, int count,
                 int flags, void *region, size_t region_bytes, hipStream_t stream,
                 sgm_batch **out, int recompute, int planes);
const void *batch_kernel_of(const sgm_batch *b);
int batch_launch(sgm_batch *b);
int batch_finish(sgm_batch *b);

}  // namespace sgm

struct sgm_batch {
  sgm_context *ctx = nullptr;
  hipStream_t stream = nullptr;
  bool external = false;  // arena and tables belong to a serving pool
  sgm_params params{};
  int flags = 0;
  std::vector<sgm_pair> pairs;
  std::vector<sgm::PairSlots> slots;
  int words = 1, sum_bytes = 2;
  bool fused = false;
  bool recompute = false;
  bool cost8 = false;  // uint8 cost volume
  int waves_per_block = 4, lines_per_item = 1, blocks = 0;
  void *arena = nullptr;
  size_t arena_bytes = 0;
  void *tables = nullptr;
  size_t tables_bytes = 0;
  unsigned *d_control = nullptr;
  size_t control_words = 0;
  unsigned long long *d_stats = nullptr;
  int total_items = 0;
  sgm::BatchArgs args{};
  hipGraphExec_t graph_exec = nullptr;
  // host-buffer batches of a server: the kernels wait for `gate_wait` (the
  // previous batch's kernels) after the uploads and record `gate_done` before
  // the downloads, so batches overlap their transfers, not their kernels
  hipEvent_t gate_wait = nullptr, gate_done = nullptr;
  // phase-major plane runs: a second stream for every other lane group of a
  // sub-batch (the groups' pairs are disjoint), forked from and joined back
  // into `stream` around each sub-batch
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool side_owned = true;  // false: the serving slot's stream and events
  // phase-major runs (SGM_BATCH_PHASED): per lane group R the pairs' path jobs
  // and, per line class (horizontal, vertical, diagonal), the first line of
  // each job in the group's launch
  struct Group {
    int R = 0, njobs = 0, sub = 0;
    bool nt = false;
    const sgm::PathJob *jobs = nullptr;
    // per line class (horizontal, vertical, diagonal, cost-pass row segments)
    const int *line_begin[4] = {nullptr, nullptr, nullptr, nullptr};
    int total_lines[4] = {0, 0, 0, 0};
    // plane runs of non-serving batches: longest-first line orders of the
    // seven-direction launch and of the last sweep (null: line order)
    const int *order_multi = nullptr, *order_last = nullptr;
  };
  bool phased = false;
  // path-cost planes: the scratch region of a sub-batch's cost volumes and
  // plane blocks, each pair's sub-batch, and the sub-batch count
  bool planes = false;
  size_t scratch_bytes = 0;
  std::vector<int> pair_sub;
  int nsub = 1;
  std::vector<Group> groups;
  std::vector<int> phase_begin;  // host copy of the queue's phase boundaries
  void *group_tables = nullptr;
  size_t group_tables_bytes = 0;
};
