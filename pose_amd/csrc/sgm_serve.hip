// Serving mode (continuous batching) and the launch-shape autotuner.
//
// A server thread admits submitted pairs in arrival order as soon as their
// device footprint fits the free part of a budgeted memory pool and a stream
// slot is free, launches every admitted pair waiting at that moment as one
// phase-major batch on that stream (census and L-R grids over the batch, the
// sweep launches per lane group), and keeps up to `streams` batches in
// flight.  A batch's host buffers go through the slot's pinned staging, so its
// copies overlap other batches' kernels; its kernels wait for the previous
// batch's kernels.  Volumes stay device-resident for the life of their batch.  With max_batch == 1 each slot is a lane instead: a context
// of its own that runs every admitted pair through the plain per-pair kernels
// on its stream, the lanes overlapping on the device, admission by the device
// memory the lanes hold.  Neither mode has a counterpart in the reference,
// which matches pairs sequentially (DenseReconstruction.java:64-72).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sgm_batch_internal.h"

using namespace sgm;

namespace {

// First-fit allocator over one device region (offsets, 256-byte granules).
class Pool {
 public:
  void reset(size_t bytes) {
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (bytes + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes, o = it->first + bytes;
        free_.erase(it);
        if (rest) free_[o] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Request {
  uint64_t ticket;
  sgm_pair pair;
  sgm_stats stats;
  size_t bytes = 0, off = 0;
  int64_t pool_bytes = 0;  // device memory the server held once this request started
  double arrival = 0, admitted = 0, started = 0, finished = 0;
  int status = 1;  // 1 pending, 0 done, < 0 error
  int batch_id = -1;
};

// Offsets of one request's host inputs and outputs in its slot's pinned
// staging buffer (SIZE_MAX: not requested).
struct Staged {
  size_t left, right, vl, vr, checked, dl, dr, reasons, boundary;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t kernels_done = nullptr;  // batching mode: the batch's kernels completed
  // batching mode: the second stream of the batches' plane sweeps and its
  // fork / join events
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  sgm_batch *batch = nullptr;
  sgm_context *lane = nullptr;  // lane mode: the slot's own context (and stream)
  std::vector<Request *> reqs;
  size_t off = 0, bytes = 0;
  bool busy = false;
  // batching mode: pinned staging of the batch's host inputs and outputs, so
  // the batch's copies run asynchronously on the slot's stream
  sgm::HostBuffer staging;
  std::vector<Staged> staged;
  // outgrown staging buffers, freed when the server closes: freeing pinned
  // memory waits for the whole device, which would stall the other streams
  std::vector<sgm::HostBuffer> retired;
};

// Lays out a request's host buffers in a staging buffer from offset `at`.
static Staged stage_layout(const sgm_pair &pr, size_t &at) {
  const size_t n = (size_t)pr.width * pr.height;
  auto take = [&](bool used, size_t bytes) {
    if (!used) return SIZE_MAX;
    const size_t o = at;
    at = (at + bytes + 255) / 256 * 256;
    return o;
  };
  Staged st;
  st.left = take(true, n);
  st.right = take(true, n);
  st.vl = take(pr.valid_left != nullptr, n);
  st.vr = take(pr.valid_right != nullptr, n);
  st.checked = take(true, 4 * n);
  st.dl = take(pr.dl != nullptr, 4 * n);
  st.dr = take(pr.dr != nullptr, 4 * n);
  st.reasons = take(pr.reasons != nullptr, n);
  st.boundary = take(pr.boundary != nullptr, n);
  return st;
}

}  // namespace

struct sgm_server {
  sgm_context *ctx = nullptr;
  sgm_params params{};
  int max_batch = 16;
  bool lanes = false;  // max_batch == 1: plain per-pair lanes
  int volume_policy = 0;  // the context's SGM_VOLUME_* at creation
  int sweeps = 0;         // the context's SGM_SWEEPS_* at creation
  void *region = nullptr;
  size_t region_bytes = 0;
  Pool pool;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending;
  std::map<uint64_t, Request *> all;
  uint64_t next_ticket = 1;
  int next_batch = 0;
  // batching mode: the slot whose batch was launched last; the next batch's
  // kernels wait for its kernels (batches overlap transfers, not kernels:
  // concurrent phase-major batches slow each other down, measured)
  Slot *last_launched = nullptr;
  bool stop = false;
  std::thread worker;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void run();
  bool admit_lane(Slot &s);
};

// Smallest device footprint of a request served alone (admission test).
static size_t request_bytes(const sgm_context *ctx, const sgm_params *p, const sgm_pair &pr) {
  const bool recompute = batch_recompute_possible(p, ctx->keep_volumes) &&
                         ctx->volume_policy != SGM_VOLUME_RESIDENT;
  return batch_arena_bytes(p, &pr, 1, 0, recompute, false) + batch_table_bytes(1, p);
}

void sgm_server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    // retire finished batches
    for (Slot &s : slots) {
      if (!s.busy) continue;
      if (hipEventQuery(s.done) != hipSuccess) continue;
      lk.unlock();
      int rc;
      if (lanes) {
        rc = lane_collect(s.lane, s.reqs[0]->pair, &s.reqs[0]->stats);
      } else {
        rc = batch_finish(s.batch);
        sgm_batch_destroy(s.batch);
        // outputs from the pinned staging to the callers' buffers
        const uint8_t *sb = (const uint8_t *)s.staging.ptr;
        std::vector<sgm::HostCopy> copies;
        for (size_t i = 0; rc == SGM_OK && i < s.reqs.size(); i++) {
          const sgm_pair &pr = s.reqs[i]->pair;
          const Staged &st = s.staged[i];
          const size_t n = (size_t)pr.width * pr.height;
          copies.push_back({pr.checked, sb + st.checked, 4 * n});
          if (pr.dl) copies.push_back({pr.dl, sb + st.dl, 4 * n});
          if (pr.dr) copies.push_back({pr.dr, sb + st.dr, 4 * n});
          if (pr.reasons) copies.push_back({pr.reasons, sb + st.reasons, n});
          if (pr.boundary) copies.push_back({pr.boundary, sb + st.boundary, n});
        }
        sgm::parallel_copy(copies);
      }
      lk.lock();
      const double t = now();
      for (Request *r : s.reqs) {
        r->finished = t;
        r->status = rc == SGM_OK ? 0 : rc;
        if (r->pair.stats) *r->pair.stats = r->stats;
      }
      if (!lanes) pool.release(s.off, s.bytes);
      s.busy = false;
      s.batch = nullptr;
      s.reqs.clear();
      cv_done.notify_all();
    }
    if (stop && pending.empty()) {
      bool any = false;
      for (Slot &s : slots) any |= s.busy;
      if (!any) break;
    }
    // admit and launch: one batch per free stream slot (lanes: one pair)
    for (Slot &s : slots) {
      if (s.busy || pending.empty()) continue;
      if (lanes) {
        if (!admit_lane(s)) break;  // wait for memory to free up
        continue;
      }
      // A request whose footprint exceeds an eighth of the pool is batched
      // alone when it reaches the head of the queue; the batches formed
      // before that take the small requests behind it in arrival order, so
      // they do not wait for its completion.
      const size_t large = region_bytes / 8;
      std::vector<Request *> take;
      size_t bytes = 0;
      if (pending.front()->bytes > large) {
        take.push_back(pending.front());
      } else {
        for (Request *r : pending) {
          if ((int)take.size() >= max_batch) break;
          if (r->bytes > large) continue;
          take.push_back(r);
        }
      }
      // Shrink until the group fits the pool (FIFO: keep the oldest).  Each
      // group first tries path-cost planes, then resident cost and aggregate
      // volumes, then recomputed census costs.
      size_t off = 0;
      int recompute = 0, planes = 0;
      const bool can_recompute = batch_recompute_possible(&params, ctx->keep_volumes);
      const bool can_planes = ctx->volume_policy != SGM_VOLUME_RECOMPUTE &&
                              batch_planes_possible(ctx, &params, SGM_BATCH_PHASED);
      while (!take.empty()) {
        std::vector<sgm_pair> prs;
        for (Request *r : take) prs.push_back(r->pair);
        const int n = (int)prs.size();
        const size_t tables = batch_table_bytes(n, &params);
        if (can_planes) {
          recompute = 0;
          planes = 1;
          bytes = batch_arena_bytes(&params, prs.data(), n, 0, false, true) + tables;
          if (pool.alloc(bytes, &off)) break;
        }
        planes = 0;
        if (ctx->volume_policy != SGM_VOLUME_RECOMPUTE || !can_recompute) {
          recompute = 0;
          bytes = batch_arena_bytes(&params, prs.data(), n, 0, false, false) + tables;
          if (pool.alloc(bytes, &off)) break;
        }
        if (can_recompute && ctx->volume_policy != SGM_VOLUME_RESIDENT) {
          recompute = 1;
          bytes = batch_arena_bytes(&params, prs.data(), n, 0, true, false) + tables;
          if (pool.alloc(bytes, &off)) break;
        }
        take.pop_back();
      }
      if (take.empty()) break;  // wait for memory to free up
      const double t = now();
      for (size_t i = 0; i < take.size(); i++) {
        take[i]->admitted = t;
        take[i]->batch_id = next_batch;
        pending.erase(std::find(pending.begin(), pending.end(), take[i]));
      }
      std::vector<sgm_pair> prs;
      for (Request *r : take) {
        sgm_pair p = r->pair;
        p.stats = &r->stats;
        prs.push_back(p);
      }
      s.off = off;
      s.bytes = bytes;
      s.reqs = take;
      s.busy = true;
      const int id = next_batch++;
      (void)id;
      lk.unlock();
      // host inputs into the slot's pinned staging; the batch copies from and
      // to it, so its transfers do not block this thread
      size_t staged_bytes = 0;
      s.staged.clear();
      for (const sgm_pair &pr : prs) s.staged.push_back(stage_layout(pr, staged_bytes));
      int rc = SGM_OK;
      if (staged_bytes > s.staging.cap) {
        // grow by at least half, keeping the old buffer until the server closes
        const size_t grow = std::max(staged_bytes, s.staging.cap + s.staging.cap / 2);
        if (s.staging.ptr) s.retired.push_back(s.staging);
        s.staging = sgm::HostBuffer{};
        rc = s.staging.ensure(grow, "serving staging");
      }
      if (rc == SGM_OK) {
        uint8_t *sb = (uint8_t *)s.staging.ptr;
        for (size_t i = 0; i < prs.size(); i++) {
          sgm_pair &pr = prs[i];
          const Staged &st = s.staged[i];
          const size_t n = (size_t)pr.width * pr.height;
          memcpy(sb + st.left, pr.left, n);
          memcpy(sb + st.right, pr.right, n);
          pr.left = sb + st.left;
          pr.right = sb + st.right;
          if (pr.valid_left) {
            memcpy(sb + st.vl, pr.valid_left, n);
            memcpy(sb + st.vr, pr.valid_right, n);
            pr.valid_left = sb + st.vl;
            pr.valid_right = sb + st.vr;
          }
          pr.checked = (float *)(sb + st.checked);
          if (pr.dl) pr.dl = (float *)(sb + st.dl);
          if (pr.dr) pr.dr = (float *)(sb + st.dr);
          if (pr.reasons) pr.reasons = sb + st.reasons;
          if (pr.boundary) pr.boundary = sb + st.boundary;
        }
      }
      sgm_batch *b = nullptr;
      if (rc == SGM_OK)
        rc = batch_create(ctx, &params, prs.data(), (int)prs.size(), SGM_BATCH_PHASED,
                          (uint8_t *)region + off, bytes, s.stream, &b, recompute, planes);
      if (rc == SGM_OK) {
        b->gate_wait = last_launched && last_launched != &s ? last_launched->kernels_done : nullptr;
        b->gate_done = s.kernels_done;
        if (b->planes) {
          b->side = s.side;
          b->fork = s.fork;
          b->join = s.join;
          b->side_owned = false;
        }
        rc = batch_launch(b);
        if (rc == SGM_OK) last_launched = &s;
      }
      if (rc == SGM_OK) {
        hipError_t e = hipEventRecord(s.done, s.stream);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
      }
      lk.lock();
      const double ts = now();
      if (rc != SGM_OK) {
        if (b) {
          lk.unlock();
          sgm_batch_destroy(b);
          lk.lock();
        }
        for (Request *r : take) {
          r->status = rc;
          r->finished = ts;
        }
        pool.release(s.off, s.bytes);
        s.busy = false;
        s.reqs.clear();
        cv_done.notify_all();
        continue;
      }
      s.batch = b;
      for (Request *r : take) {
        r->started = ts;
        r->pool_bytes = (int64_t)pool.used();
      }
    }
    bool any_busy = false;
    for (Slot &s : slots) any_busy |= s.busy;
    if (any_busy) {
      cv_work.wait_for(lk, std::chrono::microseconds(200));
    } else {
      cv_work.wait(lk, [&] { return stop || !pending.empty(); });
    }
  }
}

// Lane mode: the pair at the head of the queue on free lane `s` when the
// device bytes the lane will hold after serving it, plus what the other lanes
// hold, fit the budget.  The lane's cached buffers are reused where they are
// large enough, so that footprint counts each of the pair's buffers at the
// larger of its cached and required size.  Plans, preferred first: path-cost
// planes, paired sweeps, one aggregate with the cost volume, one aggregate
// with recomputed census costs (as the sweep schedule and the volume policy
// allow).  Tried in order: every plan on the lane's cached buffers as they
// are, then on its buffers released, then with the cached buffers of idle
// lanes released too.  The lane's budget becomes what the other lanes leave
// and its schedule and volume policy the chosen plan, so it runs exactly the
// admitted footprint.  Called and returning with `mu` held.
bool sgm_server::admit_lane(Slot &s) {
  Request *r = pending.front();
  sgm_params q = params;
  q.disparities = r->pair.disparities;
  q.min_disparity = r->pair.min_disparity;
  const int w = r->pair.width, h = r->pair.height;
  sgm_context *lane = s.lane;
  auto others = [&]() {
    int64_t c = 0;
    for (Slot &o : slots)
      if (&o != &s) c += o.lane->allocated;
    return c;
  };
  // plans the volume policy and the sweep schedule allow, preferred first
  struct Plan {
    bool cost_volume;
    int schedule;
  };
  const bool possible = census_recompute_possible(&q) && !lane->keep_volumes;
  const bool with_cv = !possible || volume_policy != SGM_VOLUME_RECOMPUTE;
  Plan plans[4];
  int nplan = 0;
  if (with_cv) {
    if (sweeps == SGM_SWEEPS_AUTO || sweeps == SGM_SWEEPS_PLANES) {
      if (schedule_applies(lane, &q, SGM_SWEEPS_PLANES, true))
        plans[nplan++] = {true, SGM_SWEEPS_PLANES};
    }
    if (sweeps != SGM_SWEEPS_SEQUENTIAL && schedule_applies(lane, &q, SGM_SWEEPS_PAIRED, true))
      plans[nplan++] = {true, SGM_SWEEPS_PAIRED};
    plans[nplan++] = {true, SGM_SWEEPS_SEQUENTIAL};
  }
  if (possible && volume_policy != SGM_VOLUME_RESIDENT)
    plans[nplan++] = {false, SGM_SWEEPS_SEQUENTIAL};
  auto pick = [&](bool fresh) {
    const int64_t avail = (int64_t)region_bytes - others();
    for (int i = 0; i < nplan; i++)
      if (resident_footprint(lane, &q, w, h, plans[i].cost_volume, plans[i].schedule, fresh) <=
          avail)
        return i;
    return -1;
  };
  int chosen;
  if ((chosen = pick(false)) < 0 && (chosen = pick(true)) >= 0) {
    (void)sgm_release_buffers(lane);
  } else if (chosen < 0) {
    bool released = false;
    for (Slot &o : slots)
      if (!o.busy && &o != &s && o.lane->allocated) {
        (void)sgm_release_buffers(o.lane);
        released = true;
      }
    if (!released) return false;  // wait for a busy lane to finish
    if ((chosen = pick(false)) < 0) {
      if ((chosen = pick(true)) < 0) return false;
      (void)sgm_release_buffers(lane);
    }
  }
  lane->budget = (int64_t)region_bytes - others();
  if (possible)
    lane->volume_policy = plans[chosen].cost_volume ? SGM_VOLUME_RESIDENT : SGM_VOLUME_RECOMPUTE;
  lane->sweeps = plans[chosen].schedule;
  pending.pop_front();
  const double t = now();
  r->admitted = t;
  r->batch_id = next_batch++;
  s.reqs.assign(1, r);
  s.busy = true;
  std::unique_lock<std::mutex> lk(mu, std::adopt_lock);
  lk.unlock();
  int rc = lane_enqueue(s.lane, &params, r->pair);
  if (rc == SGM_OK) {
    hipError_t e = hipEventRecord(s.done, s.lane->stream);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord");
  }
  lk.lock();
  lk.release();
  const double ts = now();
  if (rc != SGM_OK) {
    r->status = rc;
    r->finished = ts;
    s.busy = false;
    s.reqs.clear();
    cv_done.notify_all();
    return true;
  }
  r->started = ts;
  r->pool_bytes = others() + s.lane->allocated;
  return true;
}

extern "C" int sgm_server_create(sgm_context *ctx, const sgm_params *params, int64_t budget_bytes,
                                 int max_batch, int streams, sgm_server **out) {
  if (!ctx || !params || !out || streams < 1 || streams > 8 || max_batch < 1)
    return fail(SGM_ERR_INVALID_ARGUMENT, "bad server arguments");
  *out = nullptr;
  int rc = validate_params(params, params->window_w + 2, params->window_h + 2);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  const int64_t avail = ctx->budget - ctx->allocated;
  const int64_t bytes = budget_bytes > 0 ? std::min(budget_bytes, avail) : avail;
  if (bytes <= 0) return fail(SGM_ERR_NO_PLAN, "no device memory left for the serving pool");
  sgm_server *s = new sgm_server();
  s->ctx = ctx;
  s->params = *params;
  s->max_batch = max_batch;
  s->lanes = max_batch == 1;
  s->volume_policy = ctx->volume_policy;
  s->sweeps = ctx->sweeps;
  s->region_bytes = (size_t)bytes;
  if (!s->lanes) {
    rc = context_alloc(ctx, &s->region, s->region_bytes, "serving pool");
    if (rc) {
      delete s;
      return rc;
    }
    s->pool.reset(s->region_bytes);
  }
  s->slots.resize(streams);
  for (Slot &sl : s->slots) {
    hipError_t e = hipSuccess;
    if (s->lanes) {
      // each lane is a context of its own: buffers, stream, autotuned shapes
      if ((rc = sgm_open(ctx->device, bytes, &sl.lane))) break;
      sl.lane->tuned_waves = ctx->tuned_waves;
      sl.lane->volume_policy = ctx->volume_policy;
      sl.stream = sl.lane->stream;
    } else {
      e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess && !s->lanes)
      e = hipEventCreateWithFlags(&sl.kernels_done, hipEventDisableTiming);
    if (e == hipSuccess && !s->lanes) e = hipStreamCreateWithFlags(&sl.side, hipStreamNonBlocking);
    if (e == hipSuccess && !s->lanes) e = hipEventCreateWithFlags(&sl.fork, hipEventDisableTiming);
    if (e == hipSuccess && !s->lanes) e = hipEventCreateWithFlags(&sl.join, hipEventDisableTiming);
    if (e != hipSuccess) {
      rc = hip_fail(e, "server stream");
      break;
    }
  }
  if (rc) {
    for (Slot &sl : s->slots) {
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.kernels_done) (void)hipEventDestroy(sl.kernels_done);
      if (sl.fork) (void)hipEventDestroy(sl.fork);
      if (sl.join) (void)hipEventDestroy(sl.join);
      if (sl.side) (void)hipStreamDestroy(sl.side);
      if (sl.lane) (void)sgm_close(sl.lane);
      else if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->region) {
      (void)hipFree(s->region);
      ctx->allocated -= s->region_bytes;
    }
    delete s;
    return rc;
  }
  s->t0 = std::chrono::steady_clock::now();
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return SGM_OK;
}

extern "C" int sgm_server_submit(sgm_server *srv, const sgm_pair *pair, uint64_t *ticket) {
  if (!srv || !pair || !ticket) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  sgm_params q = srv->params;
  q.disparities = pair->disparities;
  q.min_disparity = pair->min_disparity;
  int rc = validate_params(&q, pair->width, pair->height);
  if (rc) return rc;
  if (!pair->left || !pair->right || !pair->checked)
    return fail(SGM_ERR_INVALID_ARGUMENT, "pair needs left, right and checked buffers");
  if (lanes_group(pair->disparities) > 8)
    return fail(SGM_ERR_INVALID_ARGUMENT, "serving supports up to 512 disparities per pair");
  const size_t bytes = request_bytes(srv->ctx, &srv->params, *pair);
  if (bytes > srv->region_bytes)
    return fail(SGM_ERR_REJECTED, "pair needs %zu MiB, more than the serving budget (%zu MiB)",
                bytes >> 20, srv->region_bytes >> 20);
  Request *r = new Request();
  r->pair = *pair;
  memset(&r->stats, 0, sizeof r->stats);
  r->bytes = bytes;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) {
      delete r;
      return fail(SGM_ERR_REJECTED, "server is stopping");
    }
    r->ticket = srv->next_ticket++;
    r->arrival = srv->now();
    srv->pending.push_back(r);
    srv->all[r->ticket] = r;
    *ticket = r->ticket;
  }
  srv->cv_work.notify_all();
  return SGM_OK;
}

extern "C" int sgm_server_wait(sgm_server *srv, uint64_t ticket, double timeout_s,
                               sgm_request_info *info) {
  if (!srv) return fail(SGM_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->all.find(ticket);
  if (it == srv->all.end()) return fail(SGM_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
  Request *r = it->second;
  auto ready = [&] { return r->status != 1; };
  if (timeout_s < 0) srv->cv_done.wait(lk, ready);
  else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return fail(SGM_ERR_TIMEOUT, "request %llu not finished", (unsigned long long)ticket);
  if (info) {
    info->arrival_s = r->arrival;
    info->admitted_s = r->admitted;
    info->started_s = r->started;
    info->finished_s = r->finished;
    info->device_bytes = (int64_t)r->bytes;
    info->pool_bytes = r->pool_bytes;
    info->status = r->status;
    info->batch_id = r->batch_id;
  }
  const int status = r->status;
  srv->all.erase(it);
  delete r;
  if (status != 0) return fail(status, "request failed (%d)", status);
  return SGM_OK;
}

extern "C" double sgm_server_now(sgm_server *srv) { return srv ? srv->now() : 0.0; }

extern "C" int sgm_server_destroy(sgm_server *srv) {
  if (!srv) return SGM_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  (void)hipSetDevice(srv->ctx->device);
  for (Slot &sl : srv->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.kernels_done) (void)hipEventDestroy(sl.kernels_done);
    if (sl.fork) (void)hipEventDestroy(sl.fork);
    if (sl.join) (void)hipEventDestroy(sl.join);
    if (sl.side) (void)hipStreamDestroy(sl.side);
    if (sl.lane) (void)sgm_close(sl.lane);
    else if (sl.stream) (void)hipStreamDestroy(sl.stream);
    sl.staging.release();
    for (sgm::HostBuffer &hb : sl.retired) hb.release();
  }
  for (auto &kv : srv->all) delete kv.second;
  if (srv->region) {
    std::lock_guard<std::mutex> g(srv->ctx->lock);
    (void)hipFree(srv->region);
    srv->ctx->allocated -= srv->region_bytes;
  }
  delete srv;
  return SGM_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: times candidate launch shapes on this device and keeps the best.
namespace {

uint32_t lcg(uint32_t &s) {
  s = s * 1664525u + 1013904223u;
  return s >> 8;
}

// A textured synthetic pair (host buffers): smooth random texture shifted by a
// constant disparity, enough to exercise realistic path-sweep traffic.
void tune_pair(int w, int h, int shift, uint32_t seed, std::vector<uint8_t> &l,
               std::vector<uint8_t> &r) {
  l.resize((size_t)w * h);
  r.resize((size_t)w * h);
  for (size_t i = 0; i < l.size(); i++) l[i] = (uint8_t)(lcg(seed) & 255);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) r[(size_t)y * w + x] = l[(size_t)y * w + std::min(w - 1, x + shift)];
}

double time_plain(sgm_context *ctx, const sgm_params *p, const std::vector<uint8_t> &l,
                  const std::vector<uint8_t> &r, int w, int h, std::vector<float> &out) {
  sgm_stats st;
  sgm_match_pair(ctx, p, l.data(), r.data(), nullptr, nullptr, w, h, out.data(), nullptr, nullptr,
                 &st, nullptr, nullptr);  // warm-up
  const auto t0 = std::chrono::steady_clock::now();
  const int reps = 3;
  for (int i = 0; i < reps; i++)
    if (sgm_match_pair(ctx, p, l.data(), r.data(), nullptr, nullptr, w, h, out.data(), nullptr,
                       nullptr, &st, nullptr, nullptr) != SGM_OK)
      return 1e30;
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
}

}  // namespace

extern "C" int sgm_autotune(sgm_context *ctx, const sgm_params *params, const char *path) {
  if (!ctx || !params) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  int rc = validate_params(params, 64, 64);
  if (rc) return rc;
  // 1. waves per workgroup of the plain path sweeps, for this D's lane group
  const int R = path_lanes(params);
  const int w = std::max(512, params->disparities + 256), h = 384;
  std::vector<uint8_t> l, r;
  tune_pair(w, h, std::min(params->disparities / 2, 40), 7, l, r);
  std::vector<float> out((size_t)w * h);
  double best = 1e30;
  int best_waves = 4;
  for (int waves : {1, 2, 4}) {
    ctx->tuned_waves[R] = waves;
    const double t = time_plain(ctx, params, l, r, w, h, out);
    if (t < best) {
      best = t;
      best_waves = waves;
    }
  }
  ctx->tuned_waves[R] = best_waves;
  // 2. batched mode: waves per block x lines per work item x blocks per CU
  const int npairs = 6;
  std::vector<std::vector<uint8_t>> L(npairs), Rr(npairs);
  std::vector<std::vector<float>> outs(npairs);
  std::vector<sgm_pair> pairs(npairs);
  uint32_t seed = 11;
  for (int i = 0; i < npairs; i++) {
    const int pw = 640 + (int)(lcg(seed) % 640), ph = pw * 3 / 4;
    const int D = std::min(params->disparities, 64 + (int)(lcg(seed) % 192));
    tune_pair(pw, ph, D / 3, 100 + i, L[i], Rr[i]);
    outs[i].resize((size_t)pw * ph);
    memset(&pairs[i], 0, sizeof(sgm_pair));
    pairs[i].width = pw;
    pairs[i].height = ph;
    pairs[i].disparities = std::max(3, D);
    pairs[i].min_disparity = 0;
    pairs[i].left = L[i].data();
    pairs[i].right = Rr[i].data();
    pairs[i].checked = outs[i].data();
  }
  double bbest = 1e30;
  int bw = ctx->batch_waves_per_block, bl = ctx->batch_lines_per_item, bb = ctx->batch_blocks_per_cu;
  for (int waves : {4, 8})
    for (int lines : {1, 2, 4})
      for (int bpc : {0, 1, 2}) {  // 0: all resident workgroups
        ctx->batch_waves_per_block = waves;
        ctx->batch_lines_per_item = lines;
        ctx->batch_blocks_per_cu = bpc;
        sgm_batch *b = nullptr;
        if (sgm_batch_create(ctx, params, pairs.data(), npairs, 0, &b) != SGM_OK) continue;
        double t = 1e30;
        if (sgm_batch_run(b) == SGM_OK) {
          const auto t0 = std::chrono::steady_clock::now();
          if (sgm_batch_run(b) == SGM_OK && sgm_batch_run(b) == SGM_OK)
            t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 2;
        }
        sgm_batch_destroy(b);
        if (t < bbest) {
          bbest = t;
          bw = waves;
          bl = lines;
          bb = bpc;
        }
      }
  ctx->batch_waves_per_block = bw;
  ctx->batch_lines_per_item = bl;
  ctx->batch_blocks_per_cu = bb;
  if (path) {
    FILE *f = fopen(path, "w");
    if (!f) return fail(SGM_ERR_INVALID_ARGUMENT, "cannot write %s", path);
    fprintf(f,
            "{\"device\": \"%s\", \"arch\": \"%s\", \"disparities\": %d, \"lane_group\": %d,\n"
            " \"path_waves\": %d, \"path_pair_seconds\": %.6f,\n"
            " \"batch_waves_per_block\": %d, \"batch_lines_per_item\": %d, "
            "\"batch_blocks_per_cu\": %d, \"batch_seconds\": %.6f}\n",
            ctx->info.name, ctx->info.arch, params->disparities, R, best_waves, best, bw, bl, bb,
            bbest);
    fclose(f);
  }
  return SGM_OK;
}

extern "C" int sgm_autotune_load(sgm_context *ctx, const char *path) {
  if (!ctx || !path) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  FILE *f = fopen(path, "r");
  if (!f) return fail(SGM_ERR_NOT_FOUND, "no tuning file %s", path);
  std::string s;
  char buf[4096];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
  fclose(f);
  auto field = [&](const char *name, int *v) {
    const std::string key = std::string("\"") + name + "\":";
    const size_t at = s.find(key);
    if (at == std::string::npos) return false;
    *v = atoi(s.c_str() + at + key.size());
    return true;
  };
  int R = 0, waves = 0, bw = 0, bl = 0, bb = 0;
  if (field("lane_group", &R) && field("path_waves", &waves) && waves >= 1 && waves <= 4)
    ctx->tuned_waves[R] = waves;
  if (field("batch_waves_per_block", &bw) && bw >= 1 && bw <= 8) ctx->batch_waves_per_block = bw;
  if (field("batch_lines_per_item", &bl) && bl >= 1 && bl <= 64) ctx->batch_lines_per_item = bl;
  if (field("batch_blocks_per_cu", &bb) && bb >= 0 && bb <= 16) ctx->batch_blocks_per_cu = bb;
  return SGM_OK;
}
