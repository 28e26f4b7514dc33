// libsgm_hip: C-ABI host side of the MI355X SGM engine (plain per-pair path).
//
// Replaces the reference's OpenCL session / Compute runtime
// (src/main/java/sgm/GpuSgm.java:163-619) and the resident branch of
// GpuStereo.executeTargets (src/main/java/sgm/GpuStereo.java:212-479) with one
// HIP stream per context, flat 64-bit-indexed volumes instead of 8 banks, and
// device-side census, selection and left-right check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <fcntl.h>
#include <sys/statvfs.h>
#include <unistd.h>

#include "../../include/sgm_hip.h"
#include "sgm_internal.h"
#include "sgm_filters.h"
#include "sgm_strips.h"
#include "sgm_compact.h"

using namespace sgm;

// Fixed workspace of sgm_match_pair_device_noalloc: every buffer of a resident
// pair is carved from this array, so that entry point allocates nothing.  Each
// device holds its own copy for the life of the process.  In host code the
// name `scratch` is only a stand-in for the symbol; device code gets the
// per-device address that sgm_open looks up (scratch_address).
constexpr size_t kScratchBytes = SGM_NOALLOC_WORKSPACE_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch_ptr[kMaxScratchDevices];
bool g_scratch_done[kMaxScratchDevices];
hipError_t g_scratch_error[kMaxScratchDevices];

// Device address of `scratch` on `device` (which must be current), looked up
// once per device.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_lock);
  if (!g_scratch_done[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
    }
    g_scratch_ptr[device] = ptr;
    g_scratch_error[device] = e;
    g_scratch_done[device] = true;
  }
  return g_scratch_error[device];
}

void *scratch_address(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_lock);
  return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

namespace sgm {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? SGM_ERR_OUT_OF_MEMORY : SGM_ERR_HIP, "%s failed (HIP %d: %s)",
              what, (int)e, hipGetErrorString(e));
}

int context_alloc(sgm_context *ctx, void **ptr, size_t bytes, const char *what) {
  if (ctx->fail_allocations > 0) {
    ctx->fail_allocations--;
    *ptr = nullptr;
    return fail(SGM_ERR_OUT_OF_MEMORY, "GPU buffer allocation (%zu MiB) for %s failed (simulated)",
                bytes >> 20, what);
  }
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return fail(e == hipErrorOutOfMemory ? SGM_ERR_OUT_OF_MEMORY : SGM_ERR_HIP,
                "GPU buffer allocation (%zu MiB) for %s failed (HIP %d)", bytes >> 20, what, (int)e);
  }
  ctx->allocated += bytes;
  return SGM_OK;
}

int Buffer::ensure(sgm_context *ctx, size_t bytes, const char *what) {
  if (bytes <= cap && ptr) return SGM_OK;
  release(ctx);
  int rc = context_alloc(ctx, &ptr, bytes, what);
  if (rc == SGM_OK) cap = bytes;
  return rc;
}

void Buffer::release(sgm_context *ctx) {
  if (ptr) {
    (void)hipFree(ptr);
    ctx->allocated -= cap;
  }
  ptr = nullptr;
  cap = 0;
}

int HostBuffer::ensure(size_t bytes, const char *what, int *fail_counter) {
  if (bytes <= cap && ptr) return SGM_OK;
  release();
  if (fail_counter && *fail_counter > 0) {
    (*fail_counter)--;
    return fail(SGM_ERR_OUT_OF_MEMORY, "pinned host allocation (%zu MiB) for %s failed (simulated)",
                bytes >> 20, what);
  }
  hipError_t e = hipHostMalloc(&ptr, bytes ? bytes : 8, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    return fail(SGM_ERR_OUT_OF_MEMORY, "pinned host allocation (%zu MiB) for %s failed (HIP %d)",
                bytes >> 20, what, (int)e);
  }
  cap = bytes;
  return SGM_OK;
}

void HostBuffer::release() {
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  cap = 0;
}

int validate_params(const sgm_params *p, int w, int h) {
  if (!p) return fail(SGM_ERR_INVALID_ARGUMENT, "null params");
  if (p->window_w < 3 || p->window_w > 11 || p->window_w % 2 == 0 || p->window_h < 3 ||
      p->window_h > 11 || p->window_h % 2 == 0)
    return fail(SGM_ERR_INVALID_ARGUMENT, "census window must be odd, 3..11 (got %dx%d)",
                p->window_w, p->window_h);
  if (p->disparities < 3 || p->disparities > 4096)
    return fail(SGM_ERR_INVALID_ARGUMENT, "disparities must be 3..4096 (got %d)", p->disparities);
  if (p->cost < 0 || p->cost > 4) return fail(SGM_ERR_INVALID_ARGUMENT, "unknown matching cost %d", p->cost);
  if (p->directions != 4 && p->directions != 8)
    return fail(SGM_ERR_INVALID_ARGUMENT, "directions must be 4 or 8 (got %d)", p->directions);
  if (p->p1 < 0 || p->p2 < 0) return fail(SGM_ERR_INVALID_ARGUMENT, "penalties must be >= 0");
  if (p->edge_scale < 0 || p->edge_scale > 255)
    return fail(SGM_ERR_INVALID_ARGUMENT, "edge scale must be 0..255");
  if ((p->cost == 1 && p->ad_truncation < 1) || (p->cost == 2 && p->gradient_truncation < 1))
    return fail(SGM_ERR_INVALID_ARGUMENT, "truncation must be >= 1");
  if (p->census_weight < 0 || p->census_weight > 1000)
    return fail(SGM_ERR_INVALID_ARGUMENT, "census weight must be 0..1000");
  if (!(p->uniqueness >= 0 && p->uniqueness < 1))
    return fail(SGM_ERR_INVALID_ARGUMENT, "uniqueness must be in [0, 1)");
  if (!(p->lr_threshold > 0)) return fail(SGM_ERR_INVALID_ARGUMENT, "lr threshold must be > 0");
  if (aggregate_bytes(p) < 0)
    return fail(SGM_ERR_INVALID_ARGUMENT, "16-bit SGM aggregate cannot hold proved bound %lld",
                (long long)aggregate_bound(p));
  if (aggregate_bound(p) > 0x7FFFFFFFLL)
    return fail(SGM_ERR_INVALID_ARGUMENT, "aggregate bound %lld exceeds 32 bits",
                (long long)aggregate_bound(p));
  // CensusSgm.java:56-60
  if (w < p->window_w + 2 || h < p->window_h + 2)
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo image dimensions %dx%d", w, h);
  // descriptor planes are addressed with 32-bit byte offsets (8 B per pixel)
  if ((int64_t)w * h > (int64_t)1 << 28)
    return fail(SGM_ERR_INVALID_ARGUMENT, "image %dx%d exceeds 2^28 pixels", w, h);
  // the cost kernels launch one grid row per image row (grid y <= 65535)
  if (h > 65535)
    return fail(SGM_ERR_INVALID_ARGUMENT, "image height %d exceeds 65535 rows", h);
  return SGM_OK;
}

int effective_p2(const sgm_params *p) { return std::max(p->p1 + 1, p->p2); }

int64_t aggregate_bound(const sgm_params *p) {
  const int64_t maxc = p->cost == SGM_COST_ZNCC_FP32 ? 2048 : 512;
  return (int64_t)p->directions * (maxc + effective_p2(p));
}

int aggregate_bytes(const sgm_params *p) {
  const bool fits = aggregate_bound(p) <= 65535;
  if (p->aggregate_bytes == 2) return fits ? 2 : -1;
  if (p->aggregate_bytes == 4) return 4;
  return fits ? 2 : 4;
}

int lanes_group(int D) {
  int R = 1;
  while (R * 64 < D) R <<= 1;
  return R;
}

// Disparities per pixel row of the volumes: D rounded up to a multiple of 64,
// so every row is whole 64-byte HBM write granules (a uint8 row) and the sweeps
// store whole rows, pad cells included.  Rows of D bytes would leave granules
// part-written, which HBM completes by read-modify-write (measured: 1.23x over
// rows of D rounded to 4 for the batched-distribution pairs, D 122-356).
// Never above one wave's 64 * lanes_group(D) disparities.
int padded_disparities(int D) { return (D + 63) / 64 * 64; }

int words_of(const sgm_params *p) { return (p->window_w * p->window_h - 1 + 63) / 64; }

int64_t direction_bytes(const sgm_params *p, int w, int h, bool cost_volume) {
  const int64_t n = (int64_t)w * h, Dp = padded_disparities(p->disparities);
  const int sb = std::max(aggregate_bytes(p), 2);
  const bool zncc = p->cost == SGM_COST_ZNCC || p->cost == SGM_COST_ZNCC_FP32;
  const int cb = cost_volume ? (cost_u8(p) ? 1 : 2) : 0;
  // (cost +) aggregate volumes, descriptors, mask, gray|mask, gray copy, disparity, window sums
  return n * Dp * (cb + sb) + n * (8 * words_of(p) + 1 + 2 + 1 + 4 + (zncc ? 8 : 0));
}

bool census_recompute_possible(const sgm_params *p) {
  return p->cost == SGM_COST_CENSUS && words_of(p) == 1 && aggregate_bytes(p) == 2 &&
         path_lanes(p) <= 8;
}

int64_t direction_bytes(const sgm_params *p, int w, int h) {
  return direction_bytes(p, w, h, !census_recompute_possible(p));
}

int64_t pair_bytes(const sgm_params *p, int w, int h, bool cost_volume) {
  const int64_t n = (int64_t)w * h;
  // both directions + validity copies (2n), reasons / boundary in and out and
  // selection codes (5n), checked (4n), counters
  return 2 * direction_bytes(p, w, h, cost_volume) + n * (2 + 5 + 4) + 256;
}

int64_t pair_bytes(const sgm_params *p, int w, int h) {
  // smallest footprint of the pair (the fit test of the plan)
  return pair_bytes(p, w, h, !census_recompute_possible(p));
}

// Streamed plan (GpuStereo.java:56-83): whole-image inputs and outputs of both
// images, the cost (uint16) volume of one strip and two strip aggregate slots
// (one computed while the other moves over PCIe), and the ping-pong carry rows
// of up to three vertical/diagonal directions per pass.
int64_t streamed_pair_bytes(const sgm_params *p, int w, int h, int rows) {
  const int64_t n = (int64_t)w * h, Dp = padded_disparities(p->disparities);
  const int sb = std::max(aggregate_bytes(p), 2);
  const int64_t inputs = direction_bytes(p, w, h, false) - n * Dp * sb;
  // both matching directions: strip cost + two aggregate slots each, carry rows each
  const int64_t strip = 2 * (int64_t)w * rows * Dp * (2 + 2 * sb);
  const int64_t carries = 2 * 2 * 3 * (int64_t)w * Dp * 4;
  return 2 * inputs + strip + carries + n * (2 + 5 + 4) + 256;
}

// The aggregates of both matching directions, parked between the two passes.
int64_t streamed_host_bytes(const sgm_params *p, int w, int h) {
  return 2 * (int64_t)w * h * padded_disparities(p->disparities) * std::max(aggregate_bytes(p), 2);
}

// Host memory the streamed plan may use (GpuStereo.plan's hostBudget,
// Options.memoryMb): the context's setting, else 65% of physical memory.
int64_t host_budget(const sgm_context *ctx) {
  if (ctx->host_budget > 0) return ctx->host_budget;
  const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
  return pages > 0 && page > 0 ? (int64_t)pages * page / 100 * 65 : (int64_t)1 << 30;
}

// Free host memory (GpuStereo.availableRam): MemAvailable of /proc/meminfo.
static int64_t available_ram() {
  FILE *f = fopen("/proc/meminfo", "r");
  if (!f) return INT64_MAX;
  char line[256];
  int64_t kb = -1;
  while (fgets(line, sizeof line, f))
    if (sscanf(line, "MemAvailable: %lld kB", (long long *)&kb) == 1) break;
  fclose(f);
  return kb > 0 ? kb * 1024 : INT64_MAX;
}

// GpuStereo.plan's streamed branch (GpuStereo.java:56-83) for aggregate
// storage `storage` (SGM_AGGREGATE_*): the largest strip height (gpuStripRows,
// halved) whose device footprint fits the budget and whose host side fits
// the host budget.  RAM parks both directions' whole aggregates in pinned host
// memory; DISK in a file, through two pinned strip slots per direction;
// RECOMPUTE holds a block of as many strips as the host budget allows.  AUTO
// is RAM when the aggregate fits the host budget (and three quarters of the
// free memory), else RECOMPUTE.  rows = 0: no streamed plan fits.
StreamedPlan streamed_plan(sgm_context *ctx, const sgm_params *p, int w, int h, int storage) {
  StreamedPlan sp;
  const int64_t Dp = padded_disparities(p->disparities), sb = std::max(aggregate_bytes(p), 2);
  const int64_t full = 2 * (int64_t)w * h * Dp * sb, budget = host_budget(ctx);
  const bool fits_ram = full <= budget && full <= available_ram() / 4 * 3;
  for (int rows = std::min(h, ctx->strip_rows); rows >= 1; rows /= 2) {
    const int64_t device = streamed_pair_bytes(p, w, h, rows);
    if (device > ctx->budget) continue;
    const int64_t stripe = 2 * (int64_t)w * rows * Dp * sb;  // one strip, both directions
    sp.rows = rows;
    sp.device_bytes = device;
    sp.block_rows = h;
    sp.disk_bytes = 0;
    if (storage == SGM_AGGREGATE_RAM && !fits_ram) continue;
    if (storage == SGM_AGGREGATE_DISK) {
      if (2 * stripe > budget) continue;  // two pinned staging slots
      sp.storage = SGM_AGGREGATE_DISK;
      sp.host_bytes = 2 * stripe;
      sp.disk_bytes = full;
      return sp;
    }
    if (storage != SGM_AGGREGATE_RECOMPUTE && fits_ram) {
      sp.storage = SGM_AGGREGATE_RAM;
      sp.host_bytes = full;
      return sp;
    }
    const int64_t stripes = budget / stripe;
    if (stripes < 1) continue;
    sp.storage = SGM_AGGREGATE_RECOMPUTE;
    sp.block_rows = (int)std::min<int64_t>(h, stripes * rows);
    sp.host_bytes = 2 * (int64_t)w * sp.block_rows * Dp * sb;
    return sp;
  }
  return StreamedPlan();
}

int streamed_rows(sgm_context *ctx, const sgm_params *p, int w, int h) {
  return streamed_plan(ctx, p, w, h, ctx->aggregate_storage).rows;
}

// Compact plan (GpuCandidateStereo.plan / deviceBytes, GpuCandidateStereo.java:
// 19-47, 174-182): whole-image inputs and outputs of both images, per matching
// direction the proposals (2 B), their sums (6 B) per proposal, the off-proposal
// bound, verification slot, pixel list and winner summary (2 + 4 + 4 + 12 B per
// pixel), one strip's uint16 cost volume and six carry rows (three carried
// paths per pass, ping-pong); plus one pixel of
// verification per direction (the group size grows with the budget left).
// With 1-word windows and R <= 8 the sweeps compute their census costs on the
// fly over the whole image: no strip cost volume, no carries.
bool esgm_computed(const sgm_params *p) { return words_of(p) == 1 && path_lanes(p) <= 8; }

int64_t esgm_pair_bytes(const sgm_params *p, int w, int h, int rows, int pool) {
  const int64_t n = (int64_t)w * h, Dp = padded_disparities(p->disparities);
  const int64_t C = (int64_t)p->directions * pool;
  const int sb = std::max(aggregate_bytes(p), 2);
  const int64_t inputs = direction_bytes(p, w, h, false) - n * Dp * sb;
  const int64_t strip = esgm_computed(p) ? 0 : (int64_t)w * rows * Dp * 2 + 6 * (int64_t)w * Dp * 4;
  const int64_t per = n * (C * 8 + 2 + 4 + 4 + 12) + strip;
  return 2 * inputs + 2 * per + n * (2 + 5 + 4) + 2 * Dp * 2 + 512;
}

// Why the compact plan cannot run these parameters (GpuCandidateStereo.java:
// 19-30, 198-216), or nullptr.
const char *esgm_unsupported(const sgm_params *p, int w, int h) {
  if (p->cost != SGM_COST_CENSUS) return "the compact plan takes Census costs only";
  if (aggregate_bytes(p) != 2) return "the compact plan needs bounded 16-bit aggregates";
  if (aggregate_bound(p) * (p->disparities + 1LL) >= 0x7FFFFFFFLL)
    return "aggregate bound times the range exceeds 32 bits";
  if (p->disparities < 3 || p->disparities > 4096) return "the compact plan takes 3..4096 disparities";
  if (p->directions != 4 && p->directions != 8) return "the compact plan takes 4 or 8 paths";
  if (w < p->window_w + 2 || h < p->window_h + 2) return "Invalid research stereo raster dimensions";
  return nullptr;
}

// GpuCandidateStereo.plan: proposals per path from the configured count down
// to 1, strip rows from gpuStripRows halving, first combination that fits
// (rows = h where the costs are computed on the fly: no strips).
int esgm_plan(sgm_context *ctx, const sgm_params *p, int w, int h, int *rows, int *pool) {
  *rows = 0;
  *pool = 0;
  if (esgm_unsupported(p, w, h)) return SGM_OK;
  const int top = esgm_computed(p) ? h : std::min(h, ctx->strip_rows);
  for (int q = ctx->esgm_candidates; q >= 1; q--)
    for (int r = top; r >= 1; r = esgm_computed(p) ? 0 : r / 2)
      if (esgm_pair_bytes(p, w, h, r, q) <= ctx->budget) {
        *rows = r;
        *pool = q;
        return SGM_OK;
      }
  return SGM_OK;
}

}  // namespace sgm

// ---------------------------------------------------------------------------
extern "C" const char *sgm_last_error(void) { return g_last_error.c_str(); }

extern "C" int sgm_device_count(int *count) {
  if (!count) return fail(SGM_ERR_INVALID_ARGUMENT, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return SGM_OK;  // no runtime / no device: empty list, as GpuSgm.devices (GpuSgm.java:79-86)
  }
  *count = n;
  return SGM_OK;
}

extern "C" int sgm_device_info_get(int index, sgm_device_info *out) {
  if (!out) return fail(SGM_ERR_INVALID_ARGUMENT, "null output");
  int n = 0;
  sgm_device_count(&n);
  if (index < 0 || index >= n) return fail(SGM_ERR_NO_DEVICE, "No GPU at index %d", index);
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, index);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  memset(out, 0, sizeof *out);
  snprintf(out->name, sizeof out->name, "%s", prop.name);
  snprintf(out->arch, sizeof out->arch, "%s", prop.gcnArchName);
  out->memory = (int64_t)prop.totalGlobalMem;
  out->allocation = (int64_t)prop.totalGlobalMem;
  out->local_memory = (int64_t)prop.sharedMemPerBlock;
  out->group_size = prop.maxThreadsPerBlock;
  out->compute_units = prop.multiProcessorCount;
  return SGM_OK;
}

extern "C" int sgm_open(int device, int64_t budget_bytes, sgm_context **out) {
  if (!out) return fail(SGM_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  sgm_device_info info;
  int rc = sgm_device_info_get(device, &info);
  if (rc) return rc;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  // the no-allocation entry point's workspace on this device; a failed lookup
  // only disables that entry point (it reports the error)
  (void)scratch_setup(device);
  sgm_context *ctx = new sgm_context();
  ctx->device = device;
  ctx->info = info;
  // GpuSgm.budget (GpuSgm.java:629-633): 70% of VRAM by default, at most 95%.
  ctx->budget = budget_bytes > 0 ? std::min(budget_bytes, info.memory * 95 / 100) : info.memory * 7 / 10;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = ctx;
  return SGM_OK;
}

extern "C" int sgm_close(sgm_context *ctx) {
  if (!ctx) return SGM_OK;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->buffers) kv.second.release(ctx);
  for (auto &kv : ctx->host_buffers) kv.second.release();
  for (auto &pe : ctx->pending_events) {
    (void)hipEventDestroy(pe.start);
    (void)hipEventDestroy(pe.stop);
  }
  if (ctx->copy_stream) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    (void)hipStreamDestroy(ctx->copy_stream);
  }
  for (hipEvent_t e : ctx->strip_events)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SGM_OK;
}

extern "C" void *sgm_stream(sgm_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int sgm_plan_get(sgm_context *ctx, const sgm_params *params, int w, int h, sgm_plan *out) {
  if (!ctx || !out) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  memset(out, 0, sizeof *out);
  out->aggregate_bytes = aggregate_bytes(params);
  out->padded_disparities = padded_disparities(params->disparities);
  out->lanes_per_disparity_group = lanes_group(params->disparities);
  out->device_bytes = direction_bytes(params, w, h);
  out->pair_bytes = pair_bytes(params, w, h);
  out->budget_bytes = ctx->budget;
  out->aggregate_bound = aggregate_bound(params);
  out->resident = out->pair_bytes <= ctx->budget;
  const StreamedPlan sp = streamed_plan(ctx, params, w, h, ctx->aggregate_storage);
  out->strip_rows = sp.rows;
  out->streamed_bytes = sp.device_bytes;
  out->host_bytes = sp.host_bytes;
  out->aggregate_storage = sp.rows ? sp.storage : 0;
  out->block_rows = sp.rows ? sp.block_rows : 0;
  out->disk_bytes = sp.disk_bytes;
  out->streamed = ctx->plan_mode == SGM_PLAN_STREAMED ? out->strip_rows > 0
                  : ctx->plan_mode == SGM_PLAN_AUTO   ? !out->resident && out->strip_rows > 0
                                                      : 0;
  int rows, pool = 0;
  esgm_plan(ctx, params, w, h, &rows, &pool);
  out->esgm_candidates = pool;
  out->esgm_rows = rows;
  out->esgm_bytes = pool ? esgm_pair_bytes(params, w, h, rows, pool) : 0;
  // GpuStereo.execute (GpuStereo.java:176-193): the compact plan under
  // --gpu-sgm esgm, or under auto when only a bounded recomputation would fit
  out->compact = ctx->plan_mode == SGM_PLAN_ESGM ||
                 (ctx->plan_mode == SGM_PLAN_AUTO && ctx->aggregate_storage == SGM_AGGREGATE_AUTO &&
                  out->streamed && sp.storage == SGM_AGGREGATE_RECOMPUTE && pool > 0 &&
                  !esgm_unsupported(params, w, h));
  if (out->compact && ctx->plan_mode == SGM_PLAN_AUTO) out->streamed = 0;
  out->sweeps = !out->compact && !out->streamed && out->resident
                    ? sweep_schedule(ctx, params, w, h, !recompute_cost(ctx, params, w, h),
                                     ctx->budget)
                    : SGM_SWEEPS_SEQUENTIAL;
  return SGM_OK;
}

extern "C" int sgm_set_plan_mode(sgm_context *ctx, int mode, int strip_rows) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (mode < SGM_PLAN_AUTO || mode > SGM_PLAN_ESGM)
    return fail(SGM_ERR_INVALID_ARGUMENT, "plan mode %d not in 0..3", mode);
  if (strip_rows < 1) return fail(SGM_ERR_INVALID_ARGUMENT, "--gpu-strip-rows must be positive");
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->plan_mode = mode;
  ctx->strip_rows = strip_rows;
  return SGM_OK;
}

extern "C" int sgm_set_aggregate_storage(sgm_context *ctx, int storage, const char *directory,
                                         int64_t host_budget_bytes) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (storage < SGM_AGGREGATE_AUTO || storage > SGM_AGGREGATE_RECOMPUTE)
    return fail(SGM_ERR_INVALID_ARGUMENT, "aggregate storage %d not in 0..3", storage);
  if (host_budget_bytes < 0) return fail(SGM_ERR_INVALID_ARGUMENT, "negative host budget");
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->aggregate_storage = storage;
  if (directory && *directory) ctx->aggregate_dir = directory;
  ctx->host_budget = host_budget_bytes;
  return SGM_OK;
}

extern "C" int sgm_interrupt(sgm_context *ctx) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->interrupt.store(1);
  return SGM_OK;
}

extern "C" int sgm_debug_fail_host_allocations(sgm_context *ctx, int count) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->fail_host_allocations = count;
  return SGM_OK;
}

extern "C" int sgm_set_esgm_candidates(sgm_context *ctx, int candidates) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (candidates < 1 || candidates > 4)
    return fail(SGM_ERR_INVALID_ARGUMENT, "--esgm-candidates must be 1..4");
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->esgm_candidates = candidates;
  return SGM_OK;
}

extern "C" int sgm_esgm_counts(sgm_context *ctx, int64_t counts[4]) {
  if (!ctx || !counts) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(ctx->lock);
  for (int i = 0; i < 4; i++) counts[i] = ctx->esgm_last[i];
  return SGM_OK;
}

extern "C" int sgm_disparity_limit(sgm_context *ctx, const sgm_params *params, int w, int h,
                                   int *limit) {
  // GpuStereo.limit (GpuStereo.java:86-94): binary search over the plan.
  if (!ctx || !params || !limit) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  sgm_params q = *params;
  int low = 2, high = 4096;
  while (low < high) {
    int mid = (low + high + 1) / 2;
    q.disparities = mid;
    sgm_plan plan;
    // under the compact plan: the largest range it serves (its verification
    // runs inside its own footprint)
    bool ok = mid >= 3 && sgm_plan_get(ctx, &q, w, h, &plan) == SGM_OK &&
              (ctx->plan_mode == SGM_PLAN_ESGM
                   ? plan.esgm_candidates > 0
                   : (plan.resident && ctx->plan_mode != SGM_PLAN_STREAMED) || plan.streamed);
    if (ok) low = mid;
    else high = mid - 1;
  }
  *limit = low == 2 ? 0 : low;
  return SGM_OK;
}

extern "C" int sgm_set_keep_volumes(sgm_context *ctx, int keep) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->keep_volumes = keep != 0;
  return SGM_OK;
}

extern "C" int sgm_set_cost_volume(sgm_context *ctx, int policy) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (policy < SGM_VOLUME_AUTO || policy > SGM_VOLUME_RECOMPUTE)
    return fail(SGM_ERR_INVALID_ARGUMENT, "cost volume policy %d not in 0..2", policy);
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->volume_policy = policy;
  return SGM_OK;
}

extern "C" int sgm_set_sweeps(sgm_context *ctx, int schedule) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (schedule < SGM_SWEEPS_AUTO || schedule > SGM_SWEEPS_PLANES)
    return fail(SGM_ERR_INVALID_ARGUMENT, "sweep schedule %d not in 0..3", schedule);
  std::lock_guard<std::mutex> g(ctx->lock);
  ctx->sweeps = schedule;
  return SGM_OK;
}

extern "C" int sgm_set_profiling(sgm_context *ctx, int enable) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->profiling = enable != 0;
  return SGM_OK;
}

extern "C" int sgm_debug_fail_allocations(sgm_context *ctx, int count) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->fail_allocations = count;
  return SGM_OK;
}

extern "C" int sgm_release_buffers(sgm_context *ctx) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  for (auto &kv : ctx->buffers) kv.second.release(ctx);
  ctx->buffers.clear();
  for (auto &kv : ctx->host_buffers) kv.second.release();
  ctx->host_buffers.clear();
  ctx->last_w = ctx->last_h = 0;
  return SGM_OK;
}

extern "C" int sgm_synchronize(sgm_context *ctx) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  collect_profile(ctx);
  return SGM_OK;
}

namespace sgm {

void collect_profile(sgm_context *ctx) {
  for (auto &pe : ctx->pending_events) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pe.start, pe.stop) == hipSuccess) {
      auto &k = ctx->kernel_stats[pe.name];
      k.first += ms;
      k.second += 1;
    }
    (void)hipEventDestroy(pe.start);
    (void)hipEventDestroy(pe.stop);
  }
  ctx->pending_events.clear();
}

ProfileScope::ProfileScope(sgm_context *c, const char *n) : ctx(c), name(n) {
  if (ctx->profiling) {
    (void)hipEventCreate(&ev.start);
    (void)hipEventCreate(&ev.stop);
    ev.name = n;
    (void)hipEventRecord(ev.start, ctx->stream);
  }
}
ProfileScope::~ProfileScope() {
  if (ctx->profiling) {
    (void)hipEventRecord(ev.stop, ctx->stream);
    ctx->pending_events.push_back(ev);
  }
}

}  // namespace sgm

extern "C" int sgm_profile_json(sgm_context *ctx, char *buffer, size_t length) {
  if (!ctx || !buffer) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  std::string s = "{\"device_kernel_milliseconds\":{";
  bool first = true;
  for (auto &kv : ctx->kernel_stats) {
    char b[256];
    snprintf(b, sizeof b, "%s\"%s\":{\"total_ms\":%.6f,\"launches\":%lld}", first ? "" : ",",
             kv.first.c_str(), kv.second.first, (long long)kv.second.second);
    s += b;
    first = false;
  }
  char b[256];
  static const char *storages[4] = {"auto", "ram", "disk", "recompute"};
  static const char *sweep_names[4] = {"auto", "sequential", "paired", "planes"};
  snprintf(b, sizeof b,
           "},\"allocated_bytes\":%lld,\"budget_bytes\":%lld,\"last_plan\":{\"plan\":\"%s\","
           "\"storage\":\"%s\",\"block_rows\":%d,\"sweeps\":\"%s\",\"verify_groups\":%d}}",
           (long long)ctx->allocated, (long long)ctx->budget,
           ctx->last_compact ? "compact" : ctx->last_streamed ? "streamed" : "resident",
           ctx->last_streamed ? storages[ctx->splan.storage & 3] : "none",
           ctx->last_streamed ? ctx->splan.block_rows : 0,
           ctx->last_compact || ctx->last_streamed ? "sequential"
                                                   : sweep_names[ctx->last_sweeps & 3],
           ctx->last_compact ? ctx->esgm_groups : 0);
  s += b;
  if (s.size() + 1 > length) return fail(SGM_ERR_INVALID_ARGUMENT, "buffer too small (%zu)", s.size() + 1);
  memcpy(buffer, s.c_str(), s.size() + 1);
  ctx->kernel_stats.clear();
  return SGM_OK;
}

// ---------------------------------------------------------------------------
// Launchers
namespace sgm {

void parallel_copy(const std::vector<HostCopy> &copies) {
  constexpr size_t kChunk = size_t(4) << 20;
  std::vector<HostCopy> chunks;
  size_t total = 0;
  for (const HostCopy &c : copies) {
    for (size_t o = 0; o < c.bytes; o += kChunk)
      chunks.push_back({(uint8_t *)c.dst + o, (const uint8_t *)c.src + o, std::min(kChunk, c.bytes - o)});
    total += c.bytes;
  }
  const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
  const size_t threads = std::min<size_t>({(size_t)8, hw / 2, chunks.size()});
  if (threads <= 1 || total < (size_t(16) << 20)) {
    for (const HostCopy &c : chunks) memcpy(c.dst, c.src, c.bytes);
    return;
  }
  std::atomic<size_t> next{0};
  auto work = [&] {
    for (size_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < chunks.size();)
      memcpy(chunks[i].dst, chunks[i].src, chunks[i].bytes);
  };
  std::vector<std::thread> pool;
  try {
    for (size_t t = 1; t < threads; t++) pool.emplace_back(work);
  } catch (...) {
    // no more threads: the ones started and this one share the chunks
  }
  work();
  for (std::thread &t : pool) t.join();
}

int launch_check(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  return SGM_OK;
}

int launch_fill_u8(sgm_context *ctx, uint8_t *p, uint8_t v, size_t n) {
  if (!n) return SGM_OK;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  fill_u8_kernel<<<blocks, 256, 0, ctx->stream>>>(p, v, n);
  return launch_check("fill kernel");
}

int launch_fill_u64(sgm_context *ctx, unsigned long long *p, unsigned long long v, size_t n) {
  if (!n) return SGM_OK;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  fill_u64_kernel<<<blocks, 256, 0, ctx->stream>>>(p, v, n);
  return launch_check("fill kernel");
}

int launch_census(sgm_context *ctx, const CensusJobs &jobs, int njobs, int w, int h,
                  const sgm_params *p) {
  ProfileScope ps(ctx, "census");
  dim3 grid((w + kCensusTx - 1) / kCensusTx, (h + kCensusTy - 1) / kCensusTy, njobs);
  dim3 block(kCensusTx, kCensusTy);
  if (words_of(p) == 1)
    census_kernel<1><<<grid, block, 0, ctx->stream>>>(jobs, p->window_w, p->window_h);
  else
    census_kernel<2><<<grid, block, 0, ctx->stream>>>(jobs, p->window_w, p->window_h);
  return launch_check("census kernel");
}

CostArgs cost_args(const sgm_params *p) {
  CostArgs A;
  A.mode = p->cost;
  A.ww = p->window_w;
  A.wh = p->window_h;
  A.words = words_of(p);
  A.weight = p->census_weight;
  A.adcap = p->ad_truncation;
  A.gradcap = p->gradient_truncation;
  A.inv_code = invalid_code(p);
  A.u8 = cost_u8(p) ? 1 : 0;
  return A;
}

// Cost rows [J.y0, J.y0 + rows) of every job (rows < 0: the whole image).
int launch_cost(sgm_context *ctx, const CostJobs &jobs, int njobs, int w, int h, int Dp,
                const sgm_params *p, int rows, bool force_u16) {
  ProfileScope ps(ctx, "cost");
  CostArgs A = cost_args(p);
  if (force_u16) A.u8 = 0;
  const int gy = rows < 0 ? h : rows;
  if (p->cost == SGM_COST_ZNCC || p->cost == SGM_COST_ZNCC_FP32) {
    const int ntx = (w + kZnccTX - 1) / kZnccTX, ntd = (Dp + 63) / 64;
    dim3 grid(ntx * ntd, (gy + kZnccBand - 1) / kZnccBand, njobs);
    const bool out8 = A.u8 != 0;
#define ZNCC_LAUNCH(RR)                                                                   \
  case RR:                                                                                \
    if (p->cost == SGM_COST_ZNCC_FP32)                                                    \
      zncc_band_kernel<RR, 4, false><<<grid, 64, 0, ctx->stream>>>(jobs, A, gy);          \
    else if (out8)                                                                        \
      zncc_band_kernel<RR, 3, true><<<grid, 64, 0, ctx->stream>>>(jobs, A, gy);           \
    else                                                                                  \
      zncc_band_kernel<RR, 3, false><<<grid, 64, 0, ctx->stream>>>(jobs, A, gy);          \
    break;
    switch (p->window_w / 2) {
      ZNCC_LAUNCH(1) ZNCC_LAUNCH(2) ZNCC_LAUNCH(3) ZNCC_LAUNCH(4) ZNCC_LAUNCH(5)
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported ZNCC window");
    }
#undef ZNCC_LAUNCH
    if (int rc = launch_check("zncc cost kernel")) return rc;
    return SGM_OK;
  }
  const int quads = w * (Dp / 4);
  dim3 grid((quads + 255) / 256, gy, njobs);
  cost_kernel<<<grid, 256, 0, ctx->stream>>>(jobs, A);
  return launch_check("cost kernel");
}

PathArgs path_args(const sgm_params *p, int dir, int njobs, int total_lines, int keep_sum) {
  static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  PathArgs A;
  A.dx = dirs[dir][0];
  A.dy = dirs[dir][1];
  A.cls = A.dy == 0 ? 0 : A.dx == 0 ? 1 : 2;
  A.p1 = p->p1;
  A.p2 = effective_p2(p);
  A.edge = p->edge_scale;
  A.inv_code = invalid_code(p);
  A.val_mask = p->cost == SGM_COST_ZNCC_FP32 ? 0x7FFF : 0xFFFF;
  A.keep_sum = keep_sum;
  A.njobs = njobs;
  A.total_lines = total_lines;
  A.split = 0;
  A.uniq = p->uniqueness;
  A.order = nullptr;
  return A;
}

int lines_of(int cls, int w, int h) { return cls == 0 ? h : cls == 1 ? w : w + h - 1; }

int launch_path(sgm_context *ctx, int R, int sum_bytes, int mode, const PathJobs &jobs,
                const PathArgs &A, int phase, int waves, hipStream_t stream) {
  ProfileScope ps(ctx, phase == PHASE_FIRST ? "path_first"
                       : phase == PHASE_MID ? "path_mid"
                       : phase == PHASE_COST ? "path_cost"
                                             : "path_last");
  if (A.total_lines <= 0) return SGM_OK;
  hipStream_t st = stream ? stream : ctx->stream;
  if (phase == PHASE_FIRST) return launch_path_first(st, R, sum_bytes, mode, jobs, A, waves);
  if (phase == PHASE_COST) return launch_path_cost(st, R, sum_bytes, mode, jobs, A, waves);
  if (phase == PHASE_MID) return launch_path_mid(st, R, sum_bytes, mode, jobs, A, waves);
  return launch_path_last(st, R, sum_bytes, mode, jobs, A, waves);
}

int launch_lr(sgm_context *ctx, const LrArgs &A) {
  ProfileScope ps(ctx, A.dr ? "lr_check" : "count");
  const size_t n = (size_t)A.w * A.h;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
  lr_kernel<<<blocks, 256, 0, ctx->stream>>>(A);
  return launch_check("lr kernel");
}

int invalid_code(const sgm_params *p) { return p->cost == SGM_COST_ZNCC_FP32 ? 0x8800 : 512; }

bool cost_u8(const sgm_params *p) {
  // the packed 16-bit sweeps read census-family costs (<= 120, invalid 512 stored
  // as 255) from a uint8 volume; FP32 ZNCC costs (0..2048) and the 32-bit
  // aggregate sweeps keep uint16
  return aggregate_bytes(p) == 2 && p->cost != SGM_COST_ZNCC_FP32;
}

int path_lanes(const sgm_params *p) {
  // 16-bit aggregates run packed pairs: at least two disparities per lane.
  const int R = lanes_group(p->disparities);
  return aggregate_bytes(p) == 2 ? std::max(R, 2) : R;
}

bool recompute_cost(sgm_context *ctx, const sgm_params *p, int w, int h) {
  // census cost recomputed in every sweep (no cost volume): forced by the
  // policy, or when the resident cost volume would not fit the budget.  Kept
  // volumes for inspection always use the resident volume.
  if (ctx->keep_volumes || !census_recompute_possible(p)) return false;
  if (ctx->volume_policy == SGM_VOLUME_RECOMPUTE) return true;
  if (ctx->volume_policy == SGM_VOLUME_RESIDENT) return false;
  return pair_bytes(p, w, h, true) > ctx->budget;
}

bool fused_cost(const sgm_params *p) {
  // census cost computed inside the first (left-to-right) sweep
  return p->cost == SGM_COST_CENSUS && words_of(p) == 1 && aggregate_bytes(p) == 2 &&
         path_lanes(p) <= 8;
}

int path_waves(sgm_context *ctx, int R) {
  auto it = ctx->tuned_waves.find(R);
  return it != ctx->tuned_waves.end() ? it->second : 4;
}

}  // namespace sgm

// ---------------------------------------------------------------------------
// Plain per-pair execution.
namespace {

struct PairBuffers {
  uint8_t *gray[2], *valid[2], *mask[2];
  uint16_t *gm[2];
  uint64_t *desc[2];
  int32_t *wsum[2], *wsum2[2];
  uint16_t *cost[2];
  void *sum[2];
  // paired sweeps: the second partial aggregate volume of each matching
  // direction (null: one aggregate, directions swept one per launch)
  void *sum_b[2];
  // path-cost plane schedule: kPlanes planes of Dp bytes per pixel (interleaved
  // per pixel) for each matching direction in place of the aggregate (null: an
  // aggregate volume)
  uint8_t *planes[2];
  float *disp[2], *checked;
  uint8_t *reasons, *boundary, *reasons_out, *boundary_out, *code;
  unsigned long long *stats;
  // colour-distance census: interleaved BGR of both images on the device, or null
  const uint8_t *bgr[2];
  // filter options whose subpixel_refinement runs before the L-R check, or null
  const sgm_filter_params *refine;
  // streamed plan (strip_rows > 0): one strip's volumes, carry rows, host aggregate
  int strip_rows;
  uint16_t *strip_cost;
  void *strip_sum;
  int *carry;
  // host side of the streamed aggregate: the whole aggregate (RAM), a block of
  // rows (RECOMPUTE) or two staging slots per direction (DISK)
  uint8_t *host_sum;
  int storage, block_rows;
  // compact plan (esgm_pool > 0): proposals, their sums and the per-pixel
  // bound / slot / pixel list / summary of both matching directions
  int esgm_pool, esgm_rows;
  uint16_t *e_cand, *e_sums, *e_outside, *e_strip_cost;
  int *e_slot, *e_pixels, *e_count, *e_carry;
  uint32_t *e_summary;
  // resident plan without a cost volume: the census cost is recomputed in
  // every sweep
  bool recompute;
};

}  // namespace

// Device bytes a sweep schedule adds to pair_bytes(p, w, h, true): the second
// partial aggregate volumes of paired sweeps, or the path-cost planes of both
// matching directions (kPlanes bytes per cell) in place of the aggregates.
int64_t sgm::schedule_bytes(const sgm_params *p, int w, int h, int schedule) {
  const int64_t cells = 2 * (int64_t)w * h * padded_disparities(p->disparities);
  if (schedule == SGM_SWEEPS_PAIRED) return cells * aggregate_bytes(p);
  if (schedule == SGM_SWEEPS_PLANES) return cells * (kPlanes - aggregate_bytes(p));
  return 0;
}

// Whether a multi-direction schedule applies to the pair (budget aside).  Both
// need packed 16-bit sweeps of lane groups 2..8 over a cost volume.  Path-cost
// planes also need 8 paths, uint8 costs, no kept volumes (the planes replace
// the aggregate), and path costs that fit a byte: every valid cell's path cost
// is at most its cost (<= window bits - 1 for the uint8 costs) plus P2.
bool sgm::schedule_applies(const sgm_context *ctx, const sgm_params *p, int schedule,
                           bool cost_volume) {
  if (schedule == SGM_SWEEPS_SEQUENTIAL) return true;
  if (!cost_volume || aggregate_bytes(p) != 2 || path_lanes(p) > 8) return false;
  if (schedule == SGM_SWEEPS_PAIRED) return true;
  return schedule == SGM_SWEEPS_PLANES && !ctx->keep_volumes && cost_u8(p) &&
         p->directions == kPlanes + 1 && p->window_w * p->window_h - 1 + effective_p2(p) <= 255;
}

// Sweep schedule of a resident pair under `budget` (SGM_SWEEPS_SEQUENTIAL,
// _PAIRED or _PLANES): AUTO and PLANES take the planes, then the pairs, where
// they apply and fit the budget; PAIRED the pairs where they fit.
int sgm::sweep_schedule(const sgm_context *ctx, const sgm_params *p, int w, int h,
                        bool cost_volume, int64_t budget) {
  const int want = ctx->sweeps;
  if (want == SGM_SWEEPS_SEQUENTIAL) return SGM_SWEEPS_SEQUENTIAL;
  const int64_t base = pair_bytes(p, w, h, true);
  for (int sched : {SGM_SWEEPS_PLANES, SGM_SWEEPS_PAIRED}) {
    if (sched == SGM_SWEEPS_PLANES && want == SGM_SWEEPS_PAIRED) continue;
    if (schedule_applies(ctx, p, sched, cost_volume) &&
        base + schedule_bytes(p, w, h, sched) <= budget)
      return sched;
  }
  return SGM_SWEEPS_SEQUENTIAL;
}

namespace {

// Whole-image device buffers of a pair, by name and size: inputs, descriptors
// and disparities of both matching directions, the outputs, and (`volumes`)
// the aggregate volumes (path-cost planes under SGM_SWEEPS_PLANES, a second
// partial aggregate under SGM_SWEEPS_PAIRED) with (`cost_volume`) the cost
// volumes.  Their sum is pair_bytes(p, w, h, cost_volume) (+ schedule_bytes)
// when `volumes` is set.
static std::vector<std::pair<std::string, size_t>> pair_buffer_list(const sgm_params *p, int w,
                                                                    int h, bool volumes,
                                                                    bool cost_volume,
                                                                    int schedule) {
  const size_t n = (size_t)w * h, Dp = padded_disparities(p->disparities);
  const int sb = aggregate_bytes(p), words = words_of(p);
  const bool zncc = p->cost == SGM_COST_ZNCC || p->cost == SGM_COST_ZNCC_FP32;
  std::vector<std::pair<std::string, size_t>> L;
  for (int d = 0; d < 2; d++) {
    const std::string s = std::to_string(d);
    L.emplace_back("gray" + s, n);
    L.emplace_back("valid" + s, n);
    L.emplace_back("mask" + s, n);
    L.emplace_back("gm" + s, n * 2);
    L.emplace_back("desc" + s, n * 8 * words);
    if (zncc) {
      L.emplace_back("wsum" + s, n * 4);
      L.emplace_back("wsum2" + s, n * 4);
    }
    if (volumes && cost_volume) L.emplace_back("cost" + s, n * Dp * (cost_u8(p) ? 1 : 2));
    if (volumes && schedule == SGM_SWEEPS_PLANES) L.emplace_back("planes" + s, n * Dp * kPlanes);
    else if (volumes) L.emplace_back("sum" + s, n * Dp * sb);
    if (volumes && schedule == SGM_SWEEPS_PAIRED) L.emplace_back("sumb" + s, n * Dp * sb);
    L.emplace_back("disp" + s, n * 4);
  }
  L.emplace_back("checked", n * 4);
  L.emplace_back("reasons", n);
  L.emplace_back("boundary", n);
  L.emplace_back("reasons_out", n);
  L.emplace_back("boundary_out", n);
  L.emplace_back("stats", 16 * 8);
  L.emplace_back("code", n);
  return L;
}

}  // namespace

// Device bytes the context holds once ensure_pair_buffers has served the pair
// on the resident plan with sweep schedule `schedule`: each of the pair's
// buffers grows to its size when its cached capacity is smaller, every other
// cached buffer stays as it is (`fresh`: no cached buffers).
int64_t sgm::resident_footprint(sgm_context *ctx, const sgm_params *p, int w, int h,
                                bool cost_volume, int schedule, bool fresh) {
  int64_t total = fresh ? 0 : ctx->allocated;
  for (const auto &b : pair_buffer_list(p, w, h, true, cost_volume, schedule)) {
    if (fresh) {
      total += (int64_t)b.second;
      continue;
    }
    auto it = ctx->buffers.find(b.first);
    const size_t cap = it != ctx->buffers.end() && it->second.ptr ? it->second.cap : 0;
    if (b.second > cap) total += (int64_t)(b.second - cap);
  }
  return total;
}

namespace {

// Resident (rows = 0), strip-streamed (rows > 0) or compact (pool > 0, strips
// of `rows` rows) buffers of a pair.
int ensure_pair_buffers(sgm_context *ctx, const sgm_params *p, int w, int h, PairBuffers &B,
                        int rows = 0, int pool = 0) {
  const size_t n = (size_t)w * h, Dp = padded_disparities(p->disparities);
  const int sb = aggregate_bytes(p);
  const bool zncc = p->cost == SGM_COST_ZNCC || p->cost == SGM_COST_ZNCC_FP32;
  const bool recompute = !rows && recompute_cost(ctx, p, w, h);
  const int schedule =
      rows ? SGM_SWEEPS_SEQUENTIAL : sweep_schedule(ctx, p, w, h, !recompute, ctx->budget);
  int rc;
#define ENSURE(name, bytes)                                                   \
  if ((rc = ctx->buffers[name].ensure(ctx, (bytes), name)) != SGM_OK) return rc;
  // volumes the plan and schedule do not use are released before allocating
  if (rows) {
    // the streamed and compact plans hold no whole-image volume
    for (const char *v : {"cost0", "cost1", "sum0", "sum1"}) ctx->buffers[v].release(ctx);
  }
  if (schedule != SGM_SWEEPS_PAIRED)
    for (const char *v : {"sumb0", "sumb1"}) ctx->buffers[v].release(ctx);
  if (schedule != SGM_SWEEPS_PLANES)
    for (const char *v : {"planes0", "planes1"}) ctx->buffers[v].release(ctx);
  else
    for (const char *v : {"sum0", "sum1"}) ctx->buffers[v].release(ctx);
  for (const auto &b : pair_buffer_list(p, w, h, rows == 0, !recompute, schedule))
    ENSURE(b.first.c_str(), b.second);
  for (int d = 0; d < 2; d++) {
    const std::string s = std::to_string(d);
    B.gray[d] = (uint8_t *)ctx->buffers["gray" + s].ptr;
    B.valid[d] = (uint8_t *)ctx->buffers["valid" + s].ptr;
    B.mask[d] = (uint8_t *)ctx->buffers["mask" + s].ptr;
    B.gm[d] = (uint16_t *)ctx->buffers["gm" + s].ptr;
    B.desc[d] = (uint64_t *)ctx->buffers["desc" + s].ptr;
    B.wsum[d] = zncc ? (int32_t *)ctx->buffers["wsum" + s].ptr : nullptr;
    B.wsum2[d] = zncc ? (int32_t *)ctx->buffers["wsum2" + s].ptr : nullptr;
    B.cost[d] = rows || recompute ? nullptr : (uint16_t *)ctx->buffers["cost" + s].ptr;
    B.recompute = recompute;
    B.sum[d] = rows || schedule == SGM_SWEEPS_PLANES ? nullptr : ctx->buffers["sum" + s].ptr;
    B.sum_b[d] = schedule == SGM_SWEEPS_PAIRED ? ctx->buffers["sumb" + s].ptr : nullptr;
    B.planes[d] =
        schedule == SGM_SWEEPS_PLANES ? (uint8_t *)ctx->buffers["planes" + s].ptr : nullptr;
    B.disp[d] = (float *)ctx->buffers["disp" + s].ptr;
  }
  B.bgr[0] = B.bgr[1] = nullptr;
  B.refine = nullptr;
  B.strip_rows = pool ? 0 : rows;
  B.strip_cost = nullptr;
  B.strip_sum = nullptr;
  B.carry = nullptr;
  B.host_sum = nullptr;
  B.esgm_pool = pool;
  B.esgm_rows = pool ? rows : 0;
  B.e_cand = B.e_sums = B.e_outside = B.e_strip_cost = nullptr;
  B.e_slot = B.e_pixels = B.e_count = B.e_carry = nullptr;
  B.e_summary = nullptr;
  if (pool) {
    const size_t C = (size_t)p->directions * pool;
    ENSURE("esgm_cand", 2 * n * C * 2);
    ENSURE("esgm_sums", 2 * n * C * 6);
    ENSURE("esgm_outside", 2 * n * 2);
    ENSURE("esgm_slot", 2 * n * 4);
    ENSURE("esgm_pixels", 2 * n * 4);
    ENSURE("esgm_summary", 2 * n * 12);
    const bool computed = esgm_computed(p);
    ENSURE("esgm_strip_cost", computed ? 8 : 2 * (size_t)w * rows * Dp * 2);
    ENSURE("esgm_carry", computed ? 8 : 2 * 6 * (size_t)w * Dp * 4);
    ENSURE("esgm_count", 256);
    B.e_cand = (uint16_t *)ctx->buffers["esgm_cand"].ptr;
    B.e_sums = (uint16_t *)ctx->buffers["esgm_sums"].ptr;
    B.e_outside = (uint16_t *)ctx->buffers["esgm_outside"].ptr;
    B.e_slot = (int *)ctx->buffers["esgm_slot"].ptr;
    B.e_pixels = (int *)ctx->buffers["esgm_pixels"].ptr;
    B.e_summary = (uint32_t *)ctx->buffers["esgm_summary"].ptr;
    B.e_strip_cost = (uint16_t *)ctx->buffers["esgm_strip_cost"].ptr;
    B.e_carry = (int *)ctx->buffers["esgm_carry"].ptr;
    B.e_count = (int *)ctx->buffers["esgm_count"].ptr;
  }
  B.storage = 0;
  B.block_rows = h;
  if (rows && !pool) {
    // host side first: a RAM aggregate that cannot be allocated falls back to
    // bounded recomputation under --gpu-aggregate auto (GpuStereo.java:198-208)
    StreamedPlan &sp = ctx->splan;
    auto host_of = [&](const StreamedPlan &q) {
      const size_t stripe = 2 * (size_t)w * q.rows * Dp * sb;  // both directions
      return q.storage == SGM_AGGREGATE_RAM ? 2 * n * Dp * sb
             : q.storage == SGM_AGGREGATE_RECOMPUTE ? 2 * (size_t)w * q.block_rows * Dp * sb
                                                    : 2 * stripe;
    };
    rc = ctx->host_buffers["sum"].ensure(host_of(sp), "streamed aggregate",
                                         &ctx->fail_host_allocations);
    if (rc == SGM_ERR_OUT_OF_MEMORY && sp.storage == SGM_AGGREGATE_RAM &&
        ctx->aggregate_storage == SGM_AGGREGATE_AUTO) {
      const StreamedPlan retry = streamed_plan(ctx, p, w, h, SGM_AGGREGATE_RECOMPUTE);
      if (retry.rows) {
        sp = retry;
        rc = ctx->host_buffers["sum"].ensure(host_of(sp), "streamed aggregate",
                                             &ctx->fail_host_allocations);
      }
    }
    if (rc) return rc;
    B.host_sum = (uint8_t *)ctx->host_buffers["sum"].ptr;
    B.storage = sp.storage;
    B.block_rows = sp.block_rows;
    B.strip_rows = sp.rows;
    const size_t cells = (size_t)w * sp.rows * Dp;
    // up to two matching directions per strip launch
    ENSURE("strip_cost", 2 * cells * 2);
    ENSURE("strip_sum", 2 * 2 * cells * sb);
    ENSURE("carry", 2 * 2 * 3 * (size_t)w * Dp * 4);
    B.strip_cost = (uint16_t *)ctx->buffers["strip_cost"].ptr;
    B.strip_sum = ctx->buffers["strip_sum"].ptr;
    B.carry = (int *)ctx->buffers["carry"].ptr;
  }
#undef ENSURE
  B.checked = (float *)ctx->buffers["checked"].ptr;
  B.reasons = (uint8_t *)ctx->buffers["reasons"].ptr;
  B.boundary = (uint8_t *)ctx->buffers["boundary"].ptr;
  B.reasons_out = (uint8_t *)ctx->buffers["reasons_out"].ptr;
  B.boundary_out = (uint8_t *)ctx->buffers["boundary_out"].ptr;
  B.stats = (unsigned long long *)ctx->buffers["stats"].ptr;
  B.code = (uint8_t *)ctx->buffers["code"].ptr;
  return SGM_OK;
}

template <int R, typename SumT>
void launch_strip_t(sgm_context *ctx, const StripJobs &J, const PathArgs &A, int phase) {
  constexpr int waves = 4;
  const unsigned blocks = (unsigned)((A.total_lines + waves - 1) / waves);
  if (phase == PHASE_FIRST)
    strip_kernel<R, SumT, PHASE_FIRST><<<blocks, waves * 64, 0, ctx->stream>>>(J, A);
  else if (phase == PHASE_MID)
    strip_kernel<R, SumT, PHASE_MID><<<blocks, waves * 64, 0, ctx->stream>>>(J, A);
  else
    strip_kernel<R, SumT, PHASE_LAST><<<blocks, waves * 64, 0, ctx->stream>>>(J, A);
}

int launch_strip(sgm_context *ctx, int R, int sum_bytes, const StripJobs &J, const PathArgs &A,
                 int phase) {
  ProfileScope ps(ctx, "strip");
#define CASE_STRIP(RR, T)                                                   \
  case RR:                                                                  \
    launch_strip_t<RR, T>(ctx, J, A, phase);                                \
    break;
  if (sum_bytes == 4) {
    switch (R) {
      CASE_STRIP(1, int32_t) CASE_STRIP(2, int32_t) CASE_STRIP(4, int32_t) CASE_STRIP(8, int32_t)
      CASE_STRIP(16, int32_t) CASE_STRIP(32, int32_t) CASE_STRIP(64, int32_t)
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  } else {
    switch (R) {
      CASE_STRIP(2, uint16_t) CASE_STRIP(4, uint16_t) CASE_STRIP(8, uint16_t)
      CASE_STRIP(16, uint16_t) CASE_STRIP(32, uint16_t) CASE_STRIP(64, uint16_t)
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  }
#undef CASE_STRIP
  return launch_check("strip kernel");
}

// Temporary file of the disk aggregate storage (GpuAggregate.java:26-88): made
// in the storage directory and unlinked at once, so it disappears when closed
// (or when the process ends); written and read in whole strips.
struct AggregateFile {
  int fd = -1;
  ~AggregateFile() {
    if (fd >= 0) close(fd);
  }
  int open(const std::string &dir, int64_t bytes) {
    struct statvfs sv;
    if (statvfs(dir.c_str(), &sv) == 0) {
      const int64_t avail = (int64_t)sv.f_bavail * (int64_t)sv.f_frsize;
      if (avail > 0 && bytes + (16 << 20) > avail)
        return fail(SGM_ERR_OUT_OF_MEMORY,
                    "Streamed GPU SGM needs %lld MiB temporary aggregate storage; insufficient "
                    "disk space", (long long)(bytes >> 20));
    }
    std::string path = dir + "/gpu-sgm-aggregate-XXXXXX";
    std::vector<char> name(path.begin(), path.end());
    name.push_back('\0');
    fd = mkstemp(name.data());
    if (fd < 0) return fail(SGM_ERR_INVALID_ARGUMENT, "cannot create a temporary aggregate in %s", dir.c_str());
    unlink(name.data());
    return SGM_OK;
  }
  int write(const uint8_t *src, size_t bytes, size_t off) {
    for (size_t done = 0; done < bytes;) {
      const ssize_t r = pwrite(fd, src + done, bytes - done, (off_t)(off + done));
      if (r <= 0) return fail(SGM_ERR_HIP, "Streamed aggregate I/O made no progress");
      done += (size_t)r;
    }
    return SGM_OK;
  }
  int read(uint8_t *dst, size_t bytes, size_t off) {
    for (size_t done = 0; done < bytes;) {
      const ssize_t r = pread(fd, dst + done, bytes - done, (off_t)(off + done));
      if (r <= 0) return fail(SGM_ERR_HIP, "Incomplete streamed SGM aggregate");
      done += (size_t)r;
    }
    return SGM_OK;
  }
};

// Strip-streamed counterpart of run_directions (the streamed loop of
// GpuStereo.executeTargets, GpuStereo.java:336-451).  The reference runs every
// direction over every strip and moves the strip's aggregate through host
// memory between directions.  Here the directions are split by vertical sense:
// a top-down pass over the strips runs all directions with dy >= 0 (horizontal
// ones included) while the strip is on the device, parks the strip's partial
// aggregate in pinned host memory, and a bottom-up pass brings it back, adds
// the dy < 0 directions and selects.  Aggregates are integer sums, so the
// order of the directions does not change them; each direction's path is the
// same recurrence over the same values as in one whole-image sweep.  One host
// round trip of the aggregate per pair instead of one per direction.
int run_directions_streamed(sgm_context *ctx, const sgm_params *p, int w, int h, PairBuffers &B,
                            const int *slots, int nd, int out_slot, uint8_t *reasons,
                            uint8_t *boundary) {
  static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  int rc;
  const int D = p->disparities, Dp = padded_disparities(D), R = path_lanes(p);
  const int sb = aggregate_bytes(p), rows = B.strip_rows;
  const int strips = (h + rows - 1) / rows;
  int down[8], up[8], ndown = 0, nup = 0;
  for (int dir = 0; dir < p->directions; dir++) {
    if (dirs[dir][1] >= 0) down[ndown++] = dir;
    else up[nup++] = dir;
  }
  // per matching direction i of the call: strip cost i, aggregate slots (slot, i),
  // six carry rows, the host aggregate i
  const size_t carry_row = (size_t)w * Dp;  // ints
  const size_t cells = (size_t)w * rows * Dp, slot_bytes = cells * sb;
  const size_t host_bytes = (size_t)w * h * Dp * sb;
  hipError_t e = hipSuccess;
  if (!ctx->copy_stream) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
  for (int k = 0; k < 4 && e == hipSuccess; k++)
    if (!ctx->strip_events[k])
      e = hipEventCreateWithFlags(&ctx->strip_events[k], hipEventDisableTiming);
  if (e != hipSuccess) return hip_fail(e, "streamed plan stream/events");
  hipStream_t cs = ctx->copy_stream;
  hipEvent_t *computed = ctx->strip_events, *copied = ctx->strip_events + 2;
  // both slots start free: earlier work on the compute stream is all they wait for
  for (int k = 0; k < 4; k++)
    if ((e = hipEventRecord(ctx->strip_events[k], ctx->stream)) != hipSuccess)
      return hip_fail(e, "hipEventRecord");
  CostJobs cj;
  memset(&cj, 0, sizeof cj);
  StripJobs sj;
  memset(&sj, 0, sizeof sj);
  for (int i = 0; i < nd; i++) {
    const int d = slots[i], o = 1 - d;
    CostJob &C = cj.job[i];
    C.desc_a = B.desc[d];
    C.desc_b = B.desc[o];
    C.mask_a = B.mask[d];
    C.mask_b = B.mask[o];
    C.gray_a = B.gray[d];
    C.gray_b = B.gray[o];
    C.sum_a = B.wsum[d];
    C.sum2_a = B.wsum2[d];
    C.sum_b = B.wsum[o];
    C.sum2_b = B.wsum2[o];
    C.cost = B.strip_cost + i * cells;
    C.w = w;
    C.h = h;
    C.D = D;
    C.Dp = Dp;
    C.minD = p->min_disparity;
    C.reverse = d;
    StripJob &J = sj.job[i];
    J.cost = C.cost;
    J.mask = B.mask[d];
    J.gray = B.gray[d];
    J.disp = B.disp[d];
    J.reasons = d == out_slot ? reasons : nullptr;
    J.boundary = d == out_slot ? boundary : nullptr;
    J.code = d == out_slot ? B.code : nullptr;
    J.w = w;
    J.h = h;
    J.D = D;
    J.Dp = Dp;
    J.minD = p->min_disparity;
    J.rx = p->window_w / 2;
    J.ry = p->window_h / 2;
  }
  auto dev_slot = [&](int slot, int i) {
    return (uint8_t *)B.strip_sum + ((size_t)slot * nd + i) * slot_bytes;
  };
  const int passes = nup ? 2 : 1;
  // Host side of the aggregate.  RAM: direction i's whole aggregate at
  // host_sum + i * host_bytes.  RECOMPUTE: the current block's rows only.
  // DISK: a temporary file (direction i at i * host_bytes), written and read
  // through the pinned staging slots host_sum + (slot * nd + i) * slot_bytes.
  const bool disk = B.storage == SGM_AGGREGATE_DISK;
  const int block_rows = B.storage == SGM_AGGREGATE_RECOMPUTE ? B.block_rows : h;
  const size_t dir_bytes = B.storage == SGM_AGGREGATE_RECOMPUTE
                               ? (size_t)w * block_rows * Dp * sb : host_bytes;
  AggregateFile file;
  if (disk && passes == 2 && (rc = file.open(ctx->aggregate_dir, (int64_t)nd * host_bytes)))
    return rc;
  auto stage = [&](int slot, int i) { return B.host_sum + ((size_t)slot * nd + i) * slot_bytes; };
  // pending disk writes of pass 0: the strip's slot and offset, written once
  // its download is done
  int pend_slot = -1;
  size_t pend_off = 0, pend_bytes = 0;
  auto flush_pending = [&]() -> int {
    if (pend_slot < 0) return SGM_OK;
    hipError_t e2 = hipEventSynchronize(copied[pend_slot]);
    if (e2 != hipSuccess) return hip_fail(e2, "hipEventSynchronize");
    for (int i = 0; i < nd; i++)
      if (int r2 = file.write(stage(pend_slot, i), pend_bytes, (size_t)i * host_bytes + pend_off))
        return r2;
    pend_slot = -1;
    return SGM_OK;
  };
  auto interrupted = [&]() {
    if (!ctx->interrupt.exchange(0)) return false;
    (void)hipStreamSynchronize(cs);
    (void)hipStreamSynchronize(ctx->stream);
    return true;
  };
  for (int b0 = 0; b0 < h; b0 += block_rows) {
    const int b1 = std::min(h, b0 + block_rows);
    const int s_first = b0 / rows, s_last = (b1 - 1) / rows;
    for (int pass = 0; pass < passes; pass++) {
      const int *list = pass == 0 ? down : up;
      const int count = pass == 0 ? ndown : nup;
      // strips of the pass: top-down to the block's last strip, bottom-up to
      // its first; strips outside the block only carry the paths into it
      const int nstrips = pass == 0 ? s_last + 1 : strips - s_first;
      for (int ord = 0; ord < nstrips; ord++) {
        if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
        const int s = pass == 0 ? ord : strips - 1 - ord, slot = ord & 1;
        const bool target = s >= s_first && s <= s_last;
        const int y0 = s * rows, act = std::min(rows, h - y0);
        const size_t bytes = (size_t)w * act * Dp * sb;
        const size_t hoff = (size_t)(y0 - (B.storage == SGM_AGGREGATE_RECOMPUTE ? b0 : 0)) * w * Dp * sb;
        if (pass == 1 && target) {
          // upload on the copy stream once the slot's previous strip is done;
          // it overlaps the sweeps of the strip before
          if (disk) {
            // the staging slot is free once its previous upload is done
            if ((e = hipEventSynchronize(copied[slot])) != hipSuccess)
              return hip_fail(e, "hipEventSynchronize");
            for (int i = 0; i < nd; i++)
              if ((rc = file.read(stage(slot, i), bytes, (size_t)i * host_bytes + hoff))) return rc;
          }
          if ((e = hipStreamWaitEvent(cs, computed[slot], 0)) != hipSuccess)
            return hip_fail(e, "hipStreamWaitEvent");
          for (int i = 0; i < nd; i++) {
            const uint8_t *src = disk ? stage(slot, i) : B.host_sum + i * dir_bytes + hoff;
            if ((e = hipMemcpyAsync(dev_slot(slot, i), src, bytes, hipMemcpyHostToDevice, cs)) !=
                hipSuccess)
              return hip_fail(e, "strip aggregate upload");
          }
          if ((e = hipEventRecord(copied[slot], cs)) != hipSuccess) return hip_fail(e, "hipEventRecord");
        }
        // the slot is free (pass 0: its last download is done) or filled (pass 1)
        if ((e = hipStreamWaitEvent(ctx->stream, copied[slot], 0)) != hipSuccess)
          return hip_fail(e, "hipStreamWaitEvent");
        for (int i = 0; i < nd; i++) cj.job[i].y0 = y0;
        if ((rc = launch_cost(ctx, cj, nd, w, h, Dp, p, act, true))) return rc;
        for (int i = 0; i < nd; i++) {
          sj.job[i].sum = dev_slot(slot, i);
          sj.job[i].y0 = y0;
          sj.job[i].rows = act;
        }
        int carried = 0;
        for (int t = 0; t < count; t++) {
          const int dir = list[t];
          PathArgs A = path_args(p, dir, nd, 0, 0);
          // outside the block only the paths that cross strips run, each
          // starting its strip aggregate afresh (never read back)
          if (!target && A.dy == 0) continue;
          const bool first = pass == 0 && t == 0, last = pass == passes - 1 && t == count - 1;
          const int phase = !target ? PHASE_FIRST : first ? PHASE_FIRST : last ? PHASE_LAST
                                                                             : PHASE_MID;
          sj.split = segments_of(A.cls, w, act);
          A.total_lines = nd * sj.split;
          for (int i = 0; i < nd; i++) {
            StripJob &J = sj.job[i];
            J.carry_in = J.carry_out = nullptr;
            if (A.dy != 0) {
              int *base = B.carry + (size_t)i * 6 * carry_row;
              J.carry_in = ord > 0 ? base + ((size_t)(ord & 1) * 3 + carried) * carry_row : nullptr;
              J.carry_out = ord < nstrips - 1
                                ? base + ((size_t)((ord + 1) & 1) * 3 + carried) * carry_row
                                : nullptr;
            }
          }
          if (A.dy != 0) carried++;
          if ((rc = launch_strip(ctx, R, sb, sj, A, phase))) return rc;
        }
        if ((e = hipEventRecord(computed[slot], ctx->stream)) != hipSuccess)
          return hip_fail(e, "hipEventRecord");
        if (pass == 0 && passes == 2 && target) {
          // download on the copy stream, overlapping the sweeps of the next strip
          if (disk && pend_slot == slot && (rc = flush_pending())) return rc;
          if ((e = hipStreamWaitEvent(cs, computed[slot], 0)) != hipSuccess)
            return hip_fail(e, "hipStreamWaitEvent");
          for (int i = 0; i < nd; i++) {
            uint8_t *dst = disk ? stage(slot, i) : B.host_sum + i * dir_bytes + hoff;
            if ((e = hipMemcpyAsync(dst, dev_slot(slot, i), bytes, hipMemcpyDeviceToHost, cs)) !=
                hipSuccess)
              return hip_fail(e, "strip aggregate download");
          }
          if ((e = hipEventRecord(copied[slot], cs)) != hipSuccess) return hip_fail(e, "hipEventRecord");
          if (disk) {
            // write the previous strip (other slot) while this one downloads
            if ((rc = flush_pending())) return rc;
            pend_slot = slot;
            pend_off = hoff;
            pend_bytes = bytes;
          }
        }
      }
      if (disk && pass == 0 && (rc = flush_pending())) return rc;
    }
  }
  return SGM_OK;
}

template <int R, int PH, bool HREC>
void launch_compact_t(sgm_context *ctx, const CompactJobs &J, const PathArgs &A,
                      const CompactArgs &E, int Dp) {
  // two Dp-int LDS rows per wave, at most 32 KiB per workgroup
  const int waves = Dp <= 1024 ? 4 : Dp <= 2048 ? 2 : 1;
  const unsigned blocks = (unsigned)((A.total_lines + waves - 1) / waves);
  const size_t lds = (size_t)waves * 2 * Dp * sizeof(int);
  compact_kernel<R, PH, HREC><<<blocks, waves * 64, lds, ctx->stream>>>(J, A, E);
}

template <int R, bool HREC>
void launch_compact_phase(sgm_context *ctx, int phase, const CompactJobs &J, const PathArgs &A,
                          const CompactArgs &E, int Dp) {
  if (phase == ESGM_CANDIDATES) launch_compact_t<R, ESGM_CANDIDATES, HREC>(ctx, J, A, E, Dp);
  else if (phase == ESGM_SUMS) launch_compact_t<R, ESGM_SUMS, HREC>(ctx, J, A, E, Dp);
  else launch_compact_t<R, ESGM_VERIFY, HREC>(ctx, J, A, E, Dp);
}

// hrec: whole-image lines with the census cost computed on the fly (1-word
// windows, R <= 8); otherwise lines of a strip of the cost volume.
int launch_compact(sgm_context *ctx, int R, int phase, bool hrec, const CompactJobs &J,
                   const PathArgs &A, const CompactArgs &E, int Dp) {
  static const char *names[2][3] = {
      {"compact proposals (strips)", "compact sums (strips)", "compact verify (strips)"},
      {"compact proposals", "compact sums", "compact verify"}};
  ProfileScope ps(ctx, names[hrec ? 1 : 0][phase]);
  if (hrec) {
    switch (R) {
      case 2: launch_compact_phase<2, true>(ctx, phase, J, A, E, Dp); break;
      case 4: launch_compact_phase<4, true>(ctx, phase, J, A, E, Dp); break;
      case 8: launch_compact_phase<8, true>(ctx, phase, J, A, E, Dp); break;
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  } else {
    switch (R) {
      case 2: launch_compact_phase<2, false>(ctx, phase, J, A, E, Dp); break;
      case 4: launch_compact_phase<4, false>(ctx, phase, J, A, E, Dp); break;
      case 8: launch_compact_phase<8, false>(ctx, phase, J, A, E, Dp); break;
      case 16: launch_compact_phase<16, false>(ctx, phase, J, A, E, Dp); break;
      case 32: launch_compact_phase<32, false>(ctx, phase, J, A, E, Dp); break;
      case 64: launch_compact_phase<64, false>(ctx, phase, J, A, E, Dp); break;
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  }
  return launch_check("compact sweep kernel");
}

int launch_esgm_verify(sgm_context *ctx, int R, const SummaryJobs &J, int slots) {
  ProfileScope ps(ctx, "esgm verify");
  const dim3 grid((unsigned)((slots + 3) / 4), 2);
  switch (R) {
    case 2: esgm_verify_kernel<2><<<grid, 256, 0, ctx->stream>>>(J); break;
    case 4: esgm_verify_kernel<4><<<grid, 256, 0, ctx->stream>>>(J); break;
    case 8: esgm_verify_kernel<8><<<grid, 256, 0, ctx->stream>>>(J); break;
    case 16: esgm_verify_kernel<16><<<grid, 256, 0, ctx->stream>>>(J); break;
    case 32: esgm_verify_kernel<32><<<grid, 256, 0, ctx->stream>>>(J); break;
    case 64: esgm_verify_kernel<64><<<grid, 256, 0, ctx->stream>>>(J); break;
    default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
  }
  return launch_check("compact verification kernel");
}

// Compact plan for `nd` matching directions (GpuCandidateStereo.execute/run,
// GpuCandidateStereo.java:49-171, 190-335): proposal sweeps and sum sweeps
// over every path and strip, certification, exact verification of the rest in
// groups that fit the budget, then selection of every pixel from its summary.
// Same contract as run_directions.
}  // namespace

namespace sgm {

int line_length(int dx, int dy, int line, int w, int h) {
  if (dy == 0) return w;
  if (dx == 0) return h;
  // the diagonal line enumeration of line_start
  const int x = line < w ? line : (dx > 0 ? 0 : w - 1);
  const int y = line < w ? (dy > 0 ? 0 : h - 1) : (dy > 0 ? line - w + 1 : h - 2 - (line - w));
  return std::min(dx > 0 ? w - x : x + 1, dy > 0 ? h - y : y + 1);
}

void longest_first(const std::vector<int> &len, int maxlen, std::vector<int> &order) {
  // counting sort by length, descending; lines of equal length in line order
  std::vector<int> start((size_t)maxlen + 2, 0);
  for (int n : len) start[maxlen - n + 1]++;  // bucket maxlen - n: longest first
  for (int i = 1; i <= maxlen + 1; i++) start[i] += start[i - 1];
  order.assign(len.size(), 0);
  for (size_t g = 0; g < len.size(); g++) order[start[maxlen - len[g]]++] = (int)g;
}

}  // namespace sgm

namespace {

// Longest-first dispatch order of the lines of a launch over a w x h pair (nd
// matching directions; lines [begin[k], begin[k] + nd * split[k]) sweep
// direction (dx[k], dy[k]), matching direction m's from begin[k] + m *
// split[k]), in a device buffer cached by shape.  A launch with several times
// more lines than resident waves otherwise ends on long lines dispatched last
// (the seven-direction launch at 1600x1067 D=256: 2.48 -> 2.30 ms).  Null, for
// the natural order, when the buffer does not fit the budget.
static const int *line_order(sgm_context *ctx, const char *name, int ndirs, const int *dx,
                             const int *dy, const int *split, const int *begin, int total, int w,
                             int h, int nd) {
  sgm_context::LineOrder &lo = ctx->line_orders[name];
  const int key[6] = {w, h, nd, ndirs, dx[0], dy[0]};
  Buffer &buf = ctx->buffers[name];
  if (buf.ptr && memcmp(lo.key, key, sizeof key) == 0) return (const int *)buf.ptr;
  const size_t bytes = (size_t)total * sizeof(int);
  if (bytes > buf.cap && ctx->allocated - (int64_t)buf.cap + (int64_t)bytes > ctx->budget)
    return nullptr;
  if (buf.ensure(ctx, bytes, "line order")) return nullptr;
  std::vector<int> len((size_t)total);
  for (int k = 0; k < ndirs; k++)
    for (int m = 0; m < nd; m++)
      for (int l = 0; l < split[k]; l++)
        len[(size_t)begin[k] + (size_t)m * split[k] + l] = line_length(dx[k], dy[k], l, w, h);
  longest_first(len, std::max(w, h), lo.host);
  hipError_t e = hipMemcpyAsync(buf.ptr, lo.host.data(), bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  memcpy(lo.key, key, sizeof key);
  return (const int *)buf.ptr;
}

int run_directions_esgm(sgm_context *ctx, const sgm_params *p, int w, int h, PairBuffers &B,
                        const int *slots, int nd, int out_slot, uint8_t *reasons,
                        uint8_t *boundary) {
  int rc;
  const int D = p->disparities, Dp = padded_disparities(D), R = path_lanes(p);
  const int rows = B.esgm_rows, pool = B.esgm_pool, C = p->directions * pool;
  const int strips = (h + rows - 1) / rows;
  const bool computed = esgm_computed(p);
  const size_t n = (size_t)w * h, cells = computed ? 0 : (size_t)w * rows * Dp;
  const size_t carry_row = (size_t)w * Dp;
  CostJobs cj;
  memset(&cj, 0, sizeof cj);
  CompactJobs kj;
  memset(&kj, 0, sizeof kj);
  SummaryJobs sj;
  memset(&sj, 0, sizeof sj);
  for (int i = 0; i < nd; i++) {
    const int d = slots[i], o = 1 - d;
    CostJob &Cj = cj.job[i];
    Cj.desc_a = B.desc[d];
    Cj.desc_b = B.desc[o];
    Cj.mask_a = B.mask[d];
    Cj.mask_b = B.mask[o];
    Cj.gray_a = B.gray[d];
    Cj.gray_b = B.gray[o];
    Cj.cost = B.e_strip_cost + i * cells;
    Cj.w = w;
    Cj.h = h;
    Cj.D = D;
    Cj.Dp = Dp;
    Cj.minD = p->min_disparity;
    Cj.reverse = d;
    CompactJob &K = kj.job[i];
    K.cost = Cj.cost;
    K.mask = B.mask[d];
    K.gray = B.gray[d];
    K.cand = B.e_cand + i * n * C;
    K.sums = B.e_sums + i * n * 3 * C;
    K.outside = B.e_outside + i * n;
    K.slot = B.e_slot + i * n;
    K.w = w;
    K.h = h;
    K.D = D;
    K.Dp = Dp;
    K.desc_a = B.desc[d];
    K.desc_b = B.desc[o];
    K.minD = p->min_disparity;
    K.reverse = d;
    K.rx = p->window_w / 2;
    K.ry = p->window_h / 2;
    SummaryJob &S = sj.job[i];
    S.cand = K.cand;
    S.sums = K.sums;
    S.outside = K.outside;
    S.mask_a = B.mask[d];
    S.mask_b = B.mask[o];
    S.summary = B.e_summary + i * 3 * n;
    S.slot = B.e_slot + i * n;
    S.pixels = B.e_pixels + i * n;
    S.count = B.e_count + i;
    S.w = w;
    S.h = h;
    S.D = D;
    S.Dp = Dp;
    S.minD = p->min_disparity;
    S.rx = p->window_w / 2;
    S.ry = p->window_h / 2;
    S.reverse = d;
  }
  if (nd == 1) {
    // the verification grid always spans two jobs: the second one is empty
    sj.job[1] = sj.job[0];
    sj.job[1].hi = 0;
  }
  // With 1-word windows and R <= 8 every path runs over the whole image in one
  // launch, computing its census costs on the fly.  Otherwise strip passes:
  // pass 0 runs the paths with dy >= 0 top-down, pass 1 the paths with dy < 0
  // bottom-up; up to three carried paths per pass, each with a ping-pong pair
  // of carry rows.
  static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  auto interrupted = [&]() {
    if (!ctx->interrupt.exchange(0)) return false;
    (void)hipStreamSynchronize(ctx->stream);
    return true;
  };
  auto sweep = [&](int phase) -> int {
    if (computed) {
      for (int dir = 0; dir < p->directions; dir++) {
        if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
        PathArgs A = path_args(p, dir, nd, 0, 0);
        const CompactArgs E{C, pool, dir};
        kj.split = lines_of(A.cls, w, h);
        A.total_lines = nd * kj.split;
        for (int i = 0; i < nd; i++) {
          CompactJob &K = kj.job[i];
          K.y0 = 0;
          K.rows = h;
          K.carry_in = K.carry_out = nullptr;
        }
        int rc2;
        if ((rc2 = launch_compact(ctx, R, phase, true, kj, A, E, Dp))) return rc2;
      }
      return SGM_OK;
    }
    for (int pass = 0; pass < 2; pass++) {
      for (int ord = 0; ord < strips; ord++) {
        if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
        const int s = pass ? strips - 1 - ord : ord;
        const int y0 = s * rows, act = std::min(rows, h - y0);
        bool costed = false;
        int carried = 0;
        for (int dir = 0; dir < p->directions; dir++) {
          if ((dirs[dir][1] < 0) != (pass == 1)) continue;
          int rc2;
          if (!costed) {
            for (int i = 0; i < nd; i++) cj.job[i].y0 = y0;
            if ((rc2 = launch_cost(ctx, cj, nd, w, h, Dp, p, act, true))) return rc2;
            costed = true;
          }
          PathArgs A = path_args(p, dir, nd, 0, 0);
          const CompactArgs E{C, pool, dir};
          kj.split = segments_of(A.cls, w, act);
          A.total_lines = nd * kj.split;
          for (int i = 0; i < nd; i++) {
            CompactJob &K = kj.job[i];
            K.y0 = y0;
            K.rows = act;
            K.carry_in = K.carry_out = nullptr;
            if (A.dy != 0) {
              int *base = B.e_carry + (size_t)i * 6 * carry_row;
              K.carry_in = ord > 0 ? base + ((size_t)(ord & 1) * 3 + carried) * carry_row : nullptr;
              K.carry_out = ord < strips - 1
                                ? base + ((size_t)((ord + 1) & 1) * 3 + carried) * carry_row
                                : nullptr;
            }
          }
          if (A.dy != 0) carried++;
          if ((rc2 = launch_compact(ctx, R, phase, false, kj, A, E, Dp))) return rc2;
        }
      }
    }
    return SGM_OK;
  };
  // every proposal entry is written by the proposal sweeps; sums and bounds accumulate
  if ((rc = launch_fill_u8(ctx, (uint8_t *)B.e_sums, 0, (size_t)nd * n * 3 * C * 2))) return rc;
  if ((rc = launch_fill_u8(ctx, (uint8_t *)B.e_outside, 0, (size_t)nd * n * 2))) return rc;
  if ((rc = launch_fill_u8(ctx, (uint8_t *)B.e_count, 0, 8))) return rc;
  if ((rc = sweep(ESGM_CANDIDATES))) return rc;
  if ((rc = sweep(ESGM_SUMS))) return rc;
  {
    ProfileScope ps(ctx, "esgm summary");
    esgm_summary_kernel<<<dim3((unsigned)((n + 255) / 256), nd), 256, 0, ctx->stream>>>(
        sj, C, p->uniqueness);
    if ((rc = launch_check("compact summary kernel"))) return rc;
  }
  int counts[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(counts, B.e_count, nd * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "compact verification count");
  for (int k = 0; k < 4; k++) ctx->esgm_last[k] = 0;
  ctx->esgm_groups = 0;
  int most = 0;
  for (int i = 0; i < nd; i++) {
    ctx->esgm_last[2 * slots[i]] = (int64_t)n - counts[i];
    ctx->esgm_last[2 * slots[i] + 1] = counts[i];
    most = std::max(most, counts[i]);
  }
  if (most > 0) {
    // verification groups: the budget left by the compact plan, at least one
    // pixel.  Each group sweeps every path over the whole image again, so the
    // verification time grows with the group count (reported in the profile
    // as `verify_groups`); the reference verifies in one streamed pass.
    const int64_t room = ctx->budget - esgm_pair_bytes(p, w, h, rows, pool) + 2 * Dp * 2;
    int64_t group = std::max<int64_t>(1, room / ((int64_t)nd * Dp * 2));
    group = std::min<int64_t>(group, most);
    ctx->esgm_groups = (int)((most + group - 1) / group);
    const size_t vbytes = (size_t)nd * group * Dp * 2;
    if ((rc = ctx->buffers["esgm_verify"].ensure(ctx, vbytes, "esgm_verify"))) return rc;
    uint16_t *verify = (uint16_t *)ctx->buffers["esgm_verify"].ptr;
    for (int i = 0; i < nd; i++) {
      kj.job[i].verify = verify + (size_t)i * group * Dp;
      sj.job[i].verify = kj.job[i].verify;
    }
    for (int64_t lo = 0; lo < most; lo += group) {
      const int hi = (int)std::min<int64_t>(most, lo + group);
      if ((rc = launch_fill_u8(ctx, (uint8_t *)verify, 0, vbytes))) return rc;
      for (int i = 0; i < nd; i++) {
        kj.job[i].lo = (int)lo;
        kj.job[i].hi = std::min(hi, counts[i]);
        sj.job[i].lo = (int)lo;
        sj.job[i].hi = std::min(hi, counts[i]);
      }
      if (nd == 1) sj.job[1].hi = 0;
      if ((rc = sweep(ESGM_VERIFY))) return rc;
      if ((rc = launch_esgm_verify(ctx, R, sj, hi - (int)lo))) return rc;
    }
  }
  SelectJobs xj;
  memset(&xj, 0, sizeof xj);
  for (int i = 0; i < nd; i++) {
    const int d = slots[i];
    SelectJob &X = xj.job[i];
    X.summary = B.e_summary + i * 3 * n;
    X.out.disp = B.disp[d];
    X.out.reasons = d == out_slot ? reasons : nullptr;
    X.out.boundary = d == out_slot ? boundary : nullptr;
    X.out.code = d == out_slot ? B.code : nullptr;
    X.out.w = w;
    X.out.h = h;
    X.out.rx = p->window_w / 2;
    X.out.ry = p->window_h / 2;
    X.out.D = D;
    X.out.minD = p->min_disparity;
  }
  ProfileScope ps(ctx, "esgm select");
  const PathArgs A = path_args(p, 0, nd, 0, 0);
  const unsigned chunks = (unsigned)((w + 63) / 64);
  esgm_select_kernel<<<dim3((chunks + 3) / 4, h, nd), 256, 0, ctx->stream>>>(xj, A);
  return launch_check("compact selection kernel");
}

// Cost + path sweeps + fused selection for `nd` matching directions
// (d = 0: left reference, target x - d; d = 1: right reference, target x + d).
// The selection codes of slot `out_slot` go to B.code (counted later), its
// reason / boundary updates to `reasons` / `boundary`.
int run_directions(sgm_context *ctx, const sgm_params *p, int w, int h, PairBuffers &B,
                   const int *slots, int nd, int out_slot, uint8_t *reasons, uint8_t *boundary) {
  if (B.esgm_pool)
    return run_directions_esgm(ctx, p, w, h, B, slots, nd, out_slot, reasons, boundary);
  if (B.strip_rows)
    return run_directions_streamed(ctx, p, w, h, B, slots, nd, out_slot, reasons, boundary);
  int rc;
  const int D = p->disparities, Dp = padded_disparities(D), R = path_lanes(p);
  const int sb = aggregate_bytes(p);
  const bool recompute = B.recompute;
  const bool fused = !recompute && fused_cost(p);
  int mode = recompute ? 3 : sb == 2 ? (cost_u8(p) ? 4 : 1) : 0;
  // stream the uint8 cost and uint16 aggregate volumes past the caches when
  // they far exceed the 256 MB MALL (measured: +5% at 1600x1067 D=256); rows
  // are whole 64-byte granules (padded disparities), which streaming needs
  const int vol_bytes = mode == 4 ? 3 : 4;  // cost + aggregate bytes per cell
  const bool nt = (mode == 4 || mode == 1) && R <= 8 &&
                  (int64_t)nd * w * h * Dp * vol_bytes > (int64_t)768 << 20;
  if (nt) mode |= MODE_NT;
  if (!fused && !recompute) {
    CostJobs cj;
    memset(&cj, 0, sizeof cj);
    for (int i = 0; i < nd; i++) {
      const int d = slots[i], o = 1 - d;
      CostJob &J = cj.job[i];
      J.desc_a = B.desc[d];
      J.desc_b = B.desc[o];
      J.mask_a = B.mask[d];
      J.mask_b = B.mask[o];
      J.gray_a = B.gray[d];
      J.gray_b = B.gray[o];
      J.sum_a = B.wsum[d];
      J.sum2_a = B.wsum2[d];
      J.sum_b = B.wsum[o];
      J.sum2_b = B.wsum2[o];
      J.cost = B.cost[d];
      J.w = w;
      J.h = h;
      J.D = D;
      J.Dp = Dp;
      J.minD = p->min_disparity;
      J.reverse = d;
    }
    if ((rc = launch_cost(ctx, cj, nd, w, h, Dp, p))) return rc;
  }
  PathJobs pj;
  memset(&pj, 0, sizeof pj);
  pj.table = nullptr;
  for (int i = 0; i < nd; i++) {
    const int d = slots[i], o = 1 - d;
    PathJob &J = pj.local[i];
    J.cost = B.cost[d];
    J.sum = B.sum[d];
    J.mask = B.mask[d];
    J.gray = B.gray[d];
    J.gm = B.gm[d];
    J.desc_a = B.desc[d];
    J.desc_b = B.desc[o];
    J.mask_b = B.mask[o];
    J.cost_out = B.cost[d];
    J.disp = B.disp[d];
    J.reasons = d == out_slot ? reasons : nullptr;
    J.boundary = d == out_slot ? boundary : nullptr;
    J.code = d == out_slot ? B.code : nullptr;
    J.w = w;
    J.h = h;
    J.D = D;
    J.Dp = Dp;
    J.minD = p->min_disparity;
    J.rx = p->window_w / 2;
    J.ry = p->window_h / 2;
    J.reverse = d;
    J.R = R;
  }
  const int waves = path_waves(ctx, R);
  auto dir_args = [&](int dir) {
    PathArgs A = path_args(p, dir, nd, 0, ctx->keep_volumes ? 1 : 0);
    A.total_lines = nd * lines_of(A.cls, w, h);
    A.split = lines_of(A.cls, w, h);
    return A;
  };
  auto interrupted = [&]() {
    if (!ctx->interrupt.exchange(0)) return false;
    (void)hipStreamSynchronize(ctx->stream);
    return true;
  };
  if (B.planes[slots[0]] && (mode & 7) == 4) {
    // Path-cost planes: the census cost pass (fused census cost of the
    // left-to-right lines; other costs come from the cost kernel above), every
    // direction but the last in one launch, each into its plane 0..6, and the
    // last sweep, which adds the seven planes to its own path costs before the
    // winner search.
    ctx->last_sweeps = SGM_SWEEPS_PLANES;
    if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
    if (fused) {
      PathArgs Ac = dir_args(0);
      Ac.split = h * cost_segments(w);
      Ac.total_lines = nd * Ac.split;
      if ((rc = launch_path(ctx, R, sb, 2 | (mode & MODE_NT), pj, Ac, PHASE_COST, waves))) return rc;
    }
    MultiDir M;
    memset(&M, 0, sizeof M);
    M.ndirs = p->directions - 1;
    for (int k = 0; k < M.ndirs; k++) {
      const PathArgs Ak = dir_args(k);
      M.dx[k] = Ak.dx;
      M.dy[k] = Ak.dy;
      M.cls[k] = Ak.cls;
      M.split[k] = Ak.split;
      M.begin[k] = M.total_lines;
      M.total_lines += Ak.total_lines;
      for (int i = 0; i < nd; i++) M.plane[k][i] = B.planes[slots[i]] + (size_t)k * Dp;
    }
    M.order = line_order(ctx, "multi_order", M.ndirs, M.dx, M.dy, M.split, M.begin, M.total_lines,
                         w, h, nd);
    if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
    if ((rc = launch_path_multi(ctx, R, mode | MODE_PLANE, pj, dir_args(0), M, waves))) return rc;
    PathJobs jl = pj;
    for (int i = 0; i < nd; i++) jl.local[i].sum = B.planes[slots[i]];
    if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
    PathArgs AL = dir_args(p->directions - 1);
    const int zero = 0;
    AL.order = line_order(ctx, "last_order", 1, &AL.dx, &AL.dy, &AL.split, &zero, AL.total_lines, w,
                          h, nd);
    return launch_path(ctx, R, sb, mode | MODE_PLANE, jl, AL, PHASE_LAST, waves);
  }
  // Paired sweeps: after the first direction, the middle directions run two
  // per launch, the first of each pair adding into the aggregate volume and
  // the second into a second partial volume; the last sweep adds both.
  const bool paired = B.sum_b[slots[0]] != nullptr && (mode & 7) != 3;
  ctx->last_sweeps = paired ? SGM_SWEEPS_PAIRED : SGM_SWEEPS_SEQUENTIAL;
  PathJobs pjb = pj;
  if (paired)
    for (int i = 0; i < nd; i++) pjb.local[i].sum = B.sum_b[slots[i]];
  const int last = p->directions - 1;
  for (int dir = 0; dir <= last;) {
    if (interrupted()) return fail(SGM_ERR_INTERRUPTED, "GPU SGM interrupted");
    if (paired && dir > 0 && dir + 1 < last) {
      // dir into the aggregate, dir + 1 into the second partial aggregate
      // (its first direction when dir == 1)
      static const char *pair_names[8] = {"", "pair_order1", "", "pair_order3", "", "pair_order5"};
      PathArgs Aa = dir_args(dir), Ab = dir_args(dir + 1);
      const int dxs[2] = {Aa.dx, Ab.dx}, dys[2] = {Aa.dy, Ab.dy}, splits[2] = {Aa.split, Ab.split};
      const int begins[2] = {0, Aa.total_lines};
      Aa.order = line_order(ctx, pair_names[dir], 2, dxs, dys, splits, begins,
                            Aa.total_lines + Ab.total_lines, w, h, nd);
      if ((rc = launch_path_pair(ctx, R, mode, dir == 1 ? PHASE_FIRST : PHASE_MID, pj, Aa, pjb, Ab,
                                 waves)))
        return rc;
      dir += 2;
      continue;
    }
    const int phase = dir == 0 ? PHASE_FIRST : dir == last ? PHASE_LAST : PHASE_MID;
    int m = (dir == 0 && fused) ? (2 | (mode & MODE_NT)) : mode;
    PathJobs jobs = pj;
    if (paired && dir == last) {
      for (int i = 0; i < nd; i++) jobs.local[i].sum2 = B.sum_b[slots[i]];
      m |= MODE_SUM2;
    }
    PathArgs Ad = dir_args(dir);
    if (Ad.cls == 2) {  // diagonal lines differ in length: longest first
      static const char *dir_names[8] = {"", "", "", "", "dir_order4", "dir_order5", "dir_order6",
                                         "dir_order7"};
      const int zero = 0;
      Ad.order = line_order(ctx, dir_names[dir], 1, &Ad.dx, &Ad.dy, &Ad.split, &zero,
                            Ad.total_lines, w, h, nd);
    }
    if ((rc = launch_path(ctx, R, sb, m, jobs, Ad, phase, waves))) return rc;
    dir++;
  }
  return SGM_OK;
}

int run_pair(sgm_context *ctx, const sgm_params *p, int w, int h, PairBuffers &B,
             const uint8_t *left, const uint8_t *right, const uint8_t *vl, const uint8_t *vr,
             float *checked, uint8_t *reasons_out, uint8_t *boundary_out,
             unsigned long long *stats) {
  int rc;
  const size_t n = (size_t)w * h;
  if ((rc = launch_fill_u64(ctx, stats, 0ULL, 16))) return rc;
  if ((rc = launch_fill_u8(ctx, B.reasons, 1, n))) return rc;
  if ((rc = launch_fill_u8(ctx, B.boundary, 0, n))) return rc;
  CensusJobs cj;
  memset(&cj, 0, sizeof cj);
  const uint8_t *src[2] = {left, right}, *val[2] = {vl, vr};
  for (int d = 0; d < 2; d++) {
    CensusJob &J = cj.job[d];
    J.gray = src[d];
    J.valid = val[d];
    J.mask_in = nullptr;
    J.gray_copy = src[d] == B.gray[d] ? nullptr : B.gray[d];
    J.desc = B.desc[d];
    J.mask = B.mask[d];
    J.gm = B.gm[d];
    J.wsum = B.wsum[d];
    J.wsum2 = B.wsum2[d];
    J.w = w;
    J.h = h;
  }
  if ((rc = launch_census(ctx, cj, 2, w, h, p))) return rc;
  if (B.bgr[0]) {
    ColorJobs xj;
    memset(&xj, 0, sizeof xj);
    for (int d = 0; d < 2; d++) {
      ColorJob &J = xj.job[d];
      J.bgr = B.bgr[d];
      J.gray = B.gray[d];
      J.valid = val[d];
      J.desc = B.desc[d];
      J.mask = B.mask[d];
      J.gm = B.gm[d];
      J.w = w;
      J.h = h;
    }
    dim3 grid((w + kCensusTx - 1) / kCensusTx, (h + kCensusTy - 1) / kCensusTy, 2);
    dim3 block(kCensusTx, kCensusTy);
    ProfileScope ps(ctx, "census");
    if (words_of(p) == 1)
      color_census_kernel<1><<<grid, block, 0, ctx->stream>>>(xj, p->window_w, p->window_h);
    else
      color_census_kernel<2><<<grid, block, 0, ctx->stream>>>(xj, p->window_w, p->window_h);
    if ((rc = launch_check("colour census kernel"))) return rc;
  }
  static const int both[2] = {0, 1};
  if ((rc = run_directions(ctx, p, w, h, B, both, 2, 0, B.reasons, B.boundary))) return rc;
  if (B.refine && B.refine->subpixel_refinement) {
    // PhotometricRefinement of both directions on the gray rasters (CensusSgm.java:168-170)
    const size_t n = (size_t)w * h;
    if ((rc = ctx->buffers["refine0"].ensure(ctx, n * 4, "refine0"))) return rc;
    if ((rc = ctx->buffers["refine1"].ensure(ctx, n * 4, "refine1"))) return rc;
    for (int d = 0; d < 2; d++) {
      float *orig = (float *)ctx->buffers[d ? "refine1" : "refine0"].ptr;
      hipError_t e = hipMemcpyAsync(orig, B.disp[d], n * 4, hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) return hip_fail(e, "refinement copy");
      RefineArgs G;
      G.w = w;
      G.h = h;
      G.r = B.refine->correlation_window / 2;
      G.min_d = p->min_disparity;
      G.count = p->disparities;
      G.reverse = d;
      G.correlation = B.refine->correlation;
      ProfileScope ps(ctx, "refine");
      refine_kernel<<<dim3((unsigned)((w + 255) / 256), (unsigned)h), 256, 0, ctx->stream>>>(
          orig, B.disp[d], B.gray[d], B.gray[1 - d], G);
      if ((rc = launch_check("refinement kernel"))) return rc;
    }
  }
  LrArgs L;
  L.code = B.code;
  L.dl = B.disp[0];
  L.dr = B.disp[1];
  L.reasons_in = B.reasons;
  L.boundary_in = B.boundary;
  L.checked = checked;
  L.reasons_out = reasons_out;
  L.boundary_out = boundary_out;
  L.stats = stats;
  L.w = w;
  L.h = h;
  L.rx = p->window_w / 2;
  L.ry = p->window_h / 2;
  L.threshold = p->lr_threshold;
  return launch_lr(ctx, L);
}

void stats_from_device(const unsigned long long *c, sgm_stats *s) {
  memset(s, 0, sizeof *s);
  s->lower_boundary_rejected = (int64_t)c[0];
  s->full_interval_pixels = (int64_t)c[1];
  s->boundary_rejected = (int64_t)c[2];
  s->upper_boundary_rejected = (int64_t)c[3];
  s->ambiguity_rejected = (int64_t)c[4];
  s->selected = (int64_t)c[5];
  s->lr_rejected = (int64_t)c[6];
  s->valid = (int64_t)c[10];
}

int copy_h2d(sgm_context *ctx, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  return e == hipSuccess ? SGM_OK : hip_fail(e, "host to device copy");
}

int copy_d2h(sgm_context *ctx, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  return e == hipSuccess ? SGM_OK : hip_fail(e, "device to host copy");
}

int finish(sgm_context *ctx) {
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  collect_profile(ctx);
  return SGM_OK;
}

// GpuStereo.plan (GpuStereo.java:33-84) under the context's plan mode: *rows =
// 0 for the resident plan, else the strip height of the streamed plan.
int pick_plan(sgm_context *ctx, const sgm_params *p, int w, int h, int *rows,
              int *pool = nullptr) {
  *rows = 0;
  if (pool) *pool = 0;
  if (ctx->plan_mode == SGM_PLAN_ESGM) {
    // GpuStereo.execute (GpuStereo.java:185-197): an explicit esgm request
    // that no compact plan serves is an error
    const char *why = pool ? esgm_unsupported(p, w, h) : "the compact plan takes gray Census only";
    if (!why) {
      esgm_plan(ctx, p, w, h, rows, pool);
      if (*pool) return SGM_OK;
      why = "its buffers exceed the budget";
    }
    return fail(SGM_ERR_NO_PLAN,
                "No compact Census/16-bit GPU plan fits (%s); use --gpu-sgm streamed for exact "
                "recomputation", why);
  }
  const bool fits = pair_bytes(p, w, h) <= ctx->budget;
  if (ctx->plan_mode != SGM_PLAN_STREAMED && fits) return SGM_OK;
  if (ctx->plan_mode != SGM_PLAN_RESIDENT) {
    ctx->splan = streamed_plan(ctx, p, w, h, ctx->aggregate_storage);
    *rows = ctx->splan.rows;
  }
  if (*rows && ctx->plan_mode == SGM_PLAN_AUTO && ctx->aggregate_storage == SGM_AGGREGATE_AUTO &&
      ctx->splan.storage == SGM_AGGREGATE_RECOMPUTE && pool && !esgm_unsupported(p, w, h)) {
    // GpuStereo.execute (GpuStereo.java:176-184): auto with only a bounded
    // recomputation left runs the compact plan when it serves the pair
    int erows = 0, epool = 0;
    esgm_plan(ctx, p, w, h, &erows, &epool);
    if (epool) {
      *rows = erows;
      *pool = epool;
      return SGM_OK;
    }
  }
  if (*rows) return SGM_OK;
  return fail(SGM_ERR_NO_PLAN, "No GPU SGM memory plan fits the requested range (%lld > %lld bytes)",
              (long long)pair_bytes(p, w, h), (long long)ctx->budget);
}

}  // namespace

extern "C" int sgm_match_pair(sgm_context *ctx, const sgm_params *params, const uint8_t *left,
                              const uint8_t *right, const uint8_t *valid_left,
                              const uint8_t *valid_right, int w, int h, float *checked, float *dl,
                              float *dr, sgm_stats *stats, uint8_t *reasons, uint8_t *boundary) {
  if (!ctx || !left || !right || !checked) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((valid_left == nullptr) != (valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  int rows, pool = 0;
  if ((rc = pick_plan(ctx, params, w, h, &rows, &pool))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, params, w, h, B, rows, pool))) return rc;
  ctx->last_streamed = rows > 0 && !pool;
  ctx->last_compact = pool > 0;
  const size_t n = (size_t)w * h;
  if ((rc = copy_h2d(ctx, B.gray[0], left, n))) return rc;
  if ((rc = copy_h2d(ctx, B.gray[1], right, n))) return rc;
  const uint8_t *vl = nullptr, *vr = nullptr;
  if (valid_left) {
    if ((rc = copy_h2d(ctx, B.valid[0], valid_left, n))) return rc;
    if ((rc = copy_h2d(ctx, B.valid[1], valid_right, n))) return rc;
    vl = B.valid[0];
    vr = B.valid[1];
  }
  if ((rc = run_pair(ctx, params, w, h, B, B.gray[0], B.gray[1], vl, vr, B.checked, B.reasons_out,
                     B.boundary_out, B.stats)))
    return rc;
  unsigned long long counters[16];
  if ((rc = copy_d2h(ctx, checked, B.checked, n * 4))) return rc;
  if (dl && (rc = copy_d2h(ctx, dl, B.disp[0], n * 4))) return rc;
  if (dr && (rc = copy_d2h(ctx, dr, B.disp[1], n * 4))) return rc;
  if (reasons && (rc = copy_d2h(ctx, reasons, B.reasons_out, n))) return rc;
  if (boundary && (rc = copy_d2h(ctx, boundary, B.boundary_out, n))) return rc;
  if ((rc = copy_d2h(ctx, counters, B.stats, sizeof counters))) return rc;
  if ((rc = finish(ctx))) return rc;
  if (stats) stats_from_device(counters, stats);
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = *params;
  return SGM_OK;
}

namespace {

int validate_filters(const sgm_filter_params *f) {
  if (!f) return fail(SGM_ERR_INVALID_ARGUMENT, "null filter params");
  if (f->median > kMedianMaxWindow)
    return fail(SGM_ERR_INVALID_ARGUMENT, "median window %d above %d", f->median, kMedianMaxWindow);
  if (f->correlation_window < 1 || f->correlation_window > 31)
    return fail(SGM_ERR_INVALID_ARGUMENT, "correlation window must be 1..31");
  if (f->speckle_connectivity != 4 && f->speckle_connectivity != 8)
    return fail(SGM_ERR_INVALID_ARGUMENT, "speckle connectivity must be 4 or 8");
  if (!(f->median_max_change >= 0) || !(f->speckle_range >= 0))
    return fail(SGM_ERR_INVALID_ARGUMENT, "median max change and speckle range must be >= 0");
  if (f->subpixel_refinement != 0 && f->subpixel_refinement != 1)
    return fail(SGM_ERR_INVALID_ARGUMENT, "--subpixel-refinement must be none (0) or zncc (1)");
  if (f->subpixel_refinement && f->correlation_window < 3)
    return fail(SGM_ERR_INVALID_ARGUMENT, "subpixel refinement needs a correlation window >= 3");
  return SGM_OK;
}

// 03_median .. 05_final_stereo on the device (CensusSgm.java:203-271).  `work`
// holds the L-R-validated map on entry and the final map on return.
int run_filters(sgm_context *ctx, const sgm_params *p, const sgm_filter_params *f, int w, int h,
                PairBuffers &B, float *&work) {
  int rc;
  const size_t n = (size_t)w * h;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  FilterArgs F;
  F.w = w;
  F.h = h;
  F.median = f->median;
  F.median_classic = f->median_classic;
  F.median_edge = f->median_edge_threshold;
  F.median_max_change = f->median_max_change;
  F.bx = p->window_w / 2;
  F.by = p->window_h / 2;
  F.lr_threshold = p->lr_threshold;
  F.corr_r = f->correlation_window / 2;
  F.channels = B.bgr[0] ? 3 : 1;  // CensusSgm.java:223-225
  F.correlation = f->correlation;
  F.speckle_size = f->speckle_size;
  F.speckle_conn = f->speckle_connectivity;
  F.speckle_range = f->speckle_range;
#define ENSURE(name, bytes)                                                   \
  if ((rc = ctx->buffers[name].ensure(ctx, (bytes), name)) != SGM_OK) return rc;
  ENSURE("fmed", n * 4);
  ENSURE("cc_parent", n * 4);
  ENSURE("cc_size", n * 4);
  ENSURE("cc_error", 64);
#undef ENSURE
  float *fmed = (float *)ctx->buffers["fmed"].ptr;
  int *parent = (int *)ctx->buffers["cc_parent"].ptr, *size = (int *)ctx->buffers["cc_size"].ptr;
  int *error = (int *)ctx->buffers["cc_error"].ptr;
  ProfileScope ps(ctx, "filters");
  // the final valid count replaces the L-R checkpoint's
  if ((rc = launch_fill_u64(ctx, B.stats + 10, 0ULL, 1))) return rc;
  if ((rc = launch_fill_u8(ctx, (uint8_t *)error, 0, 64))) return rc;
  work = B.checked;
  if (f->median > 1) {
    median_kernel<<<blocks, 256, 0, ctx->stream>>>(B.checked, fmed, B.gray[0], F);
    if ((rc = launch_check("median kernel"))) return rc;
    work = fmed;
  }
  photo_kernel<<<blocks, 256, 0, ctx->stream>>>(work, B.disp[1], B.bgr[0] ? B.bgr[0] : B.gray[0],
                                                B.bgr[1] ? B.bgr[1] : B.gray[1],
                                                B.reasons_out, B.stats, F);
  if ((rc = launch_check("photometric kernel"))) return rc;
  if (f->speckle_size > 1) {
    const dim3 tiles((w + kCcTx - 1) / kCcTx, (h + kCcTy - 1) / kCcTy), tile(kCcTx, kCcTy);
    cc_local_kernel<<<tiles, tile, 0, ctx->stream>>>(work, parent, size, error, F);
    cc_border_kernel<<<tiles, tile, 0, ctx->stream>>>(work, parent, error, F);
    cc_flatten_kernel<<<blocks, 256, 0, ctx->stream>>>(parent, error, n);
    cc_size_kernel<<<blocks, 256, 0, ctx->stream>>>(parent, size, n);
    cc_remove_kernel<<<blocks, 256, 0, ctx->stream>>>(work, parent, size, B.reasons_out, B.stats, F);
    if ((rc = launch_check("speckle kernels"))) return rc;
  } else {
    count_valid_kernel<<<blocks, 256, 0, ctx->stream>>>(work, B.stats, n);
    if ((rc = launch_check("count kernel"))) return rc;
  }
  return SGM_OK;
}

}  // namespace

extern "C" int sgm_match_pair_filtered(sgm_context *ctx, const sgm_params *params,
                                       const sgm_filter_params *filters, const uint8_t *left,
                                       const uint8_t *right, const uint8_t *valid_left,
                                       const uint8_t *valid_right, int w, int h, float *disparity,
                                       sgm_stats *stats, uint8_t *reasons) {
  if (!ctx || !left || !right || !disparity) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((valid_left == nullptr) != (valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  if ((rc = validate_filters(filters))) return rc;
  int rows, pool = 0;
  if ((rc = pick_plan(ctx, params, w, h, &rows, &pool))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, params, w, h, B, rows, pool))) return rc;
  ctx->last_streamed = rows > 0 && !pool;
  ctx->last_compact = pool > 0;
  const size_t n = (size_t)w * h;
  if ((rc = copy_h2d(ctx, B.gray[0], left, n))) return rc;
  if ((rc = copy_h2d(ctx, B.gray[1], right, n))) return rc;
  const uint8_t *vl = nullptr, *vr = nullptr;
  if (valid_left) {
    if ((rc = copy_h2d(ctx, B.valid[0], valid_left, n))) return rc;
    if ((rc = copy_h2d(ctx, B.valid[1], valid_right, n))) return rc;
    vl = B.valid[0];
    vr = B.valid[1];
  }
  B.refine = filters;
  if ((rc = run_pair(ctx, params, w, h, B, B.gray[0], B.gray[1], vl, vr, B.checked, B.reasons_out,
                     B.boundary_out, B.stats)))
    return rc;
  float *work = nullptr;
  if ((rc = run_filters(ctx, params, filters, w, h, B, work))) return rc;
  unsigned long long counters[16];
  int error = 0;
  if ((rc = copy_d2h(ctx, disparity, work, n * 4))) return rc;
  if (reasons && (rc = copy_d2h(ctx, reasons, B.reasons_out, n))) return rc;
  if ((rc = copy_d2h(ctx, counters, B.stats, sizeof counters))) return rc;
  if ((rc = copy_d2h(ctx, &error, ctx->buffers["cc_error"].ptr, sizeof error))) return rc;
  if ((rc = finish(ctx))) return rc;
  if (error) return fail(SGM_ERR_HIP, "speckle component search did not converge");
  if (stats) {
    stats_from_device(counters, stats);
    stats->post_median_lr_rejected = (int64_t)counters[7];
    stats->photometric_rejected = (int64_t)counters[8];
    stats->speckle_rejected = (int64_t)counters[9];
  }
  return SGM_OK;
}

extern "C" int sgm_match_pair_color(sgm_context *ctx, const sgm_params *params,
                                    const sgm_filter_params *filters, const uint8_t *left,
                                    const uint8_t *right, const uint8_t *left_bgr,
                                    const uint8_t *right_bgr, const uint8_t *valid_left,
                                    const uint8_t *valid_right, int w, int h, float *disparity,
                                    float *dl, float *dr, sgm_stats *stats, uint8_t *reasons,
                                    uint8_t *boundary) {
  if (!ctx || !left || !right || !left_bgr || !right_bgr || !disparity)
    return fail(SGM_ERR_INVALID_ARGUMENT, "Color-distance census requires interleaved BGR inputs");
  if ((valid_left == nullptr) != (valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  if (filters && (rc = validate_filters(filters))) return rc;
  int rows;
  if ((rc = pick_plan(ctx, params, w, h, &rows))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, params, w, h, B, rows))) return rc;
  ctx->last_streamed = rows > 0;
  ctx->last_compact = false;
  const size_t n = (size_t)w * h;
  if ((rc = ctx->buffers["bgr0"].ensure(ctx, 3 * n, "bgr0"))) return rc;
  if ((rc = ctx->buffers["bgr1"].ensure(ctx, 3 * n, "bgr1"))) return rc;
  uint8_t *bgr0 = (uint8_t *)ctx->buffers["bgr0"].ptr, *bgr1 = (uint8_t *)ctx->buffers["bgr1"].ptr;
  B.bgr[0] = bgr0;
  B.bgr[1] = bgr1;
  if ((rc = copy_h2d(ctx, B.gray[0], left, n))) return rc;
  if ((rc = copy_h2d(ctx, B.gray[1], right, n))) return rc;
  if ((rc = copy_h2d(ctx, bgr0, left_bgr, 3 * n))) return rc;
  if ((rc = copy_h2d(ctx, bgr1, right_bgr, 3 * n))) return rc;
  const uint8_t *vl = nullptr, *vr = nullptr;
  if (valid_left) {
    if ((rc = copy_h2d(ctx, B.valid[0], valid_left, n))) return rc;
    if ((rc = copy_h2d(ctx, B.valid[1], valid_right, n))) return rc;
    vl = B.valid[0];
    vr = B.valid[1];
  }
  B.refine = filters;
  if ((rc = run_pair(ctx, params, w, h, B, B.gray[0], B.gray[1], vl, vr, B.checked, B.reasons_out,
                     B.boundary_out, B.stats)))
    return rc;
  float *work = B.checked;
  if (filters && (rc = run_filters(ctx, params, filters, w, h, B, work))) return rc;
  unsigned long long counters[16];
  int error = 0;
  if ((rc = copy_d2h(ctx, disparity, work, n * 4))) return rc;
  if (dl && (rc = copy_d2h(ctx, dl, B.disp[0], n * 4))) return rc;
  if (dr && (rc = copy_d2h(ctx, dr, B.disp[1], n * 4))) return rc;
  if (reasons && (rc = copy_d2h(ctx, reasons, B.reasons_out, n))) return rc;
  if (boundary && (rc = copy_d2h(ctx, boundary, B.boundary_out, n))) return rc;
  if ((rc = copy_d2h(ctx, counters, B.stats, sizeof counters))) return rc;
  if (filters && (rc = copy_d2h(ctx, &error, ctx->buffers["cc_error"].ptr, sizeof error))) return rc;
  if ((rc = finish(ctx))) return rc;
  if (error) return fail(SGM_ERR_HIP, "speckle component search did not converge");
  if (stats) {
    stats_from_device(counters, stats);
    if (filters) {
      stats->post_median_lr_rejected = (int64_t)counters[7];
      stats->photometric_rejected = (int64_t)counters[8];
      stats->speckle_rejected = (int64_t)counters[9];
    }
  }
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = *params;
  return SGM_OK;
}

extern "C" int sgm_match_pair_device(sgm_context *ctx, const sgm_params *params, const void *d_left,
                                     const void *d_right, const void *d_valid_left,
                                     const void *d_valid_right, int w, int h, void *d_checked,
                                     void *d_dl, void *d_dr, void *d_stats, void *d_reasons,
                                     void *d_boundary) {
  if (!ctx || !d_left || !d_right || !d_checked) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((d_valid_left == nullptr) != (d_valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  int rows, pool = 0;
  if ((rc = pick_plan(ctx, params, w, h, &rows, &pool))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, params, w, h, B, rows, pool))) return rc;
  ctx->last_streamed = rows > 0 && !pool;
  ctx->last_compact = pool > 0;
  const size_t n = (size_t)w * h;
  unsigned long long *stats = d_stats ? (unsigned long long *)d_stats : B.stats;
  if ((rc = run_pair(ctx, params, w, h, B, (const uint8_t *)d_left, (const uint8_t *)d_right,
                     (const uint8_t *)d_valid_left, (const uint8_t *)d_valid_right,
                     (float *)d_checked, d_reasons ? (uint8_t *)d_reasons : B.reasons_out,
                     d_boundary ? (uint8_t *)d_boundary : B.boundary_out, stats)))
    return rc;
  if (d_dl) {
    hipError_t e = hipMemcpyAsync(d_dl, B.disp[0], n * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "device copy");
  }
  if (d_dr) {
    hipError_t e = hipMemcpyAsync(d_dr, B.disp[1], n * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "device copy");
  }
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = *params;
  return SGM_OK;
}

namespace {

// The resident plan's pair buffers laid out back to back (256-byte aligned)
// from `base` (nullptr: only count).  Returns the bytes the layout takes.  The
// caller's valid masks are read in place, so no valid copies are laid out.
size_t carve_pair_buffers(const sgm_params *p, int w, int h, bool cost_volume, uint8_t *base,
                          PairBuffers &B) {
  const size_t n = (size_t)w * h, Dp = padded_disparities(p->disparities);
  const int sb = aggregate_bytes(p), words = words_of(p);
  const bool zncc = p->cost == SGM_COST_ZNCC || p->cost == SGM_COST_ZNCC_FP32;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    void *q = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  memset(&B, 0, sizeof B);
  for (int d = 0; d < 2; d++) {
    B.gray[d] = (uint8_t *)take(n);
    B.valid[d] = nullptr;
    B.mask[d] = (uint8_t *)take(n);
    B.gm[d] = (uint16_t *)take(2 * n);
    B.desc[d] = (uint64_t *)take(8 * words * n);
    B.wsum[d] = zncc ? (int32_t *)take(4 * n) : nullptr;
    B.wsum2[d] = zncc ? (int32_t *)take(4 * n) : nullptr;
    B.cost[d] = cost_volume ? (uint16_t *)take(n * Dp * (cost_u8(p) ? 1 : 2)) : nullptr;
    B.sum[d] = take(n * Dp * sb);
    B.disp[d] = (float *)take(4 * n);
  }
  B.checked = (float *)take(4 * n);
  B.reasons = (uint8_t *)take(n);
  B.boundary = (uint8_t *)take(n);
  B.reasons_out = (uint8_t *)take(n);
  B.boundary_out = (uint8_t *)take(n);
  B.stats = (unsigned long long *)take(16 * 8);
  B.code = (uint8_t *)take(n);
  B.recompute = !cost_volume;
  return off;
}

}  // namespace

extern "C" int sgm_match_pair_device_noalloc(sgm_context *ctx, const sgm_params *params,
                                             const void *d_left, const void *d_right,
                                             const void *d_valid_left, const void *d_valid_right,
                                             int w, int h, void *d_checked, void *d_dl, void *d_dr,
                                             void *d_stats, void *d_reasons, void *d_boundary) {
  if (!ctx || !d_left || !d_right || !d_checked) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((d_valid_left == nullptr) != (d_valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  if (ctx->keep_volumes)
    return fail(SGM_ERR_INVALID_ARGUMENT,
                "sgm_match_pair_device_noalloc keeps no volumes (disable sgm_set_keep_volumes)");
  uint8_t *base = (uint8_t *)scratch_address(ctx->device);
  if (!base)
    return fail(SGM_ERR_HIP, "the no-allocation workspace of device %d is unavailable (HIP %d)",
                ctx->device, (int)g_scratch_error[ctx->device < kMaxScratchDevices ? ctx->device : 0]);
  PairBuffers B;
  // the resident cost volume when it fits, else census costs recomputed per sweep
  bool cost_volume = !(census_recompute_possible(params) && ctx->volume_policy == SGM_VOLUME_RECOMPUTE);
  size_t need = carve_pair_buffers(params, w, h, cost_volume, nullptr, B);
  if (need > kScratchBytes && cost_volume && census_recompute_possible(params)) {
    cost_volume = false;
    need = carve_pair_buffers(params, w, h, cost_volume, nullptr, B);
  }
  if (need > kScratchBytes)
    return fail(SGM_ERR_INVALID_ARGUMENT,
                "a %dx%d pair with %d disparities needs %zu bytes of workspace; "
                "sgm_match_pair_device_noalloc holds at most %zu (use sgm_match_pair_device)",
                w, h, params->disparities, need, kScratchBytes);
  carve_pair_buffers(params, w, h, cost_volume, base, B);
  (void)hipSetDevice(ctx->device);
  ctx->last_streamed = false;
  ctx->last_compact = false;
  const size_t n = (size_t)w * h;
  unsigned long long *stats = d_stats ? (unsigned long long *)d_stats : B.stats;
  if ((rc = run_pair(ctx, params, w, h, B, (const uint8_t *)d_left, (const uint8_t *)d_right,
                     (const uint8_t *)d_valid_left, (const uint8_t *)d_valid_right,
                     (float *)d_checked, d_reasons ? (uint8_t *)d_reasons : B.reasons_out,
                     d_boundary ? (uint8_t *)d_boundary : B.boundary_out, stats)))
    return rc;
  if (d_dl) {
    hipError_t e = hipMemcpyAsync(d_dl, B.disp[0], n * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "device copy");
  }
  if (d_dr) {
    hipError_t e = hipMemcpyAsync(d_dr, B.disp[1], n * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "device copy");
  }
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = *params;
  return SGM_OK;
}

namespace {

// GpuSgm.match (GpuSgm.java:267-303) for one matching direction with the
// caller's texture&valid masks, census descriptors computed from the gray
// rasters or (desc_a / desc_b non-null) the caller's word-major planes.
// Internally the reference image is slot 0 and the target slot 1 with the
// direction's `reverse` flag.
int match_one(sgm_context *ctx, const sgm_params *params, const uint64_t *desc_a,
              const uint64_t *desc_b, const uint8_t *gray_a, const uint8_t *gray_b,
              const uint8_t *mask_a, const uint8_t *mask_b, int w, int h, int reverse, float *disp,
              sgm_stats *stats, uint8_t *reasons, uint8_t *boundary) {
  if (!ctx || !gray_a || !gray_b || !mask_a || !mask_b || !disp)
    return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(params, w, h);
  if (rc) return rc;
  int rows, pool = 0;
  if ((rc = pick_plan(ctx, params, w, h, &rows, desc_a ? nullptr : &pool))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, params, w, h, B, rows, pool))) return rc;
  ctx->last_streamed = rows > 0 && !pool;
  ctx->last_compact = pool > 0;
  const size_t n = (size_t)w * h;
  // slot d = reverse: the reference image sits in the slot whose cost job has J.reverse = reverse.
  const int ref = reverse ? 1 : 0, tgt = 1 - ref;
  if ((rc = copy_h2d(ctx, B.gray[ref], gray_a, n))) return rc;
  if ((rc = copy_h2d(ctx, B.gray[tgt], gray_b, n))) return rc;
  if ((rc = copy_h2d(ctx, B.valid[ref], mask_a, n))) return rc;
  if ((rc = copy_h2d(ctx, B.valid[tgt], mask_b, n))) return rc;
  if ((rc = launch_fill_u64(ctx, B.stats, 0ULL, 16))) return rc;
  if (reasons) {
    if ((rc = copy_h2d(ctx, B.reasons, reasons, n))) return rc;
  }
  if (boundary) {
    if ((rc = copy_h2d(ctx, B.boundary, boundary, n))) return rc;
  }
  CensusJobs cj;
  memset(&cj, 0, sizeof cj);
  for (int d = 0; d < 2; d++) {
    CensusJob &J = cj.job[d];
    J.gray = B.gray[d];
    J.mask_in = B.valid[d];
    J.desc = B.desc[d];
    J.mask = B.mask[d];
    J.gm = B.gm[d];
    J.wsum = B.wsum[d];
    J.wsum2 = B.wsum2[d];
    J.w = w;
    J.h = h;
  }
  if ((rc = launch_census(ctx, cj, 2, w, h, params))) return rc;
  if (desc_a) {
    const int words = words_of(params);
    if ((rc = copy_h2d(ctx, B.desc[ref], desc_a, n * 8 * words))) return rc;
    if ((rc = copy_h2d(ctx, B.desc[tgt], desc_b, n * 8 * words))) return rc;
    if (words == 1)
      for (int d = 0; d < 2; d++) {
        desc_flag_kernel<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(B.desc[d], B.mask[d], n);
        if ((rc = launch_check("descriptor flag kernel"))) return rc;
      }
  }
  rc = run_directions(ctx, params, w, h, B, &ref, 1, ref, reasons ? B.reasons : nullptr,
                      boundary ? B.boundary : nullptr);
  if (rc == SGM_OK) {
    LrArgs L;
    memset(&L, 0, sizeof L);
    L.code = B.code;
    L.stats = B.stats;
    L.w = w;
    L.h = h;
    rc = launch_lr(ctx, L);
  }
  if (rc) return rc;
  unsigned long long counters[16];
  if ((rc = copy_d2h(ctx, disp, B.disp[ref], n * 4))) return rc;
  if (reasons && (rc = copy_d2h(ctx, reasons, B.reasons, n))) return rc;
  if (boundary && (rc = copy_d2h(ctx, boundary, B.boundary, n))) return rc;
  if ((rc = copy_d2h(ctx, counters, B.stats, sizeof counters))) return rc;
  if ((rc = finish(ctx))) return rc;
  if (stats) {
    // GpuSgm.java:289-294: the completed counters are added to the caller's.
    sgm_stats s;
    stats_from_device(counters, &s);
    stats->lower_boundary_rejected += s.lower_boundary_rejected;
    stats->full_interval_pixels += s.full_interval_pixels;
    stats->boundary_rejected += s.boundary_rejected;
    stats->upper_boundary_rejected += s.upper_boundary_rejected;
    stats->ambiguity_rejected += s.ambiguity_rejected;
    stats->selected += s.selected;
  }
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = *params;
  return SGM_OK;
}

}  // namespace

extern "C" int sgm_match(sgm_context *ctx, const sgm_params *params, const uint8_t *gray_a,
                         const uint8_t *gray_b, const uint8_t *mask_a, const uint8_t *mask_b, int w,
                         int h, int reverse, float *disp, sgm_stats *stats, uint8_t *reasons,
                         uint8_t *boundary) {
  return match_one(ctx, params, nullptr, nullptr, gray_a, gray_b, mask_a, mask_b, w, h, reverse,
                   disp, stats, reasons, boundary);
}

extern "C" int sgm_match_desc(sgm_context *ctx, const sgm_params *params, const uint64_t *desc_a,
                              const uint64_t *desc_b, const uint8_t *gray_a, const uint8_t *gray_b,
                              const uint8_t *mask_a, const uint8_t *mask_b, int w, int h,
                              int reverse, float *disp, sgm_stats *stats, uint8_t *reasons,
                              uint8_t *boundary) {
  if (!desc_a || !desc_b) return fail(SGM_ERR_INVALID_ARGUMENT, "null descriptors");
  return match_one(ctx, params, desc_a, desc_b, gray_a, gray_b, mask_a, mask_b, w, h, reverse,
                   disp, stats, reasons, boundary);
}

extern "C" int sgm_read_volumes(sgm_context *ctx, int direction, uint16_t *cost, uint32_t *aggregate) {
  if (!ctx || direction < 0 || direction > 1) return fail(SGM_ERR_INVALID_ARGUMENT, "bad argument");
  std::lock_guard<std::mutex> g(ctx->lock);
  if (!ctx->keep_volumes || ctx->last_w == 0)
    return fail(SGM_ERR_NOT_FOUND, "no retained volumes (enable sgm_set_keep_volumes first)");
  if (ctx->last_streamed)
    return fail(SGM_ERR_NOT_FOUND, "no retained volumes: the last pair ran strip-streamed");
  const sgm_params *p = &ctx->last_params;
  const size_t n = (size_t)ctx->last_w * ctx->last_h, D = p->disparities, Dp = padded_disparities(p->disparities);
  const int sb = aggregate_bytes(p);
  std::string s = std::to_string(direction);
  const void *dc = ctx->buffers["cost" + s].ptr, *ds = ctx->buffers["sum" + s].ptr;
  if (!dc || !ds) return fail(SGM_ERR_NOT_FOUND, "no volumes");
  // uint8 volumes hold 255 for an invalid cell (512)
  const bool u8 = cost_u8(p);
  std::vector<uint8_t> hc(n * Dp * (u8 ? 1 : 2));
  std::vector<uint8_t> hs(n * Dp * sb);
  int rc;
  if ((rc = copy_d2h(ctx, hc.data(), dc, hc.size()))) return rc;
  if ((rc = copy_d2h(ctx, hs.data(), ds, hs.size()))) return rc;
  if ((rc = finish(ctx))) return rc;
  for (size_t i = 0; i < n; i++)
    for (size_t k = 0; k < D; k++) {
      if (cost) {
        if (u8) cost[i * D + k] = hc[i * Dp + k] == 255 ? 512 : hc[i * Dp + k];
        else cost[i * D + k] = ((const uint16_t *)hc.data())[i * Dp + k];
      }
      if (aggregate) {
        if (sb == 2) aggregate[i * D + k] = ((const uint16_t *)hs.data())[i * Dp + k];
        else aggregate[i * D + k] = (uint32_t)((const int32_t *)hs.data())[i * Dp + k];
      }
    }
  return SGM_OK;
}

extern "C" int sgm_device_alloc(sgm_context *ctx, size_t bytes, void **out) {
  if (!ctx || !out) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  return context_alloc(ctx, out, bytes, "caller buffer");
}

extern "C" int sgm_device_free(sgm_context *ctx, void *ptr) {
  if (!ctx) return fail(SGM_ERR_INVALID_ARGUMENT, "null context");
  if (!ptr) return SGM_OK;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? SGM_OK : hip_fail(e, "hipFree");
}

extern "C" int sgm_device_copy(sgm_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || !dst || !src) return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
    return fail(SGM_ERR_INVALID_ARGUMENT, "device copy needs 16-byte aligned pointers and size");
  if (bytes == 0) return SGM_OK;
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  ProfileScope ps(ctx, "copy");
  const size_t n = bytes / 16;
  const size_t want = (n + 255) / 256;
  const unsigned blocks = (unsigned)std::min<size_t>(want, (size_t)ctx->info.compute_units * 16);
  copy_kernel<<<blocks, 256, 0, ctx->stream>>>((const uint4 *)src, (uint4 *)dst, n);
  return launch_check("copy kernel");
}

extern "C" int sgm_memcpy(sgm_context *ctx, void *dst, const void *src, size_t bytes, int kind) {
  if (!ctx || (!dst && bytes) || (!src && bytes) || kind < 0 || kind > 2)
    return fail(SGM_ERR_INVALID_ARGUMENT, "bad copy arguments");
  std::lock_guard<std::mutex> g(ctx->lock);
  (void)hipSetDevice(ctx->device);
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, k, ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "sgm_memcpy");
  e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? SGM_OK : hip_fail(e, "sgm_memcpy synchronize");
}

// ---------------------------------------------------------------------------
// Asynchronous plain pair for the serving lanes: sgm_match_pair enqueued on the
// context's stream through pinned staging buffers (so neither the uploads nor
// the downloads block the calling thread), completed by lane_collect once the
// stream has passed the work.
namespace sgm {

int lane_enqueue(sgm_context *ctx, const sgm_params *base, const sgm_pair &pr) {
  if (!ctx || !pr.left || !pr.right || !pr.checked)
    return fail(SGM_ERR_INVALID_ARGUMENT, "null argument");
  if ((pr.valid_left == nullptr) != (pr.valid_right == nullptr))
    return fail(SGM_ERR_INVALID_ARGUMENT, "Invalid stereo mask dimensions");
  sgm_params p = *base;
  p.disparities = pr.disparities;
  p.min_disparity = pr.min_disparity;
  const int w = pr.width, h = pr.height;
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc = validate_params(&p, w, h);
  if (rc) return rc;
  int rows;
  if ((rc = pick_plan(ctx, &p, w, h, &rows))) return rc;
  (void)hipSetDevice(ctx->device);
  PairBuffers B;
  if ((rc = ensure_pair_buffers(ctx, &p, w, h, B, rows))) return rc;
  ctx->last_streamed = rows > 0;
  ctx->last_compact = false;
  const size_t n = (size_t)w * h;
  if ((rc = ctx->host_buffers["lane_in"].ensure(4 * n, "lane inputs"))) return rc;
  if ((rc = ctx->host_buffers["lane_out"].ensure(14 * n + 128, "lane outputs"))) return rc;
  uint8_t *in = (uint8_t *)ctx->host_buffers["lane_in"].ptr;
  uint8_t *out = (uint8_t *)ctx->host_buffers["lane_out"].ptr;
  memcpy(in, pr.left, n);
  memcpy(in + n, pr.right, n);
  if ((rc = copy_h2d(ctx, B.gray[0], in, n))) return rc;
  if ((rc = copy_h2d(ctx, B.gray[1], in + n, n))) return rc;
  const uint8_t *vl = nullptr, *vr = nullptr;
  if (pr.valid_left) {
    memcpy(in + 2 * n, pr.valid_left, n);
    memcpy(in + 3 * n, pr.valid_right, n);
    if ((rc = copy_h2d(ctx, B.valid[0], in + 2 * n, n))) return rc;
    if ((rc = copy_h2d(ctx, B.valid[1], in + 3 * n, n))) return rc;
    vl = B.valid[0];
    vr = B.valid[1];
  }
  if ((rc = run_pair(ctx, &p, w, h, B, B.gray[0], B.gray[1], vl, vr, B.checked, B.reasons_out,
                     B.boundary_out, B.stats)))
    return rc;
  // staging layout: checked 4n | dl 4n | dr 4n | reasons n | boundary n | counters 128
  if ((rc = copy_d2h(ctx, out, B.checked, n * 4))) return rc;
  if (pr.dl && (rc = copy_d2h(ctx, out + 4 * n, B.disp[0], n * 4))) return rc;
  if (pr.dr && (rc = copy_d2h(ctx, out + 8 * n, B.disp[1], n * 4))) return rc;
  if (pr.reasons && (rc = copy_d2h(ctx, out + 12 * n, B.reasons_out, n))) return rc;
  if (pr.boundary && (rc = copy_d2h(ctx, out + 13 * n, B.boundary_out, n))) return rc;
  if ((rc = copy_d2h(ctx, out + 14 * n, B.stats, 128))) return rc;
  ctx->last_w = w;
  ctx->last_h = h;
  ctx->last_params = p;
  return SGM_OK;
}

int lane_collect(sgm_context *ctx, const sgm_pair &pr, sgm_stats *stats) {
  std::lock_guard<std::mutex> g(ctx->lock);
  int rc;
  if ((rc = finish(ctx))) return rc;  // the stream has already passed the work
  const size_t n = (size_t)pr.width * pr.height;
  const uint8_t *out = (const uint8_t *)ctx->host_buffers["lane_out"].ptr;
  std::vector<HostCopy> copies = {{pr.checked, out, n * 4}};
  if (pr.dl) copies.push_back({pr.dl, out + 4 * n, n * 4});
  if (pr.dr) copies.push_back({pr.dr, out + 8 * n, n * 4});
  if (pr.reasons) copies.push_back({pr.reasons, out + 12 * n, n});
  if (pr.boundary) copies.push_back({pr.boundary, out + 13 * n, n});
  parallel_copy(copies);
  unsigned long long counters[16];
  memcpy(counters, out + 14 * n, sizeof counters);
  if (stats) stats_from_device(counters, stats);
  return SGM_OK;
}

}  // namespace sgm
