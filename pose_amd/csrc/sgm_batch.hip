// Batched and served modes (work in progress).
#include <hip/hip_runtime.h>

#include "sgm_internal.h"

using namespace sgm;

extern "C" int sgm_batch_create(sgm_context *, const sgm_params *, const sgm_pair *, int, int,
                                sgm_batch **out) {
  if (out) *out = nullptr;
  return fail(SGM_ERR_INVALID_ARGUMENT, "batched mode not built yet");
}
extern "C" int sgm_batch_run(sgm_batch *) { return fail(SGM_ERR_INVALID_ARGUMENT, "batched mode not built yet"); }
extern "C" int sgm_batch_destroy(sgm_batch *) { return SGM_OK; }
extern "C" int64_t sgm_batch_device_bytes(sgm_batch *) { return 0; }
extern "C" int sgm_autotune(sgm_context *, const sgm_params *, const char *) {
  return fail(SGM_ERR_INVALID_ARGUMENT, "autotuner not built yet");
}
extern "C" int sgm_autotune_load(sgm_context *, const char *) {
  return fail(SGM_ERR_INVALID_ARGUMENT, "autotuner not built yet");
}
extern "C" int sgm_server_create(sgm_context *, const sgm_params *, int64_t, int, int, sgm_server **out) {
  if (out) *out = nullptr;
  return fail(SGM_ERR_INVALID_ARGUMENT, "serving mode not built yet");
}
extern "C" int sgm_server_submit(sgm_server *, const sgm_pair *, uint64_t *) {
  return fail(SGM_ERR_INVALID_ARGUMENT, "serving mode not built yet");
}
extern "C" int sgm_server_wait(sgm_server *, uint64_t, double, sgm_request_info *) {
  return fail(SGM_ERR_INVALID_ARGUMENT, "serving mode not built yet");
}
extern "C" double sgm_server_now(sgm_server *) { return 0; }
extern "C" int sgm_server_destroy(sgm_server *) { return SGM_OK; }
