// vvclf_batch.cpp — batches of jobs filtered by one persistent fused launch.
#include <hip/hip_runtime.h>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

extern "C" {

int vvclf_batch_create(vvclf_ctx*, vvclf_job* const*, uint32_t, const vvclf_tune*, vvclf_batch** out) {
  if (out) *out = nullptr;
  return fail(VVCLF_ERR_UNSUPPORTED, "batch: not built yet");
}
int vvclf_batch_destroy(vvclf_batch*) { return VVCLF_OK; }
int vvclf_batch_run(vvclf_batch*, void*) { return fail(VVCLF_ERR_UNSUPPORTED, "batch: not built yet"); }
int vvclf_batch_graph_launch(vvclf_batch*, void*) { return fail(VVCLF_ERR_UNSUPPORTED, "batch: not built yet"); }
int vvclf_batch_get_tune(const vvclf_batch*, vvclf_tune*) { return fail(VVCLF_ERR_UNSUPPORTED, "batch: not built yet"); }
int vvclf_autotune(vvclf_ctx*, vvclf_job* const*, uint32_t, vvclf_tune*) {
  return fail(VVCLF_ERR_UNSUPPORTED, "autotune: not built yet");
}

}  // extern "C"
