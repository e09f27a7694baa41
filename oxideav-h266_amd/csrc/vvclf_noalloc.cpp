// vvclf_noalloc.cpp — the picture-level driver without device allocation:
// vvclf_apply_in_loop_filters_noalloc (include/vvclf.h).
//
// Every device buffer of the call (the three sample sets of the plain form and
// the per-picture side tables) is carved from `scratch`, a file-scope device
// array of VVCLF_NOALLOC_SCRATCH_BYTES that each GPU holds for the life of the
// process.  Samples move through pageable copies (no pinned staging buffer is
// allocated either).  Host code never names the array except to ask the
// runtime for its address on each device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

static __device__ __align__(256) uint8_t scratch[VVCLF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

struct device_scratch {
  std::mutex mu;        // one call at a time per device: every call shares the array
  uint8_t* ptr = nullptr;  // the array's address on this device
};

device_scratch g_scratch[kMaxDevices];

vvclf_ctx* noalloc_ctx(int device) {
  static std::mutex mu;
  static vvclf_ctx* ctx[kMaxDevices] = {};
  std::lock_guard<std::mutex> g(mu);
  if (!ctx[device] && vvclf_ctx_create(device, &ctx[device]) != VVCLF_OK) ctx[device] = nullptr;
  return ctx[device];
}

// The scratch array's device address on `device` (current at the call),
// looked up once per device.  Caller holds g_scratch[device].mu.
int scratch_ptr(int device, uint8_t** out) {
  device_scratch& ds = g_scratch[device];
  if (!ds.ptr) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "noalloc: hipGetSymbolAddress(scratch)");
    ds.ptr = (uint8_t*)p;
  }
  *out = ds.ptr;
  return VVCLF_OK;
}

}  // namespace

extern "C" int vvclf_apply_in_loop_filters_noalloc(vvclf_frame* f) {
  clear_error();
  if (!f) return fail(VVCLF_ERR_INVALID, "apply_in_loop_filters_noalloc: NULL frame");
  // the sample sets alone must fit; checked before any device work
  const uint64_t W = f->pic.luma.width, H = f->pic.luma.height;
  const uint64_t Wc = f->chroma_format_idc ? f->pic.cb.width : 0, Hc = f->chroma_format_idc ? f->pic.cb.height : 0;
  const uint64_t set = (((W + 63) / 64 * 64 * H + 2 * ((Wc + 63) / 64 * 64) * Hc) * 2 + 255) / 256 * 256;
  if (3 * set > VVCLF_NOALLOC_SCRATCH_BYTES)
    return fail(VVCLF_ERR_UNSUPPORTED,
                "apply_in_loop_filters_noalloc: a %llux%llu picture needs %llu bytes of sample buffers, more than "
                "the %llu-byte scratch",
                (unsigned long long)W, (unsigned long long)H, (unsigned long long)(3 * set),
                (unsigned long long)VVCLF_NOALLOC_SCRATCH_BYTES);
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return hip_fail(e, "noalloc: hipGetDevice");
  if (device < 0 || device >= kMaxDevices) return fail(VVCLF_ERR_UNSUPPORTED, "noalloc: device %d", device);
  vvclf_ctx* ctx = noalloc_ctx(device);
  if (!ctx) return VVCLF_ERR_HIP;
  std::lock_guard<std::mutex> g(g_scratch[device].mu);
  uint8_t* base = nullptr;
  int rc = scratch_ptr(device, &base);
  if (rc) return rc;
  job_opts o;
  o.stream = ctx->stream;
  o.host_threads = host_walk_threads();
  o.arena = base;
  o.arena_bytes = VVCLF_NOALLOC_SCRATCH_BYTES;
  vvclf_job* j = nullptr;
  // derives the tables, places planes and tables in the scratch, uploads the
  // tables and the input samples (every byte a kernel reads is written here,
  // so nothing left by an earlier call is read)
  if ((rc = job_create_stages(ctx, f, STAGE_ALL, &j, nullptr, &o))) return rc;
  if (j->d.do_lmcs || j->d.do_dbk || j->d.do_sao || j->d.do_alf) {
    rc = vvclf_job_run(j, ctx->stream);
    if (!rc) {
      e = hipStreamSynchronize(ctx->stream);
      rc = e == hipSuccess ? vvclf_job_download(j, &f->pic) : hip_fail(e, "noalloc: run");
    }
  }
  vvclf_job_destroy(j);  // host record only: the scratch is not freed
  return rc;
}
