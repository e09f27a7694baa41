// vvclf_batch.cpp — batches: many jobs filtered by one persistent fused
// launch over a tile list, optionally captured once in a HIP graph and
// replayed; plus the tile-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

namespace vvclf {
size_t fused_lds_bytes(int tw, int th);
}  // namespace vvclf

struct vvclf_batch {
  vvclf_ctx* ctx = nullptr;
  std::vector<vvclf_job*> jobs;
  vvclf_tune tune{};
  dev_job* d_jobs = nullptr;
  tile_ref* d_tiles = nullptr;
  uint32_t n_tiles = 0, grid = 0;
  bool hbd = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

bool shape_supported(uint32_t tw, uint32_t th) {
  for (auto& s : kFusedShapes)
    if ((uint32_t)s[0] == tw && (uint32_t)s[1] == th) return true;
  return false;
}

bool shape_fits(const std::vector<vvclf_job*>& jobs, uint32_t tw, uint32_t th) {
  for (auto* j : jobs)
    if (!job_fits_shape(j->d, tw, th)) return false;
  return true;
}

int cu_count(int device) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return 256;
  return p.multiProcessorCount;
}

}  // namespace

extern "C" {

int vvclf_batch_create(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n_jobs, const vvclf_tune* tune,
                       vvclf_batch** out) {
  clear_error();
  if (!ctx || !jobs || !n_jobs || !out) return fail(VVCLF_ERR_INVALID, "batch_create: bad argument");
  *out = nullptr;
  std::vector<vvclf_job*> js(jobs, jobs + n_jobs);
  for (auto* j : js) {
    if (!j || j->ctx->device != ctx->device) return fail(VVCLF_ERR_INVALID, "batch_create: job of another device");
  }
  vvclf_tune t{};
  if (tune && tune->tile_w) {
    t = *tune;
  } else {
    int rc = vvclf_autotune(ctx, jobs, n_jobs, &t);
    if (rc) return rc;
  }
  if (!shape_supported(t.tile_w, t.tile_h)) return fail(VVCLF_ERR_INVALID, "batch: tile %ux%u", t.tile_w, t.tile_h);
  if (!shape_fits(js, t.tile_w, t.tile_h))
    return fail(VVCLF_ERR_INVALID, "batch: tile %ux%u does not fit the jobs' ALF / SAO CTBs", t.tile_w, t.tile_h);
  hipSetDevice(ctx->device);
  bool hbd = false;
  for (auto* j : js) hbd = hbd || job_needs_hbd(j->d);
  int occ = fused_max_blocks_per_cu((int)t.tile_w, (int)t.tile_h, hbd);
  if (occ <= 0) return fail(VVCLF_ERR_HIP, "batch: fused kernel does not fit on a CU");
  if (t.wgs_per_cu == 0 || (int)t.wgs_per_cu > occ) t.wgs_per_cu = (uint32_t)occ;
  t.threads = (uint32_t)fused_threads((int)t.tile_w, (int)t.tile_h);
  // tile list: largest pictures first, raster order within a picture
  std::vector<uint32_t> order(n_jobs);
  for (uint32_t i = 0; i < n_jobs; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)js[a]->w * js[a]->h > (uint64_t)js[b]->w * js[b]->h;
  });
  std::vector<tile_ref> tiles;
  for (uint32_t oi : order) {
    const vvclf_job* j = js[oi];
    for (uint32_t y = 0; y < j->h; y += t.tile_h)
      for (uint32_t x = 0; x < j->w; x += t.tile_w) tiles.push_back(make_tile(j, oi, x, y));
  }
  std::vector<dev_job> dj(n_jobs);
  for (uint32_t i = 0; i < n_jobs; i++) dj[i] = js[i]->d;
  vvclf_batch* b = new vvclf_batch();
  b->ctx = ctx;
  b->jobs = js;
  b->tune = t;
  b->hbd = hbd;
  b->n_tiles = (uint32_t)tiles.size();
  uint32_t grid = (uint32_t)cu_count(ctx->device) * t.wgs_per_cu;
  grid = std::min(grid, (b->n_tiles + 7) & ~7u);
  b->grid = std::max(8u, grid & ~7u);
  hipError_t e = hipMalloc(&b->d_jobs, dj.size() * sizeof(dev_job));
  if (e == hipSuccess) e = hipMalloc(&b->d_tiles, tiles.size() * sizeof(tile_ref));
  if (e == hipSuccess) e = hipMemcpy(b->d_jobs, dj.data(), dj.size() * sizeof(dev_job), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(tile_ref), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    vvclf_batch_destroy(b);
    return hip_fail(e, "batch upload");
  }
  *out = b;
  return VVCLF_OK;
}

int vvclf_batch_destroy(vvclf_batch* b) {
  if (!b) return VVCLF_OK;
  hipSetDevice(b->ctx->device);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->d_jobs) hipFree(b->d_jobs);
  if (b->d_tiles) hipFree(b->d_tiles);
  delete b;
  return VVCLF_OK;
}

int vvclf_batch_run(vvclf_batch* b, void* stream) {
  if (!b) return fail(VVCLF_ERR_INVALID, "batch_run: NULL");
  hipSetDevice(b->ctx->device);
  hipError_t e = launch_fused((int)b->tune.tile_w, (int)b->tune.tile_h, b->hbd, b->d_jobs, b->d_tiles, b->n_tiles,
                              b->grid, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "fused launch");
  for (auto* j : b->jobs) j->fused_out = true;
  return VVCLF_OK;
}

int vvclf_batch_graph_launch(vvclf_batch* b, void* stream) {
  if (!b) return fail(VVCLF_ERR_INVALID, "batch_graph_launch: NULL");
  hipSetDevice(b->ctx->device);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (!b->exec) {
    // capture on a private stream so the caller's stream may be the legacy one
    hipStream_t cs;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "graph stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = launch_fused((int)b->tune.tile_w, (int)b->tune.tile_h, b->hbd, b->d_jobs, b->d_tiles,
                                   b->n_tiles, b->grid, cs);
      e = hipStreamEndCapture(cs, &b->graph);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    hipStreamDestroy(cs);
    if (e != hipSuccess) return hip_fail(e, "graph capture");
  }
  e = hipGraphLaunch(b->exec, s);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  for (auto* j : b->jobs) j->fused_out = true;
  return VVCLF_OK;
}

int vvclf_batch_get_tune(const vvclf_batch* b, vvclf_tune* out) {
  if (!b || !out) return fail(VVCLF_ERR_INVALID, "batch_get_tune: NULL");
  *out = b->tune;
  return VVCLF_OK;
}

int vvclf_batch_time(vvclf_batch* b, void* stream, uint32_t warmup, uint32_t reps, int use_graph, double* ms) {
  if (!b || !ms || reps == 0) return fail(VVCLF_ERR_INVALID, "batch_time: bad argument");
  hipSetDevice(b->ctx->device);
  auto once = [&]() { return use_graph ? vvclf_batch_graph_launch(b, stream) : vvclf_batch_run(b, stream); };
  for (uint32_t i = 0; i < warmup; i++)
    if (int rc = once()) return rc;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return fail(VVCLF_ERR_HIP, "events");
  hipError_t e = hipEventRecord(e0, (hipStream_t)stream);
  int rc = VVCLF_OK;
  for (uint32_t i = 0; i < reps && !rc && e == hipSuccess; i++) rc = once();
  if (e == hipSuccess) e = hipEventRecord(e1, (hipStream_t)stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "batch_time events");
  *ms = (double)t / reps;
  return VVCLF_OK;
}

// Times every supported tile shape and workgroups-per-CU count on the given
// jobs (one warm launch, then the mean of three timed launches) and returns
// the fastest.
int vvclf_autotune(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n_jobs, vvclf_tune* out) {
  clear_error();
  if (!ctx || !jobs || !n_jobs || !out) return fail(VVCLF_ERR_INVALID, "autotune: bad argument");
  std::vector<vvclf_job*> js(jobs, jobs + n_jobs);
  hipSetDevice(ctx->device);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(VVCLF_ERR_HIP, "autotune: events");
  float best = 1e30f;
  vvclf_tune bt{};
  bool hbd = false;
  for (auto* j : js) hbd = hbd || job_needs_hbd(j->d);
  for (auto& sh : kFusedShapes) {
    if (!shape_fits(js, sh[0], sh[1])) continue;
    int occ = fused_max_blocks_per_cu(sh[0], sh[1], hbd);
    for (int w = 1; w <= occ; w++) {
      vvclf_tune t{(uint32_t)sh[0], (uint32_t)sh[1], (uint32_t)fused_threads(sh[0], sh[1]), (uint32_t)w};
      vvclf_batch* b = nullptr;
      if (vvclf_batch_create(ctx, jobs, n_jobs, &t, &b) != VVCLF_OK) continue;
      float ms = 0.f;
      bool ok = vvclf_batch_run(b, ctx->stream) == VVCLF_OK;
      ok = ok && hipEventRecord(e0, ctx->stream) == hipSuccess;
      for (int r = 0; r < 3 && ok; r++) ok = vvclf_batch_run(b, ctx->stream) == VVCLF_OK;
      ok = ok && hipEventRecord(e1, ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
           hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      vvclf_batch_destroy(b);
      if (ok && ms / 3 < best) {
        best = ms / 3;
        bt = t;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (bt.tile_w == 0) return fail(VVCLF_ERR_UNSUPPORTED, "autotune: no tile shape fits these jobs");
  *out = bt;
  return VVCLF_OK;
}

}  // extern "C"
