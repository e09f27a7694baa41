// vvclf_host_internal.h — host-only declarations shared by the C ABI sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>

#include <cstdlib>
#include <functional>
#include <new>
#include <utility>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_internal.h"

namespace vvclf {

enum : uint32_t { STAGE_LMCS = 1, STAGE_DBK = 2, STAGE_SAO = 4, STAGE_ALF = 8, STAGE_ALL = 15 };

// Host copies of the derived tables, before upload.
// Storage of the large per-picture maps, recycled between pictures: blocks
// of 1 MB and more return to a process-wide free list (up to 512 MB) when
// released and are handed out again for a request of the same or up to 25 %
// smaller size, so that a picture size that repeats finds its memory mapped
// and cached.  New blocks hold zeros (anonymous mappings populated by the
// kernel in one call; calloc below 1 MB); a recycled block holds what its
// last user left, and pooled_alloc_recycled() tells the allocating thread
// which of the two its last allocation was.  Elements are
// default-initialised.
void* pooled_alloc_bytes(size_t bytes);
void pooled_free_bytes(void* p, size_t bytes);
bool pooled_alloc_recycled();
template <class T>
struct pooled_alloc {
  using value_type = T;
  pooled_alloc() = default;
  template <class U>
  pooled_alloc(const pooled_alloc<U>&) {}
  T* allocate(size_t n) {
    void* p = pooled_alloc_bytes(n * sizeof(T));
    if (!p) throw std::bad_alloc();
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t n) { pooled_free_bytes(p, n * sizeof(T)); }
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
  bool operator==(const pooled_alloc&) const { return true; }
  bool operator!=(const pooled_alloc&) const { return false; }
};
// A deblocking descriptor map.
using map_vec = std::vector<uint32_t, pooled_alloc<uint32_t>>;

struct host_tables {
  // deblock
  uint32_t gw = 0, gh = 0, cvw = 0, cvh = 0, chw = 0, chh = 0;
  map_vec e_lv, e_lh, e_cv[2], e_ch[2];
  // nonzero descriptors: luma gy << 16 | gx; chroma comp << 31 | sy << 16 | sx
  std::vector<uint32_t> l_lv, l_lh, l_cv, l_ch;
  bool has_chroma_dbk = false;
  uint64_t bs_hist[3] = {0, 0, 0};
  uint64_t chroma_filtered = 0;
  // SAO
  bool do_sao = false, sao_luma = false, sao_chroma = false;
  uint32_t swctb = 0, shctb = 0, sao_ctb_log2 = 7, sao_bd = 10;
  std::vector<uint64_t> sao, sao_rect;
  // ALF
  bool do_alf = false;
  uint32_t wctb = 0, hctb = 0, alf_ctb_log2 = 7, alf_bd = 10;
  std::vector<uint32_t> alf_ctb, alf_bank, alf_cbank;
  std::vector<int32_t> alf_ccbank;
  std::vector<uint64_t> alf_rect;
  // LMCS
  std::vector<uint16_t> lmcs_lut;
};

// The context of the drop-in calls (device 0), created on first use; NULL
// when no device is usable.
vvclf_ctx* default_ctx();

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
const char* last_error();
void clear_error();

// threads: host threads walking the CU records (pictures of >= 4096 records)
int derive_deblock_maps(const vvclf_frame& f, uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, host_tables& t,
                        int threads = 1);
int derive_sao_alf(const vvclf_frame& f, uint32_t W, uint32_t H, host_tables& t, uint32_t stages);
void build_lmcs_lut(const vvclf_lmcs_binding& b, std::vector<uint16_t>& lut);
// Threads of the host CU walk (up to 8; only pictures of >= 4096 CU records
// use more than one).
int host_walk_threads();
// Runs fn(0) .. fn(n - 1) on the calling thread and a process-wide pool of
// min(16, CPUs) - 1 workers (started on first use), returning when all have
// finished.  Concurrent callers share the workers.
void parallel_for(int n, const std::function<void(int)>& fn);
void build_alf_banks(const vvclf_alf_config& ac, const vvclf_alf_aps_binding& b, host_tables& t);
// Plane copies between a caller's host plane and a pitched device plane: one
// linear copy when both sides share the row pitch (towards the host only when
// the rows carry no padding, which belongs to the caller), else a 2D copy.
// (A 2D copy from pageable host memory goes row by row: 20 ms for a 4K plane
// against 1-2 ms linear.)
inline hipError_t copy_rows(void* dst, size_t dp, const void* src, size_t sp, size_t wb, size_t rows,
                            hipMemcpyKind k, hipStream_t st) {
  if (rows && dp == sp && (k == hipMemcpyHostToDevice || wb == dp))
    return hipMemcpyAsync(dst, src, dp * (rows - 1) + wb, k, st);
  return hipMemcpy2DAsync(dst, dp, src, sp, wb, rows, k, st);
}
inline hipError_t copy_rows_sync(void* dst, size_t dp, const void* src, size_t sp, size_t wb, size_t rows,
                                 hipMemcpyKind k) {
  if (rows && dp == sp && (k == hipMemcpyHostToDevice || wb == dp)) return hipMemcpy(dst, src, dp * (rows - 1) + wb, k);
  return hipMemcpy2D(dst, dp, src, sp, wb, rows, k);
}

// The transposeIdx permutations (perm[t][k]: the APS tap read at bank
// position k) and the four clip words of a bit depth, as the luma banks use
void alf_luma_bank_tables(uint32_t bd, uint8_t perm[4][12], uint32_t clipw[4]);
// The ALF / SAO fetch rectangle of CTB (rx, ry) (pack_rect), the tile's
// rectangle with tile bounds, else the picture; 0 when empty.
uint64_t ctb_fetch_rect(const vvclf_bounds* tb, uint32_t rx, uint32_t ry, uint32_t ctb, uint32_t W, uint32_t H);

static constexpr int PLAIN_STAGES = 6;  // prep, deblock V, deblock H, SAO, ALF luma, ALF chroma
// Device tables shared by many jobs (a serving stream's resident LMCS / ALF
// state); a job built with them uploads only its own per-picture tables.
struct resident_tables {
  const uint16_t* lut = nullptr;     // 65536 entries
  const uint32_t* bank = nullptr;    // luma banks
  const uint32_t* cbank = nullptr;   // chroma bank
  const int32_t* ccbank = nullptr;   // CC banks
};

// Pinned host staging buffer, grown on demand; the owner frees it.
struct pinned_buf {
  void* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t n);
  void release();
};

struct job_opts {
  hipStream_t stream = nullptr;  // stream-ordered allocation and copies (memory pool)
  const resident_tables* res = nullptr;
  pinned_buf* stage = nullptr;   // upload through this pinned buffer
  pinned_buf* side_stage = nullptr;  // side tables assembled in this pinned buffer (else on the heap)
  int host_threads = 1;          // threads deriving the deblocking descriptors (host walk)
  // derive the deblocking descriptors on the GPU (derive_deblock_gpu); no
  // caller sets it: the path is switched off until a device memory fault it
  // raised is explained
  bool gpu_derive = false;
  // Device memory supplied by the caller instead of allocated: the job's
  // sample sets (three of them, the output sharing the input's, since the
  // input is dead after the first kernel of the plain form) and side tables
  // are carved from [arena, arena + arena_bytes); job creation fails with
  // VVCLF_ERR_UNSUPPORTED when they do not fit.
  uint8_t* arena = nullptr;
  size_t arena_bytes = 0;
};

// Device destinations of the GPU deblocking derivation, inside a job's side
// allocation: maps zero-filled, lists with room for every slot of their maps.
struct gpu_maps {
  uint32_t *e_lv, *e_lh, *e_cv[2], *e_ch[2];
  uint32_t gw, gh, cvw, cvh, chw, chh;
  uint32_t *l_lv, *l_lh, *l_cv, *l_ch;
  bool chroma;
};
// The descriptors and work lists of derive_deblock_maps, computed on the GPU
// on stream s (the call waits for it); n_out receives the list lengths
// (luma V, luma H, chroma V, chroma H).
int derive_deblock_gpu(const vvclf_frame& f, uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, const gpu_maps& o,
                       hipStream_t s, uint32_t n_out[4]);

// Host planes <-> a job's contiguous sample set (luma, Cb, Cr at the device
// pitches) through a pinned buffer: CPU row copies plus linear DMAs.
// upload_staged returns after the DMAs completed (the buffer may be reused);
// download_staged_issue only queues the DMA into `st`, download_staged_finish
// copies the rows out once `s` has reached it; download_staged does both in
// bands of rows (the row copies of one band overlapping the DMA of the next)
// and returns with the caller's planes written.
const uint16_t* job_final(const vvclf_job* j, int comp);  // the final sample set's plane
hipError_t upload_staged(vvclf_job* j, const vvclf_picture& pic, pinned_buf& st, hipStream_t s);
hipError_t download_staged(vvclf_job* j, pinned_buf& st, hipStream_t s, vvclf_picture& pic);
hipError_t download_staged_issue(vvclf_job* j, pinned_buf& st, hipStream_t s);
void download_staged_finish(const vvclf_job* j, const pinned_buf& st, vvclf_picture& pic);

int job_create_stages(vvclf_ctx* ctx, const vvclf_frame* f, uint32_t stages, vvclf_job** out, uint64_t* dry_stats,
                      const job_opts* opts = nullptr, host_tables* keep = nullptr);
// hbd selects the fused kernel variant for ALF bit depths above 12 (the
// packed 16-bit ALF arithmetic covers <= 12).
hipError_t launch_fused(int tw, int th, bool hbd, const dev_job* jobs, const tile_ref* tiles, uint32_t n_tiles,
                        uint32_t grid, hipStream_t s);
int fused_max_blocks_per_cu(int tw, int th, bool hbd);
int fused_threads(int tw, int th);  // threads per workgroup of a tile shape
inline bool job_needs_hbd(const dev_job& d) { return d.do_alf && d.alf_bd > 12; }
// A fused tile lies inside one ALF CTB (one filter bank per tile) and spans at
// most 3x3 SAO CTBs with its SAO reach (the kernel stages those parameters).
inline bool job_fits_shape(const dev_job& d, uint32_t tw, uint32_t th) {
  if (d.do_alf) {
    const uint32_t ctb = 1u << d.ctb_log2;
    if (tw > ctb || th > ctb || ctb % tw || ctb % th) return false;
  }
  if (d.do_sao) {
    const uint32_t ctb = 1u << d.sao_ctb_log2;
    if (tw > ctb || th > ctb || ctb < 8) return false;
  }
  return true;
}
// Fused tile shapes, largest first.
constexpr int kFusedShapes[4][2] = {{128, 64}, {64, 64}, {64, 32}, {32, 32}};

hipError_t launch_plain(const dev_job& j, hipStream_t s, hipEvent_t* ev = nullptr);

}  // namespace vvclf

// Opaque handle bodies (global namespace, matching the C typedefs).
struct vvclf_ctx {
  int device;
  hipStream_t stream;
};

// Per-stage timing of plain runs: PLAIN_STAGES + 1 events per recorded run.
struct vvclf_prof {
  int device;
  uint32_t cap, used;
  std::vector<hipEvent_t> ev;
};

struct vvclf_job {
  vvclf_ctx* ctx;
  vvclf::dev_job d;
  void* dev_planes;   // one allocation: in, a, b, out
  void* dev_side;     // one allocation: all side tables
  uint64_t bytes_planes, bytes_side;
  uint32_t w, h, wc, hc;
  uint32_t stages;
  bool fused_out = false;  // last run was the fused form (result in `out`)
  hipStream_t pool_stream = nullptr;  // set when allocated from the stream-ordered pool
  bool arena_backed = false;          // device memory belongs to the caller (job_opts::arena)
  hipGraph_t graph = nullptr;         // the plain sequence, captured on first graph launch
  hipGraphExec_t exec = nullptr;
  // host copies of the per-CTB ALF words and rectangles (tile records)
  std::vector<uint32_t> h_alf_ctb;
  std::vector<uint64_t> h_alf_rect;
};

namespace vvclf {
// The fused kernel's record of the tile at (x, y) of job `j` (index `ji` in
// the launch's job array).
inline tile_ref make_tile(const vvclf_job* j, uint32_t ji, uint32_t x, uint32_t y) {
  const dev_job& d = j->d;
  tile_ref t{ji, (uint16_t)x, (uint16_t)y, 0xffffffffu, 0u, ((uint64_t)(uint32_t)d.w << 32) | ((uint64_t)(uint32_t)d.h << 48)};
  const uint32_t rx = x >> d.ctb_log2, ry = y >> d.ctb_log2;
  if (d.do_alf && rx < (uint32_t)d.wctb && ry < (uint32_t)d.hctb) {
    const size_t ci = (size_t)ry * d.wctb + rx;
    t.alf_cw = j->h_alf_ctb[ci];
    t.alf_rect = j->h_alf_rect[ci];
  }
  return t;
}
}  // namespace vvclf
