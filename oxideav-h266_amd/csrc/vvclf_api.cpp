// vvclf_api.cpp — job lifetime and the synchronous drop-in entry points of
// include/vvclf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"
#include "vvclf_internal.h"

using namespace vvclf;

namespace {

inline uint32_t round_up(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

int check_plane(const vvclf_plane& p, const char* name) {
  if (!p.samples) return fail(VVCLF_ERR_INVALID, "%s: NULL samples", name);
  if (p.width == 0 || p.height == 0) return fail(VVCLF_ERR_INVALID, "%s: empty plane", name);
  if (p.stride < p.width) return fail(VVCLF_ERR_INVALID, "%s: stride %u < width %u", name, p.stride, p.width);
  if (p.width > 65535 || p.height > 65535) return fail(VVCLF_ERR_UNSUPPORTED, "%s: plane larger than 65535", name);
  if (p.bit_depth < 8 || p.bit_depth > 16) return fail(VVCLF_ERR_INVALID, "%s: bit_depth %u", name, p.bit_depth);
  return VVCLF_OK;
}

// Bands of a staged copy of `bytes`: one per copy thread from 16 MB on (a
// band's copy is its thread's share of the host work), else one DMA (whose
// row copies still split over the copy threads from 4 MB on).
int copy_bands(size_t bytes) { return bytes >= ((size_t)16 << 20) ? host_walk_threads() : 1; }

// The side tables of a job, laid out at 256-byte offsets and assembled in
// one copy pass: into a pinned buffer when one is given (no page faults on a
// fresh heap block, and the upload is a direct DMA), else on the heap.  The
// added tables must outlive materialize().
struct side_builder {
  struct part {
    const void* p;
    size_t n, off;
  };
  std::vector<part> parts;
  std::vector<uint8_t> blob;
  size_t size = 0;
  size_t add(const void* p, size_t n) {
    const size_t off = (size + 255) / 256 * 256;
    parts.push_back({p, n, off});
    size = off + std::max<size_t>(n, 4);
    return off;
  }
  template <class T, class A>
  size_t add(const std::vector<T, A>& v) {
    return add(v.data(), v.size() * sizeof(T));
  }
  // bytes [lo, hi) of the blob: the tables, and zeros in the gaps and in the
  // 4 bytes an empty table reserves
  void fill(uint8_t* dst, size_t lo, size_t hi) const {
    size_t at = 0;
    auto zero = [&](size_t a, size_t b) {
      a = std::max(a, lo);
      b = std::min(b, hi);
      if (a < b) memset(dst + a, 0, b - a);
    };
    for (const part& q : parts) {
      zero(at, q.off);
      const size_t a = std::max(q.off, lo), b = std::min(q.off + q.n, hi);
      if (a < b) memcpy(dst + a, (const uint8_t*)q.p + (a - q.off), b - a);
      at = q.off + q.n;
    }
    zero(at, size);
  }
  const uint8_t* materialize(pinned_buf* st) {
    uint8_t* dst;
    if (st) {
      if (!st->ensure(size)) return nullptr;
      dst = (uint8_t*)st->p;
    } else {
      blob.resize(size);
      dst = blob.data();
    }
    fill(dst, 0, size);
    return dst;
  }
  // Assembly in the pinned buffer and upload to `dev` on `s`, in bands of
  // 256-byte multiples on the worker pool for blobs of 4 MB and more: each
  // band's DMA is queued as soon as its bytes are in place.  Returns once the
  // stream has drained.
  hipError_t upload(pinned_buf& st, uint8_t* dev, hipStream_t s) const {
    if (!st.ensure(size)) return hipErrorOutOfMemory;
    uint8_t* dst = (uint8_t*)st.p;
    const int nb = size >= ((size_t)4 << 20) ? host_walk_threads() : 1;
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    std::atomic<int> err{(int)hipSuccess};
    parallel_for(nb, [&](int b) {
      const size_t lo = size * b / nb / 256 * 256, hi = b + 1 == nb ? size : size * (b + 1) / nb / 256 * 256;
      if (lo >= hi) return;
      fill(dst, lo, hi);
      hipError_t be = hipSetDevice(device);
      if (be == hipSuccess) be = hipMemcpyAsync(dev + lo, dst + lo, hi - lo, hipMemcpyHostToDevice, s);
      int ok = (int)hipSuccess;
      if (be != hipSuccess) err.compare_exchange_strong(ok, (int)be);
    });
    e = hipStreamSynchronize(s);
    return err.load() != (int)hipSuccess ? (hipError_t)err.load() : e;
  }
};

}  // namespace

// The context of the drop-in calls (device 0), created on first use.
vvclf_ctx* vvclf::default_ctx() {
  static std::mutex mu;
  static vvclf_ctx* ctx = nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (!ctx) {
    if (vvclf_ctx_create(0, &ctx) != VVCLF_OK) {
      ctx = nullptr;
    } else {
      // per-picture jobs allocate from the device's stream-ordered pool;
      // freed memory stays cached for the next picture
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, 0) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
      }
    }
  }
  return ctx;
}

namespace vvclf {

int job_create_stages(vvclf_ctx* ctx, const vvclf_frame* f, uint32_t stages, vvclf_job** out,
                      uint64_t* dry_stats, const job_opts* opts, host_tables* keep) {
  const resident_tables* res = opts ? opts->res : nullptr;
  hipStream_t ps = opts ? opts->stream : nullptr;
  clear_error();
  if ((!ctx && !dry_stats) || !f || (!out && !dry_stats)) return fail(VVCLF_ERR_INVALID, "job_create: NULL argument");
  if (out) *out = nullptr;
  int rc;
  if ((rc = check_plane(f->pic.luma, "luma"))) return rc;
  const uint32_t cfi = f->chroma_format_idc;
  const uint32_t W = f->pic.luma.width, H = f->pic.luma.height;
  uint32_t Wc = 0, Hc = 0;
  // Chroma is processed when any active stage reads it.
  bool want_chroma = false;
  if ((stages & STAGE_DBK) && !f->deblock.disabled && cfi != 0) want_chroma = true;
  if ((stages & STAGE_SAO) && f->sao_cfg.chroma_used && f->sao_cfg.chroma_format_idc != 0) want_chroma = true;
  if ((stages & STAGE_ALF) && f->alf_cfg.alf_enabled && f->alf_cfg.chroma_format_idc != 0) want_chroma = true;
  if (want_chroma) {
    if (cfi != 1 && (stages & STAGE_DBK) && !f->deblock.disabled)
      return fail(VVCLF_ERR_UNSUPPORTED, "chroma_format_idc %u: only 4:0:0 and 4:2:0 run on the GPU", cfi);
    if ((rc = check_plane(f->pic.cb, "cb")) || (rc = check_plane(f->pic.cr, "cr"))) return rc;
    if (f->pic.cb.width != f->pic.cr.width || f->pic.cb.height != f->pic.cr.height)
      return fail(VVCLF_ERR_INVALID, "cb and cr planes differ in size");
    if (f->pic.cb.bit_depth != f->pic.luma.bit_depth || f->pic.cr.bit_depth != f->pic.luma.bit_depth)
      return fail(VVCLF_ERR_UNSUPPORTED, "planes with different bit depths");
    Wc = f->pic.cb.width;
    Hc = f->pic.cb.height;
  }
  host_tables t;
  const bool do_lmcs = (stages & STAGE_LMCS) && f->lmcs.derived;
  const bool do_dbk = (stages & STAGE_DBK) && !f->deblock.disabled;
  const bool gpu_dbk = do_dbk && !dry_stats && opts && opts->gpu_derive;
  if (do_lmcs &&
      (f->lmcs.derived->bit_depth < 8 || f->lmcs.derived->bit_depth > 16 || f->lmcs.max_bin_idx > 15))
    return fail(VVCLF_ERR_INVALID, "lmcs: invalid derived tables");
  if (do_dbk) {
    if (f->deblock.bit_depth < 8 || f->deblock.bit_depth > 16)
      return fail(VVCLF_ERR_INVALID, "deblock: bit_depth %u", f->deblock.bit_depth);
    if (f->n_cus && !f->cus) return fail(VVCLF_ERR_INVALID, "deblock: NULL cus");
  }
  // The host tables: LMCS inverse table, deblocking descriptors, SAO / ALF
  // parameters.  Returns a status; the message stays with the thread that
  // ran it.
  auto derive_tables = [&]() -> int {
    int trc;
    if (do_lmcs) build_lmcs_lut(f->lmcs, t.lmcs_lut);
    if (do_dbk) {
      if (gpu_dbk) {  // sizes only; the descriptors are derived on the GPU after the upload
        t.gw = (W + 3) / 4 + 2;
        t.gh = (H + 3) / 4 + 2;
        t.has_chroma_dbk = f->chroma_format_idc != 0 && Wc > 0;
        if (t.has_chroma_dbk) {
          t.cvw = (Wc + 7) / 8 + 1;
          t.cvh = (Hc + 1) / 2 + 1;
          t.chw = (Wc + 1) / 2 + 1;
          t.chh = (Hc + 7) / 8 + 1;
        }
      } else if ((trc = derive_deblock_maps(*f, W, H, Wc, Hc, t, opts ? opts->host_threads : 1))) {
        return trc;
      }
    }
    return derive_sao_alf(*f, W, H, t, stages);
  };
  if (dry_stats) {
    if ((rc = derive_tables())) return rc;
    for (int i = 0; i < 3; i++) dry_stats[i] = t.bs_hist[i];
    dry_stats[3] = t.chroma_filtered;
    uint64_t sao_on = 0, alf_on = 0;
    for (size_t i = 0; i + 2 < t.sao.size(); i += 3)
      if ((t.sao[i] | t.sao[i + 1] | t.sao[i + 2]) & 3) sao_on++;
    for (uint32_t w : t.alf_ctb)
      if ((w & 0xff) != ALF_OFF8) alf_on++;
    dry_stats[4] = sao_on;
    dry_stats[5] = alf_on;
    uint64_t py = round_up(W, 64), pc = Wc ? round_up(Wc, 64) : 0;
    uint64_t side = t.lmcs_lut.size() * 2 + (t.e_lv.size() + t.e_lh.size()) * 4 +
                    (t.e_cv[0].size() + t.e_cv[1].size() + t.e_ch[0].size() + t.e_ch[1].size()) * 4 +
                    (t.l_lv.size() + t.l_lh.size() + t.l_cv.size() + t.l_ch.size()) * 4 +
                    (t.sao.size() + t.sao_rect.size() + t.alf_rect.size()) * 8 +
                    (t.alf_ctb.size() + t.alf_bank.size() + t.alf_cbank.size() + t.alf_ccbank.size()) * 4;
    dry_stats[6] = 4 * (py * H + 2 * pc * Hc) * 2 + side;
    dry_stats[7] = 0;
    return VVCLF_OK;
  }

  vvclf_job* j = new vvclf_job();
  memset(&j->d, 0, sizeof(j->d));
  j->ctx = ctx;
  j->w = W;
  j->h = H;
  j->wc = Wc;
  j->hc = Hc;
  j->stages = stages;
  dev_job& d = j->d;
  d.w = (int32_t)W;
  d.h = (int32_t)H;
  d.wc = (int32_t)Wc;
  d.hc = (int32_t)Hc;
  d.py = (int32_t)round_up(W, 64);
  d.pc = Wc ? (int32_t)round_up(Wc, 64) : 0;
  d.bd = (int32_t)f->pic.luma.bit_depth;
  d.bd_dbk = (int32_t)f->deblock.bit_depth;

  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) {
    delete j;
    return hip_fail(e, "hipSetDevice");
  }
  // sample planes: in, a, b, out (with a caller's arena: in, a, b, and out
  // sharing in's memory)
  const bool arena = opts && opts->arena;
  size_t luma_elems = (size_t)d.py * H, chroma_elems = (size_t)d.pc * Hc;
  size_t set_bytes = (luma_elems + 2 * chroma_elems) * 2;
  set_bytes = (set_bytes + 255) / 256 * 256;
  const int n_sets = arena ? 3 : 4;
  j->bytes_planes = set_bytes * n_sets;
  j->pool_stream = ps;
  j->arena_backed = arena;
  if (arena) {
    if (j->bytes_planes > opts->arena_bytes) {
      const size_t need = j->bytes_planes;
      delete j;
      return fail(VVCLF_ERR_UNSUPPORTED, "picture needs %zu bytes of sample buffers, more than the %zu-byte scratch",
                  need, opts->arena_bytes);
    }
    j->dev_planes = opts->arena;
  } else {
    e = ps ? hipMallocAsync(&j->dev_planes, j->bytes_planes, ps) : hipMalloc(&j->dev_planes, j->bytes_planes);
    if (e != hipSuccess) {
      delete j;
      return fail(VVCLF_ERR_NOMEM, "hipMalloc planes (%zu bytes): %s", (size_t)j->bytes_planes, hipGetErrorString(e));
    }
  }
  dev_planes* sets[4] = {&d.in, &d.a, &d.b, &d.out};
  for (int s = 0; s < n_sets; s++) {
    uint8_t* base = (uint8_t*)j->dev_planes + s * set_bytes;
    sets[s]->y = (uint16_t*)base;
    sets[s]->cb = Wc ? (uint16_t*)(base + luma_elems * 2) : nullptr;
    sets[s]->cr = Wc ? (uint16_t*)(base + (luma_elems + chroma_elems) * 2) : nullptr;
  }
  if (arena) d.out = d.in;
  // With a pinned staging buffer, the samples are staged and uploaded while
  // the tables are derived: both are host work on the worker pool, and the
  // upload touches only the input sample set.
  const bool early_upload = opts && opts->stage;
  if (early_upload) {
    hipError_t ue = hipSuccess;
    int trc = VVCLF_OK;
    std::string terr;
    hipStream_t us = opts->stream ? opts->stream : j->pool_stream;
    parallel_for(2, [&](int task) {
      if (task == 0) {
        trc = derive_tables();
        if (trc) terr = last_error();
      } else {
        ue = hipSetDevice(ctx->device);
        if (ue == hipSuccess) ue = upload_staged(j, f->pic, *opts->stage, us);
      }
    });
    if (trc || ue != hipSuccess) {
      vvclf_job_destroy(j);
      return trc ? fail(trc, "%s", terr.c_str()) : hip_fail(ue, "upload (staged)");
    }
  } else if ((rc = derive_tables())) {
    vvclf_job_destroy(j);
    return rc;
  }
  d.do_lmcs = do_lmcs;
  d.do_dbk = do_dbk;
  d.do_dbk_chroma = do_dbk && t.has_chroma_dbk;
  d.do_sao = t.do_sao;
  d.sao_luma = t.sao_luma;
  d.sao_chroma = t.sao_chroma && Wc > 0;
  d.sao_ctb_log2 = (int32_t)t.sao_ctb_log2;
  d.swctb = (int32_t)t.swctb;
  d.shctb = (int32_t)t.shctb;
  d.sao_bd = (int32_t)t.sao_bd;
  d.do_alf = t.do_alf;
  d.ctb_log2 = (int32_t)t.alf_ctb_log2;
  d.wctb = (int32_t)t.wctb;
  d.hctb = (int32_t)t.hctb;
  d.alf_bd = (int32_t)t.alf_bd;
  if (!t.do_sao) d.b = d.a;
  // side tables
  side_builder sb;
  if (res) {  // resident per-stream tables replace the per-job copies
    t.lmcs_lut.clear();
    t.alf_bank.clear();
    t.alf_cbank.clear();
    t.alf_ccbank.clear();
  }
  size_t o_lut = sb.add(t.lmcs_lut), o_lv = sb.add(t.e_lv), o_lh = sb.add(t.e_lh);
  size_t o_cv0 = sb.add(t.e_cv[0]), o_cv1 = sb.add(t.e_cv[1]), o_ch0 = sb.add(t.e_ch[0]), o_ch1 = sb.add(t.e_ch[1]);
  size_t o_llv = sb.add(t.l_lv), o_llh = sb.add(t.l_lh), o_lcv = sb.add(t.l_cv), o_lch = sb.add(t.l_ch);
  size_t o_sao = sb.add(t.sao), o_saor = sb.add(t.sao_rect);
  size_t o_actb = sb.add(t.alf_ctb), o_arect = sb.add(t.alf_rect), o_bank = sb.add(t.alf_bank);
  size_t o_cbank = sb.add(t.alf_cbank), o_ccbank = sb.add(t.alf_ccbank);
  const size_t host_side = sb.size;
  // GPU derivation: maps and lists live after the uploaded tables
  size_t tail = (host_side + 255) / 256 * 256;
  auto reserve = [&](size_t n) {
    const size_t off = tail;
    tail = (tail + std::max<size_t>(n, 4) + 255) / 256 * 256;
    return off;
  };
  size_t maps_begin = tail, maps_end = tail;
  if (gpu_dbk) {
    const size_t lm = (size_t)t.gw * t.gh * 4, cvm = (size_t)t.cvw * t.cvh * 4, chm = (size_t)t.chw * t.chh * 4;
    o_lv = reserve(lm);
    o_lh = reserve(lm);
    if (t.has_chroma_dbk) {
      o_cv0 = reserve(cvm);
      o_cv1 = reserve(cvm);
      o_ch0 = reserve(chm);
      o_ch1 = reserve(chm);
    }
    maps_end = tail;
    o_llv = reserve(lm);
    o_llh = reserve(lm);
    if (t.has_chroma_dbk) {
      o_lcv = reserve(2 * cvm);
      o_lch = reserve(2 * chm);
    }
  }
  j->bytes_side = gpu_dbk ? tail : host_side;
  if (arena) {
    if (gpu_dbk || j->bytes_planes + j->bytes_side > opts->arena_bytes) {
      const size_t need = j->bytes_planes + j->bytes_side;
      vvclf_job_destroy(j);
      return fail(VVCLF_ERR_UNSUPPORTED, "picture needs %zu bytes of sample buffers and tables, more than the "
                  "%zu-byte scratch", need, opts->arena_bytes);
    }
    j->dev_side = opts->arena + j->bytes_planes;  // bytes_planes is a multiple of 256
  } else {
    e = ps ? hipMallocAsync(&j->dev_side, j->bytes_side, ps) : hipMalloc(&j->dev_side, j->bytes_side);
    if (e != hipSuccess) {
      j->dev_side = nullptr;
      vvclf_job_destroy(j);
      return fail(VVCLF_ERR_NOMEM, "hipMalloc side tables: %s", hipGetErrorString(e));
    }
  }
  if (ps && opts->side_stage && !gpu_dbk) {
    e = sb.upload(*opts->side_stage, (uint8_t*)j->dev_side, ps);
  } else {
    const uint8_t* host_blob = sb.materialize(opts ? opts->side_stage : nullptr);
    if (!host_blob) {
      vvclf_job_destroy(j);
      return fail(VVCLF_ERR_NOMEM, "pinned staging of the side tables (%zu bytes)", host_side);
    }
    e = ps ? hipMemcpyAsync(j->dev_side, host_blob, host_side, hipMemcpyHostToDevice, ps)
           : hipMemcpy(j->dev_side, host_blob, host_side, hipMemcpyHostToDevice);
    if (e == hipSuccess && gpu_dbk)
      e = hipMemsetAsync((uint8_t*)j->dev_side + maps_begin, 0, maps_end - maps_begin, ps);
    if (e == hipSuccess && ps) e = hipStreamSynchronize(ps);
  }
  if (e != hipSuccess) {
    vvclf_job_destroy(j);
    return e == hipErrorOutOfMemory ? fail(VVCLF_ERR_NOMEM, "pinned staging of the side tables (%zu bytes)", host_side)
                                    : hip_fail(e, "upload side tables");
  }
  uint8_t* sd = (uint8_t*)j->dev_side;
  d.lmcs_lut = do_lmcs ? (res ? res->lut : (const uint16_t*)(sd + o_lut)) : nullptr;
  d.gw = (int32_t)t.gw;
  d.gh = (int32_t)t.gh;
  d.e_lv = (const uint32_t*)(sd + o_lv);
  d.e_lh = (const uint32_t*)(sd + o_lh);
  d.cvw = (int32_t)t.cvw;
  d.cvh = (int32_t)t.cvh;
  d.chw = (int32_t)t.chw;
  d.chh = (int32_t)t.chh;
  d.e_cv[0] = (const uint32_t*)(sd + o_cv0);
  d.e_cv[1] = (const uint32_t*)(sd + o_cv1);
  d.e_ch[0] = (const uint32_t*)(sd + o_ch0);
  d.e_ch[1] = (const uint32_t*)(sd + o_ch1);
  d.l_lv = (const uint32_t*)(sd + o_llv);
  d.l_lh = (const uint32_t*)(sd + o_llh);
  d.l_cv = (const uint32_t*)(sd + o_lcv);
  d.l_ch = (const uint32_t*)(sd + o_lch);
  d.n_lv = (int32_t)t.l_lv.size();
  d.n_lh = (int32_t)t.l_lh.size();
  d.n_cv = (int32_t)t.l_cv.size();
  d.n_ch = (int32_t)t.l_ch.size();
  if (gpu_dbk) {
    gpu_maps gm;
    gm.e_lv = (uint32_t*)(sd + o_lv);
    gm.e_lh = (uint32_t*)(sd + o_lh);
    gm.e_cv[0] = (uint32_t*)(sd + o_cv0);
    gm.e_cv[1] = (uint32_t*)(sd + o_cv1);
    gm.e_ch[0] = (uint32_t*)(sd + o_ch0);
    gm.e_ch[1] = (uint32_t*)(sd + o_ch1);
    gm.gw = t.gw;
    gm.gh = t.gh;
    gm.cvw = t.cvw;
    gm.cvh = t.cvh;
    gm.chw = t.chw;
    gm.chh = t.chh;
    gm.l_lv = (uint32_t*)(sd + o_llv);
    gm.l_lh = (uint32_t*)(sd + o_llh);
    gm.l_cv = (uint32_t*)(sd + o_lcv);
    gm.l_ch = (uint32_t*)(sd + o_lch);
    gm.chroma = t.has_chroma_dbk;
    uint32_t n[4];
    if ((rc = derive_deblock_gpu(*f, W, H, Wc, Hc, gm, ps, n))) {
      vvclf_job_destroy(j);
      return rc;
    }
    d.n_lv = (int32_t)n[0];
    d.n_lh = (int32_t)n[1];
    d.n_cv = (int32_t)n[2];
    d.n_ch = (int32_t)n[3];
  }
  d.sao = (const uint64_t*)(sd + o_sao);
  d.sao_rect = t.sao_rect.empty() ? nullptr : (const uint64_t*)(sd + o_saor);
  d.alf_ctb = (const uint32_t*)(sd + o_actb);
  d.alf_rect = (const uint64_t*)(sd + o_arect);
  d.alf_bank = res ? res->bank : (const uint32_t*)(sd + o_bank);
  d.alf_cbank = res ? res->cbank : (const uint32_t*)(sd + o_cbank);
  d.alf_ccbank = res ? res->ccbank : (const int32_t*)(sd + o_ccbank);
  j->h_alf_ctb = t.alf_ctb;
  j->h_alf_rect = t.alf_rect;
  if (keep) *keep = std::move(t);
  if (!early_upload && (rc = vvclf_job_set_input(j, &f->pic))) {
    vvclf_job_destroy(j);
    return rc;
  }
  *out = j;
  return VVCLF_OK;
}

bool pinned_buf::ensure(size_t n) {
  if (n <= cap) return true;
  release();
  if (hipHostMalloc(&p, n, 0) != hipSuccess) {
    p = nullptr;
    return false;
  }
  cap = n;
  return true;
}

void pinned_buf::release() {
  if (p) hipHostFree(p);
  p = nullptr;
  cap = 0;
}

namespace {
size_t set_samples(const vvclf_job* j) {
  return (size_t)j->d.py * j->h + 2 * (size_t)j->d.pc * j->hc;
}

// Pitched row copies of the three planes, split over up to
// host_walk_threads() threads when they move at least 4 MB (a 4K picture's
// 25 MB take several ms on one core).
struct plane_copy {
  uint8_t* dst;
  size_t dpitch;
  const uint8_t* src;
  size_t spitch, rows, bytes;
};
// rows b..e of the concatenated planes
void copy_rows(const plane_copy* pc, int n, size_t b, size_t e) {
  size_t base = 0;
  for (int i = 0; i < n; i++) {
    const size_t lo = std::max(b, base), hi = std::min(e, base + pc[i].rows);
    for (size_t r = lo; r < hi; r++)
      memcpy(pc[i].dst + (r - base) * pc[i].dpitch, pc[i].src + (r - base) * pc[i].spitch, pc[i].bytes);
    base += pc[i].rows;
  }
}
void copy_planes(const plane_copy* pc, int n) {
  size_t total = 0, rows = 0;
  for (int i = 0; i < n; i++) {
    total += pc[i].rows * pc[i].bytes;
    rows += pc[i].rows;
  }
  const int nt = total >= ((size_t)4 << 20) ? host_walk_threads() : 1;
  parallel_for(nt, [&](int ti) { copy_rows(pc, n, rows * ti / nt, rows * (ti + 1) / nt); });
}

// The staged sample set in bands of rows of the concatenated planes (luma,
// then Cb and Cr at one pitch, contiguous in the pinned buffer and on the
// device; copy_bands of them), so that the row copies of one band overlap
// the DMA of another.
struct bands {
  size_t rows, py, pc, h;
  int n;
  explicit bands(const vvclf_job* j)
      : rows((size_t)j->h + (j->wc ? 2 * (size_t)j->hc : 0)), py(j->d.py), pc(j->d.pc), h(j->h),
        n(copy_bands(set_samples(j) * 2)) {}
  size_t row_lo(int b) const { return rows * b / n; }
  // offset in samples of row r of the concatenated planes
  size_t off(size_t r) const { return r <= h ? r * py : py * h + (r - h) * pc; }
};
}  // namespace

hipError_t upload_staged(vvclf_job* j, const vvclf_picture& pic, pinned_buf& st, hipStream_t s) {
  if (pic.luma.width != j->w || pic.luma.height != j->h || !pic.luma.samples) return hipErrorInvalidValue;
  if (j->wc && (!pic.cb.samples || !pic.cr.samples || pic.cb.width != j->wc || pic.cb.height != j->hc))
    return hipErrorInvalidValue;
  const size_t n = set_samples(j);
  if (!st.ensure(n * 2)) return hipErrorOutOfMemory;
  uint16_t* h = (uint16_t*)st.p;
  uint16_t* hc = h + (size_t)j->d.py * j->h;
  plane_copy pc[3] = {{(uint8_t*)h, (size_t)j->d.py * 2, (const uint8_t*)pic.luma.samples, (size_t)pic.luma.stride * 2,
                       j->h, (size_t)j->w * 2}};
  int np = 1;
  for (int c = 0; c < 2 && j->wc; c++, np++) {
    const vvclf_plane& pl = c ? pic.cr : pic.cb;
    pc[np] = {(uint8_t*)(hc + (size_t)c * j->d.pc * j->hc), (size_t)j->d.pc * 2, (const uint8_t*)pl.samples,
              (size_t)pl.stride * 2, j->hc, (size_t)j->wc * 2};
  }
  // each band goes to the device as soon as its rows are in the pinned buffer
  const bands bd(j);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::atomic<int> err{(int)hipSuccess};
  parallel_for(bd.n, [&](int b) {
    if (bd.n == 1) copy_planes(pc, np);
    else copy_rows(pc, np, bd.row_lo(b), bd.row_lo(b + 1));
    const size_t lo = bd.off(bd.row_lo(b)), hi = bd.off(bd.row_lo(b + 1));
    hipError_t be = hipSetDevice(dev);
    if (be == hipSuccess && hi > lo) be = hipMemcpyAsync(j->d.in.y + lo, h + lo, (hi - lo) * 2, hipMemcpyHostToDevice, s);
    int ok = (int)hipSuccess;
    if (be != hipSuccess) err.compare_exchange_strong(ok, (int)be);
  });
  // the stream drains what was queued also after a failed band
  e = hipStreamSynchronize(s);
  return err.load() != (int)hipSuccess ? (hipError_t)err.load() : e;
}

hipError_t download_staged_issue(vvclf_job* j, pinned_buf& st, hipStream_t s) {
  const size_t n = set_samples(j);
  if (!st.ensure(n * 2)) return hipErrorOutOfMemory;
  return hipMemcpyAsync(st.p, job_final(j, 0), n * 2, hipMemcpyDeviceToHost, s);
}

// Bands as in upload_staged: band b's DMA is followed by an event, and the
// thread that takes band b copies its rows out once the event has completed.
// The events belong to the calling thread (created on first use, held until
// the process ends).
hipError_t download_staged(vvclf_job* j, pinned_buf& st, hipStream_t s, vvclf_picture& pic) {
  const size_t n = set_samples(j);
  if (!st.ensure(n * 2)) return hipErrorOutOfMemory;
  const bands bd(j);
  static thread_local hipEvent_t ev[16] = {};
  if (bd.n > 16) return hipErrorInvalidValue;
  const uint16_t* d = job_final(j, 0);
  uint16_t* h = (uint16_t*)st.p;
  hipError_t e = hipSuccess;
  int issued = 0;
  for (; issued < bd.n && e == hipSuccess; issued++) {
    if (!ev[issued]) e = hipEventCreateWithFlags(&ev[issued], hipEventDisableTiming);
    const size_t lo = bd.off(bd.row_lo(issued)), hi = bd.off(bd.row_lo(issued + 1));
    if (e == hipSuccess && hi > lo) e = hipMemcpyAsync(h + lo, d + lo, (hi - lo) * 2, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ev[issued], s);
  }
  if (e != hipSuccess) {
    hipStreamSynchronize(s);
    return e;
  }
  const uint16_t* hc = h + (size_t)j->d.py * j->h;
  plane_copy pc[3] = {{(uint8_t*)pic.luma.samples, (size_t)pic.luma.stride * 2, (const uint8_t*)h, (size_t)j->d.py * 2,
                       j->h, (size_t)j->w * 2}};
  int np = 1;
  for (int c = 0; c < 2 && j->wc; c++, np++) {
    vvclf_plane& pl = c ? pic.cr : pic.cb;
    pc[np] = {(uint8_t*)pl.samples, (size_t)pl.stride * 2, (const uint8_t*)(hc + (size_t)c * j->d.pc * j->hc),
              (size_t)j->d.pc * 2, j->hc, (size_t)j->wc * 2};
  }
  std::atomic<int> err{(int)hipSuccess};
  // the calling thread's events (a thread_local named inside the lambda would
  // be each worker's own)
  hipEvent_t* const evs = ev;
  parallel_for(bd.n, [&](int b) {
    const hipError_t be = hipEventSynchronize(evs[b]);
    int ok = (int)hipSuccess;
    if (be != hipSuccess) err.compare_exchange_strong(ok, (int)be);
    else if (bd.n == 1) copy_planes(pc, np);
    else copy_rows(pc, np, bd.row_lo(b), bd.row_lo(b + 1));
  });
  return (hipError_t)err.load();
}

void download_staged_finish(const vvclf_job* j, const pinned_buf& st, vvclf_picture& pic) {
  const uint16_t* h = (const uint16_t*)st.p;
  const uint16_t* hc = h + (size_t)j->d.py * j->h;
  plane_copy pc[3] = {{(uint8_t*)pic.luma.samples, (size_t)pic.luma.stride * 2, (const uint8_t*)h, (size_t)j->d.py * 2,
                       j->h, (size_t)j->w * 2}};
  int np = 1;
  for (int c = 0; c < 2 && j->wc; c++, np++) {
    vvclf_plane& pl = c ? pic.cr : pic.cb;
    pc[np] = {(uint8_t*)pl.samples, (size_t)pl.stride * 2, (const uint8_t*)(hc + (size_t)c * j->d.pc * j->hc),
              (size_t)j->d.pc * 2, j->hc, (size_t)j->wc * 2};
  }
  copy_planes(pc, np);
}

const uint16_t* job_final(const vvclf_job* j, int comp) {
  const dev_planes& p = (j->fused_out || j->d.do_alf) ? j->d.out : (j->d.do_sao ? j->d.b : j->d.a);
  return comp == 0 ? p.y : (comp == 1 ? p.cb : p.cr);
}

}  // namespace vvclf

extern "C" {

const char* vvclf_last_error(void) { return vvclf::last_error(); }

size_t vvclf_abi_sizeof(const char* n) {
  if (!n) return 0;
#define S(T) \
  if (!strcmp(n, #T)) return sizeof(T);
  S(vvclf_plane) S(vvclf_picture) S(vvclf_deblock_cu) S(vvclf_mv_cell) S(vvclf_mv_grid) S(vvclf_deblock_params)
  S(vvclf_sao_ctb) S(vvclf_sao_ctb_params) S(vvclf_sao_picture) S(vvclf_sao_config) S(vvclf_bounds)
  S(vvclf_alf_ctb) S(vvclf_alf_picture) S(vvclf_alf_config) S(vvclf_alf_aps) S(vvclf_alf_clip_rdo) S(vvclf_alf_aps_binding)
  S(vvclf_lmcs_data) S(vvclf_lmcs_derived) S(vvclf_lmcs_binding) S(vvclf_frame) S(vvclf_tune)
  S(vvclf_server_config) S(vvclf_server_stats) S(vvclf_rect) S(vvclf_crs_block) S(vvclf_crop)
  S(vvclf_ctb_syntax) S(vvclf_ctb_syntax_config)
#undef S
  return 0;
}

int vvclf_ctx_create(int device, vvclf_ctx** out) {
  if (!out) return fail(VVCLF_ERR_INVALID, "ctx_create: NULL out");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(VVCLF_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(VVCLF_ERR_INVALID, "device %d of %d", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  vvclf_ctx* c = new vvclf_ctx();
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return VVCLF_OK;
}

int vvclf_ctx_destroy(vvclf_ctx* ctx) {
  if (!ctx) return VVCLF_OK;
  hipSetDevice(ctx->device);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return VVCLF_OK;
}

void* vvclf_ctx_stream(vvclf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int vvclf_ctx_synchronize(vvclf_ctx* ctx) {
  if (!ctx) return fail(VVCLF_ERR_INVALID, "ctx_synchronize: NULL");
  hipSetDevice(ctx->device);
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? VVCLF_OK : hip_fail(e, "hipDeviceSynchronize");
}

int vvclf_job_create(vvclf_ctx* ctx, const vvclf_frame* frame, vvclf_job** out) {
  return vvclf_job_create_flags(ctx, frame, 0, out);
}

int vvclf_job_create_flags(vvclf_ctx* ctx, const vvclf_frame* frame, uint32_t flags, vvclf_job** out) {
  clear_error();
  if (flags & ~VVCLF_JOB_GPU_DERIVE) return fail(VVCLF_ERR_INVALID, "job_create_flags: unknown flags %#x", flags);
  if (flags & VVCLF_JOB_GPU_DERIVE)
    return fail(VVCLF_ERR_UNSUPPORTED, "GPU derivation of the deblocking descriptors is switched off");
  job_opts o;
  o.host_threads = host_walk_threads();
  return job_create_stages(ctx, frame, STAGE_ALL, out, nullptr, &o);
}

int vvclf_job_deblock_table(const vvclf_job* job, uint32_t which, void* out, size_t cap, size_t* size) {
  clear_error();
  if (!job || !size || which > 9) return fail(VVCLF_ERR_INVALID, "job_deblock_table: bad argument");
  const dev_job& d = job->d;
  const uint32_t* src = nullptr;
  size_t n = 0;
  if (d.do_dbk) {
    switch (which) {
      case 0: src = d.e_lv; n = (size_t)d.gw * d.gh; break;
      case 1: src = d.e_lh; n = (size_t)d.gw * d.gh; break;
      case 2: case 3: src = d.e_cv[which - 2]; n = d.do_dbk_chroma ? (size_t)d.cvw * d.cvh : 0; break;
      case 4: case 5: src = d.e_ch[which - 4]; n = d.do_dbk_chroma ? (size_t)d.chw * d.chh : 0; break;
      case 6: src = d.l_lv; n = (size_t)d.n_lv; break;
      case 7: src = d.l_lh; n = (size_t)d.n_lh; break;
      case 8: src = d.l_cv; n = (size_t)d.n_cv; break;
      default: src = d.l_ch; n = (size_t)d.n_ch; break;
    }
  }
  *size = n * 4;
  if (!n || !out || cap < n * 4) return VVCLF_OK;
  hipSetDevice(job->ctx->device);
  hipError_t e = hipMemcpy(out, src, n * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? VVCLF_OK : hip_fail(e, "job_deblock_table");
}

int vvclf_frame_check(const vvclf_frame* frame, uint64_t stats[8]) {
  uint64_t local[8];
  job_opts o;
  o.host_threads = host_walk_threads();
  return job_create_stages(nullptr, frame, STAGE_ALL, nullptr, stats ? stats : local, &o);
}

int vvclf_job_destroy(vvclf_job* j) {
  if (!j) return VVCLF_OK;
  hipSetDevice(j->ctx->device);
  if (j->exec) hipGraphExecDestroy(j->exec);
  if (j->graph) hipGraphDestroy(j->graph);
  if (j->arena_backed) {  // the caller's memory: nothing to free
    delete j;
    return VVCLF_OK;
  }
  if (j->pool_stream) {  // back to the stream-ordered pool, no device-wide sync
    if (j->dev_planes) hipFreeAsync(j->dev_planes, j->pool_stream);
    if (j->dev_side) hipFreeAsync(j->dev_side, j->pool_stream);
    delete j;
    return VVCLF_OK;
  }
  if (j->dev_planes) hipFree(j->dev_planes);
  if (j->dev_side) hipFree(j->dev_side);
  delete j;
  return VVCLF_OK;
}

int vvclf_job_set_input(vvclf_job* j, const vvclf_picture* pic) {
  if (!j || !pic) return fail(VVCLF_ERR_INVALID, "job_set_input: NULL argument");
  if (pic->luma.width != j->w || pic->luma.height != j->h || !pic->luma.samples)
    return fail(VVCLF_ERR_INVALID, "job_set_input: luma geometry differs");
  hipSetDevice(j->ctx->device);
  hipStream_t ps = j->pool_stream;
  hipError_t e = copy_rows(j->d.in.y, (size_t)j->d.py * 2, pic->luma.samples, (size_t)pic->luma.stride * 2,
                                  (size_t)j->w * 2, j->h, hipMemcpyHostToDevice, ps);
  if (e != hipSuccess) return hip_fail(e, "upload luma");
  if (j->wc) {
    if (!pic->cb.samples || !pic->cr.samples || pic->cb.width != j->wc || pic->cb.height != j->hc)
      return fail(VVCLF_ERR_INVALID, "job_set_input: chroma geometry differs");
    e = copy_rows(j->d.in.cb, (size_t)j->d.pc * 2, pic->cb.samples, (size_t)pic->cb.stride * 2,
                         (size_t)j->wc * 2, j->hc, hipMemcpyHostToDevice, ps);
    if (e == hipSuccess)
      e = copy_rows(j->d.in.cr, (size_t)j->d.pc * 2, pic->cr.samples, (size_t)pic->cr.stride * 2,
                           (size_t)j->wc * 2, j->hc, hipMemcpyHostToDevice, ps);
    if (e != hipSuccess) return hip_fail(e, "upload chroma");
  }
  // the host planes may be reused as soon as this returns
  e = hipStreamSynchronize(ps);
  if (e != hipSuccess) return hip_fail(e, "upload synchronize");
  return VVCLF_OK;
}

int vvclf_job_run(vvclf_job* j, void* stream) {
  if (!j) return fail(VVCLF_ERR_INVALID, "job_run: NULL job");
  hipSetDevice(j->ctx->device);
  hipError_t e = launch_plain(j->d, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "launch");
  j->fused_out = false;
  return VVCLF_OK;
}

int vvclf_job_graph_launch(vvclf_job* j, void* stream) {
  if (!j) return fail(VVCLF_ERR_INVALID, "job_graph_launch: NULL job");
  hipSetDevice(j->ctx->device);
  hipError_t e;
  if (!j->exec) {
    // capture on a private stream so the caller's stream may be the legacy one
    hipStream_t cs;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "graph stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = launch_plain(j->d, cs);
      e = hipStreamEndCapture(cs, &j->graph);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&j->exec, j->graph, nullptr, nullptr, 0);
    hipStreamDestroy(cs);
    if (e != hipSuccess) return hip_fail(e, "graph capture");
  }
  e = hipGraphLaunch(j->exec, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  j->fused_out = false;
  return VVCLF_OK;
}

struct vvclf_jobs_graph {
  vvclf_ctx* ctx;
  std::vector<vvclf_job*> jobs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int vvclf_jobs_graph_create(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n, uint32_t lanes,
                            vvclf_jobs_graph** out) {
  clear_error();
  if (!ctx || !out || (n && !jobs)) return fail(VVCLF_ERR_INVALID, "jobs_graph_create: NULL argument");
  if (lanes < 1 || lanes > 16) return fail(VVCLF_ERR_INVALID, "jobs_graph_create: %u lanes (1..16)", lanes);
  for (uint32_t i = 0; i < n; i++) {
    if (!jobs[i]) return fail(VVCLF_ERR_INVALID, "jobs_graph_create: NULL job %u", i);
    if (jobs[i]->ctx->device != ctx->device) return fail(VVCLF_ERR_INVALID, "jobs_graph_create: job %u on another device", i);
  }
  hipSetDevice(ctx->device);
  lanes = std::max(1u, std::min(lanes, std::max(n, 1u)));
  std::vector<hipStream_t> cs(lanes, nullptr);
  std::vector<hipEvent_t> ev(lanes, nullptr);
  hipError_t e = hipSuccess;
  for (uint32_t k = 0; k < lanes && e == hipSuccess; k++) {
    e = hipStreamCreateWithFlags(&cs[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
  }
  auto g = new vvclf_jobs_graph{ctx, std::vector<vvclf_job*>(jobs, jobs + n)};
  if (e == hipSuccess) {
    // capture on lane 0; the other lanes join through an event and are
    // joined back before the capture ends
    e = hipStreamBeginCapture(cs[0], hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t le = hipEventRecord(ev[0], cs[0]);
      for (uint32_t k = 1; k < lanes && le == hipSuccess; k++) le = hipStreamWaitEvent(cs[k], ev[0], 0);
      for (uint32_t i = 0; i < n && le == hipSuccess; i++) le = launch_plain(jobs[i]->d, cs[i % lanes]);
      for (uint32_t k = 1; k < lanes && le == hipSuccess; k++) {
        le = hipEventRecord(ev[k], cs[k]);
        if (le == hipSuccess) le = hipStreamWaitEvent(cs[0], ev[k], 0);
      }
      e = hipStreamEndCapture(cs[0], &g->graph);
      if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  }
  for (uint32_t k = 0; k < lanes; k++) {
    if (ev[k]) hipEventDestroy(ev[k]);
    if (cs[k]) hipStreamDestroy(cs[k]);
  }
  if (e != hipSuccess) {
    vvclf_jobs_graph_destroy(g);
    return hip_fail(e, "jobs graph capture");
  }
  *out = g;
  return VVCLF_OK;
}

int vvclf_jobs_graph_launch(vvclf_jobs_graph* g, void* stream) {
  if (!g) return fail(VVCLF_ERR_INVALID, "jobs_graph_launch: NULL graph");
  hipSetDevice(g->ctx->device);
  hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  for (vvclf_job* j : g->jobs) j->fused_out = false;
  return VVCLF_OK;
}

int vvclf_jobs_graph_destroy(vvclf_jobs_graph* g) {
  if (!g) return VVCLF_OK;
  hipSetDevice(g->ctx->device);
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return VVCLF_OK;
}

int vvclf_job_download(vvclf_job* j, vvclf_picture* out) {
  if (!j || !out) return fail(VVCLF_ERR_INVALID, "job_download: NULL argument");
  hipSetDevice(j->ctx->device);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "synchronize");
  if (out->luma.width != j->w || out->luma.height != j->h || !out->luma.samples)
    return fail(VVCLF_ERR_INVALID, "job_download: luma geometry differs");
  e = copy_rows_sync(out->luma.samples, (size_t)out->luma.stride * 2, job_final(j, 0), (size_t)j->d.py * 2,
                  (size_t)j->w * 2, j->h, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "download luma");
  if (j->wc) {
    e = copy_rows_sync(out->cb.samples, (size_t)out->cb.stride * 2, job_final(j, 1), (size_t)j->d.pc * 2,
                    (size_t)j->wc * 2, j->hc, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = copy_rows_sync(out->cr.samples, (size_t)out->cr.stride * 2, job_final(j, 2), (size_t)j->d.pc * 2,
                      (size_t)j->wc * 2, j->hc, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "download chroma");
  }
  return VVCLF_OK;
}

int vvclf_job_output(vvclf_job* j, uint16_t** y, uint16_t** cb, uint16_t** cr, uint32_t* py, uint32_t* pc) {
  if (!j) return fail(VVCLF_ERR_INVALID, "job_output: NULL job");
  if (y) *y = (uint16_t*)job_final(j, 0);
  if (cb) *cb = (uint16_t*)job_final(j, 1);
  if (cr) *cr = (uint16_t*)job_final(j, 2);
  if (py) *py = (uint32_t)j->d.py;
  if (pc) *pc = (uint32_t)j->d.pc;
  return VVCLF_OK;
}

int vvclf_prof_create(vvclf_ctx* ctx, uint32_t max_runs, vvclf_prof** out) {
  if (!ctx || !out || max_runs == 0) return fail(VVCLF_ERR_INVALID, "prof_create: bad argument");
  hipSetDevice(ctx->device);
  vvclf_prof* p = new vvclf_prof();
  p->device = ctx->device;
  p->cap = max_runs;
  p->used = 0;
  p->ev.resize((size_t)max_runs * (PLAIN_STAGES + 1));
  for (auto& e : p->ev) {
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      delete p;
      return hip_fail(r, "hipEventCreate");
    }
  }
  *out = p;
  return VVCLF_OK;
}

int vvclf_prof_destroy(vvclf_prof* p) {
  if (!p) return VVCLF_OK;
  hipSetDevice(p->device);
  for (auto& e : p->ev) hipEventDestroy(e);
  delete p;
  return VVCLF_OK;
}

int vvclf_prof_reset(vvclf_prof* p) {
  if (!p) return fail(VVCLF_ERR_INVALID, "prof_reset: NULL");
  p->used = 0;
  return VVCLF_OK;
}

int vvclf_job_run_prof(vvclf_job* j, void* stream, vvclf_prof* p) {
  if (!j || !p) return fail(VVCLF_ERR_INVALID, "job_run_prof: NULL argument");
  if (p->used >= p->cap) return fail(VVCLF_ERR_INVALID, "job_run_prof: event pool full");
  hipSetDevice(j->ctx->device);
  hipError_t e = launch_plain(j->d, (hipStream_t)stream, &p->ev[(size_t)p->used * (PLAIN_STAGES + 1)]);
  if (e != hipSuccess) return hip_fail(e, "launch");
  j->fused_out = false;
  p->used++;
  return VVCLF_OK;
}

int vvclf_prof_stage_ms(vvclf_prof* p, float ms[8], uint32_t* runs) {
  if (!p || !ms) return fail(VVCLF_ERR_INVALID, "prof_stage_ms: NULL argument");
  hipSetDevice(p->device);
  for (int s = 0; s < 8; s++) ms[s] = 0.f;
  for (uint32_t r = 0; r < p->used; r++) {
    hipEvent_t* ev = &p->ev[(size_t)r * (PLAIN_STAGES + 1)];
    hipError_t e = hipEventSynchronize(ev[PLAIN_STAGES]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    for (int s = 0; s < PLAIN_STAGES; s++) {
      float t = 0.f;
      e = hipEventElapsedTime(&t, ev[s], ev[s + 1]);
      if (e != hipSuccess) return hip_fail(e, "hipEventElapsedTime");
      ms[s] += t;
    }
  }
  if (runs) *runs = p->used;
  return VVCLF_OK;
}

uint64_t vvclf_job_device_bytes(const vvclf_job* j) { return j ? j->bytes_planes + j->bytes_side : 0; }
uint64_t vvclf_job_luma_pixels(const vvclf_job* j) { return j ? (uint64_t)j->w * j->h : 0; }

// ---------------------------------------------------------------------------
// Drop-in synchronous entry points: upload, run the requested stages, copy the
// result back into the caller's planes (in place, like `&mut PictureBuffer`).
// ---------------------------------------------------------------------------
// The drop-in calls: one picture, host planes in and out.  Samples move
// through a pinned buffer kept per calling thread (sized to the largest
// picture seen, held until the process ends); device memory comes from the
// context stream's pool.
static int run_sync(vvclf_frame* f, uint32_t stages) {
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  static thread_local pinned_buf stage, side;
  job_opts o;
  o.stream = ctx->stream;
  o.stage = &stage;
  o.side_stage = &side;
  o.host_threads = host_walk_threads();
  vvclf_job* j = nullptr;
  int rc = job_create_stages(ctx, f, stages, &j, nullptr, &o);
  if (rc) return rc;
  // Nothing to do: leave the planes untouched.
  if (!j->d.do_lmcs && !j->d.do_dbk && !j->d.do_sao && !j->d.do_alf) {
    vvclf_job_destroy(j);
    return VVCLF_OK;
  }
  rc = vvclf_job_run(j, ctx->stream);
  if (!rc) {
    const hipError_t e = download_staged(j, stage, ctx->stream, f->pic);
    if (e != hipSuccess) rc = hip_fail(e, "download");
  }
  vvclf_job_destroy(j);
  return rc;
}

static void frame_zero(vvclf_frame& f, vvclf_picture* pic, uint32_t cfi) {
  memset(&f, 0, sizeof(f));
  f.pic = *pic;
  f.chroma_format_idc = cfi;
  f.deblock.disabled = 1;
}

int vvclf_apply_deblocking(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                           const vvclf_deblock_params* params, uint32_t cfi) {
  return vvclf_apply_deblocking_clipped(out, cus, n_cus, nullptr, 0, nullptr, params, cfi, nullptr, 0, nullptr, 0);
}

int vvclf_apply_deblocking_clipped(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                                   const vvclf_deblock_cu* chroma_cus, uint32_t n_chroma_cus,
                                   const vvclf_mv_grid* mv_grid, const vvclf_deblock_params* params, uint32_t cfi,
                                   const uint32_t* nfc, uint32_t n_nfc, const uint32_t* nfr, uint32_t n_nfr) {
  clear_error();
  if (!out || !params) return fail(VVCLF_ERR_INVALID, "apply_deblocking: NULL argument");
  if (params->disabled) return VVCLF_OK;
  vvclf_frame f;
  frame_zero(f, out, cfi);
  f.cus = cus;
  f.n_cus = n_cus;
  f.has_chroma_cus = chroma_cus != nullptr;
  f.chroma_cus = chroma_cus;
  f.n_chroma_cus = n_chroma_cus;
  f.mv_grid = mv_grid;
  f.deblock = *params;
  f.no_filter_cols = nfc;
  f.n_no_filter_cols = n_nfc;
  f.no_filter_rows = nfr;
  f.n_no_filter_rows = n_nfr;
  int rc = run_sync(&f, STAGE_DBK);
  return rc;
}

int vvclf_apply_sao(vvclf_picture* out, const vvclf_sao_picture* sp, const vvclf_sao_config* cfg) {
  return vvclf_apply_sao_clipped(out, sp, cfg, nullptr);
}

int vvclf_apply_sao_clipped(vvclf_picture* out, const vvclf_sao_picture* sp, const vvclf_sao_config* cfg,
                            const vvclf_bounds* tb) {
  clear_error();
  if (!out || !sp || !cfg) return fail(VVCLF_ERR_INVALID, "apply_sao: NULL argument");
  vvclf_frame f;
  frame_zero(f, out, cfg->chroma_format_idc);
  f.sao = *sp;
  f.sao_cfg = *cfg;
  f.clip_bounds = tb;
  return run_sync(&f, STAGE_SAO);
}

int vvclf_apply_alf(vvclf_picture* out, const vvclf_alf_picture* ap, const vvclf_alf_config* cfg,
                    const vvclf_alf_aps_binding* b) {
  return vvclf_apply_alf_clipped(out, ap, cfg, b, nullptr);
}

int vvclf_apply_alf_clipped(vvclf_picture* out, const vvclf_alf_picture* ap, const vvclf_alf_config* cfg,
                            const vvclf_alf_aps_binding* b, const vvclf_bounds* tb) {
  clear_error();
  if (!out || !ap || !cfg) return fail(VVCLF_ERR_INVALID, "apply_alf: NULL argument");
  vvclf_frame f;
  frame_zero(f, out, cfg->chroma_format_idc);
  f.alf = *ap;
  f.alf_cfg = *cfg;
  if (b) f.alf_binding = *b;
  f.clip_bounds = tb;
  return run_sync(&f, STAGE_ALF);
}

int vvclf_lmcs_inverse_luma(vvclf_plane* luma, const vvclf_lmcs_binding* b) {
  clear_error();
  if (!luma || !b) return fail(VVCLF_ERR_INVALID, "lmcs_inverse_luma: NULL argument");
  if (!b->derived) return VVCLF_OK;
  vvclf_picture pic;
  memset(&pic, 0, sizeof(pic));
  pic.luma = *luma;
  vvclf_frame f;
  frame_zero(f, &pic, 0);
  f.lmcs = *b;
  return run_sync(&f, STAGE_LMCS);
}

int vvclf_apply_in_loop_filters(vvclf_frame* frame) {
  clear_error();
  if (!frame) return fail(VVCLF_ERR_INVALID, "apply_in_loop_filters: NULL frame");
  return run_sync(frame, STAGE_ALL);
}

// derive_lmcs, src/lmcs.rs:302-404 (host-side table derivation).
int vvclf_derive_lmcs(const vvclf_lmcs_data* d, uint32_t bd, vvclf_lmcs_derived* o) {
  clear_error();
  if (!d || !o) return fail(VVCLF_ERR_INVALID, "derive_lmcs: NULL argument");
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "LMCS: BitDepth out of range (8..=16, got %u)", bd);
  if (d->lmcs_delta_max_bin_idx > 15 || d->lmcs_min_bin_idx > 15)
    return fail(VVCLF_ERR_INVALID, "LMCS: bin index out of range");
  vvclf_lmcs_derived r;
  memset(&r, 0, sizeof(r));
  const uint32_t org = (1u << bd) / 16, lg = bd - 4;
  const int64_t lo = org >> 3, hi = (int64_t)(org << 3) - 1;
  const uint32_t mn = d->lmcs_min_bin_idx, mx = 15u - d->lmcs_delta_max_bin_idx;
  for (uint32_t i = mn; i <= mx; i++) {
    int64_t cw = (int64_t)org + (d->lmcs_delta_sign_cw_flag[i] ? -1 : 1) * (int64_t)d->lmcs_delta_abs_cw[i];
    if (cw < lo || cw > hi) return fail(VVCLF_ERR_INVALID, "LMCS: lmcsCW[%u] = %lld outside band", i, (long long)cw);
    r.lmcs_cw[i] = (uint32_t)cw;
  }
  uint64_t sum = 0;
  for (uint32_t v : r.lmcs_cw) sum += v;
  if (sum > (1u << bd) - 1) return fail(VVCLF_ERR_INVALID, "LMCS: sum of lmcsCW (%llu) exceeds budget", (unsigned long long)sum);
  for (uint32_t i = 0; i < 16; i++) {
    r.input_pivot[i] = i * org;
    r.lmcs_pivot[i + 1] = r.lmcs_pivot[i] + r.lmcs_cw[i];
    r.scale_coeff[i] = (r.lmcs_cw[i] * (1u << 11) + (1u << (lg - 1))) >> lg;
    r.inv_scale_coeff[i] = r.lmcs_cw[i] ? org * (1u << 11) / r.lmcs_cw[i] : 0;
  }
  const uint32_t sh = bd - 5;
  for (uint32_t i = mn; i <= mx; i++)
    if (r.lmcs_pivot[i] % (1u << sh) != 0 && (r.lmcs_pivot[i] >> sh) == (r.lmcs_pivot[i + 1] >> sh))
      return fail(VVCLF_ERR_INVALID, "LMCS: LmcsPivot[%u] bin-crossing clause violated", i);
  const int32_t crs = (d->lmcs_delta_sign_crs_flag ? -1 : 1) * (int32_t)d->lmcs_delta_abs_crs;
  for (uint32_t i = 0; i < 16; i++) {
    if (r.lmcs_cw[i] == 0) {
      r.chroma_scale_coeff[i] = 1u << 11;
    } else {
      int64_t joint = (int64_t)r.lmcs_cw[i] + crs;
      if (joint < lo || joint > hi) return fail(VVCLF_ERR_INVALID, "LMCS: lmcsCW[%u] + lmcsDeltaCrs outside band", i);
      r.chroma_scale_coeff[i] = org * (1u << 11) / (uint32_t)joint;
    }
  }
  r.bit_depth = bd;
  r.org_cw = org;
  r.lmcs_delta_crs = crs;
  *o = r;
  return VVCLF_OK;
}

// Batches and autotuning are implemented in vvclf_batch.cpp.

}  // extern "C"
