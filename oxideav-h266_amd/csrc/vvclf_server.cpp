// vvclf_server.cpp — serving: continuous batching of filter jobs from many
// streams on one GPU.
//
//   submit() -> pending (FIFO)
//   prep threads: admit while slots and HBM budget allow; derive tables,
//                 allocate from the stream-ordered pool, upload -> ready
//   launch thread: up to max_batch ready jobs -> one fused launch on the
//                 compute stream (two launches may be in flight)
//   retire thread: wait for a launch, copy each result into its caller's
//                 planes on the download stream, free the job, release budget
//
// Jobs that arrive while others are on the GPU are admitted as soon as a slot
// and budget free up and join the next launch; uploads, launches and
// downloads of different jobs overlap on separate HIP streams.  Each
// stream keeps its LMCS inverse table and ALF filter banks resident on the
// device and re-uploads them only when their content changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

namespace {

using clk = std::chrono::steady_clock;

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const uint8_t* b = (const uint8_t*)p;
  for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

// Content keys of the tables a stream keeps resident.
uint64_t lmcs_key(const vvclf_frame& f) {
  uint64_t h = 0xcbf29ce484222325ull;
  if (!f.lmcs.derived) return 0;
  h = fnv(h, f.lmcs.derived, sizeof(vvclf_lmcs_derived));
  h = fnv(h, &f.lmcs.min_bin_idx, 2);
  return h | 1;
}

uint64_t alf_key(const vvclf_frame& f) {
  uint64_t h = 0xcbf29ce484222325ull;
  const vvclf_alf_aps_binding& b = f.alf_binding;
  h = fnv(h, &f.alf_cfg.bit_depth, 4);
  h = fnv(h, &f.alf_cfg.chroma_format_idc, 4);
  h = fnv(h, &b.n_luma_apses, 4);
  for (uint32_t i = 0; i < b.n_luma_apses; i++) {
    const vvclf_alf_aps* a = b.luma_apses[i];
    uint8_t present = a != nullptr;
    h = fnv(h, &present, 1);
    if (a) h = fnv(h, a, sizeof(*a));
  }
  const vvclf_alf_aps* o[3] = {b.chroma_aps, b.cc_cb_aps, b.cc_cr_aps};
  for (auto* a : o) {
    uint8_t present = a != nullptr;
    h = fnv(h, &present, 1);
    if (a) h = fnv(h, a, sizeof(*a));
  }
  return h | 1;
}

struct res_obj {
  uint64_t lkey = 0, akey = 0;
  void* dev = nullptr;
  hipStream_t free_stream = nullptr;
  resident_tables t;
  ~res_obj() {
    if (dev) (void)hipFreeAsync(dev, free_stream);
  }
};

struct request {
  uint64_t ticket = 0;
  uint32_t stream = 0;
  vvclf_frame* frame = nullptr;  // the caller's frame: its planes receive the result
  // submitted as per-CTB syntax records: expanded on a prep thread into the
  // request's own parameter arrays, which `expanded` (a copy of the caller's
  // frame) points at; jobs are built from `expanded` then
  const vvclf_ctb_syntax* recs = nullptr;
  vvclf_ctb_syntax_config syn_cfg{};
  std::vector<vvclf_sao_ctb_params> sao_own;
  std::vector<vvclf_alf_ctb> alf_own;
  vvclf_frame expanded{};
  const vvclf_frame* build() const { return recs ? &expanded : frame; }
  clk::time_point t_submit, t_done;
  clk::time_point t_prep, t_ready, t_launch, t_device;  // stage boundaries
  int status = VVCLF_ERR_AGAIN;  // AGAIN while pending
  bool done = false;
  vvclf_job* job = nullptr;
  uint64_t bytes = 0;
  std::shared_ptr<res_obj> res;
};

struct srv_stream {
  bool open = true;
  std::mutex mu;  // serialises resident-table updates of this stream
  std::shared_ptr<res_obj> cur;
};

struct launch_slot {
  dev_job* d_jobs = nullptr;
  tile_ref* d_tiles = nullptr;
  dev_job* h_jobs = nullptr;  // pinned staging
  tile_ref* h_tiles = nullptr;
  uint32_t cap_jobs = 0, cap_tiles = 0;
  hipEvent_t done = nullptr;
};

struct inflight {
  std::vector<request*> reqs;
  int slot = 0;
};

// Latency distribution in log-spaced buckets (0.5 % wide, 1 us .. 100 s):
// fixed size however many jobs a server finishes, and cheap to copy.
struct lat_hist {
  static constexpr double kMinMs = 1e-3;
  static constexpr double kRatio = 1.005;
  static constexpr int kBuckets = 3700;
  uint32_t counts[kBuckets] = {};
  uint64_t n = 0;
  double sum = 0, max = 0;
  static int bucket(double ms) {
    if (!(ms > kMinMs)) return 0;
    int b = (int)(std::log(ms / kMinMs) / std::log(kRatio));
    return b < 0 ? 0 : (b >= kBuckets ? kBuckets - 1 : b);
  }
  void add(double ms) {
    counts[bucket(ms)]++;
    n++;
    sum += ms;
    max = std::max(max, ms);
  }
  // the q-quantile (nearest rank), as its bucket's geometric centre
  double quantile(double q) const {
    if (!n) return 0;
    uint64_t rank = (uint64_t)(q * (double)(n - 1) + 0.5) + 1, acc = 0;
    for (int b = 0; b < kBuckets; b++) {
      acc += counts[b];
      if (acc >= rank) return std::min(max, kMinMs * std::pow(kRatio, b + 0.5));
    }
    return max;
  }
};

}  // namespace

struct vvclf_server {
  vvclf_ctx* ctx = nullptr;
  vvclf_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_prep, cv_launch, cv_retire, cv_done;
  std::deque<request*> pending;
  std::deque<request*> ready;
  std::deque<inflight> running;
  // Requests by ticket.  A ticket is consumed by the poll or wait that
  // reports its final status; finished tickets nobody claims are dropped
  // oldest first beyond kMaxUnclaimed.
  std::unordered_map<uint64_t, std::shared_ptr<request>> tickets;
  std::deque<uint64_t> finished_order;
  static constexpr size_t kMaxUnclaimed = 1u << 16;
  std::vector<std::unique_ptr<srv_stream>> streams;
  uint64_t next_ticket = 1;
  uint64_t bytes_admitted = 0, bytes_peak = 0;
  uint32_t admitted = 0;  // admitted, not yet retired
  uint32_t inject = 0;
  bool stop = false;
  std::vector<std::thread> threads;
  std::vector<hipStream_t> prep_streams;
  // pinned staging: one buffer per prep thread (uploads), two for the
  // retire thread (downloads, double-buffered)
  std::vector<pinned_buf> prep_stage, prep_side;
  pinned_buf down_stage[2];
  hipEvent_t down_done[2] = {nullptr, nullptr};
  hipStream_t s_comp = nullptr, s_down = nullptr;
  launch_slot slots[2];
  // stats
  uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, pixels = 0, launches = 0;
  lat_hist lat;
  lat_hist stage[5];
  clk::time_point first_submit, last_done;
  bool any_submit = false;
};

namespace {

// Device bytes a frame's job will hold (sample sets + a bound on the tables).
uint64_t job_bytes_bound(const vvclf_frame& f) {
  uint64_t W = f.pic.luma.width, H = f.pic.luma.height;
  uint64_t py = (W + 63) / 64 * 64, wc = f.chroma_format_idc ? f.pic.cb.width : 0, hc = f.chroma_format_idc ? f.pic.cb.height : 0;
  uint64_t pc = (wc + 63) / 64 * 64;
  uint64_t planes = 4 * ((py * H + 2 * pc * hc) * 2 + 256);
  // deblock maps and their lists of nonzero entries (at most as long)
  uint64_t side = 2 * (((W / 4 + 2) * (H / 4 + 2) * 8) + ((wc / 2 + 1) * (hc / 2 + 1) * 16)) +
                  (W / 32 + 2) * (H / 32 + 2) * 64 + (1 << 20);
  return planes + side;
}

void finish(vvclf_server* s, request* r, int status) {
  // caller holds s->mu
  r->status = status;
  r->done = true;
  r->t_done = clk::now();
  s->last_done = r->t_done;
  if (status == VVCLF_OK) {
    s->completed++;
    s->pixels += (uint64_t)r->frame->pic.luma.width * r->frame->pic.luma.height;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    s->lat.add(ms(r->t_submit, r->t_done));
    s->stage[0].add(ms(r->t_submit, r->t_prep));
    s->stage[1].add(ms(r->t_prep, r->t_ready));
    s->stage[2].add(ms(r->t_ready, r->t_launch));
    s->stage[3].add(ms(r->t_launch, r->t_device));
    s->stage[4].add(ms(r->t_device, r->t_done));
  } else {
    s->failed++;
  }
  // bound the finished tickets nobody has claimed (consumed ones are skipped)
  s->finished_order.push_back(r->ticket);
  while (s->finished_order.size() > vvclf_server::kMaxUnclaimed) {
    s->tickets.erase(s->finished_order.front());
    s->finished_order.pop_front();
  }
  s->cv_done.notify_all();
}

// Drops a finished ticket whose final status has just been reported (caller
// holds s->mu).
void consume(vvclf_server* s, uint64_t ticket) { s->tickets.erase(ticket); }

int ensure_resident(vvclf_server* s, srv_stream& st, const vvclf_frame& f, hipStream_t ps,
                    std::shared_ptr<res_obj>& out) {
  std::lock_guard<std::mutex> g(st.mu);
  uint64_t lk = lmcs_key(f), ak = alf_key(f);
  if (st.cur && st.cur->lkey == lk && st.cur->akey == ak) {
    out = st.cur;
    return VVCLF_OK;
  }
  host_tables t;
  if (f.lmcs.derived) build_lmcs_lut(f.lmcs, t.lmcs_lut);
  if (f.alf_cfg.alf_enabled) build_alf_banks(f.alf_cfg, f.alf_binding, t);
  auto r = std::make_shared<res_obj>();
  r->lkey = lk;
  r->akey = ak;
  r->free_stream = ps;
  size_t n_lut = t.lmcs_lut.size() * 2, n_bank = t.alf_bank.size() * 4, n_cb = t.alf_cbank.size() * 4,
         n_cc = t.alf_ccbank.size() * 4;
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  size_t total = al(n_lut) + al(n_bank) + al(n_cb) + al(n_cc) + 256;
  hipError_t e = hipMallocAsync(&r->dev, total, ps);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "resident tables: %s", hipGetErrorString(e));
  uint8_t* base = (uint8_t*)r->dev;
  size_t off = 0;
  auto put = [&](const void* src, size_t n) -> void* {
    void* dst = base + off;
    if (n) e = e == hipSuccess ? hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ps) : e;
    off += al(n);
    return dst;
  };
  r->t.lut = (const uint16_t*)put(t.lmcs_lut.data(), n_lut);
  r->t.bank = (const uint32_t*)put(t.alf_bank.data(), n_bank);
  r->t.cbank = (const uint32_t*)put(t.alf_cbank.data(), n_cb);
  r->t.ccbank = (const int32_t*)put(t.alf_ccbank.data(), n_cc);
  if (e == hipSuccess) e = hipStreamSynchronize(ps);
  if (e != hipSuccess) return hip_fail(e, "resident tables upload");
  st.cur = r;
  out = r;
  return VVCLF_OK;
}

// A request submitted as syntax records: its SAO / ALF CTB parameters.
int expand_syntax(request& r) {
  const vvclf_frame& f = *r.frame;
  uint32_t w = f.sao.pic_width_in_ctbs_y, h = f.sao.pic_height_in_ctbs_y;
  if (!w || !h) {
    w = f.alf.pic_width_in_ctbs_y;
    h = f.alf.pic_height_in_ctbs_y;
  }
  const bool sao_grid = f.sao.pic_width_in_ctbs_y != 0, alf_grid = f.alf.pic_width_in_ctbs_y != 0;
  if ((sao_grid && alf_grid) && (f.alf.pic_width_in_ctbs_y != w || f.alf.pic_height_in_ctbs_y != h))
    return fail(VVCLF_ERR_INVALID, "syntax records: SAO and ALF CTB grids differ");
  r.sao_own.assign((size_t)w * h, vvclf_sao_ctb_params{});
  r.alf_own.assign((size_t)w * h, vvclf_alf_ctb{});
  if (int rc = vvclf_ctb_syntax_decode(r.recs, w, h, &r.syn_cfg, r.sao_own.data(), r.alf_own.data())) return rc;
  r.expanded = f;
  if (sao_grid) r.expanded.sao.ctbs = r.sao_own.data();
  if (alf_grid) r.expanded.alf.ctbs = r.alf_own.data();
  return VVCLF_OK;
}

void prep_loop(vvclf_server* s, int idx) {
  hipSetDevice(s->ctx->device);
  hipStream_t ps = s->prep_streams[idx];
  for (;;) {
    request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stop) return;
        if (!s->pending.empty() && s->admitted < s->cfg.max_inflight) {
          request* c = s->pending.front();
          uint64_t need = job_bytes_bound(*c->frame);
          if (need > s->cfg.hbm_budget_bytes) {  // can never fit
            s->pending.pop_front();
            s->rejected++;
            finish(s, c, fail(VVCLF_ERR_BUDGET, "job of %llu bytes exceeds the %llu-byte budget",
                              (unsigned long long)need, (unsigned long long)s->cfg.hbm_budget_bytes));
            continue;
          }
          if (s->bytes_admitted + need <= s->cfg.hbm_budget_bytes) {
            s->pending.pop_front();
            c->bytes = need;
            s->bytes_admitted += need;
            s->bytes_peak = std::max(s->bytes_peak, s->bytes_admitted);
            s->admitted++;
            r = c;
            r->t_prep = clk::now();
            break;
          }
        }
        s->cv_prep.wait(lk);
      }
    }
    int rc = VVCLF_OK;
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (s->inject) {
        s->inject--;
        rc = fail(VVCLF_ERR_HIP, "injected failure");
      }
    }
    srv_stream* st = nullptr;
    if (!rc) {
      std::lock_guard<std::mutex> g(s->mu);
      st = s->streams[r->stream].get();
    }
    if (!rc && r->recs) rc = expand_syntax(*r);
    if (!rc) rc = ensure_resident(s, *st, *r->build(), ps, r->res);
    if (!rc) {
      job_opts o;
      o.stream = ps;
      o.res = &r->res->t;
      o.stage = &s->prep_stage[idx];
      o.side_stage = &s->prep_side[idx];
      o.host_threads = host_walk_threads();  // large pictures walk their CUs on several threads
      rc = job_create_stages(s->ctx, r->build(), STAGE_ALL, &r->job, nullptr, &o);
    }
    if (!rc && !job_fits_shape(r->job->d, 32, 32)) {  // 32x32 fits every job any shape fits
      vvclf_job_destroy(r->job);
      r->job = nullptr;
      rc = fail(VVCLF_ERR_UNSUPPORTED, "served picture: ALF or SAO CTBs smaller than 32 samples");
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (rc) {
      s->bytes_admitted -= r->bytes;
      s->admitted--;
      r->res.reset();
      finish(s, r, rc);
      s->cv_prep.notify_all();
      continue;
    }
    r->t_ready = clk::now();
    s->ready.push_back(r);
    s->cv_launch.notify_one();
  }
}

bool grow_slot(launch_slot& sl, uint32_t nj, uint32_t nt) {
  if (nj > sl.cap_jobs) {
    if (sl.d_jobs) hipFree(sl.d_jobs);
    if (sl.h_jobs) hipHostFree(sl.h_jobs);
    sl.cap_jobs = std::max(nj, 2 * sl.cap_jobs);
    if (hipMalloc(&sl.d_jobs, sl.cap_jobs * sizeof(dev_job)) != hipSuccess) return false;
    if (hipHostMalloc(&sl.h_jobs, sl.cap_jobs * sizeof(dev_job), 0) != hipSuccess) return false;
  }
  if (nt > sl.cap_tiles) {
    if (sl.d_tiles) hipFree(sl.d_tiles);
    if (sl.h_tiles) hipHostFree(sl.h_tiles);
    sl.cap_tiles = std::max(nt, 2 * sl.cap_tiles);
    if (hipMalloc(&sl.d_tiles, sl.cap_tiles * sizeof(tile_ref)) != hipSuccess) return false;
    if (hipHostMalloc(&sl.h_tiles, sl.cap_tiles * sizeof(tile_ref), 0) != hipSuccess) return false;
  }
  return true;
}

void launch_loop(vvclf_server* s) {
  hipSetDevice(s->ctx->device);
  int cus = 256;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, s->ctx->device) == hipSuccess) cus = p.multiProcessorCount;
  const int cfg_tw = (int)s->cfg.tune.tile_w, cfg_th = (int)s->cfg.tune.tile_h;
  // workgroups per CU of each (shape, >12-bit variant), 0 = not yet asked
  uint32_t wpc_of[4][2] = {};
  for (;;) {
    inflight fl;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_launch.wait(lk, [&] { return s->stop || (!s->ready.empty() && s->running.size() < 2); });
      if (s->stop) return;
      uint32_t n = std::min<uint32_t>((uint32_t)s->ready.size(), s->cfg.max_batch);
      for (uint32_t i = 0; i < n; i++) {
        fl.reqs.push_back(s->ready.front());
        s->ready.pop_front();
      }
      // at most one launch is in flight here: take the slot it does not use
      fl.slot = s->running.empty() ? 0 : (s->running.back().slot ^ 1);
    }
    // tile list, largest pictures first
    std::vector<uint32_t> order(fl.reqs.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return (uint64_t)fl.reqs[a]->job->w * fl.reqs[a]->job->h > (uint64_t)fl.reqs[b]->job->w * fl.reqs[b]->job->h;
    });
    // the configured tile shape, or the largest one every picture of the
    // launch fits (32x32 fits all admitted pictures)
    auto all_fit = [&](int a, int b) {
      for (auto* r : fl.reqs)
        if (!job_fits_shape(r->job->d, (uint32_t)a, (uint32_t)b)) return false;
      return true;
    };
    int tw = cfg_tw, th = cfg_th;
    if (!all_fit(tw, th)) {
      for (auto& sh : kFusedShapes)
        if (all_fit(sh[0], sh[1])) {
          tw = sh[0];
          th = sh[1];
          break;
        }
    }
    int si = 3;
    for (int k = 0; k < 4; k++)
      if (kFusedShapes[k][0] == tw && kFusedShapes[k][1] == th) si = k;
    uint32_t nt = 0;
    bool hbd = false;
    for (auto* r : fl.reqs) {
      nt += ((r->job->w + tw - 1) / tw) * ((r->job->h + th - 1) / th);
      hbd = hbd || job_needs_hbd(r->job->d);
    }
    uint32_t& wpc = wpc_of[si][hbd ? 1 : 0];
    if (wpc == 0) {
      const int occ = std::max(1, fused_max_blocks_per_cu(tw, th, hbd));
      wpc = s->cfg.tune.wgs_per_cu ? std::min<uint32_t>(s->cfg.tune.wgs_per_cu, occ) : (uint32_t)occ;
    }
    launch_slot& sl = s->slots[fl.slot];
    hipError_t e = grow_slot(sl, (uint32_t)fl.reqs.size(), nt) ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess) {
      uint32_t k = 0;
      for (uint32_t i = 0; i < fl.reqs.size(); i++) {
        sl.h_jobs[i] = fl.reqs[i]->job->d;
        fl.reqs[i]->job->fused_out = true;
        fl.reqs[i]->t_launch = clk::now();
      }
      for (uint32_t oi : order) {
        const vvclf_job* j = fl.reqs[oi]->job;
        for (uint32_t y = 0; y < j->h; y += th)
          for (uint32_t x = 0; x < j->w; x += tw) sl.h_tiles[k++] = make_tile(j, oi, x, y);
      }
      e = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, fl.reqs.size() * sizeof(dev_job), hipMemcpyHostToDevice, s->s_comp);
      if (e == hipSuccess)
        e = hipMemcpyAsync(sl.d_tiles, sl.h_tiles, nt * sizeof(tile_ref), hipMemcpyHostToDevice, s->s_comp);
      uint32_t grid = std::min((uint32_t)cus * wpc, (nt + 7) & ~7u);
      if (e == hipSuccess) e = launch_fused(tw, th, hbd, sl.d_jobs, sl.d_tiles, nt, std::max(8u, grid), s->s_comp);
      if (e == hipSuccess) e = hipEventRecord(sl.done, s->s_comp);
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (e != hipSuccess) {
      int rc = hip_fail(e, "fused launch");
      for (auto* r : fl.reqs) {
        vvclf_job_destroy(r->job);
        r->job = nullptr;
        r->res.reset();
        s->bytes_admitted -= r->bytes;
        s->admitted--;
        finish(s, r, rc);
      }
      s->cv_prep.notify_all();
      continue;
    }
    s->launches++;
    s->running.push_back(std::move(fl));
    s->cv_retire.notify_one();
  }
}

void retire_loop(vvclf_server* s) {
  hipSetDevice(s->ctx->device);
  for (;;) {
    inflight fl;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_retire.wait(lk, [&] { return s->stop || !s->running.empty(); });
      if (s->stop) return;
      fl = s->running.front();
    }
    hipError_t e = hipEventSynchronize(s->slots[fl.slot].done);
    const clk::time_point t_dev = clk::now();
    {
      // the slot's tables are free once the kernel is done: the next launch
      // may go while this one's samples come back
      std::lock_guard<std::mutex> g(s->mu);
      s->running.pop_front();
      s->cv_launch.notify_all();
    }
    // each job is finished as soon as its own samples are back: the DMA of
    // job i+1 into one pinned buffer overlaps the row copies of job i out of
    // the other
    const size_t n = fl.reqs.size();
    std::vector<int> rc(n, e == hipSuccess ? VVCLF_OK : hip_fail(e, "fused kernel"));
    auto issue = [&](size_t i) {
      if (rc[i] != VVCLF_OK) return;
      hipError_t de = download_staged_issue(fl.reqs[i]->job, s->down_stage[i & 1], s->s_down);
      if (de == hipSuccess) de = hipEventRecord(s->down_done[i & 1], s->s_down);
      if (de != hipSuccess) rc[i] = hip_fail(de, "download");
    };
    auto complete = [&](size_t i) {
      request* r = fl.reqs[i];
      r->t_device = t_dev;
      if (rc[i] == VVCLF_OK) {
        hipError_t de = hipEventSynchronize(s->down_done[i & 1]);
        if (de == hipSuccess) download_staged_finish(r->job, s->down_stage[i & 1], r->frame->pic);
        else rc[i] = hip_fail(de, "download");
      }
      vvclf_job_destroy(r->job);
      std::lock_guard<std::mutex> g(s->mu);
      r->job = nullptr;
      r->res.reset();
      s->bytes_admitted -= r->bytes;
      s->admitted--;
      finish(s, r, rc[i]);
      s->cv_prep.notify_all();
    };
    if (n) issue(0);
    for (size_t i = 0; i < n; i++) {
      if (i + 1 < n) issue(i + 1);
      complete(i);
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->cv_launch.notify_all();
  }
}

}  // namespace

extern "C" {

int vvclf_server_create(vvclf_ctx* ctx, const vvclf_server_config* cfg, vvclf_server** out) {
  clear_error();
  if (!ctx || !out) return fail(VVCLF_ERR_INVALID, "server_create: NULL argument");
  vvclf_server* s = new vvclf_server();
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 64ull << 30;
  if (!s->cfg.max_inflight) s->cfg.max_inflight = 64;
  if (!s->cfg.max_batch) s->cfg.max_batch = 32;
  if (!s->cfg.prep_threads) s->cfg.prep_threads = 4;
  if (!s->cfg.tune.tile_w) s->cfg.tune = vvclf_tune{64, 64, 256, 0};
  bool shape_ok = false;
  for (auto& sh : kFusedShapes) shape_ok = shape_ok || (sh[0] == (int)s->cfg.tune.tile_w && sh[1] == (int)s->cfg.tune.tile_h);
  if (!shape_ok) {
    const uint32_t tw = s->cfg.tune.tile_w, th = s->cfg.tune.tile_h;
    delete s;
    return fail(VVCLF_ERR_INVALID, "server_create: tile %ux%u is not a fused tile shape", tw, th);
  }
  hipSetDevice(ctx->device);
  // keep freed job memory in the device pool for reuse
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  hipError_t e = hipStreamCreateWithFlags(&s->s_comp, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->s_down, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&s->slots[i].done, hipEventDisableTiming);
  for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&s->down_done[i], hipEventDisableTiming);
  s->prep_streams.resize(s->cfg.prep_threads);
  s->prep_stage.resize(s->cfg.prep_threads);
  s->prep_side.resize(s->cfg.prep_threads);
  for (auto& ps : s->prep_streams)
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ps, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return hip_fail(e, "server streams");
  }
  for (uint32_t i = 0; i < s->cfg.prep_threads; i++) s->threads.emplace_back(prep_loop, s, (int)i);
  s->threads.emplace_back(launch_loop, s);
  s->threads.emplace_back(retire_loop, s);
  *out = s;
  return VVCLF_OK;
}

int vvclf_server_destroy(vvclf_server* s) {
  if (!s) return VVCLF_OK;
  vvclf_server_drain(s, 60000);
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_prep.notify_all();
  s->cv_launch.notify_all();
  s->cv_retire.notify_all();
  for (auto& t : s->threads) t.join();
  hipSetDevice(s->ctx->device);
  for (auto& st : s->streams) st->cur.reset();
  for (auto& sl : s->slots) {
    if (sl.d_jobs) hipFree(sl.d_jobs);
    if (sl.d_tiles) hipFree(sl.d_tiles);
    if (sl.h_jobs) hipHostFree(sl.h_jobs);
    if (sl.h_tiles) hipHostFree(sl.h_tiles);
    if (sl.done) hipEventDestroy(sl.done);
  }
  hipDeviceSynchronize();
  for (auto& b : s->prep_stage) b.release();
  for (auto& b : s->prep_side) b.release();
  for (int k = 0; k < 2; k++) {
    s->down_stage[k].release();
    if (s->down_done[k]) hipEventDestroy(s->down_done[k]);
  }
  for (auto ps : s->prep_streams) hipStreamDestroy(ps);
  hipStreamDestroy(s->s_comp);
  hipStreamDestroy(s->s_down);
  s->tickets.clear();
  delete s;
  return VVCLF_OK;
}

int vvclf_server_open_stream(vvclf_server* s, uint32_t* id) {
  if (!s || !id) return fail(VVCLF_ERR_INVALID, "open_stream: NULL");
  std::lock_guard<std::mutex> g(s->mu);
  s->streams.emplace_back(new srv_stream());
  *id = (uint32_t)s->streams.size() - 1;
  return VVCLF_OK;
}

int vvclf_server_close_stream(vvclf_server* s, uint32_t id) {
  if (!s) return fail(VVCLF_ERR_INVALID, "close_stream: NULL");
  std::lock_guard<std::mutex> g(s->mu);
  if (id >= s->streams.size() || !s->streams[id]->open) return fail(VVCLF_ERR_INVALID, "close_stream: bad id");
  s->streams[id]->open = false;
  return VVCLF_OK;
}

static int submit(vvclf_server* s, uint32_t id, vvclf_frame* frame, const vvclf_ctb_syntax* recs,
                  const vvclf_ctb_syntax_config* cfg, uint64_t* ticket) {
  if (!s || !frame || !ticket || (recs && !cfg)) return fail(VVCLF_ERR_INVALID, "submit: NULL argument");
  auto sp = std::make_shared<request>();
  request* r = sp.get();
  r->frame = frame;
  r->recs = recs;
  if (cfg) r->syn_cfg = *cfg;
  r->stream = id;
  r->t_submit = clk::now();
  std::lock_guard<std::mutex> g(s->mu);
  if (id >= s->streams.size() || !s->streams[id]->open)
    return fail(VVCLF_ERR_INVALID, "submit: stream %u is not open", id);
  r->ticket = s->next_ticket++;
  s->tickets[r->ticket] = std::move(sp);
  if (!s->any_submit) {
    s->first_submit = r->t_submit;
    s->any_submit = true;
  }
  s->submitted++;
  s->pending.push_back(r);
  *ticket = r->ticket;
  s->cv_prep.notify_one();
  return VVCLF_OK;
}

int vvclf_server_submit(vvclf_server* s, uint32_t id, vvclf_frame* frame, uint64_t* ticket) {
  return submit(s, id, frame, nullptr, nullptr, ticket);
}

int vvclf_server_submit_syntax(vvclf_server* s, uint32_t id, vvclf_frame* frame, const vvclf_ctb_syntax* recs,
                               const vvclf_ctb_syntax_config* cfg, uint64_t* ticket) {
  if (!recs) return fail(VVCLF_ERR_INVALID, "submit_syntax: NULL records");
  return submit(s, id, frame, recs, cfg, ticket);
}

int vvclf_server_poll(vvclf_server* s, uint64_t t) {
  if (!s) return fail(VVCLF_ERR_INVALID, "poll: NULL");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->tickets.find(t);
  if (it == s->tickets.end()) return fail(VVCLF_ERR_INVALID, "poll: unknown or already reported ticket");
  if (!it->second->done) return VVCLF_ERR_AGAIN;
  int st = it->second->status;
  consume(s, t);
  return st;
}

int vvclf_server_wait(vvclf_server* s, uint64_t t, uint32_t timeout_ms) {
  if (!s) return fail(VVCLF_ERR_INVALID, "wait: NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(t);
  if (it == s->tickets.end()) return fail(VVCLF_ERR_INVALID, "wait: unknown or already reported ticket");
  std::shared_ptr<request> r = it->second;  // alive even if another caller consumes the ticket meanwhile
  if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; }))
    return VVCLF_ERR_TIMEOUT;
  consume(s, t);
  return r->status;
}

int vvclf_server_drain(vvclf_server* s, uint32_t timeout_ms) {
  if (!s) return fail(VVCLF_ERR_INVALID, "drain: NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  bool ok = s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] {
    return s->pending.empty() && s->ready.empty() && s->running.empty() && s->admitted == 0;
  });
  return ok ? VVCLF_OK : VVCLF_ERR_TIMEOUT;
}

int vvclf_server_get_stats(vvclf_server* s, vvclf_server_stats* o) {
  if (!s || !o) return fail(VVCLF_ERR_INVALID, "stats: NULL");
  // copy the fixed-size histograms under the lock, reduce them outside it
  auto h = std::make_unique<lat_hist[]>(6);
  memset(o, 0, sizeof(*o));
  {
    std::lock_guard<std::mutex> g(s->mu);
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->failed = s->failed;
    o->rejected = s->rejected;
    o->luma_pixels = s->pixels;
    o->launches = s->launches;
    o->bytes_admitted_peak = s->bytes_peak;
    if (s->any_submit && s->completed)
      o->wall_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
    h[0] = s->lat;
    for (int k = 0; k < 5; k++) h[k + 1] = s->stage[k];
  }
  if (h[0].n) {
    o->p50_ms = h[0].quantile(0.50);
    o->p99_ms = h[0].quantile(0.99);
    o->max_ms = h[0].max;
    o->mean_ms = h[0].sum / (double)h[0].n;
  }
  for (int k = 0; k < 5; k++) {
    o->stage_p50_ms[k] = h[k + 1].quantile(0.50);
    o->stage_p99_ms[k] = h[k + 1].quantile(0.99);
  }
  return VVCLF_OK;
}

int vvclf_server_reset_stats(vvclf_server* s) {
  if (!s) return fail(VVCLF_ERR_INVALID, "reset_stats: NULL");
  std::lock_guard<std::mutex> g(s->mu);
  s->submitted = s->completed = s->failed = s->rejected = s->pixels = s->launches = 0;
  s->lat = lat_hist();
  for (auto& h : s->stage) h = lat_hist();
  s->bytes_peak = s->bytes_admitted;
  s->any_submit = false;  // the wall clock restarts at the next submission
  return VVCLF_OK;
}

int vvclf_server_inject_failures(vvclf_server* s, uint32_t n) {
  if (!s) return fail(VVCLF_ERR_INVALID, "inject: NULL");
  std::lock_guard<std::mutex> g(s->mu);
  s->inject += n;
  return VVCLF_OK;
}

}  // extern "C"
