// vvclf_host.cpp — the C ABI of include/vvclf.h: descriptor validation, the
// host-side derivation of the compact per-edge / per-CTB tables the kernels
// read, job lifetime, and the synchronous drop-in entry points.
#include <hip/hip_runtime.h>
#include <sched.h>
#include <sys/mman.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"
#include "vvclf_internal.h"
#include "vvclf_walk.h"

namespace vvclf {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(VVCLF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Deblocking edge derivation (host).  Walks the CU records exactly like
// deblock_one_direction / deblock_cu_dir (src/deblock.rs:687-983) and stores,
// per segment, what the sample filter needs: bS (src/deblock.rs:1055-1119),
// maxFilterLength P/Q (:845-901 luma, :925-938 chroma), the CTB-row flag
// (:905-906), palette write skips, and the beta / tC table indices
// (:1122-1174).
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kMapThreshold = (size_t)1 << 20;
}  // namespace

namespace {
struct block_pool {
  std::mutex mu;
  std::vector<std::pair<void*, size_t>> free_blocks;  // (address, mapped size)
  std::vector<std::pair<void*, size_t>> out;          // blocks handed out larger than asked
  size_t held = 0;
};
constexpr size_t kPoolCap = (size_t)512 << 20;
block_pool& blocks() {
  static block_pool* p = new block_pool();  // lives with the process
  return *p;
}
}  // namespace

thread_local bool t_recycled = false;
bool pooled_alloc_recycled() { return t_recycled; }

void* pooled_alloc_bytes(size_t bytes) {
  t_recycled = false;
  if (bytes < kMapThreshold) return std::calloc(bytes ? bytes : 1, 1);
  block_pool& bp = blocks();
  {
    std::lock_guard<std::mutex> g(bp.mu);
    size_t best = bp.free_blocks.size();
    for (size_t i = 0; i < bp.free_blocks.size(); i++) {
      const size_t sz = bp.free_blocks[i].second;
      if (sz >= bytes && sz - bytes <= bytes / 4 && (best == bp.free_blocks.size() || sz < bp.free_blocks[best].second))
        best = i;
    }
    if (best < bp.free_blocks.size()) {
      const auto blk = bp.free_blocks[best];
      bp.free_blocks.erase(bp.free_blocks.begin() + best);
      bp.held -= blk.second;
      if (blk.second != bytes) bp.out.push_back(blk);
      t_recycled = true;
      return blk.first;
    }
  }
  void* p = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_POPULATE, -1, 0);
  return p == MAP_FAILED ? nullptr : p;
}

void pooled_free_bytes(void* p, size_t bytes) {
  if (!p) return;
  if (bytes < kMapThreshold) {
    std::free(p);
    return;
  }
  block_pool& bp = blocks();
  size_t size = bytes;
  {
    std::lock_guard<std::mutex> g(bp.mu);
    for (size_t i = 0; i < bp.out.size(); i++)
      if (bp.out[i].first == p) {
        size = bp.out[i].second;
        bp.out.erase(bp.out.begin() + i);
        break;
      }
    if (bp.held + size <= kPoolCap) {
      bp.free_blocks.push_back({p, size});
      bp.held += size;
      return;
    }
  }
  munmap(p, size);
}

namespace {

// The worker pool of parallel_for.  A call's job lives on the caller's
// stack: workers take a reference under the mutex before touching it, and
// the caller returns only once the job has left the queue, every index has
// run and no worker holds it.
struct pf_job {
  const std::function<void(int)>* fn;
  int n;
  std::atomic<int> next{0};
  int done = 0;  // indices finished (under the pool mutex)
  int refs = 0;  // workers holding the job (under the pool mutex)
};

struct host_pool {
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<pf_job*> q;
  explicit host_pool(int workers) {
    for (int i = 0; i < workers; i++) std::thread([this] { loop(); }).detach();
  }
  void loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      work_cv.wait(lk, [&] { return !q.empty(); });
      pf_job* j = q.front();
      j->refs++;
      lk.unlock();
      int ran = 0;
      for (int i; (i = j->next.fetch_add(1)) < j->n; ran++) (*j->fn)(i);
      lk.lock();
      if (!q.empty() && q.front() == j) q.pop_front();  // indices exhausted
      j->done += ran;
      j->refs--;
      done_cv.notify_all();
    }
  }
};

// CPUs this process may run on (its affinity mask, which a container or a
// scheduler share narrows below what the machine has).
int host_cpus() {
  static const int n = [] {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0 && CPU_COUNT(&set) > 0) return CPU_COUNT(&set);
    return (int)std::max(1u, std::thread::hardware_concurrency());
  }();
  return n;
}

// Up to 15 workers: enough for the callers' parallel steps to proceed side by
// side (several serving threads each walking a large picture on up to
// host_walk_threads() threads) on a 16-CPU share.
host_pool& the_pool() {
  static host_pool* p = new host_pool(std::max(0, std::min(host_cpus(), 16) - 1));  // workers live with the process
  return *p;
}

template <class F>
void run_threads(int nt, F&& fn) {
  parallel_for(nt, std::function<void(int)>(fn));
}

struct cu_grid {
  std::vector<int32_t, pooled_alloc<int32_t>> cells;  // every cell written by build
  uint32_t w = 0, h = 0;
  // CuGrid::build, src/deblock.rs:637-660: later records overwrite earlier
  // ones.  Row bands on nt threads, each visiting the records in order, so
  // that every cell still ends with the last record covering it.
  void build(uint32_t pic_w, uint32_t pic_h, const vvclf_deblock_cu* cus, uint32_t n, int nt) {
    w = (pic_w + 3) / 4;
    h = (pic_h + 3) / 4;
    cells = std::vector<int32_t, pooled_alloc<int32_t>>();
    cells.resize((size_t)w * h);
    run_threads(nt, [&](int ti) {
      const uint32_t yb = (uint32_t)((uint64_t)h * ti / nt), ye = (uint32_t)((uint64_t)h * (ti + 1) / nt);
      std::fill(cells.data() + (size_t)yb * w, cells.data() + (size_t)ye * w, -1);
      for (uint32_t i = 0; i < n; i++) {
        const vvclf_deblock_cu& c = cus[i];
        const uint32_t y0 = std::max(c.y / 4, yb), y1 = std::min(std::min((c.y + c.h + 3) / 4, h), ye);
        if (y0 >= y1) continue;
        const uint32_t x0 = c.x / 4, x1 = std::min((c.x + c.w + 3) / 4, w);
        for (uint32_t cy = y0; cy < y1; cy++)
          for (uint32_t cx = x0; cx < x1; cx++) cells[(size_t)cy * w + cx] = (int32_t)i;
      }
    });
  }
  walk::grid_view view() const { return walk::grid_view{cells.data(), w, h}; }
};

// One direction's descriptors of one plane.  The map holds zeros before the
// walk, which writes only descriptors that filter and records its claims in a bitmap of one bit per slot (a few hundred KB
// that stay in cache, where the map is megabytes).
struct edge_map {
  map_vec e;
  std::vector<uint64_t> claims;
  uint32_t w = 0, h = 0;
  bool dirty = false;  // recycled storage: derive_deblock_maps clears it
  void init(uint32_t w_, uint32_t h_) {
    w = w_;
    h = h_;
    e = map_vec();
    e.resize((size_t)w * h);
    dirty = pooled_alloc_recycled();
    claims.assign(((size_t)w * h + 63) / 64, 0);
  }
  walk::map_view view() { return walk::map_view{e.data(), w, h}; }
};

// Rows [y0, y1) of a map: lists the nonzero descriptors as
// tag | row << 16 | column (row-major), and for a luma map
// checks that consecutive same-direction edges of a line cannot reach each
// other's used samples (luma_dir 1: along a row, vertical edges; 2: down a
// column, horizontal edges, whose first and last entry per column are kept
// for the check across band borders).
struct finish_band {
  uint32_t y0 = 0, y1 = 0;
  std::vector<uint32_t> list;
  bool safe = true;
  std::vector<int32_t> fy, ly;
  std::vector<uint32_t> fe, le;
};

void finish_rows(uint32_t* e, uint32_t w, uint32_t tag, int luma_dir, finish_band& b) {
  b.list.clear();
  if (luma_dir == 2) {
    b.fy.assign(w, -1);
    b.ly.assign(w, -1);
    b.fe.assign(w, 0);
    b.le.assign(w, 0);
  }
  for (uint32_t y = b.y0; y < b.y1; y++) {
    const uint32_t* row = e + (size_t)y * w;
    int32_t px = -1;
    uint32_t pe = 0;
    for (uint32_t x = 0; x < w; x++) {
      // most slots hold no descriptor: skip four zero slots at a time
      while (x + 4 <= w && (row[x] | row[x + 1] | row[x + 2] | row[x + 3]) == 0) x += 4;
      if (x >= w) break;
      const uint32_t v = row[x];
      if (!v) continue;
      b.list.push_back(tag | (y << 16) | x);
      if (luma_dir == 1) {
        if (px >= 0 && !walk::luma_pair_safe(pe, v, 4 * (x - (uint32_t)px), true)) b.safe = false;
        px = (int32_t)x;
        pe = v;
      } else if (luma_dir == 2) {
        if (b.ly[x] >= 0) {
          if (!walk::luma_pair_safe(b.le[x], v, 4 * (y - (uint32_t)b.ly[x]), false)) b.safe = false;
        } else {
          b.fy[x] = (int32_t)y;
          b.fe[x] = v;
        }
        b.ly[x] = (int32_t)y;
        b.le[x] = v;
      }
    }
  }
}

// Joins the bands of one map: lists in row order, and the horizontal-edge
// check across band borders (each column's last entry above a band against
// its first entry in it).
bool join_bands(std::vector<finish_band>& bands, uint32_t w, int luma_dir, std::vector<uint32_t>& out) {
  bool safe = true;
  size_t n = 0;
  for (const auto& b : bands) {
    safe &= b.safe;
    n += b.list.size();
  }
  out.clear();
  out.reserve(n);
  for (const auto& b : bands) out.insert(out.end(), b.list.begin(), b.list.end());
  if (luma_dir == 2 && bands.size() > 1) {
    for (uint32_t x = 0; x < w; x++) {
      int32_t cy = -1;
      uint32_t ce = 0;
      for (const auto& b : bands) {
        if (b.fy[x] < 0) continue;
        if (cy >= 0 && !walk::luma_pair_safe(ce, b.fe[x], 4 * (uint32_t)(b.fy[x] - cy), false)) safe = false;
        cy = b.ly[x];
        ce = b.le[x];
      }
    }
  }
  return safe;
}

}  // namespace

int host_walk_threads() {
  static const int n = std::min(8, host_cpus());
  return n;
}

void parallel_for(int n, const std::function<void(int)>& fn) {
  if (n <= 0) return;
  if (n == 1) {
    fn(0);
    return;
  }
  host_pool& p = the_pool();
  pf_job j;
  j.fn = &fn;
  j.n = n;
  {
    std::lock_guard<std::mutex> g(p.mu);
    p.q.push_back(&j);
  }
  p.work_cv.notify_all();
  int ran = 0;
  for (int i; (i = j.next.fetch_add(1)) < n; ran++) fn(i);
  std::unique_lock<std::mutex> lk(p.mu);
  const auto it = std::find(p.q.begin(), p.q.end(), &j);
  if (it != p.q.end()) p.q.erase(it);  // no worker takes it from here on
  j.done += ran;
  p.done_cv.wait(lk, [&] { return j.done == n && j.refs == 0; });
}

int derive_deblock_maps(const vvclf_frame& f, uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, host_tables& t,
                        int threads) {
  const vvclf_deblock_params& prm = f.deblock;
  uint32_t ctb = 1u << prm.ctb_log2_size_y;
  // pictures of at least 4096 CU records use up to 16 threads
  const int nt = f.n_cus >= 4096 ? std::max(1, std::min(threads, 16)) : 1;
  cu_grid grid;
  grid.build(W, H, f.cus, f.n_cus, nt);
  for (uint32_t i = 0; i < f.n_cus; i++) {
    const vvclf_deblock_cu& c = f.cus[i];
    if ((c.x | c.y | c.w | c.h) & 3) return fail(VVCLF_ERR_UNSUPPORTED, "deblock: CU %u not on the 4x4 grid", i);
  }
  // the six passes of src/deblock.rs:519-604 (luma V/H, Cb V/H, Cr V/H), each
  // into its own map
  struct pass {
    walk::plane_params pp;
    const vvclf_deblock_cu* cus;
    uint32_t n;
    const cu_grid* grid;
    const vvclf_mv_grid* mv;
    bool vert;
    edge_map* m;
  };
  std::vector<pass> passes;
  edge_map lv, lh, cv[2], ch[2];
  t.gw = (W + 3) / 4 + 2;
  t.gh = (H + 3) / 4 + 2;
  struct map_dims {
    edge_map* m;
    uint32_t w, h;
  };
  map_dims dims[6] = {{&lv, t.gw, t.gh}, {&lh, t.gw, t.gh}};
  int n_maps = 2;
  const walk::plane_params pl = {0, 1, 1, prm.luma_beta_offset_div2, prm.luma_tc_offset_div2, 0, prm.bit_depth, ctb};
  passes.push_back({pl, f.cus, f.n_cus, &grid, f.mv_grid, true, &lv});
  passes.push_back({pl, f.cus, f.n_cus, &grid, f.mv_grid, false, &lh});
  t.has_chroma_dbk = false;
  cu_grid cgrid;
  if (f.chroma_format_idc != 0 && Wc > 0) {
    const vvclf_deblock_cu* ccus = f.has_chroma_cus ? f.chroma_cus : f.cus;
    uint32_t nc = f.has_chroma_cus ? f.n_chroma_cus : f.n_cus;
    const cu_grid* cg = &grid;
    if (f.has_chroma_cus) {
      cgrid.build(W, H, ccus, nc, nt);
      cg = &cgrid;
      for (uint32_t i = 0; i < nc; i++) {
        const vvclf_deblock_cu& c = ccus[i];
        if ((c.x | c.y | c.w | c.h) & 3) return fail(VVCLF_ERR_UNSUPPORTED, "deblock: chroma CU %u off grid", i);
      }
    }
    t.cvw = (Wc + 7) / 8 + 1;
    t.cvh = (Hc + 1) / 2 + 1;
    t.chw = (Wc + 1) / 2 + 1;
    t.chh = (Hc + 7) / 8 + 1;
    for (int comp = 1; comp <= 2; comp++) {
      const walk::plane_params pc = {(uint32_t)comp,
                               2,
                               2,
                               comp == 1 ? prm.cb_beta_offset_div2 : prm.cr_beta_offset_div2,
                               comp == 1 ? prm.cb_tc_offset_div2 : prm.cr_tc_offset_div2,
                               comp == 1 ? prm.chroma_qp_offset_cb : prm.chroma_qp_offset_cr,
                               prm.bit_depth,
                               ctb};
      dims[n_maps++] = {&cv[comp - 1], t.cvw, t.cvh};
      dims[n_maps++] = {&ch[comp - 1], t.chw, t.chh};
      passes.push_back({pc, ccus, nc, cg, nullptr, true, &cv[comp - 1]});
      passes.push_back({pc, ccus, nc, cg, nullptr, false, &ch[comp - 1]});
    }
    t.has_chroma_dbk = true;
  }
  // the maps' storage, one map per thread: new storage holds zeros (pages
  // populated by the kernel); recycled storage is then cleared in equal
  // shares of the recycled maps' concatenation, one per thread
  if (nt > 1) {
    std::atomic<int> next_map{0};
    run_threads(std::min(nt, n_maps), [&](int) {
      for (int k; (k = next_map.fetch_add(1)) < n_maps;) dims[k].m->init(dims[k].w, dims[k].h);
    });
  } else {
    for (int k = 0; k < n_maps; k++) dims[k].m->init(dims[k].w, dims[k].h);
  }
  size_t dirty = 0;
  for (int k = 0; k < n_maps; k++)
    if (dims[k].m->dirty) dirty += dims[k].m->e.size();
  if (dirty)
    run_threads(nt, [&](int ti) {
      const size_t b = dirty * ti / nt, e = dirty * (ti + 1) / nt;
      size_t base = 0;
      for (int k = 0; k < n_maps; k++) {
        if (!dims[k].m->dirty) continue;
        const size_t n = dims[k].m->e.size(), lo = std::max(b, base), hi = std::min(e, base + n);
        if (lo < hi) memset(dims[k].m->e.data() + (lo - base), 0, (hi - lo) * 4);
        base += n;
      }
    });
  // Every thread walks its share of the records through all passes: a
  // record is read once for every pass over the same record array (luma and,
  // without a separate chroma tree, chroma), as the passes write separate
  // maps.  Per-thread histograms sit on cache lines of their own.
  struct alignas(64) thread_hist {
    uint64_t v[4] = {0, 0, 0, 0};
  };
  std::vector<thread_hist> hists((size_t)nt);
  std::atomic<bool> bad{false};
  int32_t bad_xy[2] = {0, 0};
  std::mutex bad_mu;
  auto work = [&](int ti) {
    int32_t xy[2];
    uint64_t* hist = hists[ti].v;
    size_t g0 = 0;
    while (g0 < passes.size()) {  // groups of consecutive passes over one record array
      size_t g1 = g0 + 1;
      while (g1 < passes.size() && passes[g1].cus == passes[g0].cus && passes[g1].n == passes[g0].n) g1++;
      const uint32_t n = passes[g0].n;
      const uint32_t b = (uint32_t)((uint64_t)n * ti / nt), e = (uint32_t)((uint64_t)n * (ti + 1) / nt);
      for (uint32_t i = b; i < e; i++) {
        for (size_t k = g0; k < g1; k++) {
          const pass& ps = passes[k];
          const uint32_t* nf = ps.vert ? f.no_filter_cols : f.no_filter_rows;
          const uint32_t n_nf = ps.vert ? f.n_no_filter_cols : f.n_no_filter_rows;
          if (!walk::walk_cu(ps.pp, ps.cus, ps.grid->view(), ps.mv, i, ps.vert, nf, n_nf, W, H, ps.m->view(), hist,
                             xy, ps.m->claims.data())) {
            std::lock_guard<std::mutex> g(bad_mu);
            if (!bad.exchange(true)) {
              bad_xy[0] = xy[0];
              bad_xy[1] = xy[1];
            }
            return;
          }
        }
        if (bad.load(std::memory_order_relaxed)) return;
      }
      g0 = g1;
    }
  };
  run_threads(nt, work);
  if (bad)
    return fail(VVCLF_ERR_INVALID, "deblock: CU records overlap at (%d,%d): they must tile the picture", bad_xy[0],
                bad_xy[1]);
  uint64_t hist[4] = {0, 0, 0, 0};
  for (const auto& h : hists)
    for (int i = 0; i < 4; i++) hist[i] += h.v[i];
  // finish the maps in row bands (nt per map, all bands on nt threads): work
  // lists, and the check that lets a direction pass run concurrently in
  // place
  std::vector<uint32_t> lists[6];
  bool ok[6] = {true, true, true, true, true, true};
  {
    struct map_job {
      edge_map* m;
      uint32_t tag;
      int dir;
      std::vector<finish_band> bands;
    };
    std::vector<map_job> maps;
    maps.push_back({&lv, 0, 1, {}});
    maps.push_back({&lh, 0, 2, {}});
    if (t.has_chroma_dbk)
      for (int c = 0; c < 2; c++) {
        maps.push_back({&cv[c], (uint32_t)c << 31, 0, {}});
        maps.push_back({&ch[c], (uint32_t)c << 31, 0, {}});
      }
    std::vector<std::pair<int, int>> tasks;
    for (int k = 0; k < (int)maps.size(); k++) {
      const uint32_t h = maps[k].m->h;
      const int nb = (int)std::max(1u, std::min((uint32_t)nt, h / 16));
      maps[k].bands.resize(nb);
      for (int bi = 0; bi < nb; bi++) {
        maps[k].bands[bi].y0 = (uint32_t)((uint64_t)h * bi / nb);
        maps[k].bands[bi].y1 = (uint32_t)((uint64_t)h * (bi + 1) / nb);
        tasks.push_back({k, bi});
      }
    }
    std::atomic<size_t> next{0};
    run_threads(nt, [&](int) {
      for (size_t i; (i = next.fetch_add(1)) < tasks.size();) {
        map_job& mj = maps[tasks[i].first];
        finish_rows(mj.m->e.data(), mj.m->w, mj.tag, mj.dir, mj.bands[tasks[i].second]);
      }
    });
    // lists in the order the kernels expect: luma V, luma H, Cb then Cr V,
    // Cb then Cr H
    const int slot[6] = {0, 1, 2, 4, 3, 5};
    for (int k = 0; k < (int)maps.size(); k++)
      ok[slot[k]] = join_bands(maps[k].bands, maps[k].m->w, maps[k].dir, lists[slot[k]]);
  }
  if (!ok[0] || !ok[1])
    return fail(VVCLF_ERR_UNSUPPORTED,
                "deblock: same-direction luma edges interact (sub-block edge spacing); the concurrent pass "
                "would not reproduce the sequential result");
  t.e_lv = std::move(lv.e);
  t.e_lh = std::move(lh.e);
  t.l_lv = std::move(lists[0]);
  t.l_lh = std::move(lists[1]);
  t.l_cv.clear();
  t.l_ch.clear();
  if (t.has_chroma_dbk) {
    for (int c = 0; c < 2; c++) {
      t.e_cv[c] = std::move(cv[c].e);
      t.e_ch[c] = std::move(ch[c].e);
    }
    t.l_cv = std::move(lists[2]);
    t.l_cv.insert(t.l_cv.end(), lists[3].begin(), lists[3].end());
    t.l_ch = std::move(lists[4]);
    t.l_ch.insert(t.l_ch.end(), lists[5].begin(), lists[5].end());
  }
  for (int i = 0; i < 3; i++) t.bs_hist[i] = hist[i];
  t.chroma_filtered = hist[3];
  return VVCLF_OK;
}

// ---------------------------------------------------------------------------
// SAO / ALF tables.
// ---------------------------------------------------------------------------
namespace {
// The `span` closure of apply_sao_clipped (src/sao.rs:430-437) and
// clip_rect_for_ctb (src/alf.rs:1082-1089).
void span(const uint32_t* b, uint32_t n, uint32_t v, int32_t mx, int32_t& lo, int32_t& hi) {
  for (uint32_t i = 0; i + 1 < n; i++)
    if (v >= b[i] && v < b[i + 1]) {
      lo = (int32_t)b[i];
      hi = std::min((int32_t)b[i + 1], mx);
      return;
    }
  lo = 0;
  hi = mx;
}

uint64_t ctb_rect(const vvclf_bounds* tb, uint32_t rx, uint32_t ry, uint32_t ctb, uint32_t W, uint32_t H) {
  if (!tb || !tb->cols) return pack_rect(0, 0, W, H);
  int32_t x0, x1, y0, y1;
  span(tb->cols, tb->n_cols, rx * ctb, (int32_t)W, x0, x1);
  span(tb->rows, tb->n_rows, ry * ctb, (int32_t)H, y0, y1);
  // An empty rectangle would make the reference's clamp panic; the caller
  // rejects it (see rect_ok).
  if (x1 <= x0 || y1 <= y0) return 0;
  return pack_rect((uint32_t)x0, (uint32_t)y0, (uint32_t)x1, (uint32_t)y1);
}

// AlfFixFiltCoeff (eq. 90) and AlfClassToFiltMap (eq. 91): H.266 spec data.
const int8_t kFixCoeff[64][12] = {
    {0, 0, 2, -3, 1, -4, 1, 7, -1, 1, -1, 5},       {0, 0, 0, 0, 0, -1, 0, 1, 0, 0, -1, 2},
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0},           {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1},
    {2, 2, -7, -3, 0, -5, 13, 22, 12, -3, -3, 17},  {-1, 0, 6, -8, 1, -5, 1, 23, 0, 2, -5, 10},
    {0, 0, -1, -1, 0, -1, 2, 1, 0, 0, -1, 4},       {0, 0, 3, -11, 1, 0, -1, 35, 5, 2, -9, 9},
    {0, 0, 8, -8, -2, -7, 4, 4, 2, 1, -1, 25},      {0, 0, 1, -1, 0, -3, 1, 3, -1, 1, -1, 3},
    {0, 0, 3, -3, 0, -6, 5, -1, 2, 1, -4, 21},      {-7, 1, 5, 4, -3, 5, 11, 13, 12, -8, 11, 12},
    {-5, -3, 6, -2, -3, 8, 14, 15, 2, -7, 11, 16},  {2, -1, -6, -5, -2, -2, 20, 14, -4, 0, -3, 25},
    {3, 1, -8, -4, 0, -8, 22, 5, -3, 2, -10, 29},   {2, 1, -7, -1, 2, -11, 23, -5, 0, 2, -10, 29},
    {-6, -3, 8, 9, -4, 8, 9, 7, 14, -2, 8, 9},      {2, 1, -4, -7, 0, -8, 17, 22, 1, -1, -4, 23},
    {3, 0, -5, -7, 0, -7, 15, 18, -5, 0, -5, 27},   {2, 0, 0, -7, 1, -10, 13, 13, -4, 2, -7, 24},
    {3, 3, -13, 4, -2, -5, 9, 21, 25, -2, -3, 12},  {-5, -2, 7, -3, -7, 9, 8, 9, 16, -2, 15, 12},
    {0, -1, 0, -7, -5, 4, 11, 11, 8, -6, 12, 21},   {3, -2, -3, -8, -4, -1, 16, 15, -2, -3, 3, 26},
    {2, 1, -5, -4, -1, -8, 16, 4, -2, 1, -7, 33},   {2, 1, -4, -2, 1, -10, 17, -2, 0, 2, -11, 33},
    {1, -2, 7, -15, -16, 10, 8, 8, 20, 11, 14, 11}, {2, 2, 3, -13, -13, 4, 8, 12, 2, -3, 16, 24},
    {1, 4, 0, -7, -8, -4, 9, 9, -2, -2, 8, 29},     {1, 1, 2, -4, -1, -6, 6, 3, -1, -1, -3, 30},
    {-7, 3, 2, 10, -2, 3, 7, 11, 19, -7, 8, 10},    {0, -2, -5, -3, -2, 4, 20, 15, -1, -3, -1, 22},
    {3, -1, -8, -4, -1, -4, 22, 8, -4, 2, -8, 28},  {0, 3, -14, 3, 0, 1, 19, 17, 8, -3, -7, 20},
    {0, 2, -1, -8, 3, -6, 5, 21, 1, 1, -9, 13},     {-4, -2, 8, 20, -2, 2, 3, 5, 21, 4, 6, 1},
    {2, -2, -3, -9, -4, 2, 14, 16, 3, -6, 8, 24},   {2, 1, 5, -16, -7, 2, 3, 11, 15, -3, 11, 22},
    {1, 2, 3, -11, -2, -5, 4, 8, 9, -3, -2, 26},    {0, -1, 10, -9, -1, -8, 2, 3, 4, 0, 0, 29},
    {1, 2, 0, -5, 1, -9, 9, 3, 0, 1, -7, 20},       {-2, 8, -6, -4, 3, -9, -8, 45, 14, 2, -13, 7},
    {1, -1, 16, -19, -8, -4, -3, 2, 19, 0, 4, 30},  {1, 1, -3, 0, 2, -11, 15, -5, 1, 2, -9, 24},
    {0, 1, -2, 0, 1, -4, 4, 0, 0, 1, -4, 7},        {0, 1, 2, -5, 1, -6, 4, 10, -2, 1, -4, 10},
    {3, 0, -3, -6, -2, -6, 14, 8, -1, -1, -3, 31},  {0, 1, 0, -2, 1, -6, 5, 1, 0, 1, -5, 13},
    {3, 1, 9, -19, -21, 9, 7, 6, 13, 5, 15, 21},    {2, 4, 3, -12, -13, 1, 7, 8, 3, 0, 12, 26},
    {3, 1, -8, -2, 0, -6, 18, 2, -2, 3, -10, 23},   {1, 1, -4, -1, 1, -5, 8, 1, -1, 2, -5, 10},
    {0, 1, -1, 0, 0, -2, 2, 0, 0, 1, -2, 3},        {1, 1, -2, -7, 1, -7, 14, 18, 0, 0, -7, 21},
    {0, 1, 0, -2, 0, -7, 8, 1, -2, 0, -3, 24},      {0, 1, 1, -2, 2, -10, 10, 0, -2, 1, -7, 23},
    {0, 2, 2, -11, 2, -4, -3, 39, 7, 1, -10, 9},    {1, 0, 13, -16, -5, -6, -1, 8, 6, 0, 6, 29},
    {1, 3, 1, -6, -4, -7, 9, 6, -3, -2, 3, 33},     {4, 0, -17, -1, -1, 5, 26, 8, -2, 3, -15, 30},
    {0, 1, -2, 0, 2, -8, 12, -6, 1, 1, -6, 16},     {0, 0, 0, -1, 1, -4, 4, 0, 0, 0, -3, 11},
    {0, 1, 2, -8, 2, -6, 5, 15, 0, 2, -7, 9},       {1, -1, 12, -15, -7, -2, 3, 6, 6, -1, 7, 30},
};
const uint8_t kClassToFilt[16][25] = {
    {8, 2, 2, 2, 3, 4, 53, 9, 9, 52, 4, 4, 5, 9, 2, 8, 10, 9, 1, 3, 39, 39, 10, 9, 52},
    {11, 12, 13, 14, 15, 30, 11, 17, 18, 19, 16, 20, 20, 4, 53, 21, 22, 23, 14, 25, 26, 26, 27, 28, 10},
    {16, 12, 31, 32, 14, 16, 30, 33, 53, 34, 35, 16, 20, 4, 7, 16, 21, 36, 18, 19, 21, 26, 37, 38, 39},
    {35, 11, 13, 14, 43, 35, 16, 4, 34, 62, 35, 35, 30, 56, 7, 35, 21, 38, 24, 40, 16, 21, 48, 57, 39},
    {11, 31, 32, 43, 44, 16, 4, 17, 34, 45, 30, 20, 20, 7, 5, 21, 22, 46, 40, 47, 26, 48, 63, 58, 10},
    {12, 13, 50, 51, 52, 11, 17, 53, 45, 9, 30, 4, 53, 19, 0, 22, 23, 25, 43, 44, 37, 27, 28, 10, 55},
    {30, 33, 62, 51, 44, 20, 41, 56, 34, 45, 20, 41, 41, 56, 5, 30, 56, 38, 40, 47, 11, 37, 42, 57, 8},
    {35, 11, 23, 32, 14, 35, 20, 4, 17, 18, 21, 20, 20, 20, 4, 16, 21, 36, 46, 25, 41, 26, 48, 49, 58},
    {12, 31, 59, 59, 3, 33, 33, 59, 59, 52, 4, 33, 17, 59, 55, 22, 36, 59, 59, 60, 22, 36, 59, 25, 55},
    {31, 25, 15, 60, 60, 22, 17, 19, 55, 55, 20, 20, 53, 19, 55, 22, 46, 25, 43, 60, 37, 28, 10, 55, 52},
    {12, 31, 32, 50, 51, 11, 33, 53, 19, 45, 16, 4, 4, 53, 5, 22, 36, 18, 25, 43, 26, 27, 27, 28, 10},
    {5, 2, 44, 52, 3, 4, 53, 45, 9, 3, 4, 56, 5, 0, 2, 5, 10, 47, 52, 3, 63, 39, 10, 9, 52},
    {12, 34, 44, 44, 3, 56, 56, 62, 45, 9, 56, 56, 7, 5, 0, 22, 38, 40, 47, 52, 48, 57, 39, 10, 9},
    {35, 11, 23, 14, 51, 35, 20, 41, 56, 62, 16, 20, 41, 56, 7, 16, 21, 38, 24, 40, 26, 26, 42, 57, 39},
    {33, 34, 51, 51, 52, 41, 41, 34, 62, 0, 41, 41, 56, 7, 5, 56, 38, 38, 40, 44, 37, 42, 57, 39, 10},
    {16, 31, 32, 15, 60, 30, 4, 17, 19, 25, 22, 20, 4, 53, 19, 21, 22, 46, 25, 55, 26, 48, 63, 58, 55},
};

// transposeIdx permutations, eqs. 1442-1445 (src/alf.rs:731-742).
const uint8_t kTranspose[4][12] = {{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
                                   {9, 4, 10, 8, 1, 5, 11, 7, 3, 0, 2, 6},
                                   {0, 3, 2, 1, 8, 7, 6, 5, 4, 9, 10, 11},
                                   {9, 8, 10, 4, 3, 7, 11, 5, 1, 0, 2, 6}};

// AlfClip (Table 8, src/alf.rs:627-644) capped at 0xffff: a difference of two
// u16 samples never exceeds that, so the cap clamps identically.
uint32_t clip_word(uint32_t bd, uint8_t clip_idx) {
  static const uint8_t EXP[9][4] = {{8, 5, 3, 1},   {9, 6, 4, 2},   {10, 7, 5, 3},   {11, 8, 6, 4}, {12, 9, 7, 5},
                                    {13, 10, 8, 6}, {14, 11, 9, 7}, {15, 12, 10, 8}, {16, 13, 11, 9}};
  uint32_t row = std::clamp(bd, 8u, 16u) - 8;
  uint32_t v = 1u << EXP[row][clip_idx & 3];
  return std::min(v, 0xffffu);
}

inline uint32_t tap_word(int32_t coeff, uint32_t clip) { return ((uint32_t)(uint16_t)(int16_t)coeff) | (clip << 16); }

void build_luma_bank(const int32_t coeff[25][12], const uint8_t clip[25][12], uint32_t bd, uint32_t* out) {
  for (int c = 0; c < 25; c++)
    for (int t = 0; t < 4; t++)
      for (int k = 0; k < 12; k++) {
        int src = kTranspose[t][k];
        out[(c * 4 + t) * 12 + k] = tap_word(coeff[c][src], clip_word(bd, clip[c][src]));
      }
}
}  // namespace

void alf_luma_bank_tables(uint32_t bd, uint8_t perm[4][12], uint32_t clipw[4]) {
  memcpy(perm, kTranspose, sizeof(kTranspose));
  for (uint8_t i = 0; i < 4; i++) clipw[i] = clip_word(bd, i);
}

uint64_t ctb_fetch_rect(const vvclf_bounds* tb, uint32_t rx, uint32_t ry, uint32_t ctb, uint32_t W, uint32_t H) {
  return ctb_rect(tb, rx, ry, ctb, W, H);
}

// The filter banks of an ALF binding (src/alf.rs:564-644): 16 fixed sets,
// one per APS slot, the chroma alternatives and the CC filters.
void build_alf_banks(const vvclf_alf_config& ac, const vvclf_alf_aps_binding& b, host_tables& t) {
  uint32_t n_banks = 16 + b.n_luma_apses;
  t.alf_bank.assign((size_t)n_banks * ALF_BANK_WORDS, 0);
  for (int s = 0; s < 16; s++) {
    int32_t coeff[25][12];
    uint8_t clip[25][12];
    for (int c = 0; c < 25; c++)
      for (int k = 0; k < 12; k++) {
        coeff[c][k] = kFixCoeff[kClassToFilt[s][c]][k];
        clip[c][k] = 0;  // fixed sets: clipIdx 0 (src/alf.rs:578-580)
      }
    build_luma_bank(coeff, clip, ac.bit_depth, &t.alf_bank[(size_t)s * ALF_BANK_WORDS]);
  }
  for (uint32_t i = 0; i < b.n_luma_apses; i++) {
    const vvclf_alf_aps* a = b.luma_apses[i];
    if (!a || !a->alf_luma_filter_signal_flag) continue;
    build_luma_bank(a->luma_coeff, a->luma_clip_idx, ac.bit_depth, &t.alf_bank[(size_t)(16 + i) * ALF_BANK_WORDS]);
  }
  t.alf_cbank.assign(ALF_CBANK_WORDS, 0);
  const vvclf_alf_aps* ca = b.chroma_aps;
  bool chroma_ok = ac.chroma_format_idc != 0 && ca && ca->alf_chroma_filter_signal_flag;
  uint32_t n_alt = chroma_ok ? std::min(ca->n_chroma_coeff, (uint32_t)VVCLF_ALF_MAX_CHROMA_ALT) : 0;
  for (uint32_t a = 0; a < n_alt; a++)
    for (int k = 0; k < 6; k++)
      t.alf_cbank[a * 6 + k] = tap_word(ca->chroma_coeff[a][k], clip_word(ac.bit_depth, ca->chroma_clip_idx[a][k]));
  t.alf_ccbank.assign(2 * ALF_CCBANK_WORDS, 0);
  const vvclf_alf_aps* ccs[2] = {b.cc_cb_aps, b.cc_cr_aps};
  for (int c = 0; c < 2; c++) {
    if (!ccs[c] || ac.chroma_format_idc == 0) continue;
    uint32_t n = std::min(c == 0 ? ccs[c]->n_cc_cb_coeff : ccs[c]->n_cc_cr_coeff, (uint32_t)VVCLF_ALF_MAX_CC_FILTERS);
    for (uint32_t k = 0; k < n; k++)
      for (int j = 0; j < 7; j++)
        t.alf_ccbank[c * ALF_CCBANK_WORDS + k * 7 + j] = c == 0 ? ccs[c]->cc_cb_coeff[k][j] : ccs[c]->cc_cr_coeff[k][j];
  }
}

int derive_sao_alf(const vvclf_frame& f, uint32_t W, uint32_t H, host_tables& t, uint32_t stages) {
  // ---- SAO: src/sao.rs:389-497
  t.do_sao = false;
  if (stages & STAGE_SAO) {
    const vvclf_sao_picture& sp = f.sao;
    const vvclf_sao_config& sc = f.sao_cfg;
    size_t n = (size_t)sp.pic_width_in_ctbs_y * sp.pic_height_in_ctbs_y;
    bool any = false;
    for (size_t i = 0; i < n; i++)
      if (sp.ctbs[i].luma.sao_type_idx || sp.ctbs[i].cb.sao_type_idx || sp.ctbs[i].cr.sao_type_idx) any = true;
    if (any && (sc.luma_used || sc.chroma_used)) {
      if (sc.bit_depth < 1 || sc.bit_depth > 16) return fail(VVCLF_ERR_INVALID, "sao: bit_depth %u", sc.bit_depth);
      if (sc.chroma_used && sc.chroma_format_idc != 0 && sc.chroma_format_idc != 1)
        return fail(VVCLF_ERR_UNSUPPORTED, "sao: chroma_format_idc %u", sc.chroma_format_idc);
      // CTBs of 16..256 samples (VVC uses 32..128): the kernels take 8-sample
      // chroma runs inside one CTB
      if (sc.ctb_log2_size_y < 4 || sc.ctb_log2_size_y > 8)
        return fail(VVCLF_ERR_UNSUPPORTED, "sao: ctb_log2_size_y %u", sc.ctb_log2_size_y);
      t.do_sao = true;
      t.sao_luma = sc.luma_used != 0;
      t.sao_chroma = sc.chroma_used != 0 && sc.chroma_format_idc != 0;
      t.swctb = sp.pic_width_in_ctbs_y;
      t.shctb = sp.pic_height_in_ctbs_y;
      t.sao_ctb_log2 = sc.ctb_log2_size_y;
      t.sao_bd = sc.bit_depth;
      t.sao.assign(n * 3, 0);
      for (size_t i = 0; i < n; i++) {
        const vvclf_sao_ctb* cs[3] = {&sp.ctbs[i].luma, &sp.ctbs[i].cb, &sp.ctbs[i].cr};
        for (int c = 0; c < 3; c++) {
          const vvclf_sao_ctb& s = *cs[c];
          uint32_t type = s.sao_type_idx == 1 ? 1u : (s.sao_type_idx == 2 ? 2u : 0u);
          uint32_t param = type == 2 ? (s.eo_class & 3u) : (s.band_position & 31u);
          for (int k = 1; k <= 4; k++)
            if (s.offset_val[k] < -8192 || s.offset_val[k] > 8191)
              return fail(VVCLF_ERR_INVALID, "sao: offset %d out of range", s.offset_val[k]);
          t.sao[i * 3 + c] = pack_sao(type, param, &s.offset_val[1]);
        }
      }
      t.sao_rect.clear();
      if (f.clip_bounds && f.clip_bounds->cols) {
        t.sao_rect.resize(n);
        uint32_t ctb = 1u << sc.ctb_log2_size_y;
        for (uint32_t ry = 0; ry < t.shctb; ry++)
          for (uint32_t rx = 0; rx < t.swctb; rx++)
          {
            uint64_t r = ctb_rect(f.clip_bounds, rx, ry, ctb, W, H);
            if (r == 0) return fail(VVCLF_ERR_INVALID, "sao: empty clip rectangle at CTB (%u,%u)", rx, ry);
            t.sao_rect[(size_t)ry * t.swctb + rx] = r;
          }
      }
    }
  }
  // ---- ALF: src/alf.rs:235-434
  t.do_alf = false;
  if (stages & STAGE_ALF) {
    const vvclf_alf_picture& ap = f.alf;
    const vvclf_alf_config& ac = f.alf_cfg;
    const vvclf_alf_aps_binding& b = f.alf_binding;
    for (uint32_t i = 0; i < b.n_luma_apses; i++) {
      const vvclf_alf_aps* a = b.luma_apses[i];
      if (a && a->alf_luma_filter_signal_flag && a->n_luma_coeff < 25)
        return fail(VVCLF_ERR_INVALID, "alf: luma APS slot %u holds %u < 25 filters", i, a->n_luma_coeff);
    }
    size_t n = (size_t)ap.pic_width_in_ctbs_y * ap.pic_height_in_ctbs_y;
    bool any = false;
    for (size_t i = 0; i < n; i++) {
      const vvclf_alf_ctb& c = ap.ctbs[i];
      if (c.luma_on || c.cb_on || c.cr_on || c.cc_cb_idc || c.cc_cr_idc) any = true;
    }
    if (ac.alf_enabled && any) {
      if (ac.bit_depth < 8 || ac.bit_depth > 16) return fail(VVCLF_ERR_INVALID, "alf: bit_depth %u", ac.bit_depth);
      if (ac.chroma_format_idc != 0 && ac.chroma_format_idc != 1)
        return fail(VVCLF_ERR_UNSUPPORTED, "alf: chroma_format_idc %u", ac.chroma_format_idc);
      if (ac.ctb_log2_size_y < 3 || ac.ctb_log2_size_y > 8)
        return fail(VVCLF_ERR_UNSUPPORTED, "alf: ctb_log2_size_y %u", ac.ctb_log2_size_y);
      t.do_alf = true;
      t.wctb = ap.pic_width_in_ctbs_y;
      t.hctb = ap.pic_height_in_ctbs_y;
      t.alf_ctb_log2 = ac.ctb_log2_size_y;
      t.alf_bd = ac.bit_depth;
      if (16 + b.n_luma_apses > 255) return fail(VVCLF_ERR_INVALID, "alf: too many luma APS slots");
      // The filters hold ALF coefficients in 16-bit lanes and multiply CC-ALF
      // coefficients and their 7-tap sum in 24 bits (so each within 21 bits);
      // conformant streams stay far inside both
      // (|c| <= 128 and <= 64), larger values are refused rather than
      // filtered differently.
      for (uint32_t i = 0; i < b.n_luma_apses; i++) {
        const vvclf_alf_aps* a = b.luma_apses[i];
        if (!a || !a->alf_luma_filter_signal_flag) continue;
        for (int c = 0; c < 25; c++)
          for (int k = 0; k < 12; k++)
            if (a->luma_coeff[c][k] < -32768 || a->luma_coeff[c][k] > 32767)
              return fail(VVCLF_ERR_UNSUPPORTED, "alf: luma coefficient %d outside 16 bits", a->luma_coeff[c][k]);
      }
      if (const vvclf_alf_aps* ca = b.chroma_aps; ca && ca->alf_chroma_filter_signal_flag)
        for (uint32_t a = 0; a < std::min(ca->n_chroma_coeff, (uint32_t)VVCLF_ALF_MAX_CHROMA_ALT); a++)
          for (int k = 0; k < 6; k++)
            if (ca->chroma_coeff[a][k] < -32768 || ca->chroma_coeff[a][k] > 32767)
              return fail(VVCLF_ERR_UNSUPPORTED, "alf: chroma coefficient %d outside 16 bits", ca->chroma_coeff[a][k]);
      for (int c = 0; c < 2; c++) {
        const vvclf_alf_aps* cc = c == 0 ? b.cc_cb_aps : b.cc_cr_aps;
        if (!cc) continue;
        const uint32_t nf = std::min(c == 0 ? cc->n_cc_cb_coeff : cc->n_cc_cr_coeff, (uint32_t)VVCLF_ALF_MAX_CC_FILTERS);
        for (uint32_t k = 0; k < nf; k++)
          for (int t = 0; t < VVCLF_ALF_CC_NUM_COEFF; t++) {
            const int32_t v = c == 0 ? cc->cc_cb_coeff[k][t] : cc->cc_cr_coeff[k][t];
            if (v < -(1 << 20) || v >= (1 << 20))
              return fail(VVCLF_ERR_UNSUPPORTED, "alf: CC-ALF coefficient %d outside 21 bits", v);
          }
      }
      build_alf_banks(ac, b, t);
      std::vector<bool> slot_ok(b.n_luma_apses, false);
      for (uint32_t i = 0; i < b.n_luma_apses; i++) slot_ok[i] = b.luma_apses[i] && b.luma_apses[i]->alf_luma_filter_signal_flag;
      const vvclf_alf_aps* ca = b.chroma_aps;
      bool chroma_ok = ac.chroma_format_idc != 0 && ca && ca->alf_chroma_filter_signal_flag;
      uint32_t n_alt = chroma_ok ? std::min(ca->n_chroma_coeff, (uint32_t)VVCLF_ALF_MAX_CHROMA_ALT) : 0;
      const vvclf_alf_aps* ccs[2] = {b.cc_cb_aps, b.cc_cr_aps};
      uint32_t n_cc[2] = {0, 0};
      for (int c = 0; c < 2; c++)
        if (ccs[c] && ac.chroma_format_idc != 0)
          n_cc[c] = std::min(c == 0 ? ccs[c]->n_cc_cb_coeff : ccs[c]->n_cc_cr_coeff, (uint32_t)VVCLF_ALF_MAX_CC_FILTERS);
      t.alf_ctb.assign(n, 0xffffffffu);
      t.alf_rect.resize(n);
      uint32_t ctb = 1u << ac.ctb_log2_size_y;
      for (uint32_t ry = 0; ry < t.hctb; ry++)
        for (uint32_t rx = 0; rx < t.wctb; rx++) {
          size_t i = (size_t)ry * t.wctb + rx;
          const vvclf_alf_ctb& c = ap.ctbs[i];
          uint32_t bank = ALF_OFF8, cb = ALF_OFF8, cr = ALF_OFF8, ccb = ALF_OFF4, ccr = ALF_OFF4;
          if (c.luma_on) {
            if (c.luma_filt_set_idx < 16) bank = c.luma_filt_set_idx;
            else if ((uint32_t)(c.luma_filt_set_idx - 16) < b.n_luma_apses && slot_ok[c.luma_filt_set_idx - 16])
              bank = c.luma_filt_set_idx;
          }
          if (ac.cb_enabled && c.cb_on && c.cb_alt_idx < n_alt) cb = c.cb_alt_idx;
          if (ac.cr_enabled && c.cr_on && c.cr_alt_idx < n_alt) cr = c.cr_alt_idx;
          if (ac.cc_cb_enabled && c.cc_cb_idc != 0 && (uint32_t)(c.cc_cb_idc - 1) < n_cc[0]) ccb = c.cc_cb_idc - 1u;
          if (ac.cc_cr_enabled && c.cc_cr_idc != 0 && (uint32_t)(c.cc_cr_idc - 1) < n_cc[1]) ccr = c.cc_cr_idc - 1u;
          t.alf_ctb[i] = bank | (cb << 8) | (cr << 16) | (ccb << 24) | (ccr << 28);
          t.alf_rect[i] = ctb_rect(f.clip_bounds, rx, ry, ctb, W, H);
          if (t.alf_rect[i] == 0) return fail(VVCLF_ERR_INVALID, "alf: empty clip rectangle at CTB (%u,%u)", rx, ry);
        }
    }
  }
  return VVCLF_OK;
}

// LMCS inverse map as a 65536-entry table: idx_y_inv + inverse_map_luma_sample
// (src/lmcs.rs:209-238) evaluated for every u16 input.  idx_y_inv returns
// the first bin in min..=max whose upper pivot exceeds the sample (else
// max + 1, capped at 15); LmcsPivot is non-decreasing, so that index never
// falls as the sample grows, and one walk over the samples advances it in
// place of a scan per sample.
void build_lmcs_lut(const vvclf_lmcs_binding& b, std::vector<uint16_t>& lut) {
  const vvclf_lmcs_derived& d = *b.derived;
  lut.resize(65536);
  const int64_t hi = ((int64_t)1 << d.bit_depth) - 1;
  uint32_t scan = b.min_bin_idx;  // the scan's stopping bin (max + 1 when none)
  for (uint32_t s = 0; s < 65536;) {
    while (scan <= b.max_bin_idx && s >= d.lmcs_pivot[scan + 1]) scan++;
    // samples up to the next pivot the scan would stop at share the bin
    const uint32_t end = scan <= b.max_bin_idx ? std::min<uint32_t>(d.lmcs_pivot[scan + 1], 65536) : 65536;
    const uint32_t idx = std::min(scan, 15u);
    const int64_t isc = d.inv_scale_coeff[idx], piv = d.lmcs_pivot[idx], inp = d.input_pivot[idx];
    for (; s < end; s++) {
      const int64_t v = inp + ((isc * ((int64_t)s - piv) + (1 << 10)) >> 11);
      lut[s] = (uint16_t)(v < 0 ? 0 : (v > hi ? hi : v));
    }
  }
}

const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

}  // namespace vvclf
