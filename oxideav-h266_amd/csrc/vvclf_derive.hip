// vvclf_derive.hip — the deblocking edge descriptors derived on the GPU from
// the CU records and the motion grid (the walk of deblock_one_direction /
// deblock_cu_dir, src/deblock.rs:687-983, in vvclf_walk.h), so that a
// picture's host work is the upload of its records instead of the walk.
//
//   k_grid   CuGrid (src/deblock.rs:630-671): per 4x4 cell the largest index
//            of the records covering it (later records overwrite earlier ones)
//   k_walk   one thread per (pass, record): walk_cu into the pass's map,
//            claiming each slot it writes (two claims of one slot = records
//            that do not tile the picture)
//   k_rows   per map row: claim marks off, nonzero count, and for the luma
//            vertical map the same-direction interaction check
//   k_cols   the same check down the columns of the luma horizontal map
//   k_scan   row offsets of every map's work list
//   k_write  the work lists (tag | row << 16 | column, row-major)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <vector>

#include "vvclf_host_internal.h"
#include "vvclf_walk.h"

namespace vvclf {
namespace {

struct dpass {
  walk::plane_params pp;
  const vvclf_deblock_cu* cus;
  uint32_t n;
  uint32_t vert;
  walk::grid_view grid;
  const vvclf_mv_grid* mv;  // device copy (luma passes with a motion grid), else NULL
  const uint32_t* nf;       // no-filter columns (vertical) / rows (horizontal)
  uint32_t n_nf;
  walk::map_view m;
};

struct dmap {
  uint32_t* e;
  uint32_t w, h;
  uint32_t row0;   // first slot of this map's rows in the count / offset arrays
  uint32_t check;  // 1: luma vertical map, check along rows
  uint32_t tag;
  int32_t base;    // map whose list precedes this one in the same list, or -1
  uint32_t* list;
};

struct dstatus {
  uint32_t bad;  // two records claimed one segment
  int32_t err_x, err_y;
  uint32_t unsafe;  // same-direction luma edges interact
  uint32_t total[6];
};

// The host builds these records and the kernels read them: both compile
// passes must lay them out alike (checked for the host and the gfx950 pass).
static_assert(sizeof(walk::plane_params) == 32 && sizeof(walk::grid_view) == 16 && sizeof(walk::map_view) == 16,
              "walk views");
static_assert(sizeof(dpass) == 104 && offsetof(dpass, cus) == 32 && offsetof(dpass, grid) == 48 &&
                  offsetof(dpass, mv) == 64 && offsetof(dpass, nf) == 72 && offsetof(dpass, m) == 88,
              "dpass layout");
static_assert(sizeof(dmap) == 40 && offsetof(dmap, list) == 32, "dmap layout");
static_assert(sizeof(dstatus) == 40 && offsetof(dstatus, total) == 16, "dstatus layout");
static_assert(sizeof(vvclf_deblock_cu) == 92 && sizeof(vvclf_mv_cell) == 28 && sizeof(vvclf_mv_grid) == 16 &&
                  offsetof(vvclf_mv_grid, cells_w) == 8,
              "ABI records read by the kernels");

__global__ void __launch_bounds__(256) k_fill(int32_t* p, size_t n, int32_t v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

__global__ void __launch_bounds__(256) k_grid(const vvclf_deblock_cu* cus, uint32_t n, int32_t* cells, uint32_t w,
                                              uint32_t h) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const vvclf_deblock_cu& c = cus[i];
  const uint32_t x0 = c.x / 4, y0 = c.y / 4;
  const uint32_t x1 = walk::umin((c.x + c.w + 3) / 4, w), y1 = walk::umin((c.y + c.h + 3) / 4, h);
  for (uint32_t cy = y0; cy < y1; cy++)
    for (uint32_t cx = x0; cx < x1; cx++) atomicMax(&cells[(size_t)cy * w + cx], (int32_t)i);
}

__global__ void __launch_bounds__(256) k_walk(const dpass* passes, uint32_t W, uint32_t H, dstatus* st) {
  const dpass& p = passes[blockIdx.y];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n) return;
  int32_t xy[2];
  if (!walk::walk_cu(p.pp, p.cus, p.grid, p.mv, i, p.vert != 0, p.nf, p.n_nf, W, H, p.m, nullptr, xy)) {
    if (atomicCAS(&st->bad, 0u, 1u) == 0u) {
      st->err_x = xy[0];
      st->err_y = xy[1];
    }
  }
}

__global__ void __launch_bounds__(256) k_rows(const dmap* maps, uint32_t* counts, dstatus* st) {
  const dmap& m = maps[blockIdx.y];
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m.h) return;
  uint32_t* row = m.e + (size_t)r * m.w;
  uint32_t cnt = 0, pe = 0;
  int32_t px = -1;
  for (uint32_t x = 0; x < m.w; x++) {
    const uint32_t raw = row[x], v = raw & ~walk::SLOT_CLAIMED;
    if (raw != v) row[x] = v;
    if (!v) continue;
    cnt++;
    if (m.check) {
      if (px >= 0 && !walk::luma_pair_safe(pe, v, 4 * (x - (uint32_t)px), true)) st->unsafe = 1;
      px = (int32_t)x;
      pe = v;
    }
  }
  counts[m.row0 + r] = cnt;
}

// luma horizontal map: consecutive edges down each column (claim marks ignored)
__global__ void __launch_bounds__(256) k_cols(const uint32_t* e, uint32_t w, uint32_t h, dstatus* st) {
  const uint32_t x = blockIdx.x * 256 + threadIdx.x;
  if (x >= w) return;
  uint32_t pe = 0;
  int32_t py = -1;
  for (uint32_t y = 0; y < h; y++) {
    const uint32_t v = e[(size_t)y * w + x] & ~walk::SLOT_CLAIMED;
    if (!v) continue;
    if (py >= 0 && !walk::luma_pair_safe(pe, v, 4 * (y - (uint32_t)py), false)) st->unsafe = 1;
    py = (int32_t)y;
    pe = v;
  }
}

// exclusive offsets of one map's rows (one workgroup per map)
__global__ void __launch_bounds__(256) k_scan(const dmap* maps, const uint32_t* counts, uint32_t* offs, dstatus* st) {
  __shared__ uint32_t part[256];
  const dmap& m = maps[blockIdx.x];
  const uint32_t per = (m.h + 255) / 256, b = threadIdx.x * per, e = walk::umin(b + per, m.h);
  uint32_t sum = 0;
  for (uint32_t r = b; r < e; r++) sum += counts[m.row0 + r];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < 256; k++) {
      const uint32_t v = part[k];
      part[k] = acc;
      acc += v;
    }
    st->total[blockIdx.x] = acc;
  }
  __syncthreads();
  uint32_t acc = part[threadIdx.x];
  for (uint32_t r = b; r < e; r++) {
    offs[m.row0 + r] = acc;
    acc += counts[m.row0 + r];
  }
}

__global__ void __launch_bounds__(256) k_write(const dmap* maps, const uint32_t* offs, const dstatus* st) {
  const dmap& m = maps[blockIdx.y];
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m.h) return;
  const uint32_t* row = m.e + (size_t)r * m.w;
  uint32_t k = (m.base >= 0 ? st->total[m.base] : 0u) + offs[m.row0 + r];
  for (uint32_t x = 0; x < m.w; x++)
    if (row[x]) m.list[k++] = m.tag | (r << 16) | x;
}

// 256-byte aligned suballocation of one device buffer, with its host image
struct layout {
  size_t bytes = 0;
  size_t add(size_t n) {
    const size_t off = (bytes + 255) / 256 * 256;
    bytes = off + std::max<size_t>(n, 4);
    return off;
  }
};

}  // namespace

int derive_deblock_gpu(const vvclf_frame& f, uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, const gpu_maps& o,
                       hipStream_t s, uint32_t n_out[4]) {
  const vvclf_deblock_params& prm = f.deblock;
  const uint32_t ctb = 1u << prm.ctb_log2_size_y;
  for (uint32_t i = 0; i < f.n_cus; i++) {
    const vvclf_deblock_cu& c = f.cus[i];
    if ((c.x | c.y | c.w | c.h) & 3) return fail(VVCLF_ERR_UNSUPPORTED, "deblock: CU %u not on the 4x4 grid", i);
  }
  const bool chroma = o.chroma;
  const vvclf_deblock_cu* ccus = f.has_chroma_cus ? f.chroma_cus : f.cus;
  const uint32_t nc = f.has_chroma_cus ? f.n_chroma_cus : f.n_cus;
  const bool dual = chroma && f.has_chroma_cus;
  if (dual)
    for (uint32_t i = 0; i < nc; i++) {
      const vvclf_deblock_cu& c = ccus[i];
      if ((c.x | c.y | c.w | c.h) & 3) return fail(VVCLF_ERR_UNSUPPORTED, "deblock: chroma CU %u off grid", i);
    }
  const uint32_t g4w = (W + 3) / 4, g4h = (H + 3) / 4;
  const vvclf_mv_grid* mv = f.mv_grid;
  const size_t n_mv = mv ? (size_t)mv->cells_w * mv->cells_h : 0;
  if (mv && n_mv && !mv->cells) return fail(VVCLF_ERR_INVALID, "deblock: NULL motion cells");
  // device buffer: inputs (one upload), then work arrays
  layout L;
  const size_t o_cus = L.add((size_t)f.n_cus * sizeof(vvclf_deblock_cu));
  const size_t o_ccus = dual ? L.add((size_t)nc * sizeof(vvclf_deblock_cu)) : o_cus;
  const size_t o_mvc = L.add(n_mv * sizeof(vvclf_mv_cell));
  const size_t o_mvg = L.add(sizeof(vvclf_mv_grid));
  const size_t o_nfc = L.add((size_t)f.n_no_filter_cols * 4);
  const size_t o_nfr = L.add((size_t)f.n_no_filter_rows * 4);
  const size_t o_pass = L.add(6 * sizeof(dpass));
  const size_t o_map = L.add(6 * sizeof(dmap));
  const size_t upload = L.bytes;
  const size_t o_grid = L.add((size_t)g4w * g4h * 4);
  const size_t o_cgrid = dual ? L.add((size_t)g4w * g4h * 4) : o_grid;
  const uint32_t rows_total = 2 * o.gh + (chroma ? 2 * (o.cvh + o.chh) : 0);
  const size_t o_cnt = L.add((size_t)rows_total * 4);
  const size_t o_off = L.add((size_t)rows_total * 4);
  const size_t o_st = L.add(sizeof(dstatus));
  uint8_t* dev = nullptr;
  hipError_t e = s ? hipMallocAsync((void**)&dev, L.bytes, s) : hipMalloc((void**)&dev, L.bytes);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "deblock derivation buffer: %s", hipGetErrorString(e));
  std::vector<uint8_t> img(upload, 0);
  auto put = [&](size_t off, const void* p, size_t n) {
    if (n) memcpy(img.data() + off, p, n);
  };
  put(o_cus, f.cus, (size_t)f.n_cus * sizeof(vvclf_deblock_cu));
  if (dual) put(o_ccus, ccus, (size_t)nc * sizeof(vvclf_deblock_cu));
  put(o_mvc, mv ? mv->cells : nullptr, n_mv * sizeof(vvclf_mv_cell));
  if (mv) {
    vvclf_mv_grid g = *mv;
    g.cells = (vvclf_mv_cell*)(dev + o_mvc);
    put(o_mvg, &g, sizeof(g));
  }
  put(o_nfc, f.no_filter_cols, (size_t)f.n_no_filter_cols * 4);
  put(o_nfr, f.no_filter_rows, (size_t)f.n_no_filter_rows * 4);
  const walk::grid_view gv{(const int32_t*)(dev + o_grid), g4w, g4h};
  const walk::grid_view cgv{(const int32_t*)(dev + o_cgrid), g4w, g4h};
  const vvclf_mv_grid* dmv = mv ? (const vvclf_mv_grid*)(dev + o_mvg) : nullptr;
  const uint32_t* dnfc = (const uint32_t*)(dev + o_nfc);
  const uint32_t* dnfr = (const uint32_t*)(dev + o_nfr);
  dpass ps[6];
  dmap ms[6];
  int np = 0;
  const walk::plane_params pl = {0, 1, 1, prm.luma_beta_offset_div2, prm.luma_tc_offset_div2, 0, prm.bit_depth, ctb};
  const vvclf_deblock_cu* dcus = (const vvclf_deblock_cu*)(dev + o_cus);
  const vvclf_deblock_cu* dccus = (const vvclf_deblock_cu*)(dev + o_ccus);
  ps[np++] = dpass{pl, dcus, f.n_cus, 1, gv, dmv, dnfc, f.n_no_filter_cols, walk::map_view{o.e_lv, o.gw, o.gh}};
  ps[np++] = dpass{pl, dcus, f.n_cus, 0, gv, dmv, dnfr, f.n_no_filter_rows, walk::map_view{o.e_lh, o.gw, o.gh}};
  uint32_t row0 = 0;
  ms[0] = dmap{o.e_lv, o.gw, o.gh, row0, 1, 0, -1, o.l_lv};
  row0 += o.gh;
  ms[1] = dmap{o.e_lh, o.gw, o.gh, row0, 0, 0, -1, o.l_lh};
  row0 += o.gh;
  int nm = 2;
  if (chroma) {
    for (int c = 0; c < 2; c++) {
      const walk::plane_params pc = {(uint32_t)(c + 1),
                                     2,
                                     2,
                                     c == 0 ? prm.cb_beta_offset_div2 : prm.cr_beta_offset_div2,
                                     c == 0 ? prm.cb_tc_offset_div2 : prm.cr_tc_offset_div2,
                                     c == 0 ? prm.chroma_qp_offset_cb : prm.chroma_qp_offset_cr,
                                     prm.bit_depth,
                                     ctb};
      ps[np++] = dpass{pc, dccus, nc, 1, cgv, nullptr, dnfc, f.n_no_filter_cols, walk::map_view{o.e_cv[c], o.cvw, o.cvh}};
      ps[np++] = dpass{pc, dccus, nc, 0, cgv, nullptr, dnfr, f.n_no_filter_rows, walk::map_view{o.e_ch[c], o.chw, o.chh}};
    }
    // lists: Cb entries then Cr entries (Cr tagged in bit 31)
    ms[nm++] = dmap{o.e_cv[0], o.cvw, o.cvh, row0, 0, 0, -1, o.l_cv};
    row0 += o.cvh;
    ms[nm++] = dmap{o.e_cv[1], o.cvw, o.cvh, row0, 0, 1u << 31, 2, o.l_cv};
    row0 += o.cvh;
    ms[nm++] = dmap{o.e_ch[0], o.chw, o.chh, row0, 0, 0, -1, o.l_ch};
    row0 += o.chh;
    ms[nm++] = dmap{o.e_ch[1], o.chw, o.chh, row0, 0, 1u << 31, 4, o.l_ch};
    row0 += o.chh;
  }
  put(o_pass, ps, sizeof(dpass) * np);
  put(o_map, ms, sizeof(dmap) * nm);
  dstatus hs{};
  e = hipMemcpyAsync(dev, img.data(), upload, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(dev + o_st, 0, sizeof(dstatus), s);
  dstatus* dst = (dstatus*)(dev + o_st);
  auto blocks = [](size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); };
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fill, dim3(blocks((size_t)g4w * g4h / 4)), dim3(256), 0, s, (int32_t*)(dev + o_grid),
                       (size_t)g4w * g4h, -1);
    hipLaunchKernelGGL(k_grid, dim3(blocks(f.n_cus)), dim3(256), 0, s, dcus, f.n_cus, (int32_t*)(dev + o_grid), g4w,
                       g4h);
    if (dual) {
      hipLaunchKernelGGL(k_fill, dim3(blocks((size_t)g4w * g4h / 4)), dim3(256), 0, s, (int32_t*)(dev + o_cgrid),
                         (size_t)g4w * g4h, -1);
      hipLaunchKernelGGL(k_grid, dim3(blocks(nc)), dim3(256), 0, s, dccus, nc, (int32_t*)(dev + o_cgrid), g4w, g4h);
    }
    const uint32_t nmax = std::max(f.n_cus, chroma ? nc : 0u);
    if (nmax) hipLaunchKernelGGL(k_walk, dim3(blocks(nmax), np), dim3(256), 0, s, (const dpass*)(dev + o_pass), W, H, dst);
    uint32_t hmax = o.gh;
    if (chroma) hmax = std::max(hmax, std::max(o.cvh, o.chh));
    hipLaunchKernelGGL(k_rows, dim3(blocks(hmax), nm), dim3(256), 0, s, (const dmap*)(dev + o_map),
                       (uint32_t*)(dev + o_cnt), dst);
    hipLaunchKernelGGL(k_cols, dim3(blocks(o.gw)), dim3(256), 0, s, (const uint32_t*)o.e_lh, o.gw, o.gh, dst);
    hipLaunchKernelGGL(k_scan, dim3(nm), dim3(256), 0, s, (const dmap*)(dev + o_map), (const uint32_t*)(dev + o_cnt),
                       (uint32_t*)(dev + o_off), dst);
    hipLaunchKernelGGL(k_write, dim3(blocks(hmax), nm), dim3(256), 0, s, (const dmap*)(dev + o_map),
                       (const uint32_t*)(dev + o_off), (const dstatus*)dst);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (s) hipFreeAsync(dev, s);
  else hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "deblock derivation");
  if (hs.bad)
    return fail(VVCLF_ERR_INVALID, "deblock: CU records overlap at (%d,%d): they must tile the picture", hs.err_x,
                hs.err_y);
  if (hs.unsafe)
    return fail(VVCLF_ERR_UNSUPPORTED,
                "deblock: same-direction luma edges interact (sub-block edge spacing); the concurrent pass "
                "would not reproduce the sequential result");
  n_out[0] = hs.total[0];
  n_out[1] = hs.total[1];
  n_out[2] = chroma ? hs.total[2] + hs.total[3] : 0;
  n_out[3] = chroma ? hs.total[4] + hs.total[5] : 0;
  return VVCLF_OK;
}

}  // namespace vvclf
