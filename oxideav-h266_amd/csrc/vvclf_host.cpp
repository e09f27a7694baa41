// vvclf_host.cpp — the C ABI of include/vvclf.h: descriptor validation, the
// host-side derivation of the compact per-edge / per-CTB tables the kernels
// read, job lifetime, and the synchronous drop-in entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"
#include "vvclf_internal.h"

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

struct cu_grid {
  std::vector<int32_t> cells;
  uint32_t w = 0, h = 0;
  // CuGrid::build, src/deblock.rs:637-660: later records overwrite earlier ones.
  void build(uint32_t pic_w, uint32_t pic_h, const vvclf_deblock_cu* cus, uint32_t n) {
    w = (pic_w + 3) / 4;
    h = (pic_h + 3) / 4;
    cells.assign((size_t)w * h, -1);
    for (uint32_t i = 0; i < n; i++) {
      const vvclf_deblock_cu& c = cus[i];
      uint32_t x0 = c.x / 4, y0 = c.y / 4;
      uint32_t x1 = std::min((c.x + c.w + 3) / 4, w), y1 = std::min((c.y + c.h + 3) / 4, h);
      for (uint32_t cy = y0; cy < y1; cy++)
        for (uint32_t cx = x0; cx < x1; cx++) cells[(size_t)cy * w + cx] = (int32_t)i;
    }
  }
  int32_t at(int32_t x, int32_t y) const {
    if (x < 0 || y < 0) return -1;
    uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
    if (cx >= w || cy >= h) return -1;
    return cells[(size_t)cy * w + cx];
  }
};

inline uint32_t clampu_rel(int32_t v, uint32_t hi_excl) {
  if (v < 0) return 0;
  if ((uint32_t)v >= hi_excl) return hi_excl - 1;
  return (uint32_t)v;
}

// CU-relative offset along the split / edge-normal axis.
inline uint32_t rel_along(const vvclf_deblock_cu& c, int32_t lx, int32_t ly, bool vert) {
  return vert ? clampu_rel(lx - (int32_t)c.x, c.w) : clampu_rel(ly - (int32_t)c.y, c.h);
}

inline uint32_t tb_idx(const vvclf_deblock_cu& c, uint32_t rel) {
  uint32_t idx = 0;
  for (uint32_t i = 0; i < c.tb_n_bounds && i < 3; i++)
    if (rel >= c.tb_bounds[i]) idx = i + 1;
  return idx;
}

inline uint32_t tb_len_split(const vvclf_deblock_cu& c, uint32_t rel, uint32_t dim) {
  uint32_t s = 0, e = dim;
  for (uint32_t i = 0; i < c.tb_n_bounds && i < 3; i++) {
    if (rel < c.tb_bounds[i]) {
      e = c.tb_bounds[i];
      break;
    }
    s = c.tb_bounds[i];
  }
  return e - s;
}

// DeblockCu::tu_coded_at, src/deblock.rs:168-201
bool tu_coded(const vvclf_deblock_cu& c, uint32_t comp, int32_t lx, int32_t ly) {
  if (c.has_tu64) {
    uint32_t tx = std::min(clampu_rel(lx - (int32_t)c.x, c.w) / 64, 1u);
    uint32_t ty = std::min(clampu_rel(ly - (int32_t)c.y, c.h) / 64, 1u);
    return comp == 0 ? c.tu64_y[ty][tx] : (comp == 1 ? c.tu64_cb[ty][tx] : c.tu64_cr[ty][tx]);
  }
  if (c.has_tb_split && (comp == 0 || !c.tb_luma_only)) {
    uint32_t i = tb_idx(c, rel_along(c, lx, ly, c.tb_vertical != 0));
    return comp == 0 ? c.tb_y_coded[i] : (comp == 1 ? c.tb_cb_coded[i] : c.tb_cr_coded[i]);
  }
  return comp == 0 ? c.tu_y_coded : (comp == 1 ? c.tu_cb_coded : c.tu_cr_coded);
}

// luma_tb_len / chroma_tb_len, src/deblock.rs:207-248
uint32_t tb_len(const vvclf_deblock_cu& c, int32_t lx, int32_t ly, bool vert, bool chroma) {
  uint32_t dim = vert ? c.w : c.h;
  uint32_t rel = rel_along(c, lx, ly, vert);
  if (c.has_tu64 && dim > 64) return std::min(dim - (rel / 64) * 64, 64u);
  if (c.has_tb_split && (c.tb_vertical != 0) == vert && (!chroma || !c.tb_luma_only)) return tb_len_split(c, rel, dim);
  return dim;
}

// is_luma_tb_boundary / is_chroma_tb_boundary, src/deblock.rs:252-291
bool tb_boundary(const vvclf_deblock_cu& c, uint32_t off, bool vert, bool chroma) {
  if (off == 0) return false;
  uint32_t dim = vert ? c.w : c.h;
  if (off >= dim) return false;
  if (c.has_tu64 && dim > 64) return off % 64 == 0;
  if (c.has_tb_split && (c.tb_vertical != 0) == vert && (!chroma || !c.tb_luma_only)) {
    for (uint32_t i = 0; i < c.tb_n_bounds && i < 3; i++)
      if (c.tb_bounds[i] == off) return true;
  }
  return false;
}

vvclf_mv_cell mv_at(const vvclf_mv_grid* g, int32_t x, int32_t y) {
  vvclf_mv_cell z;
  memset(&z, 0, sizeof(z));
  if (!g || x < 0 || y < 0) return z;
  uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
  if (cx >= g->cells_w || cy >= g->cells_h) return z;
  return g->cells[(size_t)cy * g->cells_w + cx];
}

inline bool mv_far(const int32_t a[2], const int32_t b[2]) { return std::abs(a[0] - b[0]) >= 8 || std::abs(a[1] - b[1]) >= 8; }

// bs_motion_rule, src/deblock.rs:992-1045
bool motion_differs(const vvclf_mv_cell& p, const vvclf_mv_cell& q) {
  int np = (p.pred_l0 != 0) + (p.pred_l1 != 0), nq = (q.pred_l0 != 0) + (q.pred_l1 != 0);
  if (np != nq) return true;
  if (np == 0) return false;
  if (np == 1) {
    int32_t pp = p.pred_l0 ? p.poc_l0 : p.poc_l1, qp = q.pred_l0 ? q.poc_l0 : q.poc_l1;
    if (pp != qp) return true;
    return mv_far(p.pred_l0 ? p.mv_l0 : p.mv_l1, q.pred_l0 ? q.mv_l0 : q.mv_l1);
  }
  bool same = p.poc_l0 == q.poc_l0 && p.poc_l1 == q.poc_l1;
  bool swapped = p.poc_l0 == q.poc_l1 && p.poc_l1 == q.poc_l0;
  if (!same && !swapped) return true;
  if (p.poc_l0 != p.poc_l1) {
    if (same) return mv_far(p.mv_l0, q.mv_l0) || mv_far(p.mv_l1, q.mv_l1);
    return mv_far(p.mv_l0, q.mv_l1) || mv_far(p.mv_l1, q.mv_l0);
  }
  return (mv_far(p.mv_l0, q.mv_l0) || mv_far(p.mv_l1, q.mv_l1)) &&
         (mv_far(p.mv_l0, q.mv_l1) || mv_far(p.mv_l1, q.mv_l0));
}

// derive_bs, src/deblock.rs:1055-1119
int boundary_strength(uint32_t comp, const vvclf_deblock_cu& p, const vvclf_deblock_cu& q, const vvclf_mv_grid* g,
                      int32_t px, int32_t py, int32_t qx, int32_t qy, bool is_tb, bool idc2) {
  if (comp == 0 && p.bdpcm_luma && q.bdpcm_luma) return 0;
  if (comp != 0 && p.bdpcm_chroma && q.bdpcm_chroma) return 0;
  if (p.intra || q.intra || p.ciip || q.ciip) return 2;
  if (is_tb && (tu_coded(p, comp, px, py) || tu_coded(q, comp, qx, qy))) return 1;
  if (comp != 0 || !idc2) return 0;
  if ((p.ibc != 0) != (q.ibc != 0)) return 1;
  if (p.ibc && q.ibc) {
    if (!g) return 0;
    vvclf_mv_cell a = mv_at(g, px, py), b = mv_at(g, qx, qy);
    return mv_far(a.mv_l0, b.mv_l0) ? 1 : 0;
  }
  if (g && motion_differs(mv_at(g, px, py), mv_at(g, qx, qy))) return 1;
  return 0;
}

int32_t tc_value(int32_t qt, uint32_t bd) {
  static const int16_t TC[66] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,
                                 0,  3,  4,  4,  4,  4,  5,  5,  5,  5,   7,   7,   8,   9,   10,  10,  11,
                                 13, 14, 15, 17, 19, 21, 24, 25, 29, 33,  36,  41,  45,  51,  57,  64,  71,
                                 80, 89, 100, 112, 125, 141, 157, 177, 198, 222, 250, 280, 314, 352, 395};
  int32_t t = TC[qt];
  return bd < 10 ? (t + (1 << (9 - bd))) >> (10 - bd) : t * (1 << (bd - 10));
}

struct plane_params {
  uint32_t comp, sub_w, sub_h;
  int32_t beta_off, tc_off, qp_off;
  uint32_t bd, ctb;
};

int32_t chroma_qp(const plane_params& pp, const vvclf_deblock_cu& c) {
  if (c.qp_c[0] == INT32_MIN) return std::clamp(c.qp_y + pp.qp_off, 0, 63);
  if (c.joint_cbcr2) return c.qp_c[2];
  return c.qp_c[std::min(pp.comp - 1, 1u)];
}

struct edge_map {
  std::vector<uint32_t> e;
  std::vector<uint8_t> seen;
  uint32_t w = 0, h = 0;
  void init(uint32_t w_, uint32_t h_) {
    w = w_;
    h = h_;
    e.assign((size_t)w * h, 0);
    seen.assign((size_t)w * h, 0);
  }
};

// deblock_cu_dir, src/deblock.rs:707-983, emitting descriptors.  Returns
// false (with g_err set) when two records claim one segment (the records do
// not tile the picture, the precondition of src/deblock.rs:463-465).
bool walk_cu(const plane_params& pp, const vvclf_deblock_cu* cus, const cu_grid& grid, const vvclf_mv_grid* mv,
             uint32_t idx, bool vert, const uint32_t* nf, uint32_t n_nf, uint32_t pic_w, uint32_t pic_h,
             edge_map& m, uint64_t* hist) {
  const vvclf_deblock_cu& cu = cus[idx];
  const uint32_t comp = pp.comp;
  if (comp != 0) {
    uint32_t lead_c = vert ? cu.x / pp.sub_w : cu.y / pp.sub_h;
    if (lead_c % 8 != 0) return true;
  }
  uint32_t n_cb = vert ? cu.w : cu.h, n_other = vert ? cu.h : cu.w;
  uint32_t lead = vert ? cu.x : cu.y;
  bool filter_edge = lead != 0 && std::find(nf, nf + n_nf, lead) == nf + n_nf;
  uint32_t grid_step = comp == 0 ? 4u : 8u * (vert ? pp.sub_w : pp.sub_h);
  uint32_t seg_step = comp == 0 ? 4u : 2u * (vert ? pp.sub_h : pp.sub_w);
  uint32_t nsb = 1;
  if (comp == 0 && cu.has_num_sb) nsb = std::max(vert ? cu.num_sb_x : cu.num_sb_y, 1u);
  uint32_t sb = std::max(8u, n_cb / nsb);
  uint32_t n8 = n_cb / 8;
  uint32_t last_sb = std::min(std::max(n8, 1u) - 1, nsb - 1);
  uint32_t ctb_h_c = pp.ctb / std::max(pp.sub_h, 1u);
  for (uint32_t off = 0; off < n_cb; off += grid_step) {
    bool is_tb = off == 0 ? filter_edge : tb_boundary(cu, off, vert, comp != 0);
    bool is_sb = comp == 0 && off % sb == 0 && off / sb <= last_sb && (off > 0 || filter_edge);
    if (!(is_tb || is_sb)) continue;
    for (uint32_t k = 0; k < n_other; k += seg_step) {
      int32_t qx = vert ? (int32_t)(cu.x + off) : (int32_t)(cu.x + k);
      int32_t qy = vert ? (int32_t)(cu.y + k) : (int32_t)(cu.y + off);
      int32_t px = vert ? qx - 1 : qx, py = vert ? qy : qy - 1;
      int32_t pi = off == 0 ? grid.at(px, py) : (int32_t)idx;
      if (pi < 0) continue;
      if (off == 0 && (uint32_t)pi == idx) continue;
      // map slot (segment index along the edge, edge index along the normal)
      uint32_t sx, sy;
      if (comp == 0) {
        sx = (uint32_t)qx / 4;
        sy = (uint32_t)qy / 4;
        if ((uint32_t)qx >= pic_w + 8 || (uint32_t)qy >= pic_h + 8) continue;
      } else {
        int32_t ex = qx / (int32_t)pp.sub_w, ey = qy / (int32_t)pp.sub_h;
        uint32_t cw = (pic_w + pp.sub_w - 1) / pp.sub_w, chh = (pic_h + pp.sub_h - 1) / pp.sub_h;
        (void)cw;
        (void)chh;
        sx = vert ? (uint32_t)ex / 8 : (uint32_t)ex / 2;
        sy = vert ? (uint32_t)ey / 2 : (uint32_t)ey / 8;
      }
      if (sx >= m.w || sy >= m.h) continue;  // entirely outside the picture
      size_t slot = (size_t)sy * m.w + sx;
      if (m.seen[slot]) {
        fail(VVCLF_ERR_INVALID, "deblock: CU records overlap at (%d,%d): they must tile the picture", qx, qy);
        return false;
      }
      m.seen[slot] = 1;
      const vvclf_deblock_cu& p = cus[pi];
      int bs = boundary_strength(comp, p, cu, mv, px, py, qx, qy, is_tb, is_sb);
      if (hist && comp == 0) hist[bs]++;
      if (bs == 0) continue;
      uint32_t mflp, mflq;
      bool ctb_row = false;
      int32_t qpv;
      if (comp == 0) {
        mflp = mflq = 0;
        if (is_tb) {
          uint32_t wq = tb_len(cu, qx, qy, vert, false), wp = tb_len(p, px, py, vert, false);
          if (wp <= 4 || wq <= 4) {
            mflp = mflq = 1;
          } else {
            mflp = wp >= 32 ? 7 : 3;
            mflq = wq >= 32 ? 7 : 3;
          }
        }
        if (is_sb) {
          if (off == 0) {
            if (nsb > 1) mflq = std::min(mflq, 5u);
            if (p.has_num_sb) mflp = std::min(mflp, 5u);
          } else {
            auto tb_at = [&](int64_t z) -> bool {
              if (z < 0 || (uint64_t)z >= n_cb) return false;
              if (z == 0) return filter_edge;
              return tb_boundary(cu, (uint32_t)z, vert, false);
            };
            if (is_tb) {
              mflp = std::min(mflp, 5u);
              mflq = std::min(mflq, 5u);
            } else if (tb_at((int64_t)off - 4) || tb_at((int64_t)off + 4)) {
              mflp = mflq = 1;
            } else if (off / sb == 1 || off / sb == std::max(n8, 1u) - 1 || tb_at((int64_t)off - sb) ||
                       tb_at((int64_t)off + sb)) {
              mflp = mflq = 2;
            } else {
              mflp = mflq = 3;
            }
          }
        }
        ctb_row = !vert && off == 0 && pp.ctb > 0 && cu.y % pp.ctb == 0;
        qpv = (p.qp_y + cu.qp_y + 1) >> 1;
      } else {
        uint32_t sub_axis = vert ? pp.sub_w : pp.sub_h;
        uint32_t wq = tb_len(cu, qx, qy, vert, true) / std::max(sub_axis, 1u);
        uint32_t wp = tb_len(p, px, py, vert, true) / std::max(sub_axis, 1u);
        if (wp >= 8 && wq >= 8) {
          bool row = !vert && off == 0 && ctb_h_c > 0 && (cu.y / pp.sub_h) % ctb_h_c == 0;
          mflp = row ? 1 : 3;
          mflq = 3;
        } else {
          mflp = mflq = 1;
        }
        if (mflp == 1 && mflq == 1 && bs != 2) continue;  // src/deblock.rs:1823, :1888
        qpv = (chroma_qp(pp, p) + chroma_qp(pp, cu) + 1) >> 1;
      }
      int32_t qb = std::clamp(qpv + pp.beta_off * 2, 0, 63);
      int32_t qt = std::clamp(qpv + 2 * (bs - 1) + pp.tc_off * 2, 0, 65);
      if (tc_value(qt, pp.bd) == 0) continue;  // tC == 0: no-op (src/deblock.rs:1429, :1815)
      if (hist && comp != 0) hist[3]++;
      m.e[slot] = pack_edge((uint32_t)bs, mflp, mflq, p.plt != 0, cu.plt != 0, ctb_row, (uint32_t)qb, (uint32_t)qt);
    }
  }
  return true;
}

// Conservative same-direction interaction test for the concurrent pass: for
// two consecutive edges on one line, the samples one may write must not reach
// the samples the other uses.  w/r follow the filter extents of
// src/deblock.rs:1224-1776 for the given maxFilterLength pair.
inline uint32_t luma_writes(uint32_t m, uint32_t o) {
  if (m > 3) return m;
  if (o > 3 || (m > 2 && o > 2)) return 3;
  return (m > 1 && o > 1) ? 2 : 1;
}
inline uint32_t luma_reads(uint32_t m, uint32_t o) {
  if (m > 3) return m + 1;
  if (o > 3 || (m > 2 && o > 2)) return 4;
  return 3;
}

bool luma_map_concurrent_safe(const edge_map& m, bool vert) {
  // vert: consecutive edges along x within a map row; else along y within a column.
  uint32_t lines = vert ? m.h : m.w, n = vert ? m.w : m.h;
  for (uint32_t l = 0; l < lines; l++) {
    int32_t prev = -1;
    uint32_t pe = 0;
    for (uint32_t i = 0; i < n; i++) {
      uint32_t e = vert ? m.e[(size_t)l * m.w + i] : m.e[(size_t)i * m.w + l];
      if (!e) continue;
      if (prev >= 0) {
        uint32_t d = 4 * (i - (uint32_t)prev);
        uint32_t a_p = (pe >> E_MFLP_SHIFT) & 7, a_q = (pe >> E_MFLQ_SHIFT) & 7;
        uint32_t b_p = (e >> E_MFLP_SHIFT) & 7, b_q = (e >> E_MFLQ_SHIFT) & 7;
        if (!vert && (e & E_CTBROW_BIT)) b_p = std::min(b_p, 3u);
        if (luma_writes(a_q, a_p) + luma_reads(b_p, b_q) > d || luma_reads(a_q, a_p) + luma_writes(b_p, b_q) > d)
          return false;
      }
      prev = (int32_t)i;
      pe = e;
    }
  }
  return true;
}

}  // namespace

int derive_deblock_maps(const vvclf_frame& f, uint32_t W, uint32_t H, uint32_t Wc, uint32_t Hc, host_tables& t) {
  const vvclf_deblock_params& prm = f.deblock;
  uint32_t ctb = 1u << prm.ctb_log2_size_y;
  cu_grid grid;
  grid.build(W, H, f.cus, f.n_cus);
  for (uint32_t i = 0; i < f.n_cus; i++) {
    const vvclf_deblock_cu& c = f.cus[i];
    if ((c.x | c.y | c.w | c.h) & 3) return fail(VVCLF_ERR_UNSUPPORTED, "deblock: CU %u not on the 4x4 grid", i);
  }
  uint64_t hist[4] = {0, 0, 0, 0};
  edge_map lv, lh;
  t.gw = (W + 3) / 4 + 2;
  t.gh = (H + 3) / 4 + 2;
  lv.init(t.gw, t.gh);
  lh.init(t.gw, t.gh);
  plane_params pl = {0, 1, 1, prm.luma_beta_offset_div2, prm.luma_tc_offset_div2, 0, prm.bit_depth, ctb};
  for (uint32_t i = 0; i < f.n_cus; i++)
    if (!walk_cu(pl, f.cus, grid, f.mv_grid, i, true, f.no_filter_cols, f.n_no_filter_cols, W, H, lv, hist))
      return VVCLF_ERR_INVALID;
  for (uint32_t i = 0; i < f.n_cus; i++)
    if (!walk_cu(pl, f.cus, grid, f.mv_grid, i, false, f.no_filter_rows, f.n_no_filter_rows, W, H, lh, hist))
      return VVCLF_ERR_INVALID;
  if (!luma_map_concurrent_safe(lv, true) || !luma_map_concurrent_safe(lh, false))
    return fail(VVCLF_ERR_UNSUPPORTED,
                "deblock: same-direction luma edges interact (sub-block edge spacing); the concurrent pass "
                "would not reproduce the sequential result");
  t.e_lv = std::move(lv.e);
  t.e_lh = std::move(lh.e);
  t.has_chroma_dbk = false;
  if (f.chroma_format_idc != 0 && Wc > 0) {
    const vvclf_deblock_cu* ccus = f.has_chroma_cus ? f.chroma_cus : f.cus;
    uint32_t nc = f.has_chroma_cus ? f.n_chroma_cus : f.n_cus;
    cu_grid cgrid;
    const cu_grid* cg = &grid;
    if (f.has_chroma_cus) {
      cgrid.build(W, H, ccus, nc);
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
      plane_params pc = {(uint32_t)comp,
                         2,
                         2,
                         comp == 1 ? prm.cb_beta_offset_div2 : prm.cr_beta_offset_div2,
                         comp == 1 ? prm.cb_tc_offset_div2 : prm.cr_tc_offset_div2,
                         comp == 1 ? prm.chroma_qp_offset_cb : prm.chroma_qp_offset_cr,
                         prm.bit_depth,
                         ctb};
      edge_map cv, ch;
      cv.init(t.cvw, t.cvh);
      ch.init(t.chw, t.chh);
      for (uint32_t i = 0; i < nc; i++)
        if (!walk_cu(pc, ccus, *cg, nullptr, i, true, f.no_filter_cols, f.n_no_filter_cols, W, H, cv, hist))
          return VVCLF_ERR_INVALID;
      for (uint32_t i = 0; i < nc; i++)
        if (!walk_cu(pc, ccus, *cg, nullptr, i, false, f.no_filter_rows, f.n_no_filter_rows, W, H, ch, hist))
          return VVCLF_ERR_INVALID;
      t.e_cv[comp - 1] = std::move(cv.e);
      t.e_ch[comp - 1] = std::move(ch.e);
    }
    t.has_chroma_dbk = true;
  }
  // compact lists of the descriptors that filter something
  auto list_nonzero = [](const std::vector<uint32_t>& e, uint32_t w, uint32_t tag, std::vector<uint32_t>& out) {
    for (size_t i = 0; i < e.size(); i++)
      if (e[i]) out.push_back(tag | ((uint32_t)(i / w) << 16) | (uint32_t)(i % w));
  };
  t.l_lv.clear();
  t.l_lh.clear();
  t.l_cv.clear();
  t.l_ch.clear();
  list_nonzero(t.e_lv, t.gw, 0, t.l_lv);
  list_nonzero(t.e_lh, t.gw, 0, t.l_lh);
  if (t.has_chroma_dbk)
    for (int c = 0; c < 2; c++) {
      list_nonzero(t.e_cv[c], t.cvw, (uint32_t)c << 31, t.l_cv);
      list_nonzero(t.e_ch[c], t.chw, (uint32_t)c << 31, t.l_ch);
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
// (src/lmcs.rs:209-238) evaluated for every u16 input.
void build_lmcs_lut(const vvclf_lmcs_binding& b, std::vector<uint16_t>& lut) {
  const vvclf_lmcs_derived& d = *b.derived;
  lut.resize(65536);
  int64_t hi = ((int64_t)1 << d.bit_depth) - 1;
  for (uint32_t s = 0; s < 65536; s++) {
    uint32_t idx = b.min_bin_idx;
    while (idx <= b.max_bin_idx) {
      if (s < d.lmcs_pivot[idx + 1]) break;
      idx++;
    }
    idx = std::min(idx, 15u);
    int64_t prod = (int64_t)d.inv_scale_coeff[idx] * ((int64_t)s - (int64_t)d.lmcs_pivot[idx]) + (1 << 10);
    int64_t v = (int64_t)d.input_pivot[idx] + (prod >> 11);
    lut[s] = (uint16_t)std::clamp<int64_t>(v, 0, hi);
  }
}

const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

}  // namespace vvclf
