// vvclf_walk.h — derivation of the deblocking edge descriptors from the CU
// records: the walk of deblock_one_direction / deblock_cu_dir
// (src/deblock.rs:687-983), storing per segment what the sample filter needs:
// bS (src/deblock.rs:1055-1119), maxFilterLength P/Q (:845-901 luma,
// :925-938 chroma), the CTB-row flag (:905-906), palette write skips and the
// beta / tC table indices (:1122-1174).  Host and device functions: the host
// walk (vvclf_host.cpp) and the GPU derivation (vvclf_derive.hip) run this
// same code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vvclf.h"
#include "vvclf_internal.h"

namespace vvclf {
namespace walk {

#define VVCLF_HD __host__ __device__ __forceinline__

VVCLF_HD uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
VVCLF_HD uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
VVCLF_HD int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
VVCLF_HD int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

// During the walk bit 31 of a descriptor slot marks it as claimed by a record
// (descriptors use bits 0-23); the finishing pass strips it.
constexpr uint32_t SLOT_CLAIMED = 1u << 31;

// CuGrid (src/deblock.rs:630-671): the record index per 4x4 luma cell, -1 for
// none.
struct grid_view {
  const int32_t* cells;
  uint32_t w, h;
  VVCLF_HD int32_t at(int32_t x, int32_t y) const {
    if (x < 0 || y < 0) return -1;
    uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
    if (cx >= w || cy >= h) return -1;
    return cells[(size_t)cy * w + cx];
  }
};

struct map_view {
  uint32_t* e;
  uint32_t w, h;
};

struct plane_params {
  uint32_t comp, sub_w, sub_h;
  int32_t beta_off, tc_off, qp_off;
  uint32_t bd, ctb;
};

VVCLF_HD uint32_t clampu_rel(int32_t v, uint32_t hi_excl) {
  if (v < 0) return 0;
  if ((uint32_t)v >= hi_excl) return hi_excl - 1;
  return (uint32_t)v;
}

// CU-relative offset along the split / edge-normal axis.
VVCLF_HD uint32_t rel_along(const vvclf_deblock_cu& c, int32_t lx, int32_t ly, bool vert) {
  return vert ? clampu_rel(lx - (int32_t)c.x, c.w) : clampu_rel(ly - (int32_t)c.y, c.h);
}

VVCLF_HD uint32_t tb_idx(const vvclf_deblock_cu& c, uint32_t rel) {
  uint32_t idx = 0;
  for (uint32_t i = 0; i < c.tb_n_bounds && i < 3; i++)
    if (rel >= c.tb_bounds[i]) idx = i + 1;
  return idx;
}

VVCLF_HD uint32_t tb_len_split(const vvclf_deblock_cu& c, uint32_t rel, uint32_t dim) {
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
VVCLF_HD bool tu_coded(const vvclf_deblock_cu& c, uint32_t comp, int32_t lx, int32_t ly) {
  if (c.has_tu64) {
    uint32_t tx = umin(clampu_rel(lx - (int32_t)c.x, c.w) / 64, 1u);
    uint32_t ty = umin(clampu_rel(ly - (int32_t)c.y, c.h) / 64, 1u);
    return comp == 0 ? c.tu64_y[ty][tx] : (comp == 1 ? c.tu64_cb[ty][tx] : c.tu64_cr[ty][tx]);
  }
  if (c.has_tb_split && (comp == 0 || !c.tb_luma_only)) {
    uint32_t i = tb_idx(c, rel_along(c, lx, ly, c.tb_vertical != 0));
    return comp == 0 ? c.tb_y_coded[i] : (comp == 1 ? c.tb_cb_coded[i] : c.tb_cr_coded[i]);
  }
  return comp == 0 ? c.tu_y_coded : (comp == 1 ? c.tu_cb_coded : c.tu_cr_coded);
}

// luma_tb_len / chroma_tb_len, src/deblock.rs:207-248
VVCLF_HD uint32_t tb_len(const vvclf_deblock_cu& c, int32_t lx, int32_t ly, bool vert, bool chroma) {
  uint32_t dim = vert ? c.w : c.h;
  uint32_t rel = rel_along(c, lx, ly, vert);
  if (c.has_tu64 && dim > 64) return umin(dim - (rel / 64) * 64, 64u);
  if (c.has_tb_split && (c.tb_vertical != 0) == vert && (!chroma || !c.tb_luma_only)) return tb_len_split(c, rel, dim);
  return dim;
}

// is_luma_tb_boundary / is_chroma_tb_boundary, src/deblock.rs:252-291
VVCLF_HD bool tb_boundary(const vvclf_deblock_cu& c, uint32_t off, bool vert, bool chroma) {
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

// The motion cell covering luma (x, y); all-zero outside the grid.
VVCLF_HD vvclf_mv_cell mv_at(const vvclf_mv_grid* g, int32_t x, int32_t y) {
  vvclf_mv_cell z{};
  if (!g || x < 0 || y < 0) return z;
  uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
  if (cx >= g->cells_w || cy >= g->cells_h) return z;
  return g->cells[(size_t)cy * g->cells_w + cx];
}

VVCLF_HD bool mv_far(const int32_t a[2], const int32_t b[2]) { return iabs(a[0] - b[0]) >= 8 || iabs(a[1] - b[1]) >= 8; }

// bs_motion_rule, src/deblock.rs:992-1045
VVCLF_HD bool motion_differs(const vvclf_mv_cell& p, const vvclf_mv_cell& q) {
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
VVCLF_HD int boundary_strength(uint32_t comp, const vvclf_deblock_cu& p, const vvclf_deblock_cu& q,
                               const vvclf_mv_grid* g, int32_t px, int32_t py, int32_t qx, int32_t qy, bool is_tb,
                               bool idc2) {
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

// tC == 0 (src/deblock.rs:1429, :1815): tC' (Table 43) is 0 exactly for the
// indices 0..17 and at least 3 from 18 on, and the bit-depth scaling of
// src/deblock.rs:439-452 keeps 0 at 0 and a value >= 3 above 0 for every
// depth >= 8.
VVCLF_HD bool tc_is_zero(int32_t qt) { return qt < 18; }

VVCLF_HD int32_t chroma_qp(const plane_params& pp, const vvclf_deblock_cu& c) {
  if (c.qp_c[0] == INT32_MIN) return iclamp(c.qp_y + pp.qp_off, 0, 63);
  if (c.joint_cbcr2) return c.qp_c[2];
  return c.qp_c[umin(pp.comp - 1, 1u)];
}

// Claims descriptor slot `slot` for the calling record; false if another
// record claimed it first.
VVCLF_HD bool claim(uint32_t* slot) {
  return (__atomic_fetch_or(slot, SLOT_CLAIMED, __ATOMIC_RELAXED) & SLOT_CLAIMED) == 0;
}

// deblock_cu_dir, src/deblock.rs:707-983, for record idx in one direction of
// one plane, writing descriptors into m.  Returns false, with the segment's
// position in err_xy, when two records claim one segment (the records do not
// tile the picture, the precondition of src/deblock.rs:463-465).  Records of
// one pass may be walked concurrently: each writes only the slots it claims
// (in the map, or in a bitmap of claims).
// hist (may be NULL): luma bS histogram [0..2] and the chroma segment count [3].
//
// Along one edge the P-side record changes only at record borders, and what
// the descriptor takes from it — the bS arms that read the two records only
// (BDPCM, intra / CIIP), the transform length across the edge (its
// coordinate along the edge normal is fixed), hence both maxFilterLengths,
// and the averaged QP — is the same for every segment of a run with one P
// record: it is worked out once per run.  bS from coded TBs and motion is
// evaluated per segment, as the reference does.
VVCLF_HD bool walk_cu(const plane_params& pp, const vvclf_deblock_cu* cus, const grid_view& grid,
                      const vvclf_mv_grid* mv, uint32_t idx, bool vert, const uint32_t* nf, uint32_t n_nf,
                      uint32_t pic_w, uint32_t pic_h, const map_view& m, uint64_t* hist, int32_t* err_xy,
                      uint64_t* claims = nullptr) {
  const vvclf_deblock_cu& cu = cus[idx];
  // claims: NULL marks a claimed slot in the map itself (bit 31, stripped
  // afterwards); else one bit per slot in this bitmap, and the map receives
  // only the descriptors that filter
  const uint32_t mark = claims ? 0u : SLOT_CLAIMED;
  const uint32_t comp = pp.comp;
  if (comp != 0) {
    uint32_t lead_c = vert ? cu.x / pp.sub_w : cu.y / pp.sub_h;
    if (lead_c % 8 != 0) return true;
  }
  uint32_t n_cb = vert ? cu.w : cu.h, n_other = vert ? cu.h : cu.w;
  uint32_t lead = vert ? cu.x : cu.y;
  bool filter_edge = lead != 0;
  for (uint32_t i = 0; i < n_nf && filter_edge; i++)
    if (nf[i] == lead) filter_edge = false;
  uint32_t grid_step = comp == 0 ? 4u : 8u * (vert ? pp.sub_w : pp.sub_h);
  uint32_t seg_step = comp == 0 ? 4u : 2u * (vert ? pp.sub_h : pp.sub_w);
  uint32_t nsb = 1;
  if (comp == 0 && cu.has_num_sb) nsb = umax(vert ? cu.num_sb_x : cu.num_sb_y, 1u);
  uint32_t sb = umax(8u, n_cb / nsb);
  uint32_t n8 = n_cb / 8;
  uint32_t last_sb = umin(umax(n8, 1u) - 1, nsb - 1);
  uint32_t ctb_h_c = pp.ctb / umax(pp.sub_h, 1u);
  const uint32_t sub_axis = vert ? pp.sub_w : pp.sub_h;
  const int32_t qp_q = comp == 0 ? cu.qp_y : chroma_qp(pp, cu);
  // records whose coded-TB flag does not depend on the position
  const bool q_tu_fixed = !cu.has_tu64 && !(cu.has_tb_split && (comp == 0 || !cu.tb_luma_only));
  const vvclf_mv_cell zero_cell{};
  uint32_t seg_hist[4] = {0, 0, 0, 0};
  for (uint32_t off = 0; off < n_cb; off += grid_step) {
    bool is_tb = off == 0 ? filter_edge : tb_boundary(cu, off, vert, comp != 0);
    bool is_sb = comp == 0 && off % sb == 0 && off / sb <= last_sb && (off > 0 || filter_edge);
    if (!(is_tb || is_sb)) continue;
    // per edge: the Q side's transform length across the edge and, for a
    // sub-block edge inside the record, the maxFilterLength cap
    const int32_t qx_e = vert ? (int32_t)(cu.x + off) : (int32_t)cu.x;
    const int32_t qy_e = vert ? (int32_t)cu.y : (int32_t)(cu.y + off);
    const uint32_t wq = tb_len(cu, qx_e, qy_e, vert, comp != 0);
    const bool ctb_row = comp == 0 && !vert && off == 0 && pp.ctb > 0 && cu.y % pp.ctb == 0;
    const bool chroma_row = comp != 0 && !vert && off == 0 && ctb_h_c > 0 && (cu.y / pp.sub_h) % ctb_h_c == 0;
    uint32_t sb_cap = 0;  // 0: none; else the inner sub-block edge's mfl pair
    if (is_sb && off != 0 && !is_tb) {
      auto tb_at = [&](int64_t z) -> bool {
        if (z < 0 || (uint64_t)z >= n_cb) return false;
        if (z == 0) return filter_edge;
        return tb_boundary(cu, (uint32_t)z, vert, false);
      };
      if (tb_at((int64_t)off - 4) || tb_at((int64_t)off + 4)) sb_cap = 1;
      else if (off / sb == 1 || off / sb == umax(n8, 1u) - 1 || tb_at((int64_t)off - sb) || tb_at((int64_t)off + sb))
        sb_cap = 2;
      else
        sb_cap = 3;
    }
#ifndef __HIP_DEVICE_COMPILE__
    // The edge's descriptor slots (one map row apart along a vertical edge),
    // and the P-side cells of the record grid and the motion grid, requested
    // together before the segments are visited, so that their cache misses
    // overlap instead of following one another.
    if (vert) {
      for (uint32_t k = 0; k < n_other; k += seg_step) {
        const uint32_t y = cu.y + k;
        const uint32_t sy = comp == 0 ? y / 4 : y / pp.sub_h / 2;
        const uint32_t sx = comp == 0 ? (uint32_t)qx_e / 4 : (uint32_t)qx_e / pp.sub_w / 8;
        if (sy < m.h && sx < m.w) __builtin_prefetch(&m.e[(size_t)sy * m.w + sx], 1);
        if (off == 0 && qx_e > 0 && y / 4 < grid.h) __builtin_prefetch(&grid.cells[(size_t)(y / 4) * grid.w + (qx_e - 1) / 4]);
        if (comp == 0 && mv && y / 4 < mv->cells_h && (uint32_t)qx_e / 4 < mv->cells_w) {
          const vvclf_mv_cell* c = &mv->cells[(size_t)(y / 4) * mv->cells_w + (uint32_t)qx_e / 4];
          __builtin_prefetch(c);
          if (qx_e >= 4) __builtin_prefetch(c - 1);
        }
      }
    }
#endif
    // the run's P-side values
    int32_t run_pi = -1;
    const vvclf_deblock_cu* p = nullptr;
    int bs_fixed = -1;  // 0 or 2 when the records alone decide bS
    bool p_tu_fixed = false, p_tu = false, q_tu = false;
    uint32_t mflp = 0, mflq = 0, plt_bits = 0;
    int32_t qpv = 0;
    for (uint32_t k = 0; k < n_other; k += seg_step) {
      int32_t qx = vert ? qx_e : (int32_t)(cu.x + k);
      int32_t qy = vert ? (int32_t)(cu.y + k) : qy_e;
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
        sx = vert ? (uint32_t)ex / 8 : (uint32_t)ex / 2;
        sy = vert ? (uint32_t)ey / 2 : (uint32_t)ey / 8;
      }
      if (sx >= m.w || sy >= m.h) continue;  // entirely outside the picture
      const size_t slot = (size_t)sy * m.w + sx;
      if (claims ? (__atomic_fetch_or(&claims[slot >> 6], 1ull << (slot & 63), __ATOMIC_RELAXED) >> (slot & 63)) & 1
                 : !claim(&m.e[slot])) {
        err_xy[0] = qx;
        err_xy[1] = qy;
        return false;
      }
      if (pi != run_pi) {  // a new P record along this edge
        run_pi = pi;
        p = &cus[pi];
        // derive_bs, src/deblock.rs:1055-1119: the arms that read the records only
        if (comp == 0 && p->bdpcm_luma && cu.bdpcm_luma) bs_fixed = 0;
        else if (comp != 0 && p->bdpcm_chroma && cu.bdpcm_chroma) bs_fixed = 0;
        else if (p->intra || cu.intra || p->ciip || cu.ciip) bs_fixed = 2;
        else bs_fixed = -1;
        p_tu_fixed = !p->has_tu64 && !(p->has_tb_split && (comp == 0 || !p->tb_luma_only));
        p_tu = comp == 0 ? p->tu_y_coded : (comp == 1 ? p->tu_cb_coded : p->tu_cr_coded);
        q_tu = comp == 0 ? cu.tu_y_coded : (comp == 1 ? cu.tu_cb_coded : cu.tu_cr_coded);
        const uint32_t wp = tb_len(*p, px, py, vert, comp != 0);
        if (comp == 0) {
          mflp = mflq = 0;
          if (is_tb) {
            if (wp <= 4 || wq <= 4) {
              mflp = mflq = 1;
            } else {
              mflp = wp >= 32 ? 7 : 3;
              mflq = wq >= 32 ? 7 : 3;
            }
          }
          if (is_sb) {
            if (off == 0) {
              if (nsb > 1) mflq = umin(mflq, 5u);
              if (p->has_num_sb) mflp = umin(mflp, 5u);
            } else if (is_tb) {
              mflp = umin(mflp, 5u);
              mflq = umin(mflq, 5u);
            } else {
              mflp = mflq = sb_cap;
            }
          }
          qpv = (p->qp_y + qp_q + 1) >> 1;
        } else {
          const uint32_t wq_c = wq / umax(sub_axis, 1u), wp_c = wp / umax(sub_axis, 1u);
          if (wp_c >= 8 && wq_c >= 8) {
            mflp = chroma_row ? 1 : 3;
            mflq = 3;
          } else {
            mflp = mflq = 1;
          }
          qpv = (chroma_qp(pp, *p) + qp_q + 1) >> 1;
        }
        plt_bits = (p->plt != 0 ? 1u : 0u) | (cu.plt != 0 ? 2u : 0u);
      }
      int bs;
      if (bs_fixed >= 0) {
        bs = bs_fixed;
      } else if (is_tb && ((p_tu_fixed ? p_tu : tu_coded(*p, comp, px, py)) ||
                           (q_tu_fixed ? q_tu : tu_coded(cu, comp, qx, qy)))) {
        bs = 1;
      } else if (comp != 0 || !is_sb) {
        bs = 0;
      } else if ((p->ibc != 0) != (cu.ibc != 0)) {
        bs = 1;
      } else {
        // motion: the cells covering the two sides, compared in place
        const vvclf_mv_cell *a = &zero_cell, *b = &zero_cell;
        if (mv) {
          const uint32_t cw = mv->cells_w, ch = mv->cells_h;
          if (px >= 0 && py >= 0 && (uint32_t)px / 4 < cw && (uint32_t)py / 4 < ch)
            a = &mv->cells[(size_t)((uint32_t)py / 4) * cw + (uint32_t)px / 4];
          if ((uint32_t)qx / 4 < cw && (uint32_t)qy / 4 < ch) b = &mv->cells[(size_t)((uint32_t)qy / 4) * cw + (uint32_t)qx / 4];
        }
        if (p->ibc && cu.ibc) bs = mv && mv_far(a->mv_l0, b->mv_l0) ? 1 : 0;
        else bs = mv && motion_differs(*a, *b) ? 1 : 0;
      }
      if (comp == 0) seg_hist[bs]++;
      if (bs == 0) continue;
      if (comp != 0 && mflp == 1 && mflq == 1 && bs != 2) continue;  // src/deblock.rs:1823, :1888
      int32_t qb = iclamp(qpv + pp.beta_off * 2, 0, 63);
      int32_t qt = iclamp(qpv + 2 * (bs - 1) + pp.tc_off * 2, 0, 65);
      if (tc_is_zero(qt)) continue;  // tC == 0: no-op (src/deblock.rs:1429, :1815)
      if (comp != 0) seg_hist[3]++;
      m.e[slot] = mark | pack_edge((uint32_t)bs, mflp, mflq, (plt_bits & 1) != 0, (plt_bits & 2) != 0,
                                           ctb_row, (uint32_t)qb, (uint32_t)qt);
    }
  }
  if (hist)
    for (int i = 0; i < 4; i++) hist[i] += seg_hist[i];
  return true;
}

// Conservative same-direction interaction test for the concurrent pass: for
// two consecutive edges on one line, the samples one may write must not reach
// the samples the other uses.  w/r follow the filter extents of
// src/deblock.rs:1224-1776 for the given maxFilterLength pair.
VVCLF_HD uint32_t luma_writes(uint32_t m, uint32_t o) {
  if (m > 3) return m;
  if (o > 3 || (m > 2 && o > 2)) return 3;
  return (m > 1 && o > 1) ? 2 : 1;
}
VVCLF_HD uint32_t luma_reads(uint32_t m, uint32_t o) {
  if (m > 3) return m + 1;
  if (o > 3 || (m > 2 && o > 2)) return 4;
  return 3;
}
// Can consecutive same-direction luma edges a (earlier) and b, d samples
// apart on one line, reach each other's used samples?
VVCLF_HD bool luma_pair_safe(uint32_t a, uint32_t b, uint32_t d, bool vert) {
  uint32_t a_p = (a >> E_MFLP_SHIFT) & 7, a_q = (a >> E_MFLQ_SHIFT) & 7;
  uint32_t b_p = (b >> E_MFLP_SHIFT) & 7, b_q = (b >> E_MFLQ_SHIFT) & 7;
  if (!vert && (b & E_CTBROW_BIT)) b_p = umin(b_p, 3u);
  return luma_writes(a_q, a_p) + luma_reads(b_p, b_q) <= d && luma_reads(a_q, a_p) + luma_writes(b_p, b_q) <= d;
}

#undef VVCLF_HD

}  // namespace walk
}  // namespace vvclf
