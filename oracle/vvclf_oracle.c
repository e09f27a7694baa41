/*
 * vvclf_oracle.c — CPU restatement of the reference VVC in-loop filter chain
 * (LMCS inverse -> deblocking -> SAO -> ALF + CC-ALF).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The
 * product never links or calls this file.  It keeps the reference's structure
 * on purpose: sequential decode-order, in-place deblocking; whole-plane
 * snapshots for SAO and ALF; clamped per-sample fetches.  Citations are
 * reference file:line.
 *
 * Parity pins (see tests/test_oracle_kat.py): the reference's own known-answer
 * unit tests for each stage, plus closed-form invariants (flat/ramp identity,
 * tC = 0 identity, one-band BO = per-sample offset, single-tap ALF = 3-tap
 * filter) and brute force against straightforward re-derivations.
 */
#include "vvclf_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int64_t clampl(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int32_t absi(int32_t v) { return v < 0 ? -v : v; }
static inline uint32_t minu(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline uint32_t maxu(uint32_t a, uint32_t b) { return a > b ? a : b; }
/* Arithmetic right shift (floor division by 2^n), as Rust's `>>` on signed. */
static inline int32_t asr(int32_t v, int n) { return v >> n; }
static inline int64_t asr64(int64_t v, int n) { return v >> n; }

/* ======================================================================== */
/* LMCS                                                                      */
/* ======================================================================== */

/* src/lmcs.rs:302-404 — derive_lmcs with its conformance checks. */
int oracle_derive_lmcs(const vvclf_lmcs_data* d, uint32_t bd, vvclf_lmcs_derived* o) {
  if (bd < 8 || bd > 16) return VVCLF_ERR_INVALID;
  if (d->lmcs_delta_max_bin_idx > 15 || d->lmcs_min_bin_idx > 15) return VVCLF_ERR_INVALID;
  memset(o, 0, sizeof(*o));
  uint32_t org_cw = (1u << bd) / 16;
  uint32_t log2_org_cw = bd - 4;
  int64_t band_lo = org_cw >> 3, band_hi = (int64_t)(org_cw << 3) - 1;
  uint32_t mn = d->lmcs_min_bin_idx, mx = 15u - d->lmcs_delta_max_bin_idx;
  for (uint32_t i = mn; i <= mx && i < 16; i++) {
    int64_t delta = (int64_t)d->lmcs_delta_abs_cw[i] * (d->lmcs_delta_sign_cw_flag[i] ? -1 : 1);
    int64_t v = (int64_t)org_cw + delta;
    if (v < band_lo || v > band_hi) return VVCLF_ERR_INVALID;
    o->lmcs_cw[i] = (uint32_t)v;
  }
  uint64_t total = 0;
  for (int i = 0; i < 16; i++) total += o->lmcs_cw[i];
  if (total > (1u << bd) - 1) return VVCLF_ERR_INVALID;
  for (uint32_t i = 0; i < 16; i++) o->input_pivot[i] = i * org_cw;
  o->lmcs_pivot[0] = 0;
  for (int i = 0; i < 16; i++) {
    o->lmcs_pivot[i + 1] = o->lmcs_pivot[i] + o->lmcs_cw[i];
    o->scale_coeff[i] = (o->lmcs_cw[i] * (1u << 11) + (1u << (log2_org_cw - 1))) >> log2_org_cw;
    o->inv_scale_coeff[i] = o->lmcs_cw[i] ? (org_cw * (1u << 11)) / o->lmcs_cw[i] : 0;
  }
  uint32_t shift = bd - 5;
  for (uint32_t i = mn; i <= mx && i < 16; i++) {
    if (o->lmcs_pivot[i] % (1u << shift) != 0 && (o->lmcs_pivot[i] >> shift) == (o->lmcs_pivot[i + 1] >> shift))
      return VVCLF_ERR_INVALID;
  }
  int32_t crs = (int32_t)d->lmcs_delta_abs_crs * (d->lmcs_delta_sign_crs_flag ? -1 : 1);
  for (int i = 0; i < 16; i++) {
    if (o->lmcs_cw[i] == 0) {
      o->chroma_scale_coeff[i] = 1u << 11;
    } else {
      int64_t joint = (int64_t)o->lmcs_cw[i] + crs;
      if (joint < band_lo || joint > band_hi) return VVCLF_ERR_INVALID;
      o->chroma_scale_coeff[i] = (org_cw * (1u << 11)) / (uint32_t)joint;
    }
  }
  o->bit_depth = bd;
  o->org_cw = org_cw;
  o->lmcs_delta_crs = crs;
  return VVCLF_OK;
}

/* src/lmcs.rs:209-220 — linear scan over the signalled bins. */
uint32_t oracle_lmcs_idx_y_inv(const vvclf_lmcs_derived* d, uint32_t s, uint32_t mn, uint32_t mx) {
  uint32_t idx = mn;
  while (idx <= mx) {
    if (s < d->lmcs_pivot[idx + 1]) break;
    idx++;
  }
  return idx < 15 ? idx : 15;
}

/* src/lmcs.rs:230-238 — eqs. 1222/1223. */
uint32_t oracle_lmcs_inverse_map(const vvclf_lmcs_derived* d, uint32_t s, uint32_t idx) {
  int64_t prod = (int64_t)d->inv_scale_coeff[idx] * ((int64_t)s - (int64_t)d->lmcs_pivot[idx]) + (1 << 10);
  int64_t inv = (int64_t)d->input_pivot[idx] + asr64(prod, 11);
  return (uint32_t)clampl(inv, 0, ((int64_t)1 << d->bit_depth) - 1);
}

/* src/lmcs.rs:186-195 — eq. 1213 (used only by tests of the derivation). */
int64_t oracle_lmcs_forward_map(const vvclf_lmcs_derived* d, uint32_t ps) {
  uint32_t idx = ps >> (d->bit_depth - 4);
  int64_t prod = (int64_t)d->scale_coeff[idx] * ((int64_t)ps - (int64_t)d->input_pivot[idx]) + (1 << 10);
  return (int64_t)d->lmcs_pivot[idx] + asr64(prod, 11);
}

/* src/ctu.rs:11837-11849 — the picture inverse-mapping pass. */
void oracle_lmcs_inverse_luma(vvclf_plane* p, const vvclf_lmcs_binding* b) {
  if (!b || !b->derived) return;
  for (uint32_t y = 0; y < p->height; y++)
    for (uint32_t x = 0; x < p->width; x++) {
      uint16_t* s = &p->samples[(size_t)y * p->stride + x];
      uint32_t iy = oracle_lmcs_idx_y_inv(b->derived, *s, b->min_bin_idx, b->max_bin_idx);
      *s = (uint16_t)oracle_lmcs_inverse_map(b->derived, *s, iy);
    }
}

/* src/ctu.rs:1701-1725 — lmcs_forward_map_luma_rect: forward_map_luma_sample
 * (src/lmcs.rs:186-195) clamped to [0, 2^BitDepth - 1] over a rectangle
 * clipped to the plane.  A sample >= 2^BitDepth would index past the 16
 * pieces (the reference panics): -1, nothing changed. */
int oracle_lmcs_forward_luma_rect(vvclf_plane* p, const vvclf_lmcs_binding* b, uint32_t x0, uint32_t y0, uint32_t w,
                                  uint32_t h) {
  if (!b || !b->derived) return 0;
  const vvclf_lmcs_derived* d = b->derived;
  int64_t hi = ((int64_t)1 << d->bit_depth) - 1;
  uint32_t y1 = y0 + h < p->height ? y0 + h : p->height, x1 = x0 + w < p->width ? x0 + w : p->width;
  for (uint32_t y = y0; y < y1; y++)
    for (uint32_t x = x0; x < x1; x++)
      if (p->samples[(size_t)y * p->stride + x] > hi) return -1;
  for (uint32_t y = y0; y < y1; y++)
    for (uint32_t x = x0; x < x1; x++) {
      uint16_t* s = &p->samples[(size_t)y * p->stride + x];
      *s = (uint16_t)clampl(oracle_lmcs_forward_map(d, *s), 0, hi);
    }
  return 0;
}

/* src/ctu.rs:1790-1814 — varScale of a chroma block from the mapped-domain
 * luma: left column x_cu - 1 / top row y_cu - 1 of sizeY samples (clamped to
 * the plane), rounded mean (eq. 1216) or mid-grey (eq. 1217), idxYInv
 * (src/lmcs.rs:209-220), ChromaScaleCoeff (eq. 1218). */
uint32_t oracle_lmcs_chroma_var_scale(const vvclf_plane* l, const vvclf_lmcs_binding* b, uint32_t x_cu, uint32_t y_cu,
                                      int avail_l, int avail_t, uint32_t size_y) {
  const vvclf_lmcs_derived* d = b->derived;
  uint64_t sum = 0;
  uint32_t cnt = 0;
  if (avail_l) {
    for (uint32_t i = 0; i < size_y; i++) {
      uint32_t yy = y_cu + i < l->height - 1 ? y_cu + i : l->height - 1;
      sum += l->samples[(size_t)yy * l->stride + (x_cu - 1)];
    }
    cnt += size_y;
  }
  if (avail_t) {
    for (uint32_t i = 0; i < size_y; i++) {
      uint32_t xx = x_cu + i < l->width - 1 ? x_cu + i : l->width - 1;
      sum += l->samples[(size_t)(y_cu - 1) * l->stride + xx];
    }
    cnt += size_y;
  }
  uint32_t avg;
  if (cnt > 0) {
    uint32_t sh = 0; /* cnt.trailing_zeros() */
    while (!((cnt >> sh) & 1u)) sh++;
    avg = (uint32_t)((sum + (cnt >> 1)) >> sh);
  } else {
    avg = 1u << (d->bit_depth - 1);
  }
  return d->chroma_scale_coeff[oracle_lmcs_idx_y_inv(d, avg, b->min_bin_idx, b->max_bin_idx)];
}

/* src/lmcs.rs:257-279 — eqs. 1219/1220. */
uint32_t oracle_lmcs_scale_chroma_residual_sample(const vvclf_lmcs_derived* d, int64_t pred, int64_t res,
                                                  uint32_t var_scale) {
  int64_t lo = -((int64_t)1 << d->bit_depth), hi = ((int64_t)1 << d->bit_depth) - 1;
  int64_t r = clampl(res, lo, hi);
  int64_t scaled = 0;
  if (r != 0) {
    uint64_t mag = ((uint64_t)(r < 0 ? -r : r) * var_scale + (1u << 10)) >> 11;
    scaled = r < 0 ? -(int64_t)mag : (int64_t)mag;
  }
  return (uint32_t)clampl(pred + scaled, 0, hi);
}

/* Per chroma TB: oracle_lmcs_chroma_var_scale, the residual fold of
 * lmcs_scale_chroma_residuals (src/ctu.rs:734-742) and the
 * Clip1(pred + res) add, in block order. */
int oracle_lmcs_chroma_residual(vvclf_picture* pic, const vvclf_lmcs_binding* b, const vvclf_crs_block* blocks,
                                uint32_t n, const int32_t* res) {
  if (!b || !b->derived) return 0;
  for (uint32_t k = 0; k < n; k++) {
    const vvclf_crs_block* blk = &blocks[k];
    vvclf_plane* c = blk->c_idx == 1 ? &pic->cb : &pic->cr;
    uint32_t vs = oracle_lmcs_chroma_var_scale(&pic->luma, b, blk->x_cu, blk->y_cu, blk->avail_l, blk->avail_t,
                                               blk->size_y);
    for (uint32_t y = 0; y < blk->h; y++)
      for (uint32_t x = 0; x < blk->w; x++) {
        uint16_t* s = &c->samples[(size_t)(blk->y + y) * c->stride + blk->x + x];
        *s = (uint16_t)oracle_lmcs_scale_chroma_residual_sample(b->derived, *s,
                                                                res[blk->res_offset + (size_t)y * blk->w + x], vs);
      }
  }
  return 0;
}

/* ======================================================================== */
/* Deblocking                                                                */
/* ======================================================================== */

/* Table 43 (H.266 spec data), src/deblock.rs:401-418. */
static const int32_t BETA_TAB[64] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                                     6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24,
                                     26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56,
                                     58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88};
static const int32_t TC_TAB[66] = {0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,   0,   0,
                                   0,  3,  4,  4,  4,  4,  5,  5,   5,   5,   7,   7,   8,   9,   10,  10,  11,
                                   13, 14, 15, 17, 19, 21, 24, 25,  29,  33,  36,  41,  45,  51,  57,  64,  71,
                                   80, 89, 100, 112, 125, 141, 157, 177, 198, 222, 250, 280, 314, 352, 395};

int32_t oracle_beta_prime(int32_t q) { return BETA_TAB[clampi(q, 0, 63)]; }
int32_t oracle_tc_prime(int32_t q) { return TC_TAB[clampi(q, 0, 65)]; }
/* src/deblock.rs:439-446 */
int32_t oracle_scale_tc(int32_t tcp, uint32_t bd) {
  if (bd < 10) {
    int shift = 9 - (int)bd;
    return asr(tcp + (1 << shift), 10 - (int)bd);
  }
  return tcp * (1 << (bd - 10));
}
/* src/deblock.rs:449-452 */
int32_t oracle_scale_beta(int32_t bp, uint32_t bd) { return bp * (1 << (bd - 8)); }

/* TbSplit::tb_index / tb_extent, src/deblock.rs:297-320 */
static uint32_t tb_index(const vvclf_deblock_cu* cu, uint32_t rel) {
  uint32_t idx = 0;
  for (uint32_t i = 0; i < cu->tb_n_bounds && i < 3; i++)
    if (rel >= cu->tb_bounds[i]) idx = i + 1;
  return idx;
}
static void tb_extent(const vvclf_deblock_cu* cu, uint32_t rel, uint32_t dim, uint32_t* s, uint32_t* e) {
  uint32_t start = 0, end = dim;
  for (uint32_t i = 0; i < cu->tb_n_bounds && i < 3; i++) {
    uint32_t b = cu->tb_bounds[i];
    if (rel < b) {
      end = b;
      break;
    }
    start = b;
  }
  *s = start;
  *e = end;
}

static uint32_t rel_along(const vvclf_deblock_cu* cu, int32_t lx, int32_t ly, int vertical) {
  if (vertical) return (uint32_t)clampi(lx - (int32_t)cu->x, 0, (int32_t)cu->w - 1);
  return (uint32_t)clampi(ly - (int32_t)cu->y, 0, (int32_t)cu->h - 1);
}

/* DeblockCu::tu_coded_at, src/deblock.rs:168-201 */
int oracle_tu_coded_at(const vvclf_deblock_cu* cu, uint32_t c_idx, int32_t lx, int32_t ly) {
  if (cu->has_tu64) {
    uint32_t rx = (uint32_t)clampi(lx - (int32_t)cu->x, 0, (int32_t)cu->w - 1);
    uint32_t ry = (uint32_t)clampi(ly - (int32_t)cu->y, 0, (int32_t)cu->h - 1);
    uint32_t tx = minu(rx / 64, 1), ty = minu(ry / 64, 1);
    if (c_idx == 0) return cu->tu64_y[ty][tx];
    if (c_idx == 1) return cu->tu64_cb[ty][tx];
    return cu->tu64_cr[ty][tx];
  }
  if (cu->has_tb_split) {
    if (c_idx == 0 || !cu->tb_luma_only) {
      uint32_t rel = rel_along(cu, lx, ly, cu->tb_vertical);
      uint32_t idx = tb_index(cu, rel);
      if (c_idx == 0) return cu->tb_y_coded[idx];
      if (c_idx == 1) return cu->tb_cb_coded[idx];
      return cu->tb_cr_coded[idx];
    }
  }
  if (c_idx == 0) return cu->tu_y_coded;
  if (c_idx == 1) return cu->tu_cb_coded;
  return cu->tu_cr_coded;
}

/* DeblockCu::luma_tb_len, src/deblock.rs:207-223 */
uint32_t oracle_luma_tb_len(const vvclf_deblock_cu* cu, int32_t lx, int32_t ly, int vertical) {
  uint32_t dim = vertical ? cu->w : cu->h;
  uint32_t rel = rel_along(cu, lx, ly, vertical);
  if (cu->has_tu64 && dim > 64) return minu(dim - (rel / 64) * 64, 64);
  if (cu->has_tb_split && (int)cu->tb_vertical == (vertical != 0)) {
    uint32_t s, e;
    tb_extent(cu, rel, dim, &s, &e);
    return e - s;
  }
  return dim;
}

/* DeblockCu::chroma_tb_len, src/deblock.rs:229-248 */
uint32_t oracle_chroma_tb_len(const vvclf_deblock_cu* cu, int32_t lx, int32_t ly, int vertical, uint32_t sub) {
  uint32_t dim = vertical ? cu->w : cu->h;
  uint32_t rel = rel_along(cu, lx, ly, vertical);
  uint32_t luma_len;
  if (cu->has_tu64 && dim > 64) {
    luma_len = minu(dim - (rel / 64) * 64, 64);
  } else if (cu->has_tb_split) {
    if ((int)cu->tb_vertical == (vertical != 0) && !cu->tb_luma_only) {
      uint32_t s, e;
      tb_extent(cu, rel, dim, &s, &e);
      luma_len = e - s;
    } else {
      luma_len = dim;
    }
  } else {
    luma_len = dim;
  }
  return luma_len / maxu(sub, 1);
}

/* src/deblock.rs:252-270 */
int oracle_is_luma_tb_boundary(const vvclf_deblock_cu* cu, uint32_t off, int vertical) {
  if (off == 0) return 0;
  uint32_t dim = vertical ? cu->w : cu->h;
  if (off >= dim) return 0;
  if (cu->has_tu64 && dim > 64) return off % 64 == 0;
  if (cu->has_tb_split && (int)cu->tb_vertical == (vertical != 0)) {
    for (uint32_t i = 0; i < cu->tb_n_bounds && i < 3; i++)
      if (cu->tb_bounds[i] == off) return 1;
    return 0;
  }
  return 0;
}

/* src/deblock.rs:274-291 */
int oracle_is_chroma_tb_boundary(const vvclf_deblock_cu* cu, uint32_t off, int vertical) {
  if (off == 0) return 0;
  uint32_t dim = vertical ? cu->w : cu->h;
  if (off >= dim) return 0;
  if (cu->has_tu64 && dim > 64) return off % 64 == 0;
  if (cu->has_tb_split && (int)cu->tb_vertical == (vertical != 0) && !cu->tb_luma_only) {
    for (uint32_t i = 0; i < cu->tb_n_bounds && i < 3; i++)
      if (cu->tb_bounds[i] == off) return 1;
    return 0;
  }
  return 0;
}

/* DeblockMvGrid::at_luma, src/deblock.rs:350-363 */
static vvclf_mv_cell mv_at(const vvclf_mv_grid* g, int32_t x, int32_t y) {
  vvclf_mv_cell z;
  memset(&z, 0, sizeof(z));
  if (x < 0 || y < 0) return z;
  uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
  if (cx >= g->cells_w || cy >= g->cells_h) return z;
  return g->cells[(size_t)cy * g->cells_w + cx];
}

/* CuGrid, src/deblock.rs:630-671 */
typedef struct {
  int32_t* cells;
  uint32_t w, h;
} cu_grid;

static void cu_grid_build(cu_grid* g, uint32_t pic_w, uint32_t pic_h, const vvclf_deblock_cu* cus, uint32_t n) {
  g->w = (pic_w + 3) / 4;
  g->h = (pic_h + 3) / 4;
  g->cells = (int32_t*)malloc(sizeof(int32_t) * (size_t)g->w * g->h + 4);
  for (size_t i = 0; i < (size_t)g->w * g->h; i++) g->cells[i] = -1;
  for (uint32_t idx = 0; idx < n; idx++) {
    const vvclf_deblock_cu* cu = &cus[idx];
    uint32_t x0 = cu->x / 4, y0 = cu->y / 4;
    uint32_t x1 = minu((cu->x + cu->w + 3) / 4, g->w), y1 = minu((cu->y + cu->h + 3) / 4, g->h);
    for (uint32_t cy = y0; cy < y1; cy++)
      for (uint32_t cx = x0; cx < x1; cx++) g->cells[(size_t)cy * g->w + cx] = (int32_t)idx;
  }
}

static int32_t cu_grid_at(const cu_grid* g, int32_t x, int32_t y) {
  if (x < 0 || y < 0) return -1;
  uint32_t cx = (uint32_t)x / 4, cy = (uint32_t)y / 4;
  if (cx >= g->w || cy >= g->h) return -1;
  return g->cells[(size_t)cy * g->w + cx];
}

typedef struct {
  vvclf_plane* plane;
  uint32_t c_idx, sub_w, sub_h;
  int32_t beta_offset_div2, tc_offset_div2, qp_offset;
  uint32_t bit_depth, ctb_size_y;
} plane_ctx;

/* src/deblock.rs:1186-1200 */
static inline int32_t rd(const vvclf_plane* p, int32_t x, int32_t y) {
  int32_t cx = clampi(x, 0, (int32_t)p->width - 1), cy = clampi(y, 0, (int32_t)p->height - 1);
  return p->samples[(size_t)cy * p->stride + cx];
}
static inline void wr(vvclf_plane* p, int32_t x, int32_t y, int32_t v, uint32_t bd) {
  if (x < 0 || y < 0 || x >= (int32_t)p->width || y >= (int32_t)p->height) return;
  p->samples[(size_t)y * p->stride + x] = (uint16_t)clampi(v, 0, (1 << bd) - 1);
}

/* bs_motion_rule, src/deblock.rs:992-1045 */
static int diff_ge8(const int32_t a[2], const int32_t b[2]) { return absi(a[0] - b[0]) >= 8 || absi(a[1] - b[1]) >= 8; }
static int bs_motion_rule(vvclf_mv_cell p, vvclf_mv_cell q) {
  uint32_t np = (p.pred_l0 ? 1 : 0) + (p.pred_l1 ? 1 : 0);
  uint32_t nq = (q.pred_l0 ? 1 : 0) + (q.pred_l1 ? 1 : 0);
  if (np != nq) return 1;
  if (np == 0) return 0;
  if (np == 1) {
    int32_t p_poc = p.pred_l0 ? p.poc_l0 : p.poc_l1;
    const int32_t* p_mv = p.pred_l0 ? p.mv_l0 : p.mv_l1;
    int32_t q_poc = q.pred_l0 ? q.poc_l0 : q.poc_l1;
    const int32_t* q_mv = q.pred_l0 ? q.mv_l0 : q.mv_l1;
    if (p_poc != q_poc) return 1;
    return diff_ge8(p_mv, q_mv);
  }
  int same_ordered = p.poc_l0 == q.poc_l0 && p.poc_l1 == q.poc_l1;
  int same_swapped = p.poc_l0 == q.poc_l1 && p.poc_l1 == q.poc_l0;
  if (!(same_ordered || same_swapped)) return 1;
  if (p.poc_l0 != p.poc_l1) {
    if (same_ordered && (diff_ge8(p.mv_l0, q.mv_l0) || diff_ge8(p.mv_l1, q.mv_l1))) return 1;
    if (!same_ordered && (diff_ge8(p.mv_l0, q.mv_l1) || diff_ge8(p.mv_l1, q.mv_l0))) return 1;
    return 0;
  }
  int straight = diff_ge8(p.mv_l0, q.mv_l0) || diff_ge8(p.mv_l1, q.mv_l1);
  int crossed = diff_ge8(p.mv_l0, q.mv_l1) || diff_ge8(p.mv_l1, q.mv_l0);
  return straight && crossed;
}

/* derive_bs, src/deblock.rs:1055-1119 */
static int32_t derive_bs(uint32_t c_idx, const vvclf_deblock_cu* p, const vvclf_deblock_cu* q, const vvclf_mv_grid* g,
                         int32_t px, int32_t py, int32_t qx, int32_t qy, int is_tb, int idc2) {
  if (c_idx == 0 && p->bdpcm_luma && q->bdpcm_luma) return 0;
  if (c_idx != 0 && p->bdpcm_chroma && q->bdpcm_chroma) return 0;
  if (p->intra || q->intra) return 2;
  if (p->ciip || q->ciip) return 2;
  if (is_tb) {
    if (oracle_tu_coded_at(p, c_idx, px, py) || oracle_tu_coded_at(q, c_idx, qx, qy)) return 1;
  }
  if (c_idx != 0 || !idc2) return 0;
  if (p->ibc != q->ibc) return 1;
  if (p->ibc && q->ibc) {
    if (g) {
      vvclf_mv_cell a = mv_at(g, px, py), b = mv_at(g, qx, qy);
      if (absi(a.mv_l0[0] - b.mv_l0[0]) >= 8 || absi(a.mv_l0[1] - b.mv_l0[1]) >= 8) return 1;
    }
    return 0;
  }
  if (g) {
    if (bs_motion_rule(mv_at(g, px, py), mv_at(g, qx, qy))) return 1;
  }
  return 0;
}

/* compute_thresholds_luma, src/deblock.rs:1122-1136 */
static void thresholds_luma(const plane_ctx* pc, const vvclf_deblock_cu* p, const vvclf_deblock_cu* q, int32_t bs,
                            int32_t* beta, int32_t* tc) {
  int32_t qp = asr(p->qp_y + q->qp_y + 1, 1);
  int32_t qb = clampi(qp + pc->beta_offset_div2 * 2, 0, 63);
  *beta = oracle_scale_beta(oracle_beta_prime(qb), pc->bit_depth);
  int32_t qt = clampi(qp + 2 * (bs - 1) + pc->tc_offset_div2 * 2, 0, 65);
  *tc = oracle_scale_tc(oracle_tc_prime(qt), pc->bit_depth);
}

/* compute_thresholds_chroma, src/deblock.rs:1144-1174 */
static int32_t chroma_qp_pick(const plane_ctx* pc, const vvclf_deblock_cu* c) {
  if (c->qp_c[0] == INT32_MIN) return clampi(c->qp_y + pc->qp_offset, 0, 63);
  if (c->joint_cbcr2) return c->qp_c[2];
  uint32_t i = pc->c_idx == 0 ? 0 : pc->c_idx - 1;
  return c->qp_c[i > 1 ? 1 : i];
}
static void thresholds_chroma(const plane_ctx* pc, const vvclf_deblock_cu* p, const vvclf_deblock_cu* q, int32_t bs,
                              int32_t* beta, int32_t* tc) {
  int32_t qpc = asr(chroma_qp_pick(pc, p) + chroma_qp_pick(pc, q) + 1, 1);
  int32_t qb = clampi(qpc + pc->beta_offset_div2 * 2, 0, 63);
  *beta = oracle_scale_beta(oracle_beta_prime(qb), pc->bit_depth);
  int32_t qt = clampi(qpc + 2 * (bs - 1) + pc->tc_offset_div2 * 2, 0, 65);
  *tc = oracle_scale_tc(oracle_tc_prime(qt), pc->bit_depth);
}

/* Sample accessors for one luma segment: p_i,k / q_j,k (eqs. 1268-1271). */
typedef struct {
  const vvclf_plane* pl;
  int32_t cx, cy;
  int vert;
} seg_rd;
static inline int32_t RP(const seg_rd* s, int32_t i, int32_t k) {
  return s->vert ? rd(s->pl, s->cx - i - 1, s->cy + k) : rd(s->pl, s->cx + k, s->cy - i - 1);
}
static inline int32_t RQ(const seg_rd* s, int32_t j, int32_t k) {
  return s->vert ? rd(s->pl, s->cx + j, s->cy + k) : rd(s->pl, s->cx + k, s->cy + j);
}

typedef struct {
  uint32_t d_e, d_ep, d_eq, mfl_p, mfl_q;
} luma_dec;

/* luma_edge_decision, src/deblock.rs:1224-1409 */
static luma_dec luma_edge_decision(const vvclf_plane* pl, int32_t cx, int32_t cy, int32_t beta, int32_t tc,
                                   uint32_t mfl_p, uint32_t mfl_q, int vert, int ctb_row_edge) {
  seg_rd s = {pl, cx, cy, vert};
  int32_t dp0 = absi(RP(&s, 2, 0) - 2 * RP(&s, 1, 0) + RP(&s, 0, 0));
  int32_t dp3 = absi(RP(&s, 2, 3) - 2 * RP(&s, 1, 3) + RP(&s, 0, 3));
  int32_t dq0 = absi(RQ(&s, 2, 0) - 2 * RQ(&s, 1, 0) + RQ(&s, 0, 0));
  int32_t dq3 = absi(RQ(&s, 2, 3) - 2 * RQ(&s, 1, 3) + RQ(&s, 0, 3));
  int32_t sp0 = absi(RP(&s, 3, 0) - RP(&s, 0, 0));
  int32_t sq0 = absi(RQ(&s, 0, 0) - RQ(&s, 3, 0));
  int32_t spq0 = absi(RP(&s, 0, 0) - RQ(&s, 0, 0));
  int32_t sp3 = absi(RP(&s, 3, 3) - RP(&s, 0, 3));
  int32_t sq3 = absi(RQ(&s, 0, 3) - RQ(&s, 3, 3));
  int32_t spq3 = absi(RP(&s, 0, 3) - RQ(&s, 0, 3));
  int side_p_large = mfl_p > 3;
  int side_q_large = mfl_q > 3;
  if (ctb_row_edge) side_p_large = 0;
  int d_sam0 = 0, d_sam3 = 0;
  if (side_p_large || side_q_large) {
    int32_t dp0_l, dp3_l, dq0_l, dq3_l;
    if (side_p_large) {
      dp0_l = asr(dp0 + absi(RP(&s, 5, 0) - 2 * RP(&s, 4, 0) + RP(&s, 3, 0)) + 1, 1);
      dp3_l = asr(dp3 + absi(RP(&s, 5, 3) - 2 * RP(&s, 4, 3) + RP(&s, 3, 3)) + 1, 1);
    } else {
      mfl_p = 3;
      dp0_l = dp0;
      dp3_l = dp3;
    }
    if (side_q_large) {
      dq0_l = asr(dq0 + absi(RQ(&s, 5, 0) - 2 * RQ(&s, 4, 0) + RQ(&s, 3, 0)) + 1, 1);
      dq3_l = asr(dq3 + absi(RQ(&s, 5, 3) - 2 * RQ(&s, 4, 3) + RQ(&s, 3, 3)) + 1, 1);
    } else {
      dq0_l = dq0;
      dq3_l = dq3;
    }
    int32_t sp0_l = sp0, sp3_l = sp3, sq0_l = sq0, sq3_l = sq3;
    if (mfl_p == 7) {
      sp0_l = sp0 + absi(RP(&s, 7, 0) - RP(&s, 6, 0) - RP(&s, 5, 0) + RP(&s, 4, 0));
      sp3_l = sp3 + absi(RP(&s, 7, 3) - RP(&s, 6, 3) - RP(&s, 5, 3) + RP(&s, 4, 3));
    }
    if (mfl_q == 7) {
      sq0_l = sq0 + absi(RQ(&s, 4, 0) - RQ(&s, 5, 0) - RQ(&s, 6, 0) + RQ(&s, 7, 0));
      sq3_l = sq3 + absi(RQ(&s, 4, 3) - RQ(&s, 5, 3) - RQ(&s, 6, 3) + RQ(&s, 7, 3));
    }
    int32_t dl = (dp0_l + dq0_l) + (dp3_l + dq3_l);
    if (dl < beta) {
      int32_t s_thr1 = asr(3 * beta, 5), s_thr2 = asr(beta, 4);
      for (int r = 0; r < 2; r++) {
        int32_t k = r ? 3 : 0;
        int32_t dpq_l = r ? dp3_l + dq3_l : dp0_l + dq0_l;
        int32_t sp_l = r ? sp3_l : sp0_l, sq_l = r ? sq3_l : sq0_l, spq = r ? spq3 : spq0;
        int32_t sp_in = side_p_large ? asr(sp_l + absi(RP(&s, 3, k) - RP(&s, (int32_t)mfl_p, k)) + 1, 1) : sp_l;
        int32_t sq_in = side_q_large ? asr(sq_l + absi(RQ(&s, 3, k) - RQ(&s, (int32_t)mfl_q, k)) + 1, 1) : sq_l;
        int ok = 2 * dpq_l < s_thr2 && sp_in + sq_in < s_thr1 && spq < asr(5 * tc + 1, 1);
        if (r) d_sam3 = ok;
        else d_sam0 = ok;
      }
    }
  }
  luma_dec out = {0, 0, 0, mfl_p, mfl_q};
  if (d_sam0 && d_sam3) {
    out.d_e = 3;
    out.d_ep = 1;
    out.d_eq = 1;
    return out;
  }
  int32_t dpq0 = dp0 + dq0, dpq3 = dp3 + dq3, dp = dp0 + dp3, dq = dq0 + dq3, d = dpq0 + dpq3;
  int s_sam0 = 0, s_sam3 = 0;
  if (d < beta && mfl_p > 2 && mfl_q > 2) {
    int32_t t1 = asr(beta, 3), t2 = asr(beta, 2), tcl = asr(5 * tc + 1, 1);
    s_sam0 = 2 * dpq0 < t2 && sp0 + sq0 < t1 && spq0 < tcl;
    s_sam3 = 2 * dpq3 < t2 && sp3 + sq3 < t1 && spq3 < tcl;
  }
  if (d < beta) {
    out.d_e = 1;
    if (s_sam0 && s_sam3) {
      out.d_e = 2;
      mfl_p = 3;
      mfl_q = 3;
    }
    if (mfl_p > 1 && mfl_q > 1) {
      if (dp < asr(beta + asr(beta, 1), 3)) out.d_ep = 1;
      if (dq < asr(beta + asr(beta, 1), 3)) out.d_eq = 1;
    }
    if (out.d_e == 1) {
      mfl_p = 1 + out.d_ep;
      mfl_q = 1 + out.d_eq;
    }
  }
  out.mfl_p = mfl_p;
  out.mfl_q = mfl_q;
  return out;
}

/* long_luma_apply, src/deblock.rs:1519-1673 */
static void long_luma_apply(vvclf_plane* pl, int32_t cx, int32_t cy, int32_t tc, uint32_t mfl_p, uint32_t mfl_q,
                            int vert, int plt_p, int plt_q) {
  uint32_t bd = pl->bit_depth;
  int32_t p[8] = {0}, q[8] = {0};
  for (uint32_t i = 0; i <= mfl_p && i < 8; i++)
    p[i] = vert ? rd(pl, cx - (int32_t)i - 1, cy) : rd(pl, cx, cy - (int32_t)i - 1);
  for (uint32_t j = 0; j <= mfl_q && j < 8; j++)
    q[j] = vert ? rd(pl, cx + (int32_t)j, cy) : rd(pl, cx, cy + (int32_t)j);
  int32_t rm;
  if (mfl_p == 5 && mfl_q == 5) {
    rm = asr(p[4] + p[3] + 2 * (p[2] + p[1] + p[0] + q[0] + q[1] + q[2]) + q[3] + q[4] + 8, 4);
  } else if (mfl_p == mfl_q) {
    rm = asr(p[6] + p[5] + p[4] + p[3] + p[2] + p[1] + 2 * (p[0] + q[0]) + q[1] + q[2] + q[3] + q[4] + q[5] + q[6] + 8,
             4);
  } else if ((mfl_p == 7 && mfl_q == 5) || (mfl_p == 5 && mfl_q == 7)) {
    rm = asr(p[5] + p[4] + p[3] + p[2] + 2 * (p[1] + p[0] + q[0] + q[1]) + q[2] + q[3] + q[4] + q[5] + 8, 4);
  } else if ((mfl_p == 5 && mfl_q == 3) || (mfl_p == 3 && mfl_q == 5)) {
    rm = asr(p[3] + p[2] + p[1] + p[0] + q[0] + q[1] + q[2] + q[3] + 4, 3);
  } else if (mfl_p == 3 && mfl_q == 7) {
    rm = asr(2 * (p[2] + p[1] + p[0] + q[0]) + p[0] + p[1] + q[1] + q[2] + q[3] + q[4] + q[5] + q[6] + 8, 4);
  } else {
    rm = asr(p[6] + p[5] + p[4] + p[3] + p[2] + p[1] + 2 * (q[2] + q[1] + q[0] + p[0]) + q[0] + q[1] + 8, 4);
  }
  int32_t ref_p = asr(p[mfl_p] + p[mfl_p - 1] + 1, 1);
  int32_t ref_q = asr(q[mfl_q] + q[mfl_q - 1] + 1, 1);
  static const int32_t F7[7] = {59, 50, 41, 32, 23, 14, 5}, F5[5] = {58, 45, 32, 19, 6}, F3[3] = {53, 32, 11};
  static const int32_t T7[7] = {6, 5, 4, 3, 2, 1, 1}, T5[5] = {6, 5, 4, 3, 2}, T3[3] = {6, 4, 2};
  const int32_t *fi, *tp, *gj, *tq;
  if (mfl_p == 7) {
    fi = F7;
    tp = T7;
  } else if (mfl_p == 5) {
    fi = F5;
    tp = T5;
  } else {
    fi = F3;
    tp = T3;
  }
  if (mfl_q == 7) {
    gj = F7;
    tq = T7;
  } else if (mfl_q == 5) {
    gj = F5;
    tq = T5;
  } else {
    gj = F3;
    tq = T3;
  }
  if (!plt_p)
    for (uint32_t i = 0; i < mfl_p; i++) {
      int32_t lo = p[i] - asr(tc * tp[i], 1), hi = p[i] + asr(tc * tp[i], 1);
      int32_t v = clampi(asr(rm * fi[i] + ref_p * (64 - fi[i]) + 32, 6), lo, hi);
      if (vert) wr(pl, cx - (int32_t)i - 1, cy, v, bd);
      else wr(pl, cx, cy - (int32_t)i - 1, v, bd);
    }
  if (!plt_q)
    for (uint32_t j = 0; j < mfl_q; j++) {
      int32_t lo = q[j] - asr(tc * tq[j], 1), hi = q[j] + asr(tc * tq[j], 1);
      int32_t v = clampi(asr(rm * gj[j] + ref_q * (64 - gj[j]) + 32, 6), lo, hi);
      if (vert) wr(pl, cx + (int32_t)j, cy, v, bd);
      else wr(pl, cx, cy + (int32_t)j, v, bd);
    }
}

/* short_luma_apply, src/deblock.rs:1680-1776 */
static void short_luma_apply(vvclf_plane* pl, int32_t cx, int32_t cy, int32_t tc, uint32_t d_e, uint32_t d_ep,
                             uint32_t d_eq, int vert, int plt_p, int plt_q) {
  uint32_t bd = pl->bit_depth;
#define SP(i) (vert ? rd(pl, cx - (i)-1, cy) : rd(pl, cx, cy - (i)-1))
#define SQ(j) (vert ? rd(pl, cx + (j), cy) : rd(pl, cx, cy + (j)))
#define WP(i, v) (vert ? wr(pl, cx - (i)-1, cy, (v), bd) : wr(pl, cx, cy - (i)-1, (v), bd))
#define WQ(j, v) (vert ? wr(pl, cx + (j), cy, (v), bd) : wr(pl, cx, cy + (j), (v), bd))
  int32_t p3 = SP(3), p2 = SP(2), p1 = SP(1), p0 = SP(0);
  int32_t q0 = SQ(0), q1 = SQ(1), q2 = SQ(2), q3 = SQ(3);
  if (d_e == 2) {
    int32_t p0n = clampi(asr(p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4, 3), p0 - 3 * tc, p0 + 3 * tc);
    int32_t p1n = clampi(asr(p2 + p1 + p0 + q0 + 2, 2), p1 - 2 * tc, p1 + 2 * tc);
    int32_t p2n = clampi(asr(2 * p3 + 3 * p2 + p1 + p0 + q0 + 4, 3), p2 - tc, p2 + tc);
    int32_t q0n = clampi(asr(p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4, 3), q0 - 3 * tc, q0 + 3 * tc);
    int32_t q1n = clampi(asr(p0 + q0 + q1 + q2 + 2, 2), q1 - 2 * tc, q1 + 2 * tc);
    int32_t q2n = clampi(asr(p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4, 3), q2 - tc, q2 + tc);
    if (!plt_p) {
      WP(0, p0n);
      WP(1, p1n);
      WP(2, p2n);
    }
    if (!plt_q) {
      WQ(0, q0n);
      WQ(1, q1n);
      WQ(2, q2n);
    }
  } else {
    int32_t delta_raw = asr(9 * (q0 - p0) - 3 * (q1 - p1) + 8, 4);
    if (absi(delta_raw) < tc * 10) {
      int32_t delta = clampi(delta_raw, -tc, tc);
      if (!plt_p) WP(0, p0 + delta);
      if (!plt_q) WQ(0, q0 - delta);
      if (d_ep == 1 && !plt_p) {
        int32_t dp = clampi(asr(asr(p2 + p0 + 1, 1) - p1 + delta, 1), -asr(tc, 1), asr(tc, 1));
        WP(1, p1 + dp);
      }
      if (d_eq == 1 && !plt_q) {
        int32_t dq = clampi(asr(asr(q2 + q0 + 1, 1) - q1 - delta, 1), -asr(tc, 1), asr(tc, 1));
        WQ(1, q1 + dq);
      }
    }
  }
#undef SP
#undef SQ
#undef WP
#undef WQ
}

/* run_luma_filter, src/deblock.rs:1416-1508 */
static void run_luma_filter(vvclf_plane* pl, int32_t cx, int32_t cy, int32_t beta, int32_t tc, uint32_t mfl_p,
                            uint32_t mfl_q, int vert, int ctb_row_edge, int plt_p, int plt_q) {
  if (tc == 0) return;
  luma_dec dec = luma_edge_decision(pl, cx, cy, beta, tc, mfl_p, mfl_q, vert, ctb_row_edge);
  if (dec.d_e == 0) return;
  for (int32_t k = 0; k < 4; k++) {
    int32_t x = vert ? cx : cx + k, y = vert ? cy + k : cy;
    if (dec.d_e == 3) long_luma_apply(pl, x, y, tc, dec.mfl_p, dec.mfl_q, vert, plt_p, plt_q);
    else short_luma_apply(pl, x, y, tc, dec.d_e, dec.d_ep, dec.d_eq, vert, plt_p, plt_q);
  }
}

/* Chroma accessor along the edge normal: a(i) with i<0 on the P side. */
#define CR(o) (vert ? rd(pl, cx + (o), cy) : rd(pl, cx, cy + (o)))
#define CW(o, v) (vert ? wr(pl, cx + (o), cy, (v), bd) : wr(pl, cx, cy + (o), (v), bd))

/* chroma_strong_decision_v/h, src/deblock.rs:1932-1982 */
static int chroma_strong_decision(const vvclf_plane* pl, int32_t cx, int32_t cy, int32_t beta, int32_t tc, int vert,
                                  int short_p) {
  int32_t p1 = CR(-2);
  int32_t p3 = short_p ? p1 : CR(-4);
  int32_t p2 = short_p ? p1 : CR(-3);
  int32_t p0 = CR(-1), q0 = CR(0), q1 = CR(1), q2 = CR(2), q3 = CR(3);
  int32_t dp = absi(p2 - 2 * p1 + p0), dq = absi(q2 - 2 * q1 + q0);
  int32_t dpq = 2 * (dp + dq);
  int32_t edge = absi(p3 - p0) + absi(q0 - q3);
  int32_t centre = absi(p0 - q0);
  return dpq < asr(beta, 2) && edge < asr(beta, 3) && centre < asr(5 * tc + 1, 1);
}

/* chroma_strong_apply_v/h, src/deblock.rs:1986-2074 */
static void chroma_strong_apply(vvclf_plane* pl, int32_t cx, int32_t cy, int32_t tc, uint32_t bd, int vert,
                                int short_p, int plt_p, int plt_q) {
  int32_t p1 = CR(-2), p0 = CR(-1), q0 = CR(0), q1 = CR(1), q2 = CR(2), q3 = CR(3);
  if (short_p) {
    int32_t p0n = clampi(asr(3 * p1 + 2 * p0 + q0 + q1 + q2 + 4, 3), p0 - tc, p0 + tc);
    int32_t q0n = clampi(asr(2 * p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4, 3), q0 - tc, q0 + tc);
    int32_t q1n = clampi(asr(p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4, 3), q1 - tc, q1 + tc);
    int32_t q2n = clampi(asr(p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4, 3), q2 - tc, q2 + tc);
    if (!plt_p) CW(-1, p0n);
    if (!plt_q) {
      CW(0, q0n);
      CW(1, q1n);
      CW(2, q2n);
    }
    return;
  }
  int32_t p3 = CR(-4), p2 = CR(-3);
  int32_t p0n = clampi(asr(p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2 + 4, 3), p0 - tc, p0 + tc);
  int32_t p1n = clampi(asr(2 * p3 + p2 + 2 * p1 + p0 + q0 + q1 + 4, 3), p1 - tc, p1 + tc);
  int32_t p2n = clampi(asr(3 * p3 + 2 * p2 + p1 + p0 + q0 + 4, 3), p2 - tc, p2 + tc);
  int32_t q0n = clampi(asr(p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4, 3), q0 - tc, q0 + tc);
  int32_t q1n = clampi(asr(p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4, 3), q1 - tc, q1 + tc);
  int32_t q2n = clampi(asr(p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4, 3), q2 - tc, q2 + tc);
  if (!plt_p) {
    CW(-1, p0n);
    CW(-2, p1n);
    CW(-3, p2n);
  }
  if (!plt_q) {
    CW(0, q0n);
    CW(1, q1n);
    CW(2, q2n);
  }
}

/* run_chroma_filter_v / run_chroma_filter_h, src/deblock.rs:1803-1926 */
static void run_chroma_filter(vvclf_plane* pl, int32_t ex, int32_t ey, int32_t beta, int32_t tc, int32_t bs,
                              uint32_t mfl_p, uint32_t mfl_q, int vert, int plt_p, int plt_q) {
  if (tc == 0) return;
  if (mfl_p == 1 && mfl_q == 1 && bs != 2) return;
  uint32_t bd = pl->bit_depth;
  int strong_eligible, short_p;
  if (vert) {
    strong_eligible = mfl_p == 3 && mfl_q == 3;
    short_p = 0;
  } else {
    strong_eligible = mfl_q == 3 && (mfl_p == 3 || mfl_p == 1);
    short_p = mfl_p == 1;
  }
  if (strong_eligible) {
    int32_t cx, cy;
    cx = ex;
    cy = ey;
    int dec0 = chroma_strong_decision(pl, cx, cy, beta, tc, vert, short_p);
    int dec1 = vert ? chroma_strong_decision(pl, ex, ey + 1, beta, tc, vert, short_p)
                    : chroma_strong_decision(pl, ex + 1, ey, beta, tc, vert, short_p);
    if (dec0 && dec1) {
      for (int32_t k = 0; k <= 1; k++) {
        if (vert) chroma_strong_apply(pl, ex, ey + k, tc, bd, vert, short_p, plt_p, plt_q);
        else chroma_strong_apply(pl, ex + k, ey, tc, bd, vert, short_p, plt_p, plt_q);
      }
      return;
    }
  }
  for (int32_t k = 0; k < 2; k++) {
    int32_t cx = vert ? ex : ex + k, cy = vert ? ey + k : ey;
    int32_t p1 = CR(-2), p0 = CR(-1), q0 = CR(0), q1 = CR(1);
    int32_t delta = clampi(asr((q0 - p0) * 4 + p1 - q1 + 4, 3), -tc, tc);
    if (!plt_p) CW(-1, p0 + delta);
    if (!plt_q) CW(0, q0 - delta);
  }
}
#undef CR
#undef CW

static int contains_u32(const uint32_t* v, uint32_t n, uint32_t x) {
  for (uint32_t i = 0; i < n; i++)
    if (v[i] == x) return 1;
  return 0;
}

/* deblock_cu_dir, src/deblock.rs:707-983 */
static void deblock_cu_dir(plane_ctx* pc, const vvclf_deblock_cu* cus, const cu_grid* grid, const vvclf_mv_grid* mv,
                           uint32_t idx, int vert, const uint32_t* nf, uint32_t n_nf) {
  const vvclf_deblock_cu* cu = &cus[idx];
  uint32_t c_idx = pc->c_idx;
  if (c_idx != 0) {
    uint32_t lead_c = vert ? cu->x / pc->sub_w : cu->y / pc->sub_h;
    if (lead_c % 8 != 0) return;
  }
  uint32_t n_cb = vert ? cu->w : cu->h, n_other = vert ? cu->h : cu->w;
  uint32_t lead_luma = vert ? cu->x : cu->y;
  int filter_edge = lead_luma != 0 && !contains_u32(nf, n_nf, lead_luma);
  uint32_t grid_step = c_idx == 0 ? 4u : 8u * (vert ? pc->sub_w : pc->sub_h);
  uint32_t seg_step = c_idx == 0 ? 4u : 2u * (vert ? pc->sub_h : pc->sub_w);
  uint32_t num_sb_axis = 1;
  if (c_idx == 0) {
    num_sb_axis = cu->has_num_sb ? (vert ? cu->num_sb_x : cu->num_sb_y) : 1;
    num_sb_axis = maxu(num_sb_axis, 1);
  }
  uint32_t sb = maxu(8, n_cb / num_sb_axis);
  uint32_t n8 = n_cb / 8;
  uint32_t last_sb_edge = minu(maxu(n8, 1) - 1, num_sb_axis - 1);
  uint32_t ctb_h_c = pc->ctb_size_y / maxu(pc->sub_h, 1);

  for (uint32_t off = 0; off < n_cb; off += grid_step) {
    int is_tb;
    if (off == 0) is_tb = filter_edge;
    else if (c_idx == 0) is_tb = oracle_is_luma_tb_boundary(cu, off, vert);
    else is_tb = oracle_is_chroma_tb_boundary(cu, off, vert);
    int is_sb = c_idx == 0 && off % sb == 0 && off / sb <= last_sb_edge && (off > 0 || filter_edge);
    if (!(is_tb || is_sb)) continue;
    for (uint32_t k = 0; k < n_other; k += seg_step) {
      int32_t qx = vert ? (int32_t)(cu->x + off) : (int32_t)(cu->x + k);
      int32_t qy = vert ? (int32_t)(cu->y + k) : (int32_t)(cu->y + off);
      int32_t px = vert ? qx - 1 : qx, py = vert ? qy : qy - 1;
      int32_t p_idx = off == 0 ? cu_grid_at(grid, px, py) : (int32_t)idx;
      if (p_idx < 0) continue;
      if (off == 0 && (uint32_t)p_idx == idx) continue;
      const vvclf_deblock_cu* p_cu = &cus[p_idx];
      const vvclf_deblock_cu* q_cu = cu;
      int32_t bs = derive_bs(c_idx, p_cu, q_cu, mv, px, py, qx, qy, is_tb, is_sb);
      if (bs == 0) continue;
      if (c_idx == 0) {
        uint32_t mfl_p = 0, mfl_q = 0;
        if (is_tb) {
          uint32_t wq = oracle_luma_tb_len(q_cu, qx, qy, vert), wp = oracle_luma_tb_len(p_cu, px, py, vert);
          if (wp <= 4 || wq <= 4) {
            mfl_p = mfl_q = 1;
          } else {
            mfl_p = wp >= 32 ? 7 : 3;
            mfl_q = wq >= 32 ? 7 : 3;
          }
        }
        if (is_sb) {
          if (off == 0) {
            if (num_sb_axis > 1) mfl_q = minu(mfl_q, 5);
            if (p_cu->has_num_sb) mfl_p = minu(mfl_p, 5);
          } else {
            int64_t zs[4] = {(int64_t)off - 4, (int64_t)off + 4, (int64_t)off - (int64_t)sb, (int64_t)off + (int64_t)sb};
            int tb_at[4];
            for (int t = 0; t < 4; t++) {
              int64_t z = zs[t];
              if (z < 0 || (uint64_t)z >= n_cb) tb_at[t] = 0;
              else if (z == 0) tb_at[t] = filter_edge;
              else tb_at[t] = oracle_is_luma_tb_boundary(cu, (uint32_t)z, vert);
            }
            if (is_tb) {
              mfl_p = minu(mfl_p, 5);
              mfl_q = minu(mfl_q, 5);
            } else if (tb_at[0] || tb_at[1]) {
              mfl_p = mfl_q = 1;
            } else if (off / sb == 1 || off / sb == maxu(n8, 1) - 1 || tb_at[2] || tb_at[3]) {
              mfl_p = mfl_q = 2;
            } else {
              mfl_p = mfl_q = 3;
            }
          }
        }
        int ctb_row_edge = !vert && off == 0 && pc->ctb_size_y > 0 && cu->y % pc->ctb_size_y == 0;
        int32_t beta, tc;
        thresholds_luma(pc, p_cu, q_cu, bs, &beta, &tc);
        run_luma_filter(pc->plane, qx, qy, beta, tc, mfl_p, mfl_q, vert, ctb_row_edge, p_cu->plt, q_cu->plt);
      } else {
        uint32_t sub_axis = vert ? pc->sub_w : pc->sub_h;
        uint32_t wq = oracle_chroma_tb_len(q_cu, qx, qy, vert, sub_axis);
        uint32_t wp = oracle_chroma_tb_len(p_cu, px, py, vert, sub_axis);
        uint32_t mfl_p, mfl_q;
        if (wp >= 8 && wq >= 8) {
          int ctb_row = !vert && off == 0 && ctb_h_c > 0 && (cu->y / pc->sub_h) % ctb_h_c == 0;
          mfl_p = ctb_row ? 1 : 3;
          mfl_q = 3;
        } else {
          mfl_p = mfl_q = 1;
        }
        int32_t beta, tc;
        thresholds_chroma(pc, p_cu, q_cu, bs, &beta, &tc);
        int32_t ex = qx / (int32_t)pc->sub_w, ey = qy / (int32_t)pc->sub_h;
        run_chroma_filter(pc->plane, ex, ey, beta, tc, bs, mfl_p, mfl_q, vert, p_cu->plt, q_cu->plt);
      }
    }
  }
}

/* apply_deblocking_clipped, src/deblock.rs:497-617 */
void oracle_apply_deblocking_clipped(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                                     const vvclf_deblock_cu* chroma_cus, uint32_t n_chroma_cus,
                                     const vvclf_mv_grid* mv_grid, const vvclf_deblock_params* params,
                                     uint32_t cfi, const uint32_t* nfc, uint32_t n_nfc, const uint32_t* nfr,
                                     uint32_t n_nfr) {
  if (params->disabled) return;
  cu_grid grid, cgrid;
  cu_grid_build(&grid, out->luma.width, out->luma.height, cus, n_cus);
  if (chroma_cus) cu_grid_build(&cgrid, out->luma.width, out->luma.height, chroma_cus, n_chroma_cus);
  uint32_t ctb = 1u << params->ctb_log2_size_y;
  plane_ctx l = {&out->luma, 0, 1, 1, params->luma_beta_offset_div2, params->luma_tc_offset_div2, 0,
                 params->bit_depth, ctb};
  for (uint32_t i = 0; i < n_cus; i++) deblock_cu_dir(&l, cus, &grid, mv_grid, i, 1, nfc, n_nfc);
  for (uint32_t i = 0; i < n_cus; i++) deblock_cu_dir(&l, cus, &grid, mv_grid, i, 0, nfr, n_nfr);
  if (cfi != 0) {
    const vvclf_deblock_cu* ccus = chroma_cus ? chroma_cus : cus;
    uint32_t nc = chroma_cus ? n_chroma_cus : n_cus;
    const cu_grid* cg = chroma_cus ? &cgrid : &grid;
    plane_ctx cb = {&out->cb, 1, 2, 2, params->cb_beta_offset_div2, params->cb_tc_offset_div2,
                    params->chroma_qp_offset_cb, params->bit_depth, ctb};
    for (uint32_t i = 0; i < nc; i++) deblock_cu_dir(&cb, ccus, cg, NULL, i, 1, nfc, n_nfc);
    for (uint32_t i = 0; i < nc; i++) deblock_cu_dir(&cb, ccus, cg, NULL, i, 0, nfr, n_nfr);
    plane_ctx cr = {&out->cr, 2, 2, 2, params->cr_beta_offset_div2, params->cr_tc_offset_div2,
                    params->chroma_qp_offset_cr, params->bit_depth, ctb};
    for (uint32_t i = 0; i < nc; i++) deblock_cu_dir(&cr, ccus, cg, NULL, i, 1, nfc, n_nfc);
    for (uint32_t i = 0; i < nc; i++) deblock_cu_dir(&cr, ccus, cg, NULL, i, 0, nfr, n_nfr);
  }
  free(grid.cells);
  if (chroma_cus) free(cgrid.cells);
}

/* ======================================================================== */
/* SAO                                                                       */
/* ======================================================================== */

/* derive_offset_val, src/sao.rs:356-365 */
void oracle_sao_derive_offset_val(const uint32_t abs_[4], const uint32_t sign[4], uint32_t bd, int32_t out[5]) {
  int shift = (int)bd - (int)(bd < 10 ? bd : 10);
  out[0] = 0;
  for (int i = 0; i < 4; i++) out[i + 1] = (1 - 2 * (int32_t)(sign[i] & 1)) * ((int32_t)abs_[i] << shift);
}

static void sub_of(uint32_t cfi, uint32_t* sw, uint32_t* sh) {
  if (cfi == 1) {
    *sw = 2;
    *sh = 2;
  } else if (cfi == 2) {
    *sw = 2;
    *sh = 1;
  } else {
    *sw = 1;
    *sh = 1;
  }
}

/* The `span` closure of apply_sao_clipped / clip_rect_for_ctb. */
static void span(const uint32_t* bds, uint32_t n, uint32_t v, int32_t mx, int32_t* a, int32_t* b) {
  for (uint32_t i = 0; i + 1 < n; i++) {
    if (v >= bds[i] && v < bds[i + 1]) {
      *a = (int32_t)bds[i];
      *b = (int32_t)bds[i + 1] < mx ? (int32_t)bds[i + 1] : mx;
      return;
    }
  }
  *a = 0;
  *b = mx;
}

static const int32_t EO_H0[4] = {-1, 0, -1, 1}, EO_V0[4] = {0, -1, -1, -1};
static const int32_t EO_H1[4] = {1, 0, 1, -1}, EO_V1[4] = {0, 1, 1, 1};

static inline int32_t sgn(int32_t v) { return (v > 0) - (v < 0); }

/* apply_sao_ctb, src/sao.rs:504-626 */
static void sao_ctb(vvclf_plane* pl, const uint16_t* pre, const int32_t* clip, uint32_t rx, uint32_t ry, uint32_t nw,
                    uint32_t nh, uint32_t bd, const vvclf_sao_ctb* prm) {
  size_t x_ctb = (size_t)rx * nw, y_ctb = (size_t)ry * nh;
  size_t stride = pl->stride;
  int32_t pw = (int32_t)pl->width, ph = (int32_t)pl->height;
  size_t i_max = x_ctb < pl->width ? (pl->width - x_ctb < nw ? pl->width - x_ctb : nw) : 0;
  size_t j_max = y_ctb < pl->height ? (pl->height - y_ctb < nh ? pl->height - y_ctb : nh) : 0;
  int32_t maxv = (1 << bd) - 1;
  if (prm->sao_type_idx == VVCLF_SAO_EDGE) {
    uint32_t c = prm->eo_class & 3;
    for (size_t j = 0; j < j_max; j++)
      for (size_t i = 0; i < i_max; i++) {
        int32_t xs = (int32_t)(x_ctb + i), ys = (int32_t)(y_ctb + j);
        int32_t nx0 = xs + EO_H0[c], ny0 = ys + EO_V0[c], nx1 = xs + EO_H1[c], ny1 = ys + EO_V1[c];
        int outside = nx0 < 0 || ny0 < 0 || nx1 < 0 || ny1 < 0 || nx0 >= pw || ny0 >= ph || nx1 >= pw || ny1 >= ph;
        if (clip)
          outside = outside || nx0 < clip[0] || ny0 < clip[1] || nx1 < clip[0] || ny1 < clip[1] || nx0 >= clip[2] ||
                    ny0 >= clip[3] || nx1 >= clip[2] || ny1 >= clip[3];
        size_t at = (size_t)ys * stride + (size_t)xs;
        if (outside) {
          pl->samples[at] = (uint16_t)clampi(pre[at], 0, maxv);
          continue;
        }
        int32_t p = pre[at];
        int32_t n0 = pre[(size_t)ny0 * stride + nx0], n1 = pre[(size_t)ny1 * stride + nx1];
        int32_t e = 2 + sgn(p - n0) + sgn(p - n1);
        if (e <= 2) e = e == 2 ? 0 : e + 1;
        int32_t off = (e >= 1 && e <= 4) ? prm->offset_val[e] : 0;
        pl->samples[at] = (uint16_t)clampi(p + off, 0, maxv);
      }
  } else if (prm->sao_type_idx == VVCLF_SAO_BAND) {
    int32_t band_shift = (int32_t)bd - 5;
    int32_t table[32] = {0};
    for (int32_t k = 0; k < 4; k++) table[(k + prm->band_position) & 31] = k + 1;
    for (size_t j = 0; j < j_max; j++)
      for (size_t i = 0; i < i_max; i++) {
        size_t at = (y_ctb + j) * stride + (x_ctb + i);
        int32_t p = pre[at];
        int32_t b = band_shift >= 0 ? table[(p >> band_shift) & 31] : table[p & 31];
        int32_t off = (b >= 1 && b <= 4) ? prm->offset_val[b] : 0;
        pl->samples[at] = (uint16_t)clampi(p + off, 0, maxv);
      }
  }
}

static uint16_t* dup_plane(const vvclf_plane* p) {
  size_t n = (size_t)p->stride * p->height;
  uint16_t* c = (uint16_t*)malloc(n * 2 + 2);
  if (n) memcpy(c, p->samples, n * 2);
  return c;
}

/* apply_sao_clipped, src/sao.rs:389-497 */
void oracle_apply_sao_clipped(vvclf_picture* out, const vvclf_sao_picture* sp, const vvclf_sao_config* cfg,
                              const vvclf_bounds* tb) {
  size_t n = (size_t)sp->pic_width_in_ctbs_y * sp->pic_height_in_ctbs_y;
  int all_off = 1;
  for (size_t i = 0; i < n; i++) {
    const vvclf_sao_ctb_params* c = &sp->ctbs[i];
    if (c->luma.sao_type_idx || c->cb.sao_type_idx || c->cr.sao_type_idx) all_off = 0;
  }
  if (all_off) return;
  if (!cfg->luma_used && !cfg->chroma_used) return;
  uint16_t* lpre = cfg->luma_used ? dup_plane(&out->luma) : NULL;
  int chroma = cfg->chroma_used && cfg->chroma_format_idc != 0;
  uint16_t* cbpre = chroma ? dup_plane(&out->cb) : NULL;
  uint16_t* crpre = chroma ? dup_plane(&out->cr) : NULL;
  uint32_t sw, sh;
  sub_of(cfg->chroma_format_idc, &sw, &sh);
  uint32_t ctb = 1u << cfg->ctb_log2_size_y;
  for (uint32_t ry = 0; ry < sp->pic_height_in_ctbs_y; ry++)
    for (uint32_t rx = 0; rx < sp->pic_width_in_ctbs_y; rx++) {
      const vvclf_sao_ctb_params* p = &sp->ctbs[(size_t)ry * sp->pic_width_in_ctbs_y + rx];
      int32_t cy_[4], cc_[4];
      const int32_t* clip_y = NULL;
      const int32_t* clip_c = NULL;
      if (tb && tb->cols) {
        span(tb->cols, tb->n_cols, rx * ctb, (int32_t)out->luma.width, &cy_[0], &cy_[2]);
        span(tb->rows, tb->n_rows, ry * ctb, (int32_t)out->luma.height, &cy_[1], &cy_[3]);
        cc_[0] = cy_[0] / (int32_t)sw;
        cc_[1] = cy_[1] / (int32_t)sh;
        cc_[2] = (cy_[2] + (int32_t)sw - 1) / (int32_t)sw;
        cc_[3] = (cy_[3] + (int32_t)sh - 1) / (int32_t)sh;
        clip_y = cy_;
        clip_c = cc_;
      }
      if (cfg->luma_used && p->luma.sao_type_idx)
        sao_ctb(&out->luma, lpre, clip_y, rx, ry, ctb, ctb, cfg->bit_depth, &p->luma);
      if (chroma) {
        uint32_t nw = ctb / sw, nh = ctb / sh;
        if (p->cb.sao_type_idx) sao_ctb(&out->cb, cbpre, clip_c, rx, ry, nw, nh, cfg->bit_depth, &p->cb);
        if (p->cr.sao_type_idx) sao_ctb(&out->cr, crpre, clip_c, rx, ry, nw, nh, cfg->bit_depth, &p->cr);
      }
    }
  free(lpre);
  free(cbpre);
  free(crpre);
}

/* ======================================================================== */
/* ALF + CC-ALF                                                              */
/* ======================================================================== */

/* AlfFixFiltCoeff (eq. 90) and AlfClassToFiltMap (eq. 91), H.266 spec data,
 * as transcribed at src/alf_fixed.rs:32-164. */
static const int8_t FIX_COEFF[64][12] = {
    {0, 0, 2, -3, 1, -4, 1, 7, -1, 1, -1, 5},         {0, 0, 0, 0, 0, -1, 0, 1, 0, 0, -1, 2},
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0},             {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1},
    {2, 2, -7, -3, 0, -5, 13, 22, 12, -3, -3, 17},    {-1, 0, 6, -8, 1, -5, 1, 23, 0, 2, -5, 10},
    {0, 0, -1, -1, 0, -1, 2, 1, 0, 0, -1, 4},         {0, 0, 3, -11, 1, 0, -1, 35, 5, 2, -9, 9},
    {0, 0, 8, -8, -2, -7, 4, 4, 2, 1, -1, 25},        {0, 0, 1, -1, 0, -3, 1, 3, -1, 1, -1, 3},
    {0, 0, 3, -3, 0, -6, 5, -1, 2, 1, -4, 21},        {-7, 1, 5, 4, -3, 5, 11, 13, 12, -8, 11, 12},
    {-5, -3, 6, -2, -3, 8, 14, 15, 2, -7, 11, 16},    {2, -1, -6, -5, -2, -2, 20, 14, -4, 0, -3, 25},
    {3, 1, -8, -4, 0, -8, 22, 5, -3, 2, -10, 29},     {2, 1, -7, -1, 2, -11, 23, -5, 0, 2, -10, 29},
    {-6, -3, 8, 9, -4, 8, 9, 7, 14, -2, 8, 9},        {2, 1, -4, -7, 0, -8, 17, 22, 1, -1, -4, 23},
    {3, 0, -5, -7, 0, -7, 15, 18, -5, 0, -5, 27},     {2, 0, 0, -7, 1, -10, 13, 13, -4, 2, -7, 24},
    {3, 3, -13, 4, -2, -5, 9, 21, 25, -2, -3, 12},    {-5, -2, 7, -3, -7, 9, 8, 9, 16, -2, 15, 12},
    {0, -1, 0, -7, -5, 4, 11, 11, 8, -6, 12, 21},     {3, -2, -3, -8, -4, -1, 16, 15, -2, -3, 3, 26},
    {2, 1, -5, -4, -1, -8, 16, 4, -2, 1, -7, 33},     {2, 1, -4, -2, 1, -10, 17, -2, 0, 2, -11, 33},
    {1, -2, 7, -15, -16, 10, 8, 8, 20, 11, 14, 11},   {2, 2, 3, -13, -13, 4, 8, 12, 2, -3, 16, 24},
    {1, 4, 0, -7, -8, -4, 9, 9, -2, -2, 8, 29},       {1, 1, 2, -4, -1, -6, 6, 3, -1, -1, -3, 30},
    {-7, 3, 2, 10, -2, 3, 7, 11, 19, -7, 8, 10},      {0, -2, -5, -3, -2, 4, 20, 15, -1, -3, -1, 22},
    {3, -1, -8, -4, -1, -4, 22, 8, -4, 2, -8, 28},    {0, 3, -14, 3, 0, 1, 19, 17, 8, -3, -7, 20},
    {0, 2, -1, -8, 3, -6, 5, 21, 1, 1, -9, 13},       {-4, -2, 8, 20, -2, 2, 3, 5, 21, 4, 6, 1},
    {2, -2, -3, -9, -4, 2, 14, 16, 3, -6, 8, 24},     {2, 1, 5, -16, -7, 2, 3, 11, 15, -3, 11, 22},
    {1, 2, 3, -11, -2, -5, 4, 8, 9, -3, -2, 26},      {0, -1, 10, -9, -1, -8, 2, 3, 4, 0, 0, 29},
    {1, 2, 0, -5, 1, -9, 9, 3, 0, 1, -7, 20},         {-2, 8, -6, -4, 3, -9, -8, 45, 14, 2, -13, 7},
    {1, -1, 16, -19, -8, -4, -3, 2, 19, 0, 4, 30},    {1, 1, -3, 0, 2, -11, 15, -5, 1, 2, -9, 24},
    {0, 1, -2, 0, 1, -4, 4, 0, 0, 1, -4, 7},          {0, 1, 2, -5, 1, -6, 4, 10, -2, 1, -4, 10},
    {3, 0, -3, -6, -2, -6, 14, 8, -1, -1, -3, 31},    {0, 1, 0, -2, 1, -6, 5, 1, 0, 1, -5, 13},
    {3, 1, 9, -19, -21, 9, 7, 6, 13, 5, 15, 21},      {2, 4, 3, -12, -13, 1, 7, 8, 3, 0, 12, 26},
    {3, 1, -8, -2, 0, -6, 18, 2, -2, 3, -10, 23},     {1, 1, -4, -1, 1, -5, 8, 1, -1, 2, -5, 10},
    {0, 1, -1, 0, 0, -2, 2, 0, 0, 1, -2, 3},          {1, 1, -2, -7, 1, -7, 14, 18, 0, 0, -7, 21},
    {0, 1, 0, -2, 0, -7, 8, 1, -2, 0, -3, 24},        {0, 1, 1, -2, 2, -10, 10, 0, -2, 1, -7, 23},
    {0, 2, 2, -11, 2, -4, -3, 39, 7, 1, -10, 9},      {1, 0, 13, -16, -5, -6, -1, 8, 6, 0, 6, 29},
    {1, 3, 1, -6, -4, -7, 9, 6, -3, -2, 3, 33},       {4, 0, -17, -1, -1, 5, 26, 8, -2, 3, -15, 30},
    {0, 1, -2, 0, 2, -8, 12, -6, 1, 1, -6, 16},       {0, 0, 0, -1, 1, -4, 4, 0, 0, 0, -3, 11},
    {0, 1, 2, -8, 2, -6, 5, 15, 0, 2, -7, 9},         {1, -1, 12, -15, -7, -2, 3, 6, 6, -1, 7, 30},
};
static const uint8_t CLASS_TO_FILT[16][25] = {
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

/* fixed_filter_coeff_row, src/alf_fixed.rs:173-185 */
int oracle_alf_fixed_coeff_row(uint32_t set_idx, uint32_t class_idx, int32_t row[12]) {
  if (set_idx >= 16 || class_idx >= 25) return -1;
  uint32_t r = CLASS_TO_FILT[set_idx][class_idx];
  if (r >= 64) return -1;
  for (int j = 0; j < 12; j++) row[j] = FIX_COEFF[r][j];
  return 0;
}

/* resolve_clip_value, Table 8, src/alf.rs:627-644 */
int32_t oracle_alf_clip_value(uint32_t bd, uint8_t clip_idx) {
  static const uint8_t EXP[9][4] = {{8, 5, 3, 1},   {9, 6, 4, 2},   {10, 7, 5, 3},  {11, 8, 6, 4},  {12, 9, 7, 5},
                                    {13, 10, 8, 6}, {14, 11, 9, 7}, {15, 12, 10, 8}, {16, 13, 11, 9}};
  uint32_t row = (bd < 8 ? 8 : (bd > 16 ? 16 : bd)) - 8;
  return 1 << EXP[row][clip_idx & 3];
}

/* table_45, src/alf.rs:1117-1130 -> (alfShiftY, y1, y2, y3) */
void oracle_alf_table_45(int32_t y, int32_t ctb, int lbb, int32_t o[4]) {
  if (lbb) {
    if (y == ctb - 5 || y == ctb - 4) {
      o[0] = 10; o[1] = 0; o[2] = 0; o[3] = 0;
      return;
    }
    if (y == ctb - 6 || y == ctb - 3) {
      o[0] = 7; o[1] = 1; o[2] = 1; o[3] = 1;
      return;
    }
    if (y == ctb - 7 || y == ctb - 2) {
      o[0] = 7; o[1] = 1; o[2] = 2; o[3] = 2;
      return;
    }
  }
  o[0] = 7; o[1] = 1; o[2] = 2; o[3] = 3;
}

/* table_46, src/alf.rs:1133-1143 -> (alfShiftC, y1, y2) */
void oracle_alf_table_46(int32_t y, int32_t h, int lbb, int32_t o[3]) {
  if (lbb) {
    if (y == h - 2 || y == h - 3) {
      o[0] = 10; o[1] = 0; o[2] = 0;
      return;
    }
    if (y == h - 1 || y == h - 4) {
      o[0] = 7; o[1] = 1; o[2] = 1;
      return;
    }
  }
  o[0] = 7; o[1] = 1; o[2] = 2;
}

/* cc_alf_table_47, src/alf.rs:528-539 -> (yP1, yP2) */
void oracle_alf_table_47(int32_t yl, int32_t ctb, int lbb, int32_t o[2]) {
  if (lbb) {
    if (yl == ctb - 5 || yl == ctb - 4) {
      o[0] = 0; o[1] = 0;
      return;
    }
    if (yl == ctb - 6 || yl == ctb - 3) {
      o[0] = 1; o[1] = 1;
      return;
    }
  }
  o[0] = 1; o[1] = 2;
}

/* transpose_idx_table, src/alf.rs:731-742 (eqs. 1442-1445) */
void oracle_alf_transpose_idx(uint8_t t, uint8_t out[12]) {
  static const uint8_t T[4][12] = {{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
                                   {9, 4, 10, 8, 1, 5, 11, 7, 3, 0, 2, 6},
                                   {0, 3, 2, 1, 8, 7, 6, 5, 4, 9, 10, 11},
                                   {9, 8, 10, 4, 3, 7, 11, 5, 1, 0, 2, 6}};
  memcpy(out, T[t < 4 ? t : 0], 12);
}

/* sample(), src/alf.rs:1104-1109 */
static inline int32_t smp(const uint16_t* buf, size_t stride, const int32_t* r, int32_t x, int32_t y) {
  int32_t xc = clampi(x, r[0], r[2] - 1), yc = clampi(y, r[1], r[3] - 1);
  return buf[(size_t)yc * stride + xc];
}

static inline int32_t pad_y(int32_t v, int32_t vb, int top, int bot) {
  if (top && v < vb) v = vb;
  if (bot && v > vb - 1) v = vb - 1;
  return v;
}

/* derive_luma_classification, src/alf.rs:822-1010 */
void oracle_alf_classify_ctb(const uint16_t* pre, uint32_t stride, int32_t pw, int32_t ph, const int32_t* clip,
                             uint32_t rx, uint32_t ry, uint32_t ctb_size_y, uint32_t bd, uint16_t* cls) {
  (void)pw;
  int32_t x_ctb = (int32_t)(rx * ctb_size_y), y_ctb = (int32_t)(ry * ctb_size_y), ctb = (int32_t)ctb_size_y;
  uint32_t sub = ctb_size_y >> 2;
  int bottom_is_pic = y_ctb + ctb >= ph;
  int carve = !bottom_is_pic || (ph - y_ctb) > ctb - 4;
  static const uint8_t VAR_TAB[16] = {0, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 4};
  static const uint8_t TRANSPOSE_TABLE[8] = {0, 1, 0, 2, 2, 3, 1, 3};
  for (uint32_t sy = 0; sy < sub; sy++)
    for (uint32_t sx = 0; sx < sub; sx++) {
      int32_t x4 = (int32_t)sx << 2, y4 = (int32_t)sy << 2;
      int32_t min_y, max_y, ac;
      if (y4 == ctb - 8 && carve) {
        min_y = -2; max_y = 3; ac = 3;
      } else if (y4 == ctb - 4 && carve) {
        min_y = 0; max_y = 5; ac = 3;
      } else {
        min_y = -2; max_y = 5; ac = 2;
      }
      int32_t vb = y_ctb + ctb - 4;
      int has_top = y4 >= ctb - 4;
      int has_bot = (ctb - 4 - y4) > 0 && (ctb - 4 - y4) < 5;
      int32_t sh = 0, sv = 0, sd0 = 0, sd1 = 0;
      for (int32_t j = min_y; j <= max_y; j++)
        for (int32_t i = -2; i <= 5; i++) {
          if (((i % 2) + 2) % 2 != ((j % 2) + 2) % 2) continue;
          int32_t cx = x_ctb + x4 + i;
          int32_t yy = y_ctb + y4 + j;
          int32_t cy = pad_y(yy, vb, has_top, has_bot), cu = pad_y(yy - 1, vb, has_top, has_bot),
                  cd = pad_y(yy + 1, vb, has_top, has_bot);
          int32_t c2 = smp(pre, stride, clip, cx, cy) * 2;
          sh += absi(c2 - smp(pre, stride, clip, cx - 1, cy) - smp(pre, stride, clip, cx + 1, cy));
          sv += absi(c2 - smp(pre, stride, clip, cx, cu) - smp(pre, stride, clip, cx, cd));
          sd0 += absi(c2 - smp(pre, stride, clip, cx - 1, cu) - smp(pre, stride, clip, cx + 1, cd));
          sd1 += absi(c2 - smp(pre, stride, clip, cx + 1, cu) - smp(pre, stride, clip, cx - 1, cd));
        }
      int32_t hv1, hv0, dir_hv, d1, d0, dir_d;
      if (sv > sh) {
        hv1 = sv; hv0 = sh; dir_hv = 1;
      } else {
        hv1 = sh; hv0 = sv; dir_hv = 3;
      }
      if (sd0 > sd1) {
        d1 = sd0; d0 = sd1; dir_d = 0;
      } else {
        d1 = sd1; d0 = sd0; dir_d = 2;
      }
      int prefer_d = (int64_t)d1 * hv0 > (int64_t)hv1 * d0;
      int32_t hvd1 = prefer_d ? d1 : hv1, hvd0 = prefer_d ? d0 : hv0;
      int32_t dir1 = prefer_d ? dir_d : dir_hv, dir2 = prefer_d ? dir_hv : dir_d;
      int32_t dir_s = hvd1 * 2 > 9 * hvd0 ? 2 : (hvd1 > 2 * hvd0 ? 1 : 0);
      int64_t prod = (int64_t)(sh + sv) * ac;
      int64_t shifted = clampl(prod >> (bd - 1), 0, 15);
      int32_t filt = VAR_TAB[shifted];
      int32_t tr = TRANSPOSE_TABLE[dir1 * 2 + (dir2 >> 1)];
      if (dir_s != 0) filt += (((dir1 & 1) << 1) + dir_s) * 5;
      cls[sy * sub + sx] = (uint16_t)(filt | (tr << 8));
    }
}

typedef struct {
  int32_t coeff[25][12];
  uint8_t clip[25][12];
} luma_set;

/* resolve_luma_filter_set, src/alf.rs:564-595; 0 on Some, -1 on None */
static int resolve_luma_set(uint8_t idx, const vvclf_alf_aps_binding* b, luma_set* s) {
  if (idx < 16) {
    for (int c = 0; c < 25; c++) {
      int32_t row[12];
      if (oracle_alf_fixed_coeff_row(idx, c, row)) return -1;
      for (int j = 0; j < 12; j++) {
        s->coeff[c][j] = row[j];
        s->clip[c][j] = 0;
      }
    }
    return 0;
  }
  uint32_t slot = idx - 16u;
  if (!b || slot >= b->n_luma_apses || !b->luma_apses[slot]) return -1;
  const vvclf_alf_aps* a = b->luma_apses[slot];
  if (!a->alf_luma_filter_signal_flag) return -1;
  memcpy(s->coeff, a->luma_coeff, sizeof(s->coeff));
  memcpy(s->clip, a->luma_clip_idx, sizeof(s->clip));
  return 0;
}

/* apply_alf_luma_ctb, src/alf.rs:652-724 */
static void alf_luma_ctb(vvclf_plane* pl, const uint16_t* pre, const int32_t* clip, uint32_t rx, uint32_t ry,
                         uint32_t ctb, uint32_t bd, const luma_set* f, const uint16_t* cls) {
  size_t x_ctb = (size_t)rx * ctb, y_ctb = (size_t)ry * ctb, stride = pl->stride;
  size_t i_max = x_ctb < pl->width ? (pl->width - x_ctb < ctb ? pl->width - x_ctb : ctb) : 0;
  size_t j_max = y_ctb < pl->height ? (pl->height - y_ctb < ctb ? pl->height - y_ctb : ctb) : 0;
  int32_t maxv = (1 << bd) - 1;
  uint32_t sub = ctb >> 2;
  for (size_t j = 0; j < j_max; j++)
    for (size_t i = 0; i < i_max; i++) {
      int32_t xs = (int32_t)(x_ctb + i), ys = (int32_t)(y_ctb + j);
      uint16_t cl = cls[(j >> 2) * sub + (i >> 2)];
      uint32_t fidx = cl & 0xff;
      uint8_t idx[12];
      oracle_alf_transpose_idx((uint8_t)(cl >> 8), idx);
      int32_t c[12];
      for (int k = 0; k < 12; k++) c[k] = oracle_alf_clip_value(bd, f->clip[fidx][k]);
      const int32_t* fc = f->coeff[fidx];
      int32_t t[4];
      oracle_alf_table_45((int32_t)j, (int32_t)ctb, 1, t);
      int32_t y1 = t[1], y2 = t[2], y3 = t[3];
      int32_t curr = smp(pre, stride, clip, xs, ys);
#define CL(k, dx, dy) clampi(smp(pre, stride, clip, xs + (dx), ys + (dy)) - curr, -c[idx[k]], c[idx[k]])
      int32_t sum = 0;
      sum += fc[idx[0]] * (CL(0, 0, y3) + CL(0, 0, -y3));
      sum += fc[idx[1]] * (CL(1, 1, y2) + CL(1, -1, -y2));
      sum += fc[idx[2]] * (CL(2, 0, y2) + CL(2, 0, -y2));
      sum += fc[idx[3]] * (CL(3, -1, y2) + CL(3, 1, -y2));
      sum += fc[idx[4]] * (CL(4, 2, y1) + CL(4, -2, -y1));
      sum += fc[idx[5]] * (CL(5, 1, y1) + CL(5, -1, -y1));
      sum += fc[idx[6]] * (CL(6, 0, y1) + CL(6, 0, -y1));
      sum += fc[idx[7]] * (CL(7, -1, y1) + CL(7, 1, -y1));
      sum += fc[idx[8]] * (CL(8, -2, y1) + CL(8, 2, -y1));
      sum += fc[idx[9]] * (CL(9, 3, 0) + CL(9, -3, 0));
      sum += fc[idx[10]] * (CL(10, 2, 0) + CL(10, -2, 0));
      sum += fc[idx[11]] * (CL(11, 1, 0) + CL(11, -1, 0));
#undef CL
      int32_t v = curr + asr(sum + (1 << (t[0] - 1)), t[0]);
      pl->samples[(size_t)ys * stride + xs] = (uint16_t)clampi(v, 0, maxv);
    }
}

/* apply_alf_chroma_ctb, src/alf.rs:1014-1065 */
static void alf_chroma_ctb(vvclf_plane* pl, const uint16_t* pre, const int32_t* clip, uint32_t rx, uint32_t ry,
                           uint32_t ctb, uint32_t sw, uint32_t sh, uint32_t bd, const int32_t* f, const uint8_t* ci) {
  size_t w = ctb / sw, h = ctb / sh;
  size_t x0 = rx * w, y0 = ry * h, stride = pl->stride;
  size_t i_max = x0 < pl->width ? (pl->width - x0 < w ? pl->width - x0 : w) : 0;
  size_t j_max = y0 < pl->height ? (pl->height - y0 < h ? pl->height - y0 : h) : 0;
  int32_t maxv = (1 << bd) - 1;
  int32_t c[6];
  for (int k = 0; k < 6; k++) c[k] = oracle_alf_clip_value(bd, ci[k]);
  for (size_t j = 0; j < j_max; j++)
    for (size_t i = 0; i < i_max; i++) {
      int32_t xs = (int32_t)(x0 + i), ys = (int32_t)(y0 + j);
      int32_t t[3];
      oracle_alf_table_46((int32_t)j, (int32_t)h, 1, t);
      int32_t y1 = t[1], y2 = t[2];
      int32_t curr = smp(pre, stride, clip, xs, ys);
#define CL(k, dx, dy) clampi(smp(pre, stride, clip, xs + (dx), ys + (dy)) - curr, -c[k], c[k])
      int32_t sum = 0;
      sum += f[0] * (CL(0, 0, y2) + CL(0, 0, -y2));
      sum += f[1] * (CL(1, 1, y1) + CL(1, -1, -y1));
      sum += f[2] * (CL(2, 0, y1) + CL(2, 0, -y1));
      sum += f[3] * (CL(3, -1, y1) + CL(3, 1, -y1));
      sum += f[4] * (CL(4, 2, 0) + CL(4, -2, 0));
      sum += f[5] * (CL(5, 1, 0) + CL(5, -1, 0));
#undef CL
      int32_t v = curr + asr(sum + (1 << (t[0] - 1)), t[0]);
      pl->samples[(size_t)ys * stride + xs] = (uint16_t)clampi(v, 0, maxv);
    }
}

/* apply_cc_alf_ctb, src/alf.rs:447-523 */
static void cc_alf_ctb(vvclf_plane* ch, const uint16_t* lpre, size_t lstride, uint32_t lh, const int32_t* clip,
                       uint32_t rx, uint32_t ry, uint32_t ctb, uint32_t sw, uint32_t sh, uint32_t bd,
                       const int32_t* cf) {
  size_t w = ctb / sw, h = ctb / sh;
  size_t x0 = (rx * ctb) / sw, y0 = (ry * ctb) / sh;
  size_t i_max = x0 < ch->width ? (ch->width - x0 < w ? ch->width - x0 : w) : 0;
  size_t j_max = y0 < ch->height ? (ch->height - y0 < h ? ch->height - y0 : h) : 0;
  int32_t maxv = (1 << bd) - 1, half = 1 << (bd - 1);
  int32_t ph = (int32_t)lh;
  for (size_t y = 0; y < j_max; y++)
    for (size_t x = 0; x < i_max; x++) {
      int32_t xl = (int32_t)(x0 + x) * (int32_t)sw, yl = (int32_t)(y0 + y) * (int32_t)sh;
      int32_t y_line = (int32_t)y * (int32_t)sh;
      int lbb = !(yl + (int32_t)ctb >= ph && (ph - (int32_t)y0 * (int32_t)sh) <= (int32_t)ctb - 4);
      int32_t t[2];
      oracle_alf_table_47(y_line, (int32_t)ctb, lbb, t);
      int32_t yp1 = t[0], yp2 = t[1];
      size_t at = (y0 + y) * ch->stride + (x0 + x);
      int32_t cur = ch->samples[at];
      int32_t centre = smp(lpre, lstride, clip, xl, yl);
      int32_t sum = 0;
      sum += cf[0] * (smp(lpre, lstride, clip, xl, yl - yp1) - centre);
      sum += cf[1] * (smp(lpre, lstride, clip, xl - 1, yl) - centre);
      sum += cf[2] * (smp(lpre, lstride, clip, xl + 1, yl) - centre);
      sum += cf[3] * (smp(lpre, lstride, clip, xl - 1, yl + yp1) - centre);
      sum += cf[4] * (smp(lpre, lstride, clip, xl, yl + yp1) - centre);
      sum += cf[5] * (smp(lpre, lstride, clip, xl + 1, yl + yp1) - centre);
      sum += cf[6] * (smp(lpre, lstride, clip, xl, yl + yp2) - centre);
      int32_t sc = clampi(asr(sum + 64, 7), -half, half - 1);
      int suppress = sh == 1 && (y_line == (int32_t)ctb - 3 || y_line == (int32_t)ctb - 4);
      int32_t v = suppress ? cur : cur + sc;
      ch->samples[at] = (uint16_t)clampi(v, 0, maxv);
    }
}

/* clip_rect_for_ctb, src/alf.rs:1071-1095 */
static void clip_rect(uint32_t rx, uint32_t ry, uint32_t ctb, int32_t pw, int32_t ph, const vvclf_bounds* tb,
                      int32_t r[4]) {
  if (!tb || !tb->cols) {
    r[0] = 0; r[1] = 0; r[2] = pw; r[3] = ph;
    return;
  }
  span(tb->cols, tb->n_cols, rx * ctb, pw, &r[0], &r[2]);
  span(tb->rows, tb->n_rows, ry * ctb, ph, &r[1], &r[3]);
}

/* apply_alf_clipped, src/alf.rs:235-434 */
int oracle_apply_alf_clipped(vvclf_picture* out, const vvclf_alf_picture* ap, const vvclf_alf_config* cfg,
                             const vvclf_alf_aps_binding* b, const vvclf_bounds* tb) {
  if (b) {
    for (uint32_t i = 0; i < b->n_luma_apses; i++) {
      const vvclf_alf_aps* a = b->luma_apses[i];
      if (a && a->alf_luma_filter_signal_flag && a->n_luma_coeff < 25) return VVCLF_ERR_INVALID;
    }
  }
  if (!cfg->alf_enabled) return VVCLF_OK;
  size_t n = (size_t)ap->pic_width_in_ctbs_y * ap->pic_height_in_ctbs_y;
  int all_off = 1;
  for (size_t i = 0; i < n; i++) {
    const vvclf_alf_ctb* c = &ap->ctbs[i];
    if (c->luma_on || c->cb_on || c->cr_on || c->cc_cb_idc || c->cc_cr_idc) all_off = 0;
  }
  if (all_off) return VVCLF_OK;
  uint32_t ctb = 1u << cfg->ctb_log2_size_y;
  uint32_t sw, sh;
  sub_of(cfg->chroma_format_idc, &sw, &sh);
  uint16_t* lpre = dup_plane(&out->luma);
  uint16_t* cbpre = cfg->chroma_format_idc ? dup_plane(&out->cb) : NULL;
  uint16_t* crpre = cfg->chroma_format_idc ? dup_plane(&out->cr) : NULL;
  int32_t pw = (int32_t)out->luma.width, ph = (int32_t)out->luma.height;
  uint16_t* cls = (uint16_t*)malloc(sizeof(uint16_t) * (ctb / 4) * (ctb / 4) + 2);
  luma_set* set = (luma_set*)malloc(sizeof(luma_set));
  for (uint32_t ry = 0; ry < ap->pic_height_in_ctbs_y; ry++)
    for (uint32_t rx = 0; rx < ap->pic_width_in_ctbs_y; rx++) {
      const vvclf_alf_ctb* p = &ap->ctbs[(size_t)ry * ap->pic_width_in_ctbs_y + rx];
      int32_t cy_[4], cc_[4];
      clip_rect(rx, ry, ctb, pw, ph, tb, cy_);
      cc_[0] = cy_[0] / (int32_t)sw;
      cc_[1] = cy_[1] / (int32_t)sh;
      cc_[2] = (cy_[2] + (int32_t)sw - 1) / (int32_t)sw;
      cc_[3] = (cy_[3] + (int32_t)sh - 1) / (int32_t)sh;
      if (p->luma_on && resolve_luma_set(p->luma_filt_set_idx, b, set) == 0) {
        oracle_alf_classify_ctb(lpre, out->luma.stride, pw, ph, cy_, rx, ry, ctb, cfg->bit_depth, cls);
        alf_luma_ctb(&out->luma, lpre, cy_, rx, ry, ctb, cfg->bit_depth, set, cls);
      }
      if (cfg->chroma_format_idc != 0 && b && b->chroma_aps) {
        const vvclf_alf_aps* a = b->chroma_aps;
        if (cfg->cb_enabled && p->cb_on && a->alf_chroma_filter_signal_flag && p->cb_alt_idx < a->n_chroma_coeff &&
            p->cb_alt_idx < VVCLF_ALF_MAX_CHROMA_ALT)
          alf_chroma_ctb(&out->cb, cbpre, cc_, rx, ry, ctb, sw, sh, cfg->bit_depth, a->chroma_coeff[p->cb_alt_idx],
                         a->chroma_clip_idx[p->cb_alt_idx]);
        if (cfg->cr_enabled && p->cr_on && a->alf_chroma_filter_signal_flag && p->cr_alt_idx < a->n_chroma_coeff &&
            p->cr_alt_idx < VVCLF_ALF_MAX_CHROMA_ALT)
          alf_chroma_ctb(&out->cr, crpre, cc_, rx, ry, ctb, sw, sh, cfg->bit_depth, a->chroma_coeff[p->cr_alt_idx],
                         a->chroma_clip_idx[p->cr_alt_idx]);
      }
    }
  if (cfg->chroma_format_idc != 0 && b) {
    for (uint32_t ry = 0; ry < ap->pic_height_in_ctbs_y; ry++)
      for (uint32_t rx = 0; rx < ap->pic_width_in_ctbs_y; rx++) {
        const vvclf_alf_ctb* p = &ap->ctbs[(size_t)ry * ap->pic_width_in_ctbs_y + rx];
        int32_t cy_[4];
        clip_rect(rx, ry, ctb, pw, ph, tb, cy_);
        if (cfg->cc_cb_enabled && p->cc_cb_idc != 0 && b->cc_cb_aps) {
          uint32_t fi = p->cc_cb_idc - 1u;
          if (fi < b->cc_cb_aps->n_cc_cb_coeff && fi < VVCLF_ALF_MAX_CC_FILTERS)
            cc_alf_ctb(&out->cb, lpre, out->luma.stride, out->luma.height, cy_, rx, ry, ctb, sw, sh, cfg->bit_depth,
                       b->cc_cb_aps->cc_cb_coeff[fi]);
        }
        if (cfg->cc_cr_enabled && p->cc_cr_idc != 0 && b->cc_cr_aps) {
          uint32_t fi = p->cc_cr_idc - 1u;
          if (fi < b->cc_cr_aps->n_cc_cr_coeff && fi < VVCLF_ALF_MAX_CC_FILTERS)
            cc_alf_ctb(&out->cr, lpre, out->luma.stride, out->luma.height, cy_, rx, ry, ctb, sw, sh, cfg->bit_depth,
                       b->cc_cr_aps->cc_cr_coeff[fi]);
        }
      }
  }
  free(cls);
  free(set);
  free(lpre);
  free(cbpre);
  free(crpre);
  return VVCLF_OK;
}

/* ======================================================================== */
/* Driver                                                                    */
/* ======================================================================== */

/* apply_in_loop_filters_with_alf, src/ctu.rs:11822-12066 */
/* ctb_sse_y, src/alf_enc.rs:79-105: luma SSE of the w x h block at (x0, y0),
 * clipped to both planes. */
static uint64_t ctb_sse_y(const vvclf_plane* a, const vvclf_plane* b, uint32_t x0, uint32_t y0, uint32_t w,
                          uint32_t h) {
  uint64_t sse = 0;
  for (uint32_t y = 0; y < h; y++) {
    uint32_t yi = y0 + y;
    if (yi >= a->height || yi >= b->height) break;
    for (uint32_t x = 0; x < w; x++) {
      uint32_t xi = x0 + x;
      if (xi >= a->width || xi >= b->width) break;
      int64_t d = (int64_t)a->samples[(size_t)yi * a->stride + xi] - (int64_t)b->samples[(size_t)yi * b->stride + xi];
      sse += (uint64_t)(d * d);
    }
  }
  return sse;
}

/* alf_decide_and_apply_clipped, src/alf_enc.rs:185-316, and (n_aps = 1)
 * alf_decide_and_apply_with_aps_clipped, :350-470: per CTB the lowest luma
 * SSE among off, the 16 fixed sets and the APS sets (trial index 16 + i,
 * filter set 16 + i), the earlier candidate winning ties; rec's luma takes the
 * chosen trial's samples.  out: w_ctbs * h_ctbs entries; sse (optional):
 * [n_ctbs][17 + n_aps - 1 + 1] = off, sets 0..15, APS sets. */
int oracle_alf_decide_luma(const vvclf_picture* src, vvclf_picture* rec, uint32_t ctb_log2, uint32_t bd, uint32_t cfi,
                           const vvclf_alf_aps* const* apses, uint32_t n_aps, const vvclf_bounds* tb,
                           vvclf_alf_ctb* out, uint64_t* sse) {
  uint32_t ctb = 1u << ctb_log2;
  uint32_t wc = (rec->luma.width + ctb - 1) / ctb, hc = (rec->luma.height + ctb - 1) / ctb;
  size_t n = (size_t)wc * hc;
  uint32_t nt = 16 + n_aps, ncand = 1 + nt;
  uint64_t* tab = (uint64_t*)calloc(n * ncand, sizeof(uint64_t));
  for (uint32_t ry = 0; ry < hc; ry++)
    for (uint32_t rx = 0; rx < wc; rx++)
      tab[((size_t)ry * wc + rx) * ncand] = ctb_sse_y(&src->luma, &rec->luma, rx * ctb, ry * ctb, ctb, ctb);
  vvclf_alf_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.alf_enabled = 1;
  cfg.bit_depth = bd;
  cfg.ctb_log2_size_y = ctb_log2;
  cfg.chroma_format_idc = cfi;
  vvclf_alf_aps_binding b;
  memset(&b, 0, sizeof(b));
  b.luma_apses = apses;
  b.n_luma_apses = n_aps;
  size_t plane = (size_t)rec->luma.stride * rec->luma.height;
  uint16_t** trial = (uint16_t**)calloc(nt, sizeof(uint16_t*));
  vvclf_alf_ctb* pic = (vvclf_alf_ctb*)calloc(n, sizeof(vvclf_alf_ctb));
  int rc = VVCLF_OK;
  for (uint32_t t = 0; t < nt && rc == VVCLF_OK; t++) {
    vvclf_picture r = *rec;
    trial[t] = (uint16_t*)malloc(plane * 2 + 2);
    memcpy(trial[t], rec->luma.samples, plane * 2);
    r.luma.samples = trial[t];  /* chroma ALF is off in the trials: chroma is only read */
    for (size_t i = 0; i < n; i++) {
      memset(&pic[i], 0, sizeof(pic[i]));
      pic[i].luma_on = 1;
      pic[i].luma_filt_set_idx = (uint8_t)t;
    }
    vvclf_alf_picture ap = {wc, hc, pic};
    rc = oracle_apply_alf_clipped(&r, &ap, &cfg, &b, tb);
    for (uint32_t ry = 0; ry < hc; ry++)
      for (uint32_t rx = 0; rx < wc; rx++)
        tab[((size_t)ry * wc + rx) * ncand + 1 + t] = ctb_sse_y(&src->luma, &r.luma, rx * ctb, ry * ctb, ctb, ctb);
  }
  for (uint32_t ry = 0; ry < hc && rc == VVCLF_OK; ry++)
    for (uint32_t rx = 0; rx < wc; rx++) {
      size_t i = (size_t)ry * wc + rx;
      uint64_t best = tab[i * ncand];
      int best_t = -1;
      for (uint32_t t = 0; t < nt; t++)
        if (tab[i * ncand + 1 + t] < best) {
          best = tab[i * ncand + 1 + t];
          best_t = (int)t;
        }
      memset(&out[i], 0, sizeof(out[i]));
      if (best_t >= 0) {
        out[i].luma_on = 1;
        out[i].luma_filt_set_idx = (uint8_t)best_t;
        uint32_t x0 = rx * ctb, y0 = ry * ctb;
        for (uint32_t y = y0; y < y0 + ctb && y < rec->luma.height; y++)
          for (uint32_t x = x0; x < x0 + ctb && x < rec->luma.width; x++) {
            size_t at = (size_t)y * rec->luma.stride + x;
            rec->luma.samples[at] = trial[best_t][at];
          }
      }
    }
  if (sse && rc == VVCLF_OK) memcpy(sse, tab, n * ncand * sizeof(uint64_t));
  for (uint32_t t = 0; t < nt; t++) free(trial[t]);
  free(trial);
  free(pic);
  free(tab);
  return rc;
}

/* Shared tail of the chroma decisions (src/alf_enc.rs:766-806, :1092-1139):
 * per CTB the first trial strictly below the off SSE; its samples are copied
 * into the component plane and mark(ctb, k) records the choice. */
static void chroma_pick_commit(vvclf_plane* dst, uint16_t* const* trial, const uint64_t* tab, uint32_t ncand,
                               uint32_t wc, uint32_t hc, uint32_t cw, uint32_t ch, vvclf_alf_ctb* apply, uint32_t c_idx,
                               int cc) {
  for (uint32_t ry = 0; ry < hc; ry++)
    for (uint32_t rx = 0; rx < wc; rx++) {
      size_t i = (size_t)ry * wc + rx;
      uint64_t best = tab[i * ncand];
      int best_k = -1;
      for (uint32_t k = 0; k + 1 < ncand; k++)
        if (tab[i * ncand + 1 + k] < best) {
          best = tab[i * ncand + 1 + k];
          best_k = (int)k;
        }
      if (best_k < 0) continue;
      uint32_t x0 = rx * cw, y0 = ry * ch;
      for (uint32_t y = y0; y < y0 + ch && y < dst->height; y++)
        for (uint32_t x = x0; x < x0 + cw && x < dst->width; x++) {
          size_t at = (size_t)y * dst->stride + x;
          dst->samples[at] = trial[best_k][at];
        }
      vvclf_alf_ctb* e = &apply[i];
      if (cc) {
        if (c_idx == 1) e->cc_cb_idc = (uint8_t)(best_k + 1);
        else e->cc_cr_idc = (uint8_t)(best_k + 1);
      } else if (c_idx == 1) {
        e->cb_on = 1;
        e->cb_alt_idx = (uint8_t)best_k;
      } else {
        e->cr_on = 1;
        e->cr_alt_idx = (uint8_t)best_k;
      }
    }
}

/* chroma_alf_decide_and_apply_clipped, src/alf_enc.rs:628-806 (c_idx 1 Cb,
 * 2 Cr), and cc_alf_decide_and_apply_clipped, :935-1140 (cc = 1, with the
 * pre-luma-ALF plane pre_luma).  A no-op returning VVCLF_OK where the
 * reference returns early (4:0:0, no signalled filter). */
static int chroma_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma,
                         vvclf_alf_ctb* apply, const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2,
                         uint32_t bd, uint32_t cfi, const vvclf_bounds* tb, uint64_t* sse, int cc) {
  if (cfi == 0) return VVCLF_OK;
  uint32_t n;
  if (cc) {
    n = c_idx == 1 ? aps->n_cc_cb_coeff : aps->n_cc_cr_coeff;
  } else {
    if (!aps->alf_chroma_filter_signal_flag) return VVCLF_OK;
    n = aps->n_chroma_coeff;
  }
  if (n == 0) return VVCLF_OK;
  uint32_t ctb = 1u << ctb_log2, sw, sh;
  sub_of(cfi, &sw, &sh);
  uint32_t wc = (rec->luma.width + ctb - 1) / ctb, hc = (rec->luma.height + ctb - 1) / ctb;
  uint32_t cw = ctb / sw, ch = ctb / sh;
  size_t nc = (size_t)wc * hc;
  uint32_t ncand = 1 + n;
  const vvclf_plane* sp = c_idx == 1 ? &src->cb : &src->cr;
  vvclf_plane* rp = c_idx == 1 ? &rec->cb : &rec->cr;
  uint64_t* tab = (uint64_t*)calloc(nc * ncand, sizeof(uint64_t));
  for (uint32_t ry = 0; ry < hc; ry++)
    for (uint32_t rx = 0; rx < wc; rx++)
      tab[((size_t)ry * wc + rx) * ncand] = ctb_sse_y(sp, rp, rx * cw, ry * ch, cw, ch);
  vvclf_alf_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.alf_enabled = 1;
  cfg.cb_enabled = cfg.cc_cb_enabled = c_idx == 1;
  cfg.cr_enabled = cfg.cc_cr_enabled = c_idx == 2;
  cfg.bit_depth = bd;
  cfg.ctb_log2_size_y = ctb_log2;
  cfg.chroma_format_idc = cfi;
  vvclf_alf_aps_binding b;
  memset(&b, 0, sizeof(b));
  if (!cc) b.chroma_aps = aps;
  else if (c_idx == 1) b.cc_cb_aps = aps;
  else b.cc_cr_aps = aps;
  size_t plane = (size_t)rp->stride * rp->height;
  uint16_t** trial = (uint16_t**)calloc(n, sizeof(uint16_t*));
  vvclf_alf_ctb* pic = (vvclf_alf_ctb*)calloc(nc, sizeof(vvclf_alf_ctb));
  int rc = VVCLF_OK;
  for (uint32_t k = 0; k < n && rc == VVCLF_OK; k++) {
    vvclf_picture t = *rec;
    if (cc) t.luma = *pre_luma; /* the CC-ALF source is the pre-luma-ALF plane */
    trial[k] = (uint16_t*)malloc(plane * 2 + 2);
    memcpy(trial[k], rp->samples, plane * 2);
    if (c_idx == 1) t.cb.samples = trial[k];
    else t.cr.samples = trial[k];
    for (size_t i = 0; i < nc; i++) {
      memset(&pic[i], 0, sizeof(pic[i]));
      if (cc) {
        if (c_idx == 1) pic[i].cc_cb_idc = (uint8_t)(k + 1);
        else pic[i].cc_cr_idc = (uint8_t)(k + 1);
      } else if (c_idx == 1) {
        pic[i].cb_on = 1;
        pic[i].cb_alt_idx = (uint8_t)k;
      } else {
        pic[i].cr_on = 1;
        pic[i].cr_alt_idx = (uint8_t)k;
      }
    }
    vvclf_alf_picture ap = {wc, hc, pic};
    rc = oracle_apply_alf_clipped(&t, &ap, &cfg, &b, tb);
    const vvclf_plane* tp = c_idx == 1 ? &t.cb : &t.cr;
    for (uint32_t ry = 0; ry < hc; ry++)
      for (uint32_t rx = 0; rx < wc; rx++)
        tab[((size_t)ry * wc + rx) * ncand + 1 + k] = ctb_sse_y(sp, tp, rx * cw, ry * ch, cw, ch);
  }
  if (rc == VVCLF_OK) {
    chroma_pick_commit(rp, trial, tab, ncand, wc, hc, cw, ch, apply, c_idx, cc);
    if (sse) memcpy(sse, tab, nc * ncand * sizeof(uint64_t));
  }
  for (uint32_t k = 0; k < n; k++) free(trial[k]);
  free(trial);
  free(pic);
  free(tab);
  return rc;
}

int oracle_alf_decide_chroma(const vvclf_picture* src, vvclf_picture* rec, vvclf_alf_ctb* apply,
                             const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2, uint32_t bd, uint32_t cfi,
                             const vvclf_bounds* tb, uint64_t* sse) {
  return chroma_decide(src, rec, NULL, apply, aps, c_idx, ctb_log2, bd, cfi, tb, sse, 0);
}

int oracle_cc_alf_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma,
                         vvclf_alf_ctb* apply, const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2,
                         uint32_t bd, uint32_t cfi, const vvclf_bounds* tb, uint64_t* sse) {
  return chroma_decide(src, rec, pre_luma, apply, aps, c_idx, ctb_log2, bd, cfi, tb, sse, 1);
}

/* ======================================================================== */
/* SAO encoder decision (src/sao_enc.rs)                                     */
/* ======================================================================== */

/* eo_category, src/sao_enc.rs:349-383 */
static int sao_eo_category(int r, int n0, int n1) {
  int s0 = r > n0 ? 1 : (r < n0 ? -1 : 0), s1 = r > n1 ? 1 : (r < n1 ? -1 : 0);
  switch (s0 + s1) {
    case 2: return 4;
    case 1: return 3;
    case -1: return 2;
    case -2: return 1;
    default: return 0;
  }
}

static const int SAO_EO_DIRS[4][4] = {{-1, 0, 1, 0}, {0, -1, 0, 1}, {-1, -1, 1, 1}, {1, -1, -1, 1}};

/* try_band_offset, src/sao_enc.rs:77-170; returns 0 when no band window
 * improves (the reference's None), -1 for a sample of 2^BitDepth or more
 * (its band index runs past the 32 bands: the reference panics) */
static int sao_try_band(const vvclf_plane* src, const vvclf_plane* rec, uint32_t x0, uint32_t y0, uint32_t w,
                        uint32_t h, uint32_t bd, vvclf_sao_ctb* out) {
  int64_t cnt[32] = {0}, dif[32] = {0};
  uint32_t shift = bd - 5;
  uint32_t xe = x0 + w < rec->width ? x0 + w : rec->width, ye = y0 + h < rec->height ? y0 + h : rec->height;
  for (uint32_t y = y0; y < ye; y++)
    for (uint32_t x = x0; x < xe; x++) {
      if (x >= src->width || y >= src->height) continue;
      uint32_t r = rec->samples[(size_t)y * rec->stride + x], sv = src->samples[(size_t)y * src->stride + x];
      uint32_t b = r >> shift;
      if (b >= 32) return -1;
      cnt[b] += 1;
      dif[b] += (int64_t)sv - (int64_t)r;
    }
  int64_t best_delta = 0;
  int best_start = 0;
  int32_t best_off[4] = {0, 0, 0, 0};
  int64_t cmax = ((int64_t)1 << ((bd < 10 ? bd : 10) - 5)) - 1;
  for (int start = 0; start < 28; start++) {
    int64_t delta = 0;
    int32_t off[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
      int b = start + i;
      if (cnt[b] > 0) {
        int64_t opt = (dif[b] + cnt[b] / 2) / cnt[b];
        int64_t oc = opt < -cmax ? -cmax : (opt > cmax ? cmax : opt);
        off[i] = (int32_t)oc;
        delta -= dif[b] * oc;
      }
    }
    if (delta < best_delta) {
      best_delta = delta;
      best_start = start;
      memcpy(best_off, off, sizeof(off));
    }
  }
  if (best_delta >= 0) return 0;
  uint32_t ab[4], sg[4];
  for (int i = 0; i < 4; i++) {
    ab[i] = (uint32_t)(best_off[i] < 0 ? -best_off[i] : best_off[i]);
    sg[i] = best_off[i] < 0;
  }
  memset(out, 0, sizeof(*out));
  out->sao_type_idx = VVCLF_SAO_BAND;
  out->band_position = (uint8_t)(best_start & 31);
  oracle_sao_derive_offset_val(ab, sg, bd, out->offset_val);
  return 1;
}

/* try_edge_offset_class, src/sao_enc.rs:211-284: the class's candidate and
 * its distortion delta */
static int64_t sao_try_eo_class(const vvclf_plane* src, const vvclf_plane* rec, uint32_t x0, uint32_t y0, uint32_t w,
                                uint32_t h, uint32_t bd, int cls, vvclf_sao_ctb* out) {
  int64_t cnt[5] = {0}, dif[5] = {0};
  uint32_t xe = x0 + w < rec->width ? x0 + w : rec->width, ye = y0 + h < rec->height ? y0 + h : rec->height;
  for (uint32_t y = y0; y < ye; y++)
    for (uint32_t x = x0; x < xe; x++) {
      int nx0 = (int)x + SAO_EO_DIRS[cls][0], ny0 = (int)y + SAO_EO_DIRS[cls][1];
      int nx1 = (int)x + SAO_EO_DIRS[cls][2], ny1 = (int)y + SAO_EO_DIRS[cls][3];
      if (nx0 < 0 || ny0 < 0 || nx0 >= (int)rec->width || ny0 >= (int)rec->height) continue;
      if (nx1 < 0 || ny1 < 0 || nx1 >= (int)rec->width || ny1 >= (int)rec->height) continue;
      int r = rec->samples[(size_t)y * rec->stride + x];
      int c = sao_eo_category(r, rec->samples[(size_t)ny0 * rec->stride + nx0],
                              rec->samples[(size_t)ny1 * rec->stride + nx1]);
      if (x >= src->width || y >= src->height) continue;
      cnt[c] += 1;
      dif[c] += (int64_t)src->samples[(size_t)y * src->stride + x] - r;
    }
  uint32_t off[4] = {0, 0, 0, 0};
  int64_t delta = 0;
  for (int i = 0; i < 4; i++) {
    int c = i + 1;
    if (cnt[c] > 0) {
      int64_t a = dif[c] < 0 ? -dif[c] : dif[c];
      int64_t opt = (a + cnt[c] / 2) / cnt[c];
      int64_t oc = opt < 0 ? 0 : (opt > 7 ? 7 : opt);
      off[i] = (uint32_t)oc;
      delta -= a * oc;
    }
  }
  static const uint32_t eo_sign[4] = {0, 0, 1, 1};
  memset(out, 0, sizeof(*out));
  out->sao_type_idx = VVCLF_SAO_EDGE;
  out->eo_class = (uint8_t)cls;
  oracle_sao_derive_offset_val(off, eo_sign, bd, out->offset_val);
  return delta;
}

/* sao_decide_ctb, src/sao_enc.rs:44-71 (with try_best_edge_offset, :175-207):
 * edge offset when some class improves, else band offset when some band
 * window improves, else not applied */
static int sao_decide_ctb_c(const vvclf_plane* src, const vvclf_plane* rec, uint32_t x0, uint32_t y0, uint32_t w,
                            uint32_t h, uint32_t bd, vvclf_sao_ctb* out) {
  vvclf_sao_ctb bo, eo, t;
  int has_bo = sao_try_band(src, rec, x0, y0, w, h, bd, &bo);
  if (has_bo < 0) return VVCLF_ERR_INVALID;
  int64_t best = 0;
  int has_eo = 0;
  for (int c = 0; c < 4; c++) {
    int64_t d = sao_try_eo_class(src, rec, x0, y0, w, h, bd, c, &t);
    if (d < best) {
      best = d;
      eo = t;
      has_eo = 1;
    }
  }
  memset(out, 0, sizeof(*out));
  if (has_eo) *out = eo;
  else if (has_bo) *out = bo;
  return VVCLF_OK;
}

/* sao_decide_ctb_constrained, src/sao_enc.rs:286-347: Cr within Cb's type
 * (and edge class); zero offsets of that type when nothing improves */
static int sao_decide_ctb_cr(const vvclf_plane* src, const vvclf_plane* rec, uint32_t x0, uint32_t y0, uint32_t w,
                             uint32_t h, uint32_t bd, const vvclf_sao_ctb* shared, vvclf_sao_ctb* out) {
  memset(out, 0, sizeof(*out));
  if (shared->sao_type_idx == VVCLF_SAO_BAND) {
    int r = sao_try_band(src, rec, x0, y0, w, h, bd, out);
    if (r < 0) return VVCLF_ERR_INVALID;
    if (!r) {
      memset(out, 0, sizeof(*out));
      out->sao_type_idx = VVCLF_SAO_BAND;
    }
  } else if (shared->sao_type_idx == VVCLF_SAO_EDGE) {
    vvclf_sao_ctb t;
    int64_t d = sao_try_eo_class(src, rec, x0, y0, w, h, bd, shared->eo_class, &t);
    if (d < 0) {
      *out = t;
    } else {
      out->sao_type_idx = VVCLF_SAO_EDGE;
      out->eo_class = shared->eo_class;
    }
  }
  return VVCLF_OK;
}

/* sao_decide_picture, src/sao_enc.rs:385-462 (4:2:0 chroma at half the luma
 * CTB, rounded down) */
int oracle_sao_decide_picture(const vvclf_picture* src, const vvclf_picture* rec, uint32_t ctb_log2, uint32_t bd,
                              uint32_t sao_luma, uint32_t sao_chroma, vvclf_sao_ctb_params* out) {
  uint32_t ctb = 1u << ctb_log2, W = rec->luma.width, H = rec->luma.height;
  uint32_t wc = (W + ctb - 1) / ctb, hc = (H + ctb - 1) / ctb;
  for (uint32_t ry = 0; ry < hc; ry++)
    for (uint32_t rx = 0; rx < wc; rx++) {
      vvclf_sao_ctb_params* p = &out[(size_t)ry * wc + rx];
      memset(p, 0, sizeof(*p));
      uint32_t x0 = rx * ctb, y0 = ry * ctb;
      uint32_t w = ctb < W - x0 ? ctb : W - x0, h = ctb < H - y0 ? ctb : H - y0;
      if (sao_luma && sao_decide_ctb_c(&src->luma, &rec->luma, x0, y0, w, h, bd, &p->luma)) return VVCLF_ERR_INVALID;
      uint32_t cx = x0 / 2, cy = y0 / 2, cw = w / 2, ch = h / 2;
      if (sao_chroma && cw > 0 && ch > 0) {
        if (sao_decide_ctb_c(&src->cb, &rec->cb, cx, cy, cw, ch, bd, &p->cb)) return VVCLF_ERR_INVALID;
        if (sao_decide_ctb_cr(&src->cr, &rec->cr, cx, cy, cw, ch, bd, &p->cb, &p->cr)) return VVCLF_ERR_INVALID;
      }
    }
  return VVCLF_OK;
}

/* chroma_independent_bit_cost, src/sao_enc.rs:670-696 */
static int64_t sao_chroma_indep_bits(const vvclf_sao_ctb_params* p) {
  int64_t bits = 1;
  const vvclf_sao_ctb* c[2] = {&p->cb, &p->cr};
  for (int k = 0; k < 2; k++)
    bits += c[k]->sao_type_idx == VVCLF_SAO_BAND ? 16 : (c[k]->sao_type_idx == VVCLF_SAO_EDGE ? 10 : 1);
  return bits;
}

/* chroma_ctb_sse_with, src/sao_enc.rs:698-749: Cb + Cr SSE of chroma CTB
 * (rx, ry) after SAO with (cb, cr) on the pre-SAO planes; `tmp` holds copies
 * of those planes and is restored on return */
static uint64_t sao_chroma_sse_with(const vvclf_picture* src, const vvclf_picture* rec, vvclf_plane* tmp,
                                    uint32_t rx, uint32_t ry, uint32_t cctb, uint32_t bd, const vvclf_sao_ctb* cb,
                                    const vvclf_sao_ctb* cr) {
  const vvclf_sao_ctb* prm[2] = {cb, cr};
  const vvclf_plane* pre[2] = {&rec->cb, &rec->cr};
  for (int k = 0; k < 2; k++)
    if (prm[k]->sao_type_idx) sao_ctb(&tmp[k], pre[k]->samples, NULL, rx, ry, cctb, cctb, bd, prm[k]);
  uint32_t cx = rx * cctb, cy = ry * cctb, W = tmp[0].width, H = tmp[0].height;
  uint32_t cw = cx < W ? (W - cx < cctb ? W - cx : cctb) : 0, ch = cy < H ? (H - cy < cctb ? H - cy : cctb) : 0;
  uint64_t sse = 0;
  for (uint32_t y = cy; y < cy + ch; y++)
    for (uint32_t x = cx; x < cx + cw; x++) {
      int64_t d0 = (int64_t)src->cb.samples[(size_t)y * src->cb.stride + x] - tmp[0].samples[(size_t)y * tmp[0].stride + x];
      int64_t d1 = (int64_t)src->cr.samples[(size_t)y * src->cr.stride + x] - tmp[1].samples[(size_t)y * tmp[1].stride + x];
      sse += (uint64_t)(d0 * d0 + d1 * d1);
    }
  for (int k = 0; k < 2; k++) /* restore the CTB region */
    for (uint32_t y = cy; y < cy + ch; y++)
      memcpy(tmp[k].samples + (size_t)y * tmp[k].stride + cx, pre[k]->samples + (size_t)y * pre[k]->stride + cx,
             (size_t)cw * 2);
  return sse;
}

/* apply_chroma_sao_merge, src/sao_enc.rs:525-660: per CTB in raster order,
 * keep the independent chroma parameters or take the (already decided) left
 * or above neighbour's, by SSE + 8 * bits (merge left 1 bit, above 2; ties
 * keep the earlier option); map: 0 independent, 1 merge left, 2 merge above */
int oracle_sao_chroma_merge(const vvclf_picture* src, const vvclf_picture* rec, const vvclf_sao_ctb_params* indep,
                            uint32_t bd, uint32_t ctb_log2, vvclf_sao_ctb_params* merged, uint8_t* map) {
  uint32_t ctb = 1u << ctb_log2, W = rec->luma.width, H = rec->luma.height;
  uint32_t wc = (W + ctb - 1) / ctb, hc = (H + ctb - 1) / ctb, cctb = ctb / 2;
  size_t n = (size_t)wc * hc;
  memcpy(merged, indep, n * sizeof(*merged));
  vvclf_plane tmp[2] = {rec->cb, rec->cr};
  tmp[0].samples = dup_plane(&rec->cb);
  tmp[1].samples = dup_plane(&rec->cr);
  for (uint32_t ry = 0; ry < hc; ry++)
    for (uint32_t rx = 0; rx < wc; rx++) {
      vvclf_sao_ctb_params* cur = &merged[(size_t)ry * wc + rx];
      vvclf_sao_ctb_params in = *cur;
      int64_t best = (int64_t)sao_chroma_sse_with(src, rec, tmp, rx, ry, cctb, bd, &in.cb, &in.cr) +
                     8 * sao_chroma_indep_bits(&in);
      int choice = 0;
      vvclf_sao_ctb bcb = in.cb, bcr = in.cr;
      if (rx > 0) {
        const vvclf_sao_ctb_params* l = &merged[(size_t)ry * wc + rx - 1];
        int64_t c = (int64_t)sao_chroma_sse_with(src, rec, tmp, rx, ry, cctb, bd, &l->cb, &l->cr) + 8 * 1;
        if (c < best) {
          best = c;
          choice = 1;
          bcb = l->cb;
          bcr = l->cr;
        }
      }
      if (ry > 0) {
        const vvclf_sao_ctb_params* a = &merged[(size_t)(ry - 1) * wc + rx];
        int64_t c = (int64_t)sao_chroma_sse_with(src, rec, tmp, rx, ry, cctb, bd, &a->cb, &a->cr) + 8 * 2;
        if (c < best) {
          best = c;
          choice = 2;
          bcb = a->cb;
          bcr = a->cr;
        }
      }
      map[(size_t)ry * wc + rx] = (uint8_t)choice;
      cur->cb = bcb;
      cur->cr = bcr;
    }
  free(tmp[0].samples);
  free(tmp[1].samples);
  return VVCLF_OK;
}

/* total_sse_luma, src/alf_aps_design.rs:599-612 (planes of equal size) */
static uint64_t total_sse_luma(const vvclf_plane* a, const vvclf_plane* b) {
  uint32_t w = minu(a->width, b->width), h = minu(a->height, b->height);
  uint64_t sse = 0;
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      int64_t d = (int64_t)a->samples[(size_t)y * a->stride + x] - b->samples[(size_t)y * b->stride + x];
      sse += (uint64_t)(d * d);
    }
  return sse;
}

/* replay_aps_full_picture_sse, src/alf_aps_design.rs:618-664: ALF of a copy
 * of rec with every CTB on APS slot 0 (set 16), luma only, picture-wide
 * clip; the luma SSE against src */
static uint64_t clip_rdo_replay(const vvclf_picture* src, const vvclf_picture* rec, const vvclf_alf_aps* aps,
                                uint32_t ctb_log2, uint32_t bd, uint32_t cfi, uint16_t* buf, vvclf_alf_ctb* ctbs,
                                uint32_t wc, uint32_t hc) {
  size_t plane = (size_t)rec->luma.stride * rec->luma.height;
  memcpy(buf, rec->luma.samples, plane * 2);
  vvclf_picture r = *rec;
  r.luma.samples = buf;
  vvclf_alf_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.alf_enabled = 1;
  cfg.bit_depth = bd;
  cfg.ctb_log2_size_y = ctb_log2;
  cfg.chroma_format_idc = cfi;
  const vvclf_alf_aps* slot[1] = {aps};
  vvclf_alf_aps_binding b;
  memset(&b, 0, sizeof(b));
  b.luma_apses = slot;
  b.n_luma_apses = 1;
  for (size_t i = 0; i < (size_t)wc * hc; i++) {
    memset(&ctbs[i], 0, sizeof(ctbs[i]));
    ctbs[i].luma_on = 1;
    ctbs[i].luma_filt_set_idx = 16;
  }
  vvclf_alf_picture ap = {wc, hc, ctbs};
  oracle_apply_alf_clipped(&r, &ap, &cfg, &b, NULL);
  return total_sse_luma(&src->luma, &r.luma);
}

/* design_clip_rdo_for_luma_aps, src/alf_aps_design.rs:697-768: greedy
 * per-class, per-tap descent over clip indices 1..3, each trial a full
 * picture replay; a trial is kept only when it strictly lowers the SSE */
int oracle_alf_clip_rdo_luma(const vvclf_picture* src, const vvclf_picture* rec, const int32_t coeff[25][12],
                             uint32_t ctb_log2, uint32_t bd, uint32_t cfi, vvclf_alf_clip_rdo* out) {
  uint32_t ctb = 1u << ctb_log2;
  uint32_t wc = (rec->luma.width + ctb - 1) / ctb, hc = (rec->luma.height + ctb - 1) / ctb;
  memset(out, 0, sizeof(*out));
  memcpy(out->coeff, coeff, sizeof(out->coeff));
  vvclf_alf_aps* aps = (vvclf_alf_aps*)calloc(1, sizeof(vvclf_alf_aps));
  aps->alf_luma_filter_signal_flag = 1;
  aps->alf_luma_clip_flag = 1;
  aps->n_luma_coeff = 25;
  memcpy(aps->luma_coeff, coeff, sizeof(aps->luma_coeff));
  uint16_t* buf = (uint16_t*)malloc((size_t)rec->luma.stride * rec->luma.height * 2 + 2);
  vvclf_alf_ctb* ctbs = (vvclf_alf_ctb*)calloc((size_t)wc * hc, sizeof(vvclf_alf_ctb));
  uint64_t running = clip_rdo_replay(src, rec, aps, ctb_log2, bd, cfi, buf, ctbs, wc, hc);
  out->baseline_sse_y = running;
  for (int k = 0; k < 25; k++)
    for (int j = 0; j < 12; j++) {
      uint8_t best_clip = 0;
      uint64_t best = running;
      for (uint8_t c = 1; c <= 3; c++) {
        aps->luma_clip_idx[k][j] = c;
        uint64_t t = clip_rdo_replay(src, rec, aps, ctb_log2, bd, cfi, buf, ctbs, wc, hc);
        if (t < best) {
          best = t;
          best_clip = c;
        }
      }
      aps->luma_clip_idx[k][j] = best_clip;
      out->clip_idx[k][j] = best_clip;
      running = best;
    }
  out->post_clip_sse_y = running;
  for (int k = 0; k < 25; k++)
    for (int j = 0; j < 12; j++)
      if (out->clip_idx[k][j]) out->alf_luma_clip_flag = 1;
  free(ctbs);
  free(buf);
  free(aps);
  return VVCLF_OK;
}

int oracle_apply_in_loop_filters(vvclf_frame* f) {
  static const vvclf_deblock_cu empty_cus[1];
  const vvclf_deblock_cu* ccus = f->has_chroma_cus ? (f->chroma_cus ? f->chroma_cus : empty_cus) : NULL;
  oracle_lmcs_inverse_luma(&f->pic.luma, &f->lmcs);
  oracle_apply_deblocking_clipped(&f->pic, f->cus, f->n_cus, ccus,
                                  f->n_chroma_cus, f->mv_grid, &f->deblock, f->chroma_format_idc, f->no_filter_cols,
                                  f->n_no_filter_cols, f->no_filter_rows, f->n_no_filter_rows);
  oracle_apply_sao_clipped(&f->pic, &f->sao, &f->sao_cfg, f->clip_bounds);
  return oracle_apply_alf_clipped(&f->pic, &f->alf, &f->alf_cfg, &f->alf_binding, f->clip_bounds);
}

typedef struct {
  vvclf_frame* const* frames;
  uint32_t n;
  uint32_t next;
  pthread_mutex_t mu;
  int rc;
} mt_state;

static void* mt_worker(void* arg) {
  mt_state* s = (mt_state*)arg;
  for (;;) {
    pthread_mutex_lock(&s->mu);
    uint32_t i = s->next++;
    pthread_mutex_unlock(&s->mu);
    if (i >= s->n) break;
    int rc = oracle_apply_in_loop_filters(s->frames[i]);
    if (rc) s->rc = rc;
  }
  return NULL;
}

int oracle_apply_in_loop_filters_mt(vvclf_frame* const* frames, uint32_t n, uint32_t n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  mt_state s;
  s.frames = frames;
  s.n = n;
  s.next = 0;
  s.rc = 0;
  pthread_mutex_init(&s.mu, NULL);
  pthread_t th[256];
  for (uint32_t t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, mt_worker, &s);
  for (uint32_t t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&s.mu);
  return s.rc;
}
