// synth.cpp — seeded synthetic workloads (see synth.h).
#include "synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int32_t range(int32_t lo, int32_t hi) { return lo + (int32_t)(next() % (uint64_t)(hi - lo + 1)); }
  bool chance(double p) { return uni() < p; }
};

}  // namespace

struct vvclf_synth_frame {
  vvclf_synth_spec spec;
  vvclf_frame desc;
  std::vector<uint16_t> y, cb, cr, y0, cb0, cr0;
  vvclf_picture input;
  std::vector<vvclf_deblock_cu> cus, chroma_cus;
  std::vector<vvclf_mv_cell> mv;
  vvclf_mv_grid grid;
  std::vector<vvclf_sao_ctb_params> sao;
  std::vector<vvclf_alf_ctb> alf;
  std::vector<vvclf_alf_aps> luma_aps;
  std::vector<const vvclf_alf_aps*> luma_aps_ptr;
  vvclf_alf_aps chroma_aps, cc_aps;
  vvclf_lmcs_derived lmcs;
  std::vector<uint32_t> nf_cols, nf_rows, bd_cols, bd_rows;
  vvclf_bounds bounds;
  uint64_t n_intra = 0;
};

namespace {

void fill_plane(std::vector<uint16_t>& v, uint32_t w, uint32_t h, uint32_t bd, rng& r) {
  v.resize((size_t)w * h);
  const double scale = (double)(1u << bd) / 1024.0;
  const int32_t maxv = (1 << bd) - 1;
  for (uint32_t y = 0; y < h; y++)
    for (uint32_t x = 0; x < w; x++) {
      double g = 64.0 + 896.0 * ((double)x / w + (double)y / h) / 2.0;
      int32_t s = (int32_t)std::floor(g * scale) + (int32_t)std::floor(r.range(-16, 16) * scale);
      v[(size_t)y * w + x] = (uint16_t)std::clamp(s, 0, maxv);
    }
}

double split_prob(uint32_t w, uint32_t h, float scale) {
  uint32_t m = std::max(w, h);
  double p = m >= 128 ? 0.95 : (m >= 64 ? 0.8 : (m >= 32 ? 0.6 : (m >= 16 ? 0.5 : 0.4)));
  return std::min(1.0, p * scale);
}

struct builder {
  vvclf_synth_frame* f;
  rng r;
  uint32_t W, H;

  void emit(uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    const vvclf_synth_spec& s = f->spec;
    vvclf_deblock_cu c;
    memset(&c, 0, sizeof(c));
    c.x = x;
    c.y = y;
    c.w = w;
    c.h = h;
    c.qp_y = r.range((int32_t)s.qp_min, (int32_t)s.qp_max);
    c.qp_c[0] = c.qp_c[1] = c.qp_c[2] = VVCLF_QP_C_LEGACY;
    c.intra = r.chance(s.p_intra);
    if (c.intra) f->n_intra++;
    c.tu_y_coded = c.intra ? r.chance(0.8) : r.chance(s.p_coded);
    c.tu_cb_coded = r.chance(0.3);
    c.tu_cr_coded = r.chance(0.3);
    if (w > 64 || h > 64) {
      c.has_tu64 = 1;
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
          c.tu64_y[a][b] = r.chance(0.5);
          c.tu64_cb[a][b] = r.chance(0.3);
          c.tu64_cr[a][b] = r.chance(0.3);
        }
    }
    // motion: one vector per CU, or per 4x4 sub-block for affine records
    vvclf_mv_cell m;
    memset(&m, 0, sizeof(m));
    bool affine = false;
    if (s.rich) {
      if (c.intra) {
        c.plt = r.chance(0.04);
        c.bdpcm_luma = r.chance(0.06);
        c.bdpcm_chroma = r.chance(0.06);
        if (!c.plt && w >= 8 && h >= 8 && (w <= 64 && h <= 64) && r.chance(0.12)) {
          c.has_tb_split = 1;  // ISP: luma-only 1-D split into 2 or 4 parts
          c.tb_luma_only = 1;
          c.tb_vertical = r.chance(0.5);
          uint32_t dim = c.tb_vertical ? w : h;
          uint32_t parts = (dim >= 16 && r.chance(0.5)) ? 4 : 2;
          c.tb_n_bounds = (uint8_t)(parts - 1);
          for (uint32_t i = 1; i < parts; i++) c.tb_bounds[i - 1] = dim * i / parts;
          for (int i = 0; i < 4; i++) c.tb_y_coded[i] = r.chance(0.6);
        }
      } else {
        c.ibc = r.chance(0.05);
        c.ciip = !c.ibc && r.chance(0.05);
        if (!c.ibc && !c.ciip && w >= 8 && h >= 8 && r.chance(0.15)) {
          affine = true;
          c.has_num_sb = 1;
          c.num_sb_x = w / 4;
          c.num_sb_y = h / 4;
        } else if (!c.ibc && w >= 8 && h >= 8 && w <= 64 && h <= 64 && r.chance(0.12)) {
          c.has_tb_split = 1;  // SBT: half split, chroma split too
          c.tb_vertical = r.chance(0.5);
          uint32_t dim = c.tb_vertical ? w : h;
          c.tb_n_bounds = 1;
          c.tb_bounds[0] = dim / 2;
          c.tb_y_coded[0] = 1;
          c.tb_cb_coded[0] = r.chance(0.5);
          c.tb_cr_coded[0] = r.chance(0.5);
        }
      }
      if (r.chance(0.3)) {
        c.qp_c[0] = std::clamp(c.qp_y + r.range(-3, 3), 0, 63);
        c.qp_c[1] = std::clamp(c.qp_y + r.range(-3, 3), 0, 63);
        c.qp_c[2] = std::clamp(c.qp_y + r.range(-3, 3), 0, 63);
        c.joint_cbcr2 = r.chance(0.2);
      }
    }
    if (!c.intra) {
      if (c.ibc) {
        m.mv_l0[0] = r.range(-64, 0) * 16;
        m.mv_l0[1] = r.range(-64, 0) * 16;
      } else {
        m.pred_l0 = 1;
        m.poc_l0 = 0;
        if (s.rich && r.chance(0.3)) {
          m.pred_l1 = 1;
          m.poc_l1 = r.chance(0.5) ? 0 : 4;
          m.mv_l1[0] = r.range(-2, 2) * 8;
        }
        if (s.rich && r.chance(0.1)) m.poc_l0 = -4;
        if (r.chance(s.p_mv_change)) {
          m.mv_l0[0] = (r.range(1, 32)) * 16 * (r.chance(0.5) ? 1 : -1);
          m.mv_l0[1] = r.range(-32, 32) * 16;
        }
      }
    }
    uint32_t cw = f->grid.cells_w, ch = f->grid.cells_h;
    for (uint32_t cy = y / 4; cy < std::min((y + h) / 4, ch); cy++)
      for (uint32_t cx = x / 4; cx < std::min((x + w) / 4, cw); cx++) {
        vvclf_mv_cell mm = m;
        if (affine) mm.mv_l0[0] += (int32_t)((cx - x / 4) * r.range(0, 6));
        f->mv[(size_t)cy * cw + cx] = mm;
      }
    f->cus.push_back(c);
  }

  void split(uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    const vvclf_synth_spec& s = f->spec;
    if (x >= W || y >= H) return;
    bool cross_r = x + w > W, cross_b = y + h > H;
    bool can_qt = w == h && w / 2 >= std::max(s.min_cu_w, s.min_cu_h);
    bool can_bh = h / 2 >= s.min_cu_h;
    bool can_bv = w / 2 >= s.min_cu_w;
    int mode = 0;  // 0 leaf, 1 QT, 2 BH, 3 BV
    if (cross_r || cross_b) {
      if (cross_r && cross_b && can_qt) mode = 1;
      else if (cross_b && can_bh) mode = 2;
      else if (cross_r && can_bv) mode = 3;
      else if (can_qt) mode = 1;
    } else if ((can_qt || can_bh || can_bv) && r.chance(split_prob(w, h, s.split_scale))) {
      int opts[3], n = 0;
      if (can_qt) opts[n++] = 1;
      if (can_bh) opts[n++] = 2;
      if (can_bv) opts[n++] = 3;
      mode = opts[r.next() % n];
    }
    if (mode == 1) {
      split(x, y, w / 2, h / 2);
      split(x + w / 2, y, w / 2, h / 2);
      split(x, y + h / 2, w / 2, h / 2);
      split(x + w / 2, y + h / 2, w / 2, h / 2);
    } else if (mode == 2) {
      split(x, y, w, h / 2);
      split(x, y + h / 2, w, h / 2);
    } else if (mode == 3) {
      split(x, y, w / 2, h);
      split(x + w / 2, y, w / 2, h);
    } else {
      emit(x, y, w, h);
    }
  }
};

void lmcs_random(vvclf_lmcs_derived& d, uint32_t bd, rng& r) {
  // 16 monotone pieces; sum of codewords within the eq. 96 budget and every
  // piece inside the eq. 95 band, which keeps the derivation conforming.
  for (int attempt = 0; attempt < 1000; attempt++) {
    uint32_t org = (1u << bd) / 16, lo = org >> 3, hi = (org << 3) - 1;
    uint32_t cw[16];
    uint64_t sum = 0;
    for (int i = 0; i < 16; i++) {
      cw[i] = (uint32_t)std::clamp<int32_t>((int32_t)org + r.range(-(int32_t)org / 3, (int32_t)org / 3), lo, hi);
      sum += cw[i];
    }
    while (sum > (1u << bd) - 1) {
      int i = (int)(r.next() % 16);
      if (cw[i] > lo + 1) {
        cw[i]--;
        sum--;
      }
    }
    memset(&d, 0, sizeof(d));
    d.bit_depth = bd;
    d.org_cw = org;
    uint32_t lg = bd - 4;
    bool ok = true;
    for (int i = 0; i < 16; i++) {
      d.lmcs_cw[i] = cw[i];
      d.input_pivot[i] = i * org;
      d.lmcs_pivot[i + 1] = d.lmcs_pivot[i] + cw[i];
      d.scale_coeff[i] = (cw[i] * (1u << 11) + (1u << (lg - 1))) >> lg;
      d.inv_scale_coeff[i] = org * (1u << 11) / cw[i];
      d.chroma_scale_coeff[i] = org * (1u << 11) / cw[i];
    }
    uint32_t sh = bd - 5;
    for (int i = 0; i < 16; i++)
      if (d.lmcs_pivot[i] % (1u << sh) != 0 && (d.lmcs_pivot[i] >> sh) == (d.lmcs_pivot[i + 1] >> sh)) ok = false;
    if (ok) return;
  }
}

void sao_random(vvclf_sao_ctb& s, rng& r) {
  memset(&s, 0, sizeof(s));
  s.sao_type_idx = (uint8_t)r.range(0, 2);
  if (s.sao_type_idx == VVCLF_SAO_EDGE) {
    s.eo_class = (uint8_t)r.range(0, 3);
    for (int i = 0; i < 4; i++) s.offset_val[i + 1] = r.range(0, 31) * (i < 2 ? 1 : -1);
  } else if (s.sao_type_idx == VVCLF_SAO_BAND) {
    s.band_position = (uint8_t)r.range(0, 31);
    for (int i = 0; i < 4; i++) s.offset_val[i + 1] = r.range(-31, 31);
  }
}

}  // namespace

extern "C" {

void vvclf_synth_default_spec(uint32_t config, uint32_t width, uint32_t height, uint64_t seed,
                              vvclf_synth_spec* s) {
  memset(s, 0, sizeof(*s));
  s->width = width;
  s->height = height;
  s->bit_depth = 10;
  s->ctb_log2 = 7;
  s->seed = seed;
  s->min_cu_w = 8;
  s->min_cu_h = 8;
  s->split_scale = 1.0f;
  s->p_intra = 0.106f;
  s->p_mv_change = 0.21f;
  s->p_coded = 0.0f;
  s->qp_min = 22;
  s->qp_max = 37;
  s->sao = 1;
  s->alf = 1;
  s->p_alf_luma = 0.8f;
  s->n_luma_aps = 1;
  s->p_fixed_set = 0.2f;
  s->p_cc = 0.0f;
  switch (config) {
    case 1:  // full chain incl. long filters, 4 APS, CC-ALF, LMCS
      s->n_luma_aps = 4;
      s->chroma_alf = 1;
      s->cc_alf = 1;
      s->p_cc = 0.8f;
      s->lmcs = 1;
      s->split_scale = 0.8f;
      break;
    case 3:  // ALF-heavy, dense edge grid, bS 2 at p = 0.5
      s->min_cu_w = 4;
      s->min_cu_h = 8;
      s->split_scale = 1.6f;
      s->p_intra = 0.29f;
      s->p_alf_luma = 1.0f;
      s->n_luma_aps = 8;
      s->p_fixed_set = 0.0f;
      s->chroma_alf = 1;
      s->cc_alf = 1;
      s->p_cc = 1.0f;
      break;
    case 4:  // serving: full chain
      s->n_luma_aps = 4;
      s->chroma_alf = 1;
      s->cc_alf = 1;
      s->p_cc = 0.8f;
      s->lmcs = 1;
      break;
    default:
      break;
  }
}

int vvclf_synth_make(const vvclf_synth_spec* spec, vvclf_synth_frame** out) {
  if (!spec || !out || spec->width == 0 || spec->height == 0 || spec->bit_depth < 8 || spec->bit_depth > 16)
    return VVCLF_ERR_INVALID;
  vvclf_synth_frame* f = new vvclf_synth_frame();
  f->spec = *spec;
  const vvclf_synth_spec& s = f->spec;
  rng r{s.seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
  const uint32_t W = s.width, H = s.height, Wc = W / 2, Hc = H / 2, bd = s.bit_depth;
  fill_plane(f->y0, W, H, bd, r);
  fill_plane(f->cb0, Wc, Hc, bd, r);
  fill_plane(f->cr0, Wc, Hc, bd, r);
  f->y = f->y0;
  f->cb = f->cb0;
  f->cr = f->cr0;
  // CU tree
  f->grid.cells_w = (W + 3) / 4;
  f->grid.cells_h = (H + 3) / 4;
  f->mv.assign((size_t)f->grid.cells_w * f->grid.cells_h, vvclf_mv_cell{});
  const uint32_t ctb = 1u << s.ctb_log2;
  const uint32_t wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  builder b{f, rng{r.next()}, W, H};
  for (uint32_t ry = 0; ry < hctb; ry++)
    for (uint32_t rx = 0; rx < wctb; rx++) b.split(rx * ctb, ry * ctb, ctb, ctb);
  f->grid.cells = f->mv.data();
  // dual chroma tree (rich): chroma records = coarser re-split of each CTB
  if (s.rich) {
    builder cbld{f, rng{r.next()}, W, H};
    std::vector<vvclf_deblock_cu> luma = f->cus;
    std::vector<vvclf_mv_cell> mv = f->mv;
    f->cus.clear();
    vvclf_synth_spec cs = s;
    cs.min_cu_w = cs.min_cu_h = 16;
    vvclf_synth_spec saved = f->spec;
    f->spec = cs;
    for (uint32_t ry = 0; ry < hctb; ry++)
      for (uint32_t rx = 0; rx < wctb; rx++) cbld.split(rx * ctb, ry * ctb, ctb, ctb);
    f->spec = saved;
    f->chroma_cus = f->cus;
    for (auto& c : f->chroma_cus) {
      c.has_num_sb = 0;  // chroma trees carry no sub-block motion
    }
    f->cus = luma;
    f->mv = mv;
    f->grid.cells = f->mv.data();
  }
  // LMCS and APS tables: from their own seed when the stream shares them
  rng ra{s.aps_seed ? s.aps_seed * 0x9E3779B97F4A7C15ull + 17 : r.next()};
  if (s.lmcs) lmcs_random(f->lmcs, bd, ra);
  // SAO
  f->sao.resize((size_t)wctb * hctb);
  for (auto& p : f->sao) {
    if (s.sao) {
      sao_random(p.luma, r);
      sao_random(p.cb, r);
      sao_random(p.cr, r);
    } else {
      memset(&p, 0, sizeof(p));
    }
  }
  // ALF APS sets
  f->luma_aps.resize(s.n_luma_aps);
  for (auto& a : f->luma_aps) {
    memset(&a, 0, sizeof(a));
    a.alf_luma_filter_signal_flag = 1;
    a.alf_luma_clip_flag = 1;
    a.n_luma_coeff = 25;
    for (int c = 0; c < 25; c++)
      for (int k = 0; k < 12; k++) {
        a.luma_coeff[c][k] = ra.range(-128, 127);
        a.luma_clip_idx[c][k] = (uint8_t)ra.range(0, 3);
      }
  }
  for (auto& a : f->luma_aps) f->luma_aps_ptr.push_back(&a);
  memset(&f->chroma_aps, 0, sizeof(f->chroma_aps));
  if (s.chroma_alf) {
    vvclf_alf_aps& a = f->chroma_aps;
    a.alf_chroma_filter_signal_flag = 1;
    a.alf_chroma_clip_flag = 1;
    a.n_chroma_coeff = 4;
    a.alf_chroma_num_alt_filters_minus1 = 3;
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 6; k++) {
        a.chroma_coeff[i][k] = ra.range(-128, 127);
        a.chroma_clip_idx[i][k] = (uint8_t)ra.range(0, 3);
      }
  }
  memset(&f->cc_aps, 0, sizeof(f->cc_aps));
  if (s.cc_alf) {
    vvclf_alf_aps& a = f->cc_aps;
    a.alf_cc_cb_filter_signal_flag = a.alf_cc_cr_filter_signal_flag = 1;
    a.n_cc_cb_coeff = a.n_cc_cr_coeff = 4;
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 7; k++) {
        // power-of-two mapped coefficients (src/aps.rs:314-326): 0, +-1 .. +-64
        int m = ra.range(0, 7);
        int v = m == 0 ? 0 : (1 << (m - 1));
        a.cc_cb_coeff[i][k] = ra.chance(0.5) ? -v : v;
        m = ra.range(0, 7);
        v = m == 0 ? 0 : (1 << (m - 1));
        a.cc_cr_coeff[i][k] = ra.chance(0.5) ? -v : v;
      }
  }
  f->alf.resize((size_t)wctb * hctb);
  for (auto& c : f->alf) {
    memset(&c, 0, sizeof(c));
    if (!s.alf) continue;
    c.luma_on = r.chance(s.p_alf_luma);
    if (s.n_luma_aps == 0 || r.chance(s.p_fixed_set)) c.luma_filt_set_idx = (uint8_t)r.range(0, 15);
    else c.luma_filt_set_idx = (uint8_t)(16 + r.range(0, (int32_t)s.n_luma_aps - 1));
    if (s.chroma_alf) {
      c.cb_on = r.chance(0.6);
      c.cr_on = r.chance(0.6);
      c.cb_alt_idx = (uint8_t)r.range(0, 3);
      c.cr_alt_idx = (uint8_t)r.range(0, 3);
    }
    if (s.cc_alf) {
      c.cc_cb_idc = r.chance(s.p_cc) ? (uint8_t)r.range(1, 4) : 0;
      c.cc_cr_idc = r.chance(s.p_cc) ? (uint8_t)r.range(1, 4) : 0;
    }
  }
  // tiles: 2x2 grid on CTB boundaries, loop filtering across tiles off
  if (s.tiles && wctb >= 2 && hctb >= 2) {
    uint32_t cx = (wctb / 2) * ctb, cy = (hctb / 2) * ctb;
    f->nf_cols = {cx};
    f->nf_rows = {cy};
    f->bd_cols = {0, cx, W};
    f->bd_rows = {0, cy, H};
  }
  // descriptor
  vvclf_frame& d = f->desc;
  memset(&d, 0, sizeof(d));
  d.pic.luma = {f->y.data(), W, W, H, bd};
  d.pic.cb = {f->cb.data(), Wc, Wc, Hc, bd};
  d.pic.cr = {f->cr.data(), Wc, Wc, Hc, bd};
  f->input = d.pic;
  f->input.luma.samples = f->y0.data();
  f->input.cb.samples = f->cb0.data();
  f->input.cr.samples = f->cr0.data();
  d.chroma_format_idc = 1;
  if (s.lmcs) {
    d.lmcs.derived = &f->lmcs;
    d.lmcs.min_bin_idx = 0;
    d.lmcs.max_bin_idx = 15;
  }
  d.cus = f->cus.data();
  d.n_cus = (uint32_t)f->cus.size();
  if (!f->chroma_cus.empty()) {
    d.has_chroma_cus = 1;
    d.chroma_cus = f->chroma_cus.data();
    d.n_chroma_cus = (uint32_t)f->chroma_cus.size();
  }
  d.mv_grid = &f->grid;
  d.deblock.bit_depth = bd;
  d.deblock.ctb_log2_size_y = s.ctb_log2;
  if (s.rich) {
    d.deblock.luma_beta_offset_div2 = r.range(-2, 2);
    d.deblock.luma_tc_offset_div2 = r.range(-2, 2);
    d.deblock.cb_beta_offset_div2 = r.range(-2, 2);
    d.deblock.cb_tc_offset_div2 = r.range(-2, 2);
    d.deblock.cr_beta_offset_div2 = r.range(-2, 2);
    d.deblock.cr_tc_offset_div2 = r.range(-2, 2);
    d.deblock.chroma_qp_offset_cb = r.range(-3, 3);
    d.deblock.chroma_qp_offset_cr = r.range(-3, 3);
  }
  if (!f->nf_cols.empty()) {
    d.no_filter_cols = f->nf_cols.data();
    d.n_no_filter_cols = (uint32_t)f->nf_cols.size();
    d.no_filter_rows = f->nf_rows.data();
    d.n_no_filter_rows = (uint32_t)f->nf_rows.size();
    f->bounds = {f->bd_cols.data(), (uint32_t)f->bd_cols.size(), f->bd_rows.data(), (uint32_t)f->bd_rows.size()};
    d.clip_bounds = &f->bounds;
  }
  d.sao.pic_width_in_ctbs_y = wctb;
  d.sao.pic_height_in_ctbs_y = hctb;
  d.sao.ctbs = f->sao.data();
  d.sao_cfg.luma_used = 1;
  d.sao_cfg.chroma_used = 1;
  d.sao_cfg.bit_depth = bd;
  d.sao_cfg.ctb_log2_size_y = s.ctb_log2;
  d.sao_cfg.chroma_format_idc = 1;
  d.alf.pic_width_in_ctbs_y = wctb;
  d.alf.pic_height_in_ctbs_y = hctb;
  d.alf.ctbs = f->alf.data();
  d.alf_cfg.alf_enabled = s.alf ? 1 : 0;
  d.alf_cfg.cb_enabled = d.alf_cfg.cr_enabled = s.chroma_alf ? 1 : 0;
  d.alf_cfg.cc_cb_enabled = d.alf_cfg.cc_cr_enabled = s.cc_alf ? 1 : 0;
  d.alf_cfg.bit_depth = bd;
  d.alf_cfg.ctb_log2_size_y = s.ctb_log2;
  d.alf_cfg.chroma_format_idc = 1;
  d.alf_binding.luma_apses = f->luma_aps_ptr.data();
  d.alf_binding.n_luma_apses = (uint32_t)f->luma_aps_ptr.size();
  d.alf_binding.chroma_aps = s.chroma_alf ? &f->chroma_aps : nullptr;
  d.alf_binding.cc_cb_aps = s.cc_alf ? &f->cc_aps : nullptr;
  d.alf_binding.cc_cr_aps = s.cc_alf ? &f->cc_aps : nullptr;
  *out = f;
  return VVCLF_OK;
}

void vvclf_synth_free(vvclf_synth_frame* f) { delete f; }
vvclf_frame* vvclf_synth_desc(vvclf_synth_frame* f) { return f ? &f->desc : nullptr; }
const vvclf_picture* vvclf_synth_input(const vvclf_synth_frame* f) { return f ? &f->input : nullptr; }

void vvclf_synth_reset(vvclf_synth_frame* f) {
  if (!f) return;
  f->y = f->y0;
  f->cb = f->cb0;
  f->cr = f->cr0;
  f->desc.pic.luma.samples = f->y.data();
  f->desc.pic.cb.samples = f->cb.data();
  f->desc.pic.cr.samples = f->cr.data();
}

void vvclf_synth_stats(const vvclf_synth_frame* f, uint64_t out[8]) {
  memset(out, 0, 8 * sizeof(uint64_t));
  if (!f) return;
  out[0] = f->cus.size();
  out[1] = f->n_intra;
  out[2] = f->chroma_cus.size();
}

}  // extern "C"
