// vvclf_syntax.cpp — per-CTB side information as 16-byte syntax records
// (include/vvclf.h, vvclf_ctb_syntax): expansion into SAO / ALF CTB
// parameters with the inference rules of decode_sao_ctb
// (src/sao_syntax.rs:154-284) and decode_alf_ctu (src/alf_syntax.rs:170-290),
// and the inverse.  Host code.
#include <cstring>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

namespace {

// bit fields of the 128-bit record
uint32_t get(const vvclf_ctb_syntax& r, uint32_t pos, uint32_t n) {
  uint32_t v = 0;
  for (uint32_t i = 0; i < n; i++) v |= (uint32_t)((r.bytes[(pos + i) >> 3] >> ((pos + i) & 7)) & 1u) << i;
  return v;
}
void put(vvclf_ctb_syntax& r, uint32_t pos, uint32_t n, uint32_t v) {
  for (uint32_t i = 0; i < n; i++)
    if ((v >> i) & 1u) r.bytes[(pos + i) >> 3] |= (uint8_t)(1u << ((pos + i) & 7));
}

// field positions
constexpr uint32_t kMergeL = 0, kMergeU = 1, kY = 2, kCb = 33, kCr = 64, kAlf = 93, kEnd = 113;
// within a luma / Cb group: type @0, abs @2 + 5i, sign @22 + i, pos @26
// within the Cr group: abs @0 + 5i, sign @20 + i, pos @24

// Table 127 cMax of sao_offset_abs (src/sao_syntax.rs:125-136)
uint32_t offset_cmax(uint32_t bd) {
  const uint32_t m = bd < 10 ? bd : 10;
  return m < 5 ? 0 : (1u << (m - 5)) - 1;
}

// eq. 153 (offset_val_from_raw, src/sao_syntax.rs:290-300)
void offset_val(const uint32_t abs[4], const uint32_t sign[4], uint32_t bd, int32_t out[5]) {
  const int shift = (int)bd - (int)(bd < 10 ? bd : 10);
  out[0] = 0;
  for (int i = 0; i < 4; i++) out[i + 1] = (1 - 2 * (int32_t)(sign[i] & 1)) * ((int32_t)abs[i] << shift);
}

bool same_sao(const vvclf_sao_ctb_params& a, const vvclf_sao_ctb_params& b) { return !memcmp(&a, &b, sizeof(a)); }

int bad(uint32_t rx, uint32_t ry, const char* what) {
  return fail(VVCLF_ERR_INVALID, "CTB syntax record at (%u,%u): %s", rx, ry, what);
}

// One component's explicit SAO syntax (group at `g`; cr: the Cr layout with
// type / class taken from Cb).
int decode_sao_comp(const vvclf_ctb_syntax& r, uint32_t g, bool cr, const vvclf_sao_ctb& cb, uint32_t bd,
                    vvclf_sao_ctb& out, uint32_t rx, uint32_t ry) {
  const uint32_t tbits = cr ? 0 : get(r, g, 2);
  const uint32_t abs_at = cr ? g : g + 2, sign_at = cr ? g + 20 : g + 22, pos_at = cr ? g + 24 : g + 26;
  const uint32_t type = cr ? cb.sao_type_idx : tbits;
  memset(&out, 0, sizeof(out));
  if (type == 3) return bad(rx, ry, "sao_type_idx 3");
  if (type == 0) {
    // not applied: every field of the group absent
    if (get(r, abs_at, 20) || get(r, sign_at, 4) || get(r, pos_at, 5)) return bad(rx, ry, "fields of an unused SAO component");
    return VVCLF_OK;
  }
  uint32_t abs[4], sign[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    abs[i] = get(r, abs_at + 5 * i, 5);
    if (abs[i] > offset_cmax(bd)) return bad(rx, ry, "sao_offset_abs above cMax");
  }
  const uint32_t pos = get(r, pos_at, 5);
  out.sao_type_idx = (uint8_t)type;
  if (type == VVCLF_SAO_BAND) {
    for (int i = 0; i < 4; i++) {
      sign[i] = get(r, sign_at + i, 1);
      if (sign[i] && !abs[i]) return bad(rx, ry, "offset sign without magnitude");
    }
    out.band_position = (uint8_t)pos;
  } else {
    if (get(r, sign_at, 4)) return bad(rx, ry, "offset signs on an edge offset");
    if (cr ? pos != 0 : pos > 3) return bad(rx, ry, "sao_eo_class");
    out.eo_class = cr ? cb.eo_class : (uint8_t)pos;
    sign[2] = sign[3] = 1;  // edge offset: (+, +, -, -), src/sao.rs:169-179
  }
  offset_val(abs, sign, bd, out.offset_val);
  return VVCLF_OK;
}

int decode_one(const vvclf_ctb_syntax* recs, uint32_t w, uint32_t rx, uint32_t ry, const vvclf_ctb_syntax_config& c,
               vvclf_sao_ctb_params* sao, vvclf_alf_ctb* alf) {
  const vvclf_ctb_syntax& r = recs[(size_t)ry * w + rx];
  if (get(r, kEnd, 128 - kEnd)) return bad(rx, ry, "reserved bits");
  const size_t i = (size_t)ry * w + rx;
  // --- SAO (decode_sao_ctb) ---
  vvclf_sao_ctb_params p;
  memset(&p, 0, sizeof(p));
  const bool ml = get(r, kMergeL, 1), mu = get(r, kMergeU, 1);
  if (!c.sao_luma_used && !c.sao_chroma_used) {
    if (get(r, 0, kAlf)) return bad(rx, ry, "SAO fields with SAO off in the slice");
  } else if (ml || mu) {
    if (ml && mu) return bad(rx, ry, "both merge flags");
    if ((ml && rx == 0) || (mu && ry == 0)) return bad(rx, ry, "merge with no neighbour");
    if (get(r, kY, kAlf - kY)) return bad(rx, ry, "fields of a merged CTB");
    if (sao) p = sao[ml ? i - 1 : i - w];
  } else {
    const int n_comp = c.chroma_format_idc != 0 ? 3 : 1;
    if (c.sao_luma_used) {
      if (int rc = decode_sao_comp(r, kY, false, p.cb, c.bit_depth, p.luma, rx, ry)) return rc;
    } else if (get(r, kY, kCb - kY)) {
      return bad(rx, ry, "luma SAO fields with luma SAO off");
    }
    if (n_comp == 3 && c.sao_chroma_used) {
      if (int rc = decode_sao_comp(r, kCb, false, p.cb, c.bit_depth, p.cb, rx, ry)) return rc;
      if (int rc = decode_sao_comp(r, kCr, true, p.cb, c.bit_depth, p.cr, rx, ry)) return rc;
    } else if (get(r, kCb, kAlf - kCb)) {
      return bad(rx, ry, "chroma SAO fields with chroma SAO off");
    }
  }
  if (sao) sao[i] = p;
  // --- ALF (decode_alf_ctu) ---
  vvclf_alf_ctb a;
  memset(&a, 0, sizeof(a));
  const bool chroma = c.chroma_format_idc != 0;
  if (!c.alf_enabled) {
    if (get(r, kAlf, kEnd - kAlf)) return bad(rx, ry, "ALF fields with ALF off");
  } else {
    a.luma_on = (uint8_t)get(r, kAlf, 1);
    const uint32_t use_aps = get(r, kAlf + 1, 1), idx = get(r, kAlf + 2, 4);
    if (a.luma_on) {
      if (use_aps) {
        if (idx >= c.num_alf_aps_ids_luma) return bad(rx, ry, "alf_luma_prev_filter_idx beyond the slice's APS ids");
        a.luma_filt_set_idx = (uint8_t)(16 + idx);
      } else {
        a.luma_filt_set_idx = (uint8_t)idx;
      }
    } else if (use_aps || idx) {
      return bad(rx, ry, "luma filter fields with alf_ctb_flag[0] = 0");
    }
    const uint32_t f_cb = get(r, kAlf + 6, 1), alt_cb = get(r, kAlf + 7, 3);
    const uint32_t f_cr = get(r, kAlf + 10, 1), alt_cr = get(r, kAlf + 11, 3);
    const uint32_t cc_cb = get(r, kAlf + 14, 3), cc_cr = get(r, kAlf + 17, 3);
    auto chroma_flag = [&](bool enabled, uint32_t on, uint32_t alt, uint8_t& on_o, uint8_t& alt_o) -> int {
      if (!(chroma && enabled)) return (on || alt) ? bad(rx, ry, "chroma ALF fields with chroma ALF off") : VVCLF_OK;
      on_o = (uint8_t)on;
      if (on && c.alf_chroma_num_alt_filters_minus1 > 0) {
        if (alt > c.alf_chroma_num_alt_filters_minus1) return bad(rx, ry, "alf_ctb_filter_alt_idx");
        alt_o = (uint8_t)alt;
      } else if (alt) {
        return bad(rx, ry, "alf_ctb_filter_alt_idx not present");
      }
      return VVCLF_OK;
    };
    if (int rc = chroma_flag(c.alf_cb_enabled, f_cb, alt_cb, a.cb_on, a.cb_alt_idx)) return rc;
    if (int rc = chroma_flag(c.alf_cr_enabled, f_cr, alt_cr, a.cr_on, a.cr_alt_idx)) return rc;
    auto cc = [&](bool enabled, uint32_t v, uint32_t minus1, uint8_t& o) -> int {
      if (!(chroma && enabled)) return v ? bad(rx, ry, "CC-ALF index with CC-ALF off") : VVCLF_OK;
      if (v > minus1 + 1) return bad(rx, ry, "alf_ctb_cc_idc beyond the signalled filters");
      o = (uint8_t)v;
      return VVCLF_OK;
    };
    if (int rc = cc(c.alf_cc_cb_enabled, cc_cb, c.alf_cc_cb_filters_signalled_minus1, a.cc_cb_idc)) return rc;
    if (int rc = cc(c.alf_cc_cr_enabled, cc_cr, c.alf_cc_cr_filters_signalled_minus1, a.cc_cr_idc)) return rc;
  }
  if (alf) alf[i] = a;
  return VVCLF_OK;
}

// sao_offset_abs / sign of one component, or false when its offsets have no
// syntax (the inverse of eq. 153)
bool raw_offsets(const vvclf_sao_ctb& s, uint32_t bd, uint32_t abs[4], uint32_t sign[4]) {
  const int shift = (int)bd - (int)(bd < 10 ? bd : 10);
  if (s.offset_val[0] != 0) return false;
  for (int i = 0; i < 4; i++) {
    const int32_t v = s.offset_val[i + 1];
    const uint32_t m = (uint32_t)(v < 0 ? -v : v);
    if (m & ((1u << shift) - 1)) return false;
    abs[i] = m >> shift;
    sign[i] = v < 0 ? 1u : 0u;
    if (abs[i] > offset_cmax(bd)) return false;
  }
  return true;
}

bool encode_comp(const vvclf_sao_ctb& s, uint32_t g, bool cr, const vvclf_sao_ctb& cb, uint32_t bd,
                 vvclf_ctb_syntax& r) {
  if (s.pad) return false;
  if (cr ? s.sao_type_idx != cb.sao_type_idx : s.sao_type_idx > 2) return false;
  if (s.sao_type_idx == 0) {
    vvclf_sao_ctb z;
    memset(&z, 0, sizeof(z));
    return !memcmp(&s, &z, sizeof(z));
  }
  uint32_t abs[4], sign[4];
  if (!raw_offsets(s, bd, abs, sign)) return false;
  const uint32_t abs_at = cr ? g : g + 2, sign_at = cr ? g + 20 : g + 22, pos_at = cr ? g + 24 : g + 26;
  if (!cr) put(r, g, 2, s.sao_type_idx);
  for (int i = 0; i < 4; i++) put(r, abs_at + 5 * i, 5, abs[i]);
  if (s.sao_type_idx == VVCLF_SAO_BAND) {
    if (s.eo_class != 0 || s.band_position > 31) return false;
    for (int i = 0; i < 4; i++) put(r, sign_at + i, 1, abs[i] ? sign[i] : 0);
    put(r, pos_at, 5, s.band_position);
  } else {
    if (s.band_position != 0 || s.eo_class > 3 || (cr && s.eo_class != cb.eo_class)) return false;
    for (int i = 0; i < 4; i++)
      if (abs[i] && sign[i] != (i >= 2 ? 1u : 0u)) return false;
    if (!cr) put(r, pos_at, 5, s.eo_class);
  }
  return true;
}

}  // namespace

extern "C" {

int vvclf_ctb_syntax_decode(const vvclf_ctb_syntax* recs, uint32_t w, uint32_t h, const vvclf_ctb_syntax_config* cfg,
                            vvclf_sao_ctb_params* sao, vvclf_alf_ctb* alf) {
  clear_error();
  if (!recs || !cfg || (!sao && !alf)) return fail(VVCLF_ERR_INVALID, "ctb_syntax_decode: NULL argument");
  if (cfg->bit_depth < 8 || cfg->bit_depth > 16) return fail(VVCLF_ERR_INVALID, "ctb_syntax_decode: bit depth");
  // merges read the neighbour's decoded parameters
  vvclf_sao_ctb_params* sp = sao;
  std::vector<vvclf_sao_ctb_params> tmp;
  if (!sp) {
    tmp.resize((size_t)w * h);
    sp = tmp.data();
  }
  for (uint32_t ry = 0; ry < h; ry++)
    for (uint32_t rx = 0; rx < w; rx++)
      if (int rc = decode_one(recs, w, rx, ry, *cfg, sp, alf)) return rc;
  return VVCLF_OK;
}

int vvclf_ctb_syntax_encode(const vvclf_sao_ctb_params* sao, const vvclf_alf_ctb* alf, uint32_t w, uint32_t h,
                            const vvclf_ctb_syntax_config* cfg, vvclf_ctb_syntax* out) {
  clear_error();
  if (!sao || !alf || !cfg || !out) return fail(VVCLF_ERR_INVALID, "ctb_syntax_encode: NULL argument");
  const uint32_t bd = cfg->bit_depth;
  const bool chroma = cfg->chroma_format_idc != 0;
  for (uint32_t ry = 0; ry < h; ry++)
    for (uint32_t rx = 0; rx < w; rx++) {
      const size_t i = (size_t)ry * w + rx;
      vvclf_ctb_syntax& r = out[i];
      memset(&r, 0, sizeof(r));
      const vvclf_sao_ctb_params& s = sao[i];
      vvclf_sao_ctb_params z;
      memset(&z, 0, sizeof(z));
      bool ok = true;
      if (!cfg->sao_luma_used && !cfg->sao_chroma_used) {
        ok = same_sao(s, z);
      } else if (rx > 0 && same_sao(s, sao[i - 1])) {
        put(r, kMergeL, 1, 1);
      } else if (ry > 0 && same_sao(s, sao[i - w])) {
        put(r, kMergeU, 1, 1);
      } else {
        ok = cfg->sao_luma_used ? encode_comp(s.luma, kY, false, s.cb, bd, r) : !memcmp(&s.luma, &z.luma, sizeof(z.luma));
        if (ok && chroma && cfg->sao_chroma_used)
          ok = encode_comp(s.cb, kCb, false, s.cb, bd, r) && encode_comp(s.cr, kCr, true, s.cb, bd, r);
        else if (ok)
          ok = !memcmp(&s.cb, &z.cb, sizeof(z.cb)) && !memcmp(&s.cr, &z.cr, sizeof(z.cr));
      }
      if (!ok) return fail(VVCLF_ERR_INVALID, "ctb_syntax_encode: SAO parameters of CTB (%u,%u) have no syntax", rx, ry);
      // ALF
      const vvclf_alf_ctb& a = alf[i];
      vvclf_alf_ctb az;
      memset(&az, 0, sizeof(az));
      if (!cfg->alf_enabled) {
        ok = !memcmp(&a, &az, sizeof(az));
      } else {
        put(r, kAlf, 1, a.luma_on ? 1 : 0);
        if (a.luma_on > 1) ok = false;
        if (a.luma_on) {
          if (a.luma_filt_set_idx >= 16) {
            const uint32_t prev = a.luma_filt_set_idx - 16u;
            ok = ok && prev < cfg->num_alf_aps_ids_luma && prev < 16;
            put(r, kAlf + 1, 1, 1);
            put(r, kAlf + 2, 4, prev & 15);
          } else {
            put(r, kAlf + 2, 4, a.luma_filt_set_idx);
          }
        } else {
          ok = ok && a.luma_filt_set_idx == 0;
        }
        auto chroma_flag = [&](bool enabled, uint8_t on, uint8_t alt, uint32_t at) {
          if (!(chroma && enabled)) return on == 0 && alt == 0;
          if (on > 1) return false;
          put(r, at, 1, on);
          if (on && cfg->alf_chroma_num_alt_filters_minus1 > 0) {
            if (alt > cfg->alf_chroma_num_alt_filters_minus1) return false;
            put(r, at + 1, 3, alt);
            return true;
          }
          return alt == 0;
        };
        ok = ok && chroma_flag(cfg->alf_cb_enabled, a.cb_on, a.cb_alt_idx, kAlf + 6) &&
             chroma_flag(cfg->alf_cr_enabled, a.cr_on, a.cr_alt_idx, kAlf + 10);
        auto cc = [&](bool enabled, uint8_t v, uint32_t minus1, uint32_t at) {
          if (!(chroma && enabled)) return v == 0;
          if (v > minus1 + 1 || v > 7) return false;
          put(r, at, 3, v);
          return true;
        };
        ok = ok && cc(cfg->alf_cc_cb_enabled, a.cc_cb_idc, cfg->alf_cc_cb_filters_signalled_minus1, kAlf + 14) &&
             cc(cfg->alf_cc_cr_enabled, a.cc_cr_idc, cfg->alf_cc_cr_filters_signalled_minus1, kAlf + 17);
      }
      if (!ok) return fail(VVCLF_ERR_INVALID, "ctb_syntax_encode: ALF parameters of CTB (%u,%u) have no syntax", rx, ry);
    }
  return VVCLF_OK;
}

}  // extern "C"
