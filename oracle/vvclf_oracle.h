/*
 * vvclf_oracle.h — CPU restatement of the reference VVC in-loop filters.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  The
 * product library (libvvclf.so) never links, loads or calls it.
 *
 * Every function follows the cited reference lines step by step, on the same
 * POD types as the product's C ABI (include/vvclf.h).
 */
#ifndef VVCLF_ORACLE_H
#define VVCLF_ORACLE_H

#include "../include/vvclf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/lmcs.rs:302-404 */
int oracle_derive_lmcs(const vvclf_lmcs_data* data, uint32_t bit_depth, vvclf_lmcs_derived* out);
/* src/lmcs.rs:209-220 */
uint32_t oracle_lmcs_idx_y_inv(const vvclf_lmcs_derived* d, uint32_t s, uint32_t min_bin, uint32_t max_bin);
/* src/lmcs.rs:230-238 */
uint32_t oracle_lmcs_inverse_map(const vvclf_lmcs_derived* d, uint32_t s, uint32_t idx);
/* src/lmcs.rs:186-195 */
int64_t oracle_lmcs_forward_map(const vvclf_lmcs_derived* d, uint32_t pred_sample);
/* src/ctu.rs:11837-11849 */
void oracle_lmcs_inverse_luma(vvclf_plane* luma, const vvclf_lmcs_binding* b);

/* src/ctu.rs:1701-1725; -1 (the reference panics) for a sample >= 2^BitDepth */
int oracle_lmcs_forward_luma_rect(vvclf_plane* p, const vvclf_lmcs_binding* b, uint32_t x0, uint32_t y0, uint32_t w,
                                  uint32_t h);
/* src/ctu.rs:1790-1814 */
uint32_t oracle_lmcs_chroma_var_scale(const vvclf_plane* luma, const vvclf_lmcs_binding* b, uint32_t x_cu,
                                      uint32_t y_cu, int avail_l, int avail_t, uint32_t size_y);
/* src/lmcs.rs:257-279 */
uint32_t oracle_lmcs_scale_chroma_residual_sample(const vvclf_lmcs_derived* d, int64_t pred, int64_t res,
                                                  uint32_t var_scale);
/* src/ctu.rs:734-742 and the chroma TB add, per block */
int oracle_lmcs_chroma_residual(vvclf_picture* pic, const vvclf_lmcs_binding* b, const vvclf_crs_block* blocks,
                                uint32_t n, const int32_t* res);

/* src/deblock.rs:424-452 */
int32_t oracle_beta_prime(int32_t q);
int32_t oracle_tc_prime(int32_t q);
int32_t oracle_scale_tc(int32_t tc_prime, uint32_t bit_depth);
int32_t oracle_scale_beta(int32_t beta_prime, uint32_t bit_depth);
/* src/deblock.rs:168-291 (tu_coded_at, luma_tb_len, chroma_tb_len, is_*_tb_boundary) */
int oracle_tu_coded_at(const vvclf_deblock_cu* cu, uint32_t c_idx, int32_t lx, int32_t ly);
uint32_t oracle_luma_tb_len(const vvclf_deblock_cu* cu, int32_t lx, int32_t ly, int vertical);
uint32_t oracle_chroma_tb_len(const vvclf_deblock_cu* cu, int32_t lx, int32_t ly, int vertical, uint32_t sub);
int oracle_is_luma_tb_boundary(const vvclf_deblock_cu* cu, uint32_t off, int vertical);
int oracle_is_chroma_tb_boundary(const vvclf_deblock_cu* cu, uint32_t off, int vertical);
/* src/deblock.rs:497-617 */
void oracle_apply_deblocking_clipped(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                                     const vvclf_deblock_cu* chroma_cus, uint32_t n_chroma_cus,
                                     const vvclf_mv_grid* mv_grid, const vvclf_deblock_params* params,
                                     uint32_t chroma_format_idc, const uint32_t* no_filter_cols,
                                     uint32_t n_no_filter_cols, const uint32_t* no_filter_rows,
                                     uint32_t n_no_filter_rows);

/* src/sao.rs:389-497 */
void oracle_apply_sao_clipped(vvclf_picture* out, const vvclf_sao_picture* sao_pic,
                              const vvclf_sao_config* cfg, const vvclf_bounds* tile_bounds);
/* src/sao.rs:356-365 */
void oracle_sao_derive_offset_val(const uint32_t offset_abs[4], const uint32_t sign[4], uint32_t bit_depth,
                                  int32_t out[5]);

/* src/alf.rs:627-644 */
int32_t oracle_alf_clip_value(uint32_t bit_depth, uint8_t clip_idx);
/* src/alf.rs:1117-1143, :528-539 */
void oracle_alf_table_45(int32_t y, int32_t ctb, int lbb, int32_t out[4]);
void oracle_alf_table_46(int32_t y, int32_t ctb_h_c, int lbb, int32_t out[3]);
void oracle_alf_table_47(int32_t y_line, int32_t ctb, int lbb, int32_t out[2]);
/* src/alf.rs:731-742 */
void oracle_alf_transpose_idx(uint8_t t, uint8_t out[12]);
/* src/alf_fixed.rs:173-185: returns 0 and fills row, or -1 for None */
int oracle_alf_fixed_coeff_row(uint32_t set_idx, uint32_t class_idx, int32_t row[12]);
/* src/alf.rs:822-1010: cls[(sy*(ctb/4)+sx)] = filt_idx | transpose_idx << 8 */
void oracle_alf_classify_ctb(const uint16_t* pre, uint32_t stride, int32_t pw, int32_t ph, const int32_t* clip,
                             uint32_t rx, uint32_t ry, uint32_t ctb_size_y, uint32_t bit_depth, uint16_t* cls);
/* src/alf.rs:235-434; returns VVCLF_ERR_INVALID on an APS whose Vec lengths the
 * reference would index out of range, VVCLF_OK otherwise. */
int oracle_apply_alf_clipped(vvclf_picture* out, const vvclf_alf_picture* alf_pic, const vvclf_alf_config* cfg,
                             const vvclf_alf_aps_binding* binding, const vvclf_bounds* tile_bounds);

/* src/alf_enc.rs:185-316 (and :350-470 with APS sets) */
int oracle_alf_decide_luma(const vvclf_picture* src, vvclf_picture* rec, uint32_t ctb_log2, uint32_t bd, uint32_t cfi,
                           const vvclf_alf_aps* const* apses, uint32_t n_aps, const vvclf_bounds* tb,
                           vvclf_alf_ctb* out, uint64_t* sse);
/* src/alf_enc.rs:628-806: chroma ALF alternative per CTB for component c_idx
 * (1 Cb, 2 Cr); apply is the CTB array (in / out), sse [n_ctbs][1 + n_alts] */
int oracle_alf_decide_chroma(const vvclf_picture* src, vvclf_picture* rec, vvclf_alf_ctb* apply,
                             const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2, uint32_t bd, uint32_t cfi,
                             const vvclf_bounds* tb, uint64_t* sse);
/* src/alf_enc.rs:935-1140: CC-ALF filter per CTB from the pre-luma-ALF plane;
 * sse [n_ctbs][1 + n_filters] */
int oracle_cc_alf_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma,
                         vvclf_alf_ctb* apply, const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2,
                         uint32_t bd, uint32_t cfi, const vvclf_bounds* tb, uint64_t* sse);
/* src/sao_enc.rs:385-462: per-CTB SAO parameters of a picture (raster CTB
 * order), from per-CTB band and edge statistics of rec against src */
int oracle_sao_decide_picture(const vvclf_picture* src, const vvclf_picture* rec, uint32_t ctb_log2, uint32_t bd,
                              uint32_t sao_luma, uint32_t sao_chroma, vvclf_sao_ctb_params* out);
/* src/sao_enc.rs:525-660: chroma SAO merge decision over independent
 * parameters (map: 0 independent, 1 merge left, 2 merge above) */
int oracle_sao_chroma_merge(const vvclf_picture* src, const vvclf_picture* rec, const vvclf_sao_ctb_params* indep,
                            uint32_t bd, uint32_t ctb_log2, vvclf_sao_ctb_params* merged, uint8_t* map);
/* src/alf_aps_design.rs:697-768: greedy clip-index search of a 25-row luma
 * APS by full-picture replays */
int oracle_alf_clip_rdo_luma(const vvclf_picture* src, const vvclf_picture* rec, const int32_t coeff[25][12],
                             uint32_t ctb_log2, uint32_t bd, uint32_t cfi, vvclf_alf_clip_rdo* out);

/* src/ctu.rs:11822-12066 */
int oracle_apply_in_loop_filters(vvclf_frame* frame);

/* Runs oracle_apply_in_loop_filters on n frames with up to n_threads host
 * threads (one frame per thread at a time); used for the CPU baseline. */
int oracle_apply_in_loop_filters_mt(vvclf_frame* const* frames, uint32_t n, uint32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif
