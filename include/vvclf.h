/*
 * vvclf.h — C ABI of the MI355X (gfx950) VVC in-loop filter chain:
 * LMCS inverse luma mapping -> deblocking -> SAO -> ALF + CC-ALF.
 *
 * The POD types mirror the Rust types the reference filters consume, field for
 * field, with Rust `Option<T>` turned into a `has_*` flag or a NULL pointer and
 * `Vec<T>` / slices turned into (pointer, count).  Every entry point returns an
 * int: VVCLF_OK (0) on success, a negative VVCLF_ERR_* code otherwise.  The
 * reference functions are infallible; the negative codes only report
 * descriptors the reference would panic on, or device/runtime failures.
 *
 * Two layers:
 *   1. Drop-in, synchronous, host-pointer entry points: one per reference
 *      filter entry (vvclf_apply_deblocking[_clipped], vvclf_apply_sao[_clipped],
 *      vvclf_apply_alf[_clipped], vvclf_lmcs_inverse_luma) plus the picture-level
 *      driver vvclf_apply_in_loop_filters.  Planes are modified in place, like
 *      the reference's `&mut PictureBuffer`.  Work runs on the GPU.
 *   2. Device-resident additions with no reference counterpart: contexts, jobs
 *      (one picture with its side information uploaded to HBM), batches (many
 *      jobs filtered by one persistent fused launch, optionally captured in a
 *      HIP graph) and a serving loop (continuous batching over many streams).
 *
 * Only 4:2:0 (chroma_format_idc 1) and 4:0:0 (chroma_format_idc 0) pictures run
 * on the GPU; other chroma formats return VVCLF_ERR_UNSUPPORTED.
 */
#ifndef VVCLF_H
#define VVCLF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VVCLF_OK 0
#define VVCLF_ERR_INVALID (-1)     /* descriptor the reference would reject or panic on */
#define VVCLF_ERR_HIP (-2)         /* HIP runtime error (message in vvclf_last_error) */
#define VVCLF_ERR_NOMEM (-3)       /* host or device allocation failed */
#define VVCLF_ERR_BUDGET (-4)      /* serving: job larger than the admission budget */
#define VVCLF_ERR_TIMEOUT (-5)     /* serving: wait timed out */
#define VVCLF_ERR_UNSUPPORTED (-6) /* chroma format other than 4:0:0 / 4:2:0 on the GPU */
#define VVCLF_ERR_AGAIN (-7)       /* serving: result not ready yet */

/* ------------------------------------------------------------------------ */
/* Picture buffers — PicturePlane / PictureBuffer, src/reconstruct.rs:41-52,  */
/* :116-145.  Row-major u16 samples, stride >= width, bit_depth 8..16.        */
/* ------------------------------------------------------------------------ */
typedef struct vvclf_plane {
  uint16_t* samples; /* stride * height samples */
  uint32_t stride;
  uint32_t width;
  uint32_t height;
  uint32_t bit_depth;
} vvclf_plane;

typedef struct vvclf_picture {
  vvclf_plane luma;
  vvclf_plane cb;
  vvclf_plane cr;
} vvclf_picture;

/* ------------------------------------------------------------------------ */
/* Deblocking inputs — src/deblock.rs:52-161 (DeblockCu, TbSplit, Tu64CbfMap),*/
/* :330-364 (DeblockMvCell, DeblockMvGrid), :376-393 (DeblockParams).         */
/* ------------------------------------------------------------------------ */
#define VVCLF_QP_C_LEGACY INT32_MIN /* DEBLOCK_QP_C_LEGACY, src/deblock.rs:368 */

typedef struct vvclf_deblock_cu {
  uint32_t x, y, w, h;   /* luma rectangle */
  int32_t qp_y;
  int32_t qp_c[3];       /* [QpCb, QpCr, QpCbCr] or VVCLF_QP_C_LEGACY in qp_c[0] */
  uint8_t intra, tu_y_coded, tu_cb_coded, tu_cr_coded;
  uint8_t bdpcm_luma, bdpcm_chroma, plt, joint_cbcr2;
  uint8_t ciip, ibc;
  uint8_t has_num_sb;    /* num_sb: Option<(u32, u32)> */
  uint8_t has_tu64;      /* tu64: Option<Tu64CbfMap> */
  uint32_t num_sb_x, num_sb_y;
  uint8_t tu64_y[2][2];  /* [tile_y][tile_x] */
  uint8_t tu64_cb[2][2];
  uint8_t tu64_cr[2][2];
  uint8_t has_tb_split;  /* tb_split: Option<TbSplit> */
  uint8_t tb_vertical, tb_n_bounds, tb_luma_only;
  uint32_t tb_bounds[3];
  uint8_t tb_y_coded[4], tb_cb_coded[4], tb_cr_coded[4];
} vvclf_deblock_cu;

typedef struct vvclf_mv_cell {
  uint8_t pred_l0, pred_l1, pad0, pad1;
  int32_t poc_l0, poc_l1;
  int32_t mv_l0[2], mv_l1[2]; /* 1/16 luma sample units */
} vvclf_mv_cell;

typedef struct vvclf_mv_grid {
  const vvclf_mv_cell* cells; /* cells_w * cells_h, one per 4x4 luma cell */
  uint32_t cells_w, cells_h;
} vvclf_mv_grid;

typedef struct vvclf_deblock_params {
  int32_t disabled;
  int32_t luma_beta_offset_div2, luma_tc_offset_div2;
  int32_t cb_beta_offset_div2, cb_tc_offset_div2;
  int32_t cr_beta_offset_div2, cr_tc_offset_div2;
  int32_t chroma_qp_offset_cb, chroma_qp_offset_cr;
  uint32_t bit_depth;
  uint32_t ctb_log2_size_y;
} vvclf_deblock_params;

/* ------------------------------------------------------------------------ */
/* SAO inputs — src/sao.rs:53-352.                                          */
/* ------------------------------------------------------------------------ */
#define VVCLF_SAO_NOT_APPLIED 0
#define VVCLF_SAO_BAND 1
#define VVCLF_SAO_EDGE 2

typedef struct vvclf_sao_ctb {
  uint8_t sao_type_idx;  /* 0 not applied, 1 band offset, 2 edge offset */
  uint8_t eo_class;      /* 0 horizontal, 1 vertical, 2 135 deg, 3 45 deg */
  uint8_t band_position; /* 0..31 */
  uint8_t pad;
  int32_t offset_val[5]; /* SaoOffsetVal after the eq. 153 shift; [0] == 0 */
} vvclf_sao_ctb;

typedef struct vvclf_sao_ctb_params {
  vvclf_sao_ctb luma, cb, cr;
} vvclf_sao_ctb_params;

typedef struct vvclf_sao_picture {
  uint32_t pic_width_in_ctbs_y, pic_height_in_ctbs_y;
  const vvclf_sao_ctb_params* ctbs; /* raster CTB order */
} vvclf_sao_picture;

typedef struct vvclf_sao_config {
  uint8_t luma_used, chroma_used, pad0, pad1;
  uint32_t bit_depth, ctb_log2_size_y, chroma_format_idc;
} vvclf_sao_config;

/* Tile / slice clip bounds: prefix lists in luma samples, including 0 and the
 * picture extents (the `tile_bounds: Option<(&[u32], &[u32])>` argument of
 * apply_sao_clipped / apply_alf_clipped).  A NULL pointer means None. */
typedef struct vvclf_bounds {
  const uint32_t* cols;
  uint32_t n_cols;
  const uint32_t* rows;
  uint32_t n_rows;
} vvclf_bounds;

/* ------------------------------------------------------------------------ */
/* ALF inputs — src/alf.rs:67-206 and AlfApsData, src/aps.rs:21-102.         */
/* ------------------------------------------------------------------------ */
#define VVCLF_ALF_NUM_FILTERS 25
#define VVCLF_ALF_LUMA_NUM_COEFF 12
#define VVCLF_ALF_CHROMA_NUM_COEFF 6
#define VVCLF_ALF_CC_NUM_COEFF 7
#define VVCLF_ALF_MAX_CHROMA_ALT 8
#define VVCLF_ALF_MAX_CC_FILTERS 4

typedef struct vvclf_alf_ctb {
  uint8_t luma_on, cb_on, cr_on;
  uint8_t luma_filt_set_idx; /* <16 fixed set, >=16 APS slot idx-16 */
  uint8_t cb_alt_idx, cr_alt_idx;
  uint8_t cc_cb_idc, cc_cr_idc; /* 0 off, k -> filter k-1 */
} vvclf_alf_ctb;

typedef struct vvclf_alf_picture {
  uint32_t pic_width_in_ctbs_y, pic_height_in_ctbs_y;
  const vvclf_alf_ctb* ctbs; /* raster CTB order */
} vvclf_alf_picture;

typedef struct vvclf_alf_config {
  uint8_t alf_enabled, cb_enabled, cr_enabled, cc_cb_enabled, cc_cr_enabled;
  uint8_t pad0, pad1, pad2;
  uint32_t bit_depth, ctb_log2_size_y, chroma_format_idc;
} vvclf_alf_config;

/* AlfApsData with its Vec fields as fixed arrays plus their lengths. */
typedef struct vvclf_alf_aps {
  uint8_t alf_luma_filter_signal_flag, alf_chroma_filter_signal_flag;
  uint8_t alf_cc_cb_filter_signal_flag, alf_cc_cr_filter_signal_flag;
  uint8_t alf_luma_clip_flag, alf_chroma_clip_flag, pad0, pad1;
  uint32_t alf_chroma_num_alt_filters_minus1;
  uint32_t n_luma_coeff;   /* luma_coeff.len(): 25 when the luma filter is signalled */
  uint32_t n_chroma_coeff; /* chroma_coeff.len(), <= 8 */
  uint32_t n_cc_cb_coeff;  /* cc_cb_coeff.len(), <= 4 */
  uint32_t n_cc_cr_coeff;  /* cc_cr_coeff.len(), <= 4 */
  int32_t luma_coeff[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF];
  uint8_t luma_clip_idx[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF];
  int32_t chroma_coeff[VVCLF_ALF_MAX_CHROMA_ALT][VVCLF_ALF_CHROMA_NUM_COEFF];
  uint8_t chroma_clip_idx[VVCLF_ALF_MAX_CHROMA_ALT][VVCLF_ALF_CHROMA_NUM_COEFF];
  int32_t cc_cb_coeff[VVCLF_ALF_MAX_CC_FILTERS][VVCLF_ALF_CC_NUM_COEFF];
  int32_t cc_cr_coeff[VVCLF_ALF_MAX_CC_FILTERS][VVCLF_ALF_CC_NUM_COEFF];
} vvclf_alf_aps;

typedef struct vvclf_alf_aps_binding {
  const vvclf_alf_aps* const* luma_apses; /* entries may be NULL (None slots) */
  uint32_t n_luma_apses;
  uint32_t pad;
  const vvclf_alf_aps* chroma_aps; /* NULL = None */
  const vvclf_alf_aps* cc_cb_aps;
  const vvclf_alf_aps* cc_cr_aps;
} vvclf_alf_aps_binding;

/* ------------------------------------------------------------------------ */
/* LMCS — LmcsData / LmcsDerived, src/lmcs.rs:51-153; LmcsBinding,           */
/* src/ctu.rs:712-725.                                                       */
/* ------------------------------------------------------------------------ */
#define VVCLF_LMCS_NUM_BINS 16

typedef struct vvclf_lmcs_data {
  uint8_t lmcs_min_bin_idx, lmcs_delta_max_bin_idx, lmcs_delta_cw_prec_minus1, lmcs_delta_abs_crs;
  uint8_t lmcs_delta_sign_crs_flag, pad0, pad1, pad2;
  uint32_t lmcs_delta_abs_cw[VVCLF_LMCS_NUM_BINS];
  uint8_t lmcs_delta_sign_cw_flag[VVCLF_LMCS_NUM_BINS];
} vvclf_lmcs_data;

typedef struct vvclf_lmcs_derived {
  uint32_t bit_depth, org_cw;
  uint32_t lmcs_cw[VVCLF_LMCS_NUM_BINS];
  uint32_t input_pivot[VVCLF_LMCS_NUM_BINS];
  uint32_t lmcs_pivot[VVCLF_LMCS_NUM_BINS + 1];
  uint32_t scale_coeff[VVCLF_LMCS_NUM_BINS];
  uint32_t inv_scale_coeff[VVCLF_LMCS_NUM_BINS];
  int32_t lmcs_delta_crs;
  uint32_t chroma_scale_coeff[VVCLF_LMCS_NUM_BINS];
} vvclf_lmcs_derived;

typedef struct vvclf_lmcs_binding {
  const vvclf_lmcs_derived* derived; /* NULL = LMCS not active */
  uint8_t min_bin_idx, max_bin_idx, pad0, pad1;
} vvclf_lmcs_binding;

/* ------------------------------------------------------------------------ */
/* Picture-level descriptor for the driver, src/ctu.rs:11822-12066.         */
/* ------------------------------------------------------------------------ */
typedef struct vvclf_frame {
  vvclf_picture pic; /* in/out */
  uint32_t chroma_format_idc;
  uint32_t pad0;
  vvclf_lmcs_binding lmcs; /* lmcs_active() when derived != NULL */
  /* deblocking (apply_deblocking_clipped arguments) */
  const vvclf_deblock_cu* cus;
  uint32_t n_cus;
  uint32_t has_chroma_cus; /* chroma_cus: Option<&[DeblockCu]> */
  const vvclf_deblock_cu* chroma_cus;
  uint32_t n_chroma_cus;
  uint32_t pad1;
  const vvclf_mv_grid* mv_grid; /* NULL = None */
  vvclf_deblock_params deblock;
  const uint32_t* no_filter_cols;
  uint32_t n_no_filter_cols;
  uint32_t n_no_filter_rows;
  const uint32_t* no_filter_rows;
  /* SAO */
  vvclf_sao_picture sao;
  vvclf_sao_config sao_cfg;
  /* ALF */
  vvclf_alf_picture alf;
  vvclf_alf_config alf_cfg;
  vvclf_alf_aps_binding alf_binding;
  /* SAO / ALF tile-slice clip bounds (NULL = None) */
  const vvclf_bounds* clip_bounds;
} vvclf_frame;

/* ======================================================================== */
/* Layer 1: drop-in synchronous entry points (host pointers, in place).      */
/* ======================================================================== */

/* Replaces `apply_deblocking(out, cus, params, chroma_format_idc)`,
 * src/deblock.rs:480-487. */
int vvclf_apply_deblocking(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                           const vvclf_deblock_params* params, uint32_t chroma_format_idc);

/* Replaces `apply_deblocking_clipped(out, cus, chroma_cus, mv_grid, params,
 * chroma_format_idc, no_filter_cols, no_filter_rows)`, src/deblock.rs:497-617.
 * chroma_cus == NULL means None; mv_grid == NULL means None. */
int vvclf_apply_deblocking_clipped(vvclf_picture* out, const vvclf_deblock_cu* cus, uint32_t n_cus,
                                   const vvclf_deblock_cu* chroma_cus, uint32_t n_chroma_cus,
                                   const vvclf_mv_grid* mv_grid, const vvclf_deblock_params* params,
                                   uint32_t chroma_format_idc, const uint32_t* no_filter_cols,
                                   uint32_t n_no_filter_cols, const uint32_t* no_filter_rows,
                                   uint32_t n_no_filter_rows);

/* Replaces `apply_sao(out, sao_pic, cfg)`, src/sao.rs:379-381. */
int vvclf_apply_sao(vvclf_picture* out, const vvclf_sao_picture* sao_pic, const vvclf_sao_config* cfg);

/* Replaces `apply_sao_clipped(out, sao_pic, cfg, tile_bounds)`, src/sao.rs:389-497. */
int vvclf_apply_sao_clipped(vvclf_picture* out, const vvclf_sao_picture* sao_pic,
                            const vvclf_sao_config* cfg, const vvclf_bounds* tile_bounds);

/* Replaces `apply_alf(out, alf_pic, cfg, binding)`, src/alf.rs:220-227. */
int vvclf_apply_alf(vvclf_picture* out, const vvclf_alf_picture* alf_pic, const vvclf_alf_config* cfg,
                    const vvclf_alf_aps_binding* binding);

/* Replaces `apply_alf_clipped(out, alf_pic, cfg, binding, tile_bounds)`,
 * src/alf.rs:235-434. */
int vvclf_apply_alf_clipped(vvclf_picture* out, const vvclf_alf_picture* alf_pic,
                            const vvclf_alf_config* cfg, const vvclf_alf_aps_binding* binding,
                            const vvclf_bounds* tile_bounds);

/* Replaces the per-picture LMCS inverse-mapping loop of the driver,
 * src/ctu.rs:11837-11849, built on LmcsDerived::idx_y_inv /
 * inverse_map_luma_sample, src/lmcs.rs:209-238.  No-op when binding->derived
 * is NULL (lmcs_active() false, src/ctu.rs:1675-1677). */
int vvclf_lmcs_inverse_luma(vvclf_plane* luma, const vvclf_lmcs_binding* binding);

/* Replaces `CtuWalker::apply_in_loop_filters_with_alf(out, alf_binding)`,
 * src/ctu.rs:11822-12066, with the walker state (SPS/PPS/slice header fields,
 * CU records, motion grid, SAO/ALF pictures, gating lists) flattened into
 * the frame descriptor.  Order: LMCS inverse, deblock, SAO, ALF + CC-ALF. */
int vvclf_apply_in_loop_filters(vvclf_frame* frame);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture around the call, tight serving loops,
 * harnesses that forbid hipMalloc): the same result as
 * vvclf_apply_in_loop_filters, with every device buffer of the call (three
 * sample sets and the per-picture side tables) taken from a fixed scratch
 * array of VVCLF_NOALLOC_SCRATCH_BYTES that each GPU holds until the process
 * ends, and samples moved by pageable copies (no pinned buffer either); host
 * working memory of the table derivation still comes from the C++ heap.
 * Runs on the calling thread's current device.  Pictures whose sample sets
 * and tables do not fit (e.g. above 3840x2160 4:2:0) return
 * VVCLF_ERR_UNSUPPORTED before any device work.  Calls on one device are
 * serialised by a lock, since they share the array; each call writes every
 * byte it later reads. */
#define VVCLF_NOALLOC_SCRATCH_BYTES (96ull << 20)
int vvclf_apply_in_loop_filters_noalloc(vvclf_frame* frame);

/* Replaces `derive_lmcs(data, bit_depth)`, src/lmcs.rs:302-404 (host-side). */
int vvclf_derive_lmcs(const vvclf_lmcs_data* data, uint32_t bit_depth, vvclf_lmcs_derived* out);

/* ------------------------------------------------------------------------ */
/* The steps either side of the chain: LMCS forward mapping and chroma       */
/* residual scaling during reconstruction (before the filters), and the      */
/* conformance crop and plane hashes of the output (after them).             */
/* ------------------------------------------------------------------------ */
typedef struct vvclf_rect {
  uint32_t x, y, w, h;
} vvclf_rect;

/* Replaces CtuWalker::lmcs_forward_map_luma_rect(out, x0, y0, w, h),
 * src/ctu.rs:1701-1725, called once per rectangle (the inter CUs whose
 * prediction is forward-mapped before their residual is added): every luma
 * sample s inside a rectangle (clipped to the plane) becomes
 * Clip3(0, 2^BitDepth - 1, forward_map_luma_sample(s)), src/lmcs.rs:186-195.
 * Rectangles must not overlap.  No-op when binding->derived is NULL.  A
 * sample of 2^BitDepth or more inside a rectangle (the reference indexes past
 * its 16 pieces and panics) returns VVCLF_ERR_INVALID and leaves the plane
 * unchanged. */
int vvclf_lmcs_forward_luma(vvclf_plane* luma, const vvclf_lmcs_binding* binding, const vvclf_rect* rects,
                            uint32_t n_rects);

/* One chroma transform block reconstructed with LMCS chroma residual
 * scaling. */
typedef struct vvclf_crs_block {
  uint32_t c_idx;        /* 1 Cb, 2 Cr */
  uint32_t x, y, w, h;   /* the transform block, chroma samples */
  uint32_t x_cu, y_cu;   /* (xCuCb, yCuCb): luma origin of the CU holding the sizeY-aligned corner */
  uint8_t avail_l, avail_t, pad0, pad1; /* availL / availT of that CU */
  uint32_t size_y;       /* sizeY = Min(CtbSizeY, 64) */
  uint32_t pad2;
  uint64_t res_offset;   /* index of the block's first residual in `residuals` (w * h, row-major) */
} vvclf_crs_block;

/* Replaces, per block, CtuWalker::lmcs_chroma_var_scale_for_cu
 * (src/ctu.rs:1752-1815: invAvgLuma from the left column / top row of the
 * luma plane, eqs. 1216/1217, then idxYInv and ChromaScaleCoeff), the residual
 * fold lmcs_scale_chroma_residuals (src/ctu.rs:734-742, eqs. 1219/1220) and
 * the add recSamples = Clip1(predSamples + resSamples) of the chroma TB: the
 * chroma planes of `pic` hold the prediction on entry and the reconstruction
 * on return; the luma plane (mapped-domain reconstructed luma) is read only.
 * Blocks must not overlap (VVCLF_ERR_INVALID otherwise).  No-op when
 * binding->derived is NULL. */
int vvclf_lmcs_chroma_residual(vvclf_picture* pic, const vvclf_lmcs_binding* binding,
                               const vvclf_crs_block* blocks, uint32_t n_blocks, const int32_t* residuals,
                               uint64_t n_residuals);

/* The conformance crop of a decoded picture, src/stream.rs:816-842: offsets
 * (left, right, top, bottom) in chroma units of the conformance window, or
 * NULL for none, give out = (left * SubWidthC, top * SubHeightC, width,
 * height) in luma samples. */
typedef struct vvclf_crop {
  uint32_t x, y, w, h;
} vvclf_crop;
int vvclf_conformance_crop(uint32_t pic_w, uint32_t pic_h, uint32_t chroma_format_idc, const uint32_t* window,
                           vvclf_crop* out);
/* MD5 of each cropped plane in the conformance-stream convention,
 * tests/jvet_conformance.rs:400-430: rows of the crop, one byte per sample at
 * 8 bits, two bytes little-endian above; chroma crop = luma crop / 2 (4:2:0).
 * md5[c] receives plane c (luma, Cb, Cr); chroma entries are left untouched
 * when chroma_format_idc is 0.  Host-side. */
int vvclf_picture_md5(const vvclf_picture* pic, uint32_t chroma_format_idc, uint32_t bit_depth,
                      const vvclf_crop* crop, uint8_t md5[3][16]);

/* ------------------------------------------------------------------------ */
/* Encoder: per-CTB luma ALF decision                                       */
/* ------------------------------------------------------------------------ */
/* Replaces alf_decide_and_apply_clipped (src/alf_enc.rs:185-316; n_aps = 0)
 * and alf_decide_and_apply_with_aps_clipped (src/alf_enc.rs:352-470; n_aps =
 * 1), generalised to n_aps APS luma sets (apses[i] is candidate 16 + i, as
 * AlfApsBinding::luma_apses slot i).  Every CTB of `rec` (pre-ALF luma) is
 * filtered with each of the 16 fixed sets and each APS set; per CTB the luma
 * SSE against `src` (the CTB clipped to the picture, ctb_sse_y) is compared
 * with the SSE of leaving it unfiltered, and the first candidate with a
 * strictly lower SSE, in the order off, set 0 .. 15, APS 0 .. n_aps - 1,
 * is chosen.  On return out[i] (raster CTB order, ceil(W / CtbSizeY) x
 * ceil(H / CtbSizeY) entries) holds luma_on and luma_filt_set_idx of CTB i
 * (all other fields zero) and rec's luma holds the chosen samples; chroma is
 * not touched.  `sse` (optional) receives per CTB 1 + 16 + n_aps values:
 * off, the 16 fixed sets, the APS sets.  An absent or unsignalled APS leaves
 * the CTB unfiltered in its trial (its SSE equals the off SSE).  tile_bounds
 * as for the ALF stage (NULL: picture-wide clip).  Host pictures; the trials
 * run on the GPU, one pass over the picture for all candidates. */
int vvclf_alf_decide_luma(const vvclf_picture* src, vvclf_picture* rec, uint32_t ctb_log2_size_y, uint32_t bit_depth,
                          uint32_t chroma_format_idc, const vvclf_alf_aps* const* apses, uint32_t n_aps,
                          const vvclf_bounds* tile_bounds, vvclf_alf_ctb* out, uint64_t* sse);

/* Replaces chroma_alf_decide_and_apply_clipped (src/alf_enc.rs:628-806) for
 * component c_idx (1 Cb, 2 Cr): every CTB's chroma is filtered with each
 * alternative of `aps` (chroma_coeff / chroma_clip_idx rows 0 ..
 * n_chroma_coeff - 1); per CTB the first alternative whose SSE against `src`
 * (the chroma CTB clipped to the plane) is strictly below the SSE of leaving
 * it unfiltered is kept: its samples are written into rec's plane and
 * apply[i] gets cb_on / cb_alt_idx (cr_on / cr_alt_idx); the records of other
 * CTBs, and every other field, keep what they held (apply: raster CTB order,
 * ceil(W / CtbSizeY) x ceil(H / CtbSizeY)).  `sse` (optional) receives per CTB
 * 1 + n_chroma_coeff values: off, then each alternative.  A no-op returning
 * VVCLF_OK where the reference returns early (chroma_format_idc 0, no
 * signalled chroma filter); 4:2:0 only on the GPU.  Host pictures. */
int vvclf_alf_decide_chroma(const vvclf_picture* src, vvclf_picture* rec, vvclf_alf_ctb* apply,
                            const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2_size_y, uint32_t bit_depth,
                            uint32_t chroma_format_idc, const vvclf_bounds* tile_bounds, uint64_t* sse);

/* Replaces cc_alf_decide_and_apply_clipped (src/alf_enc.rs:935-1140): the
 * same decision over the CC-ALF filters of `aps` for component c_idx
 * (cc_cb_coeff / cc_cr_coeff rows 0 .. n - 1; idc k + 1), the correction
 * taken from `pre_luma`, the luma plane before luma ALF (as the reference's
 * pre_luma_alf_samples) and added to rec's chroma; apply[i].cc_cb_idc /
 * cc_cr_idc of the CTBs that take a filter.  `sse`: off, then each filter. */
int vvclf_cc_alf_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma,
                        vvclf_alf_ctb* apply, const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2_size_y,
                        uint32_t bit_depth, uint32_t chroma_format_idc, const vvclf_bounds* tile_bounds,
                        uint64_t* sse);

/* Replaces sao_decide_picture (src/sao_enc.rs:385-462): per CTB (raster
 * order, out has ceil(W / CtbSizeY) x ceil(H / CtbSizeY) entries) and
 * component, the SAO parameters the reference's encoder picks from the band
 * and edge statistics of `rec` (the pre-SAO reconstruction) against `src`
 * over the CTB (chroma: half the luma CTB clipped to the picture, rounded
 * down): edge offset (first of the four classes with the largest gain) when
 * one improves, else band offset (the first of the 4-band windows starting
 * at 0..27 with the largest gain), else not applied; Cr keeps Cb's type and
 * edge class (zero offsets when nothing improves).  sao_luma / sao_chroma
 * as the slice flags; 4:2:0 or 4:0:0 (empty chroma planes).  A
 * reconstruction sample of 2^BitDepth or more that reaches a band statistic
 * returns VVCLF_ERR_INVALID (the reference indexes past its 32 bands).
 * Host pictures; the statistics and the choice run on the GPU. */
int vvclf_sao_decide_picture(const vvclf_picture* src, const vvclf_picture* rec, uint32_t ctb_log2_size_y,
                             uint32_t bit_depth, uint32_t sao_luma, uint32_t sao_chroma, vvclf_sao_ctb_params* out);

/* Replaces apply_chroma_sao_merge (src/sao_enc.rs:525-660): per CTB in
 * raster order, keep the independent chroma parameters (`independent`, as
 * vvclf_sao_decide_picture gives them) or take those already decided for the
 * left or the above neighbour, by the smallest Cb + Cr SSE (SAO of the
 * candidate on `rec_pre_sao`'s chroma CTB against `src`) + 8 * bits (the
 * reference's bit estimates: 1 + per component 1 / 16 / 10 for not applied /
 * band / edge independently, 1 for merge left, 2 for merge above; ties keep
 * the earlier option).  merged: luma as given, chroma as decided;
 * merge_map[i]: 0 independent, 1 merge left, 2 merge above.  On the GPU the
 * CTBs are decided in anti-diagonal waves. */
int vvclf_sao_chroma_merge(const vvclf_picture* src, const vvclf_picture* rec_pre_sao,
                           const vvclf_sao_ctb_params* independent, uint32_t bit_depth, uint32_t ctb_log2_size_y,
                           vvclf_sao_ctb_params* merged, uint8_t* merge_map);

/* DesignedLumaAlfClipRdo, src/alf_aps_design.rs:552-597. */
typedef struct vvclf_alf_clip_rdo {
  int32_t coeff[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF]; /* the input rows, carried over */
  uint8_t clip_idx[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF];
  uint8_t alf_luma_clip_flag, pad[3];
  uint64_t baseline_sse_y;  /* luma SSE with clip index 0 everywhere */
  uint64_t post_clip_sse_y; /* luma SSE with clip_idx */
} vvclf_alf_clip_rdo;

/* Replaces design_clip_rdo_for_luma_aps (src/alf_aps_design.rs:697-768): a
 * greedy search of alf_luma_clip_idx for an APS whose 25 luma rows are
 * `coeff`.  From clip index 0 everywhere, for filtIdx 0 .. 24 and tap
 * j 0 .. 11 in that order, indices 1, 2, 3 are tried; the one giving the
 * lowest luma SSE against `src` (ALF on `rec_pre_alf` with every CTB on the
 * APS, luma only, picture-wide clip) is kept if it is strictly below the
 * current SSE (ties keep the earlier index), else the tap stays 0.
 * alf_luma_clip_flag: any index nonzero.  On the GPU: a 4x4 block's ALF
 * output depends only on its class's row, so the SSE splits over classes
 * and the 25 rows are searched together: per tap one pass over the picture
 * filters each block with its row's three trial indices and sums the
 * squared error per class (13 passes against the reference's 901 picture
 * replays); the picks and the next trial banks are made on the device.
 * Coefficients outside 16 bits: VVCLF_ERR_UNSUPPORTED.  Host pictures,
 * luma read. */
int vvclf_alf_clip_rdo_luma(const vvclf_picture* src, const vvclf_picture* rec_pre_alf,
                            const int32_t coeff[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF],
                            uint32_t ctb_log2_size_y, uint32_t bit_depth, uint32_t chroma_format_idc,
                            vvclf_alf_clip_rdo* out);

/* ------------------------------------------------------------------------ */
/* Per-CTB side information as a compact record: the syntax elements a      */
/* slice decoder has parsed for one CTB, from which the SAO and ALF CTB      */
/* parameters follow (the inference of decode_sao_ctb,                       */
/* src/sao_syntax.rs:154-284, and of decode_alf_ctu,                         */
/* src/alf_syntax.rs:170-290, without the CABAC engine).                     */
/* ------------------------------------------------------------------------ */
/* 16 bytes per CTB, raster order; bit i of the record is bit (i & 7) of
 * bytes[i >> 3]:
 *   0 sao_merge_left_flag, 1 sao_merge_up_flag
 *   2..32   luma:  sao_type_idx (2) | sao_offset_abs[0..3] (5 each) |
 *                  sao_offset_sign_flag[0..3] (1 each) | sao_band_position
 *                  (5; band offset) or sao_eo_class (low 2 bits; edge offset)
 *   33..63  Cb:    the same fields (type and EO class also serve Cr)
 *   64..92  Cr:    sao_offset_abs[0..3] | sao_offset_sign_flag[0..3] |
 *                  sao_band_position
 *   93 alf_ctb_flag[0], 94 alf_use_aps_flag, 95..98 alf_luma_prev_filter_idx
 *      (use_aps) or alf_luma_fixed_filter_idx, 99 alf_ctb_flag[1],
 *      100..102 alf_ctb_filter_alt_idx[0], 103 alf_ctb_flag[2],
 *      104..106 alf_ctb_filter_alt_idx[1], 107..109 alf_ctb_cc_cb_idc,
 *      110..112 alf_ctb_cc_cr_idc
 *   113..127 zero
 * Fields a decoder does not parse (absent syntax) are zero. */
typedef struct vvclf_ctb_syntax {
  uint8_t bytes[16];
} vvclf_ctb_syntax;

typedef struct vvclf_ctb_syntax_config {
  uint8_t sao_luma_used, sao_chroma_used;  /* sh_sao_luma/chroma_used_flag */
  uint8_t alf_enabled, alf_cb_enabled, alf_cr_enabled, alf_cc_cb_enabled, alf_cc_cr_enabled, pad0;
  uint32_t bit_depth, chroma_format_idc;
  uint32_t num_alf_aps_ids_luma;              /* sh_num_alf_aps_ids_luma */
  uint32_t alf_chroma_num_alt_filters_minus1; /* of the bound chroma APS */
  uint32_t alf_cc_cb_filters_signalled_minus1, alf_cc_cr_filters_signalled_minus1;
} vvclf_ctb_syntax_config;

/* Expands w_ctbs * h_ctbs records into SAO and ALF CTB parameters (either
 * output may be NULL): merge flags copy the left / upper CTB's result,
 * Cr takes Cb's type and EO class, edge-offset signs are (+, +, -, -),
 * SaoOffsetVal = (1 - 2 sign) * (abs << (BitDepth - Min(BitDepth, 10)))
 * (eq. 153), AlfCtbFiltSetIdxY = 16 + prev_idx with the APS flag, the fixed
 * index otherwise.  VVCLF_ERR_INVALID for a record the syntax cannot carry
 * (a merge at the picture edge, an offset above cMax, an index beyond its
 * range, reserved bits set). */
int vvclf_ctb_syntax_decode(const vvclf_ctb_syntax* recs, uint32_t w_ctbs, uint32_t h_ctbs,
                            const vvclf_ctb_syntax_config* cfg, vvclf_sao_ctb_params* sao_out,
                            vvclf_alf_ctb* alf_out);
/* The inverse: records that decode to the given parameters, using merge
 * flags where a CTB equals its left (else upper) neighbour.
 * VVCLF_ERR_INVALID when a parameter set has no syntax (an offset that is not
 * abs << shift with abs <= cMax, edge-offset signs other than (+, +, -, -),
 * Cr type or EO class differing from Cb's, ...). */
int vvclf_ctb_syntax_encode(const vvclf_sao_ctb_params* sao, const vvclf_alf_ctb* alf, uint32_t w_ctbs,
                            uint32_t h_ctbs, const vvclf_ctb_syntax_config* cfg, vvclf_ctb_syntax* out);

/* ======================================================================== */
/* Layer 2: device-resident jobs, batches and serving (no reference          */
/* counterpart; the reference is single-threaded CPU code).                  */
/* ======================================================================== */
typedef struct vvclf_ctx vvclf_ctx;
typedef struct vvclf_job vvclf_job;
typedef struct vvclf_batch vvclf_batch;
typedef struct vvclf_server vvclf_server;

/* Last error message of the calling thread ("" when none). */
const char* vvclf_last_error(void);
/* Size of a public struct by name ("vvclf_deblock_cu", ...), 0 if unknown;
 * lets bindings verify their layouts. */
size_t vvclf_abi_sizeof(const char* type_name);

/* Host-only dry run: validates the frame and derives every table the GPU
 * path would upload, without touching a device.  stats (may be NULL) receives
 * [0..2] luma deblock segments visited with bS 0 / 1 / 2, [3] chroma segments
 * filtered, [4] SAO CTBs with any component on, [5] ALF CTBs with luma on,
 * [6] device bytes the job would hold, [7] reserved. */
int vvclf_frame_check(const vvclf_frame* frame, uint64_t stats[8]);

/* Context: one per device. */
int vvclf_ctx_create(int device, vvclf_ctx** out);
int vvclf_ctx_destroy(vvclf_ctx* ctx);
/* The context's HIP stream (hipStream_t) and a wait for all its work. */
void* vvclf_ctx_stream(vvclf_ctx* ctx);
int vvclf_ctx_synchronize(vvclf_ctx* ctx);

/* Job: one picture plus its side information, resident in HBM.  The host
 * derives the compact per-edge and per-CTB tables once here (the CU walk runs
 * on several host threads for large pictures); input samples are copied from
 * frame->pic.  The input planes stay untouched by runs, so a job can be run
 * repeatedly with identical results. */
int vvclf_job_create(vvclf_ctx* ctx, const vvclf_frame* frame, vvclf_job** out);
/* As vvclf_job_create.  VVCLF_JOB_GPU_DERIVE asks for the deblocking
 * descriptors to be derived on the GPU from the uploaded CU records and
 * motion grid (the walk of deblock_cu_dir, src/deblock.rs:707-983, one thread
 * per record); that path is switched off in this build and the call returns
 * VVCLF_ERR_UNSUPPORTED.  No counterpart in the reference. */
#define VVCLF_JOB_GPU_DERIVE 1u
int vvclf_job_create_flags(vvclf_ctx* ctx, const vvclf_frame* frame, uint32_t flags, vvclf_job** out);
/* Copies one of a job's deblocking tables to host memory: which 0..5 the
 * descriptor maps (luma V, luma H, Cb V, Cr V, Cb H, Cr H; one u32 per
 * segment), 6..9 the work lists (luma V, luma H, chroma V, chroma H).  *size
 * receives the table's size in bytes (0 when the job does not deblock that
 * plane); nothing is copied when cap is smaller.  For tests and tools; no
 * counterpart in the reference. */
int vvclf_job_deblock_table(const vvclf_job* job, uint32_t which, void* out, size_t cap, size_t* size);
int vvclf_job_destroy(vvclf_job* job);
/* Re-upload input samples (same geometry as at creation). */
int vvclf_job_set_input(vvclf_job* job, const vvclf_picture* pic);
/* Plain per-picture form: separate stage kernels on `stream` (hipStream_t,
 * NULL = default stream).  Asynchronous. */
int vvclf_job_run(vvclf_job* job, void* stream);
/* The same per-picture kernel sequence, captured into a HIP graph on the
 * first call and replayed from then on (fewer launch gaps). */
int vvclf_job_graph_launch(vvclf_job* job, void* stream);
/* Many pictures' plain kernel sequences in one HIP graph of `lanes` parallel
 * branches (job i on branch i % lanes, in order within a branch), so that
 * independent pictures overlap on the device: one picture's latency-bound
 * stages run beside another's VALU-bound ALF.  The jobs must outlive the
 * graph.  Launches are asynchronous on `stream`; each job's result is what
 * vvclf_job_run gives.  No counterpart in the reference (its driver filters
 * one picture at a time). */
typedef struct vvclf_jobs_graph vvclf_jobs_graph;
int vvclf_jobs_graph_create(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n_jobs, uint32_t lanes,
                            vvclf_jobs_graph** out);
int vvclf_jobs_graph_launch(vvclf_jobs_graph* g, void* stream);
int vvclf_jobs_graph_destroy(vvclf_jobs_graph* g);
/* Stage timing for the plain form: a pool of HIP events for up to
 * `max_runs` runs.  vvclf_job_run_prof records events between the stage
 * kernels on the run's stream (no synchronisation); vvclf_prof_stage_ms waits
 * for the recorded runs and returns, per stage, the summed milliseconds:
 * [prep, deblock V (luma + chroma), deblock H (luma + chroma), SAO, ALF luma,
 *  ALF chroma + CC-ALF, 0, 0]. */
typedef struct vvclf_prof vvclf_prof;
int vvclf_prof_create(vvclf_ctx* ctx, uint32_t max_runs, vvclf_prof** out);
int vvclf_prof_destroy(vvclf_prof* prof);
int vvclf_prof_reset(vvclf_prof* prof);
int vvclf_job_run_prof(vvclf_job* job, void* stream, vvclf_prof* prof);
int vvclf_prof_stage_ms(vvclf_prof* prof, float stage_ms[8], uint32_t* runs);
/* Copy the filtered picture to host planes (synchronous). */
int vvclf_job_download(vvclf_job* job, vvclf_picture* out);
/* Device pointer of the filtered output planes (luma, cb, cr) and their
 * pitch in samples. */
int vvclf_job_output(vvclf_job* job, uint16_t** luma, uint16_t** cb, uint16_t** cr,
                     uint32_t* pitch_luma, uint32_t* pitch_chroma);
/* vvclf_picture_md5 of the job's filtered output: the crop is cut and packed
 * (one or two bytes per sample) on the device, so only the hashed bytes come
 * back.  Waits for the job's work on the device. */
int vvclf_job_output_md5(vvclf_job* job, uint32_t bit_depth, const vvclf_crop* crop, uint8_t md5[3][16]);
/* Device bytes a job holds (side information + sample planes). */
uint64_t vvclf_job_device_bytes(const vvclf_job* job);
/* Luma pixels of the job's picture. */
uint64_t vvclf_job_luma_pixels(const vvclf_job* job);

/* Tile shape for the fused kernel (chosen by vvclf_autotune when tile_w == 0). */
typedef struct vvclf_tune {
  uint32_t tile_w, tile_h;   /* output tile in luma samples */
  uint32_t threads;          /* workgroup size */
  uint32_t wgs_per_cu;       /* persistent workgroups per CU */
} vvclf_tune;

/* Batch: many jobs (uneven sizes) filtered by one persistent fused
 * deblock->SAO->ALF launch over a global tile queue, largest first. */
int vvclf_batch_create(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n_jobs,
                       const vvclf_tune* tune, vvclf_batch** out);
int vvclf_batch_destroy(vvclf_batch* batch);
int vvclf_batch_run(vvclf_batch* batch, void* stream);
/* Capture the batch launch in a hipGraph once, then replay it. */
int vvclf_batch_graph_launch(vvclf_batch* batch, void* stream);
int vvclf_batch_get_tune(const vvclf_batch* batch, vvclf_tune* out);
/* Times `reps` launches (graph replays when use_graph) on `stream` with HIP
 * events bracketing them, after `warmup` untimed launches; returns the mean
 * milliseconds per launch. */
int vvclf_batch_time(vvclf_batch* batch, void* stream, uint32_t warmup, uint32_t reps, int use_graph,
                     double* ms_per_launch);
/* Pick a tile shape for the given jobs by timing candidates on the device. */
int vvclf_autotune(vvclf_ctx* ctx, vvclf_job* const* jobs, uint32_t n_jobs, vvclf_tune* out);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching over many streams.                           */
/* ------------------------------------------------------------------------ */
typedef struct vvclf_server_config {
  uint64_t hbm_budget_bytes; /* admission cap on device bytes of admitted jobs */
  uint32_t max_inflight;     /* admitted jobs not yet retired (slots) */
  uint32_t max_batch;        /* jobs per fused launch */
  uint32_t prep_threads;     /* host threads deriving tables and uploading */
  uint32_t pad;
  vvclf_tune tune;           /* fused tile shape; tile_w == 0 selects 64x64 */
} vvclf_server_config;

typedef struct vvclf_server_stats {
  uint64_t submitted, completed, failed, rejected;
  uint64_t luma_pixels;        /* of completed jobs */
  uint64_t launches;           /* fused launches issued */
  uint64_t bytes_admitted_peak;
  /* submit -> result in the caller's planes; quantiles come from 0.5 %-wide
   * log-spaced buckets over every completed job */
  double p50_ms, p99_ms, max_ms, mean_ms;
  double wall_s;               /* first submit -> last completion */
  /* median time per stage of completed jobs: [0] waiting for admission,
   * [1] host table derivation + upload, [2] waiting for a launch, [3] on the
   * device, [4] download into the caller's planes */
  double stage_p50_ms[5];
  double stage_p99_ms[5];      /* the same stages at the 99th percentile */
} vvclf_server_stats;

/* A server owns three HIP streams (upload, compute, download), a pool of
 * prep threads, a launch thread and a retire thread.  Jobs submitted while
 * others run are admitted as slots and HBM budget free up and join the next
 * fused launch; uploads, launches and downloads of different jobs overlap. */
int vvclf_server_create(vvclf_ctx* ctx, const vvclf_server_config* cfg, vvclf_server** out);
int vvclf_server_destroy(vvclf_server* srv);
/* A stream: a sequence of pictures that share resident LMCS / ALF tables
 * (the APS state of src/stream.rs:90-92); tables are re-uploaded only when
 * the content of a submitted frame's LMCS binding or APS binding changes. */
int vvclf_server_open_stream(vvclf_server* srv, uint32_t* stream_id);
int vvclf_server_close_stream(vvclf_server* srv, uint32_t stream_id);
/* Queue one picture.  `frame` and everything it points to must stay valid
 * until the ticket completes; the result is written into frame->pic.
 * VVCLF_ERR_BUDGET if the job alone exceeds the budget. */
int vvclf_server_submit(vvclf_server* srv, uint32_t stream_id, vvclf_frame* frame, uint64_t* ticket);
/* As vvclf_server_submit, with the frame's SAO and ALF CTB parameters given
 * as per-CTB syntax records (frame->sao / frame->alf give only the CTB grid
 * sizes; their ctbs pointers are ignored): the records are expanded on a
 * prep thread into storage the request owns.  `recs` must stay valid until
 * the ticket completes. */
int vvclf_server_submit_syntax(vvclf_server* srv, uint32_t stream_id, vvclf_frame* frame,
                               const vvclf_ctb_syntax* recs, const vvclf_ctb_syntax_config* cfg, uint64_t* ticket);
/* VVCLF_OK when done (or the job's own error), VVCLF_ERR_AGAIN if pending.
 * A ticket's final status is reported once: the poll or wait that returns it
 * releases the ticket, and later queries of it return VVCLF_ERR_INVALID.
 * Finished tickets that nobody queries are released oldest first once more
 * than 65536 of them are held. */
int vvclf_server_poll(vvclf_server* srv, uint64_t ticket);
/* Blocks until done or timeout: VVCLF_OK / the job's error / VVCLF_ERR_TIMEOUT
 * (a timeout keeps the ticket). */
int vvclf_server_wait(vvclf_server* srv, uint64_t ticket, uint32_t timeout_ms);
int vvclf_server_drain(vvclf_server* srv, uint32_t timeout_ms);
int vvclf_server_get_stats(vvclf_server* srv, vvclf_server_stats* out);
/* Starts the statistics afresh (counters, latency quantiles, wall clock and
 * admitted-bytes peak), e.g. after a warm-up whose first-use costs (resident
 * tables, staging buffers, pool growth) the steady state does not pay.  No
 * reference counterpart (the reference has no server). */
int vvclf_server_reset_stats(vvclf_server* srv);
/* Test hook: make the next `n` jobs fail at admission with VVCLF_ERR_HIP, to
 * check that one failed job does not disturb the others. */
int vvclf_server_inject_failures(vvclf_server* srv, uint32_t n);

#ifdef __cplusplus
}
#endif

#endif /* VVCLF_H */
