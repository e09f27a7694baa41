"""MI355X (gfx950) VVC in-loop filter chain: Python binding of include/vvclf.h.

The filters run in libvvclf.so (hand-written HIP kernels plus a C++ host
layer).  This module mirrors the C structs with ctypes and exposes thin
wrappers; importing it fails loudly when the native library is missing.

Load it by path (the directory name is not a Python identifier)::

    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "vvclf_amd", os.path.join(ROOT, "oxideav-h266_amd", "__init__.py"))
    vv = importlib.util.module_from_spec(spec); spec.loader.exec_module(vv)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvvclf.so")
SYNTH_PATH = os.path.join(HERE, "libvvclf_synth.so")

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_BUDGET = -4
ERR_TIMEOUT = -5
ERR_UNSUPPORTED = -6
ERR_AGAIN = -7
QP_C_LEGACY = -(2 ** 31)

u8, u16, u32, i32, u64, f32 = C.c_uint8, C.c_uint16, C.c_uint32, C.c_int32, C.c_uint64, C.c_float


class Plane(C.Structure):
    """PicturePlane, src/reconstruct.rs:41-52."""
    _fields_ = [("samples", C.POINTER(u16)), ("stride", u32), ("width", u32), ("height", u32), ("bit_depth", u32)]


class Picture(C.Structure):
    """PictureBuffer (4:2:0), src/reconstruct.rs:116-145."""
    _fields_ = [("luma", Plane), ("cb", Plane), ("cr", Plane)]


class DeblockCu(C.Structure):
    """DeblockCu + TbSplit + Tu64CbfMap, src/deblock.rs:52-161."""
    _fields_ = [("x", u32), ("y", u32), ("w", u32), ("h", u32), ("qp_y", i32), ("qp_c", i32 * 3),
                ("intra", u8), ("tu_y_coded", u8), ("tu_cb_coded", u8), ("tu_cr_coded", u8),
                ("bdpcm_luma", u8), ("bdpcm_chroma", u8), ("plt", u8), ("joint_cbcr2", u8),
                ("ciip", u8), ("ibc", u8), ("has_num_sb", u8), ("has_tu64", u8),
                ("num_sb_x", u32), ("num_sb_y", u32),
                ("tu64_y", (u8 * 2) * 2), ("tu64_cb", (u8 * 2) * 2), ("tu64_cr", (u8 * 2) * 2),
                ("has_tb_split", u8), ("tb_vertical", u8), ("tb_n_bounds", u8), ("tb_luma_only", u8),
                ("tb_bounds", u32 * 3), ("tb_y_coded", u8 * 4), ("tb_cb_coded", u8 * 4), ("tb_cr_coded", u8 * 4)]


class MvCell(C.Structure):
    """DeblockMvCell, src/deblock.rs:330-337."""
    _fields_ = [("pred_l0", u8), ("pred_l1", u8), ("pad0", u8), ("pad1", u8), ("poc_l0", i32), ("poc_l1", i32),
                ("mv_l0", i32 * 2), ("mv_l1", i32 * 2)]


class MvGrid(C.Structure):
    """DeblockMvGrid, src/deblock.rs:344-348."""
    _fields_ = [("cells", C.POINTER(MvCell)), ("cells_w", u32), ("cells_h", u32)]


class DeblockParams(C.Structure):
    """DeblockParams, src/deblock.rs:376-393."""
    _fields_ = [("disabled", i32), ("luma_beta_offset_div2", i32), ("luma_tc_offset_div2", i32),
                ("cb_beta_offset_div2", i32), ("cb_tc_offset_div2", i32), ("cr_beta_offset_div2", i32),
                ("cr_tc_offset_div2", i32), ("chroma_qp_offset_cb", i32), ("chroma_qp_offset_cr", i32),
                ("bit_depth", u32), ("ctb_log2_size_y", u32)]


class SaoCtb(C.Structure):
    """SaoCtb, src/sao.rs:137-147."""
    _fields_ = [("sao_type_idx", u8), ("eo_class", u8), ("band_position", u8), ("pad", u8), ("offset_val", i32 * 5)]


class SaoCtbParams(C.Structure):
    _fields_ = [("luma", SaoCtb), ("cb", SaoCtb), ("cr", SaoCtb)]


class SaoPicture(C.Structure):
    _fields_ = [("pic_width_in_ctbs_y", u32), ("pic_height_in_ctbs_y", u32), ("ctbs", C.POINTER(SaoCtbParams))]


class SaoConfig(C.Structure):
    _fields_ = [("luma_used", u8), ("chroma_used", u8), ("pad0", u8), ("pad1", u8), ("bit_depth", u32),
                ("ctb_log2_size_y", u32), ("chroma_format_idc", u32)]


class Bounds(C.Structure):
    _fields_ = [("cols", C.POINTER(u32)), ("n_cols", u32), ("rows", C.POINTER(u32)), ("n_rows", u32)]


class AlfCtb(C.Structure):
    """AlfCtb, src/alf.rs:67-88."""
    _fields_ = [("luma_on", u8), ("cb_on", u8), ("cr_on", u8), ("luma_filt_set_idx", u8), ("cb_alt_idx", u8),
                ("cr_alt_idx", u8), ("cc_cb_idc", u8), ("cc_cr_idc", u8)]


class AlfPicture(C.Structure):
    _fields_ = [("pic_width_in_ctbs_y", u32), ("pic_height_in_ctbs_y", u32), ("ctbs", C.POINTER(AlfCtb))]


class AlfConfig(C.Structure):
    _fields_ = [("alf_enabled", u8), ("cb_enabled", u8), ("cr_enabled", u8), ("cc_cb_enabled", u8),
                ("cc_cr_enabled", u8), ("pad0", u8), ("pad1", u8), ("pad2", u8), ("bit_depth", u32),
                ("ctb_log2_size_y", u32), ("chroma_format_idc", u32)]


class AlfAps(C.Structure):
    """AlfApsData, src/aps.rs:61-89."""
    _fields_ = [("alf_luma_filter_signal_flag", u8), ("alf_chroma_filter_signal_flag", u8),
                ("alf_cc_cb_filter_signal_flag", u8), ("alf_cc_cr_filter_signal_flag", u8),
                ("alf_luma_clip_flag", u8), ("alf_chroma_clip_flag", u8), ("pad0", u8), ("pad1", u8),
                ("alf_chroma_num_alt_filters_minus1", u32), ("n_luma_coeff", u32), ("n_chroma_coeff", u32),
                ("n_cc_cb_coeff", u32), ("n_cc_cr_coeff", u32),
                ("luma_coeff", (i32 * 12) * 25), ("luma_clip_idx", (u8 * 12) * 25),
                ("chroma_coeff", (i32 * 6) * 8), ("chroma_clip_idx", (u8 * 6) * 8),
                ("cc_cb_coeff", (i32 * 7) * 4), ("cc_cr_coeff", (i32 * 7) * 4)]


class AlfClipRdo(C.Structure):
    """DesignedLumaAlfClipRdo, src/alf_aps_design.rs:552-597."""
    _fields_ = [("coeff", (i32 * 12) * 25), ("clip_idx", (u8 * 12) * 25), ("alf_luma_clip_flag", u8),
                ("pad", u8 * 3), ("baseline_sse_y", u64), ("post_clip_sse_y", u64)]


class AlfApsBinding(C.Structure):
    """AlfApsBinding, src/alf.rs:180-195."""
    _fields_ = [("luma_apses", C.POINTER(C.POINTER(AlfAps))), ("n_luma_apses", u32), ("pad", u32),
                ("chroma_aps", C.POINTER(AlfAps)), ("cc_cb_aps", C.POINTER(AlfAps)), ("cc_cr_aps", C.POINTER(AlfAps))]


class LmcsData(C.Structure):
    """LmcsData, src/lmcs.rs:51-76."""
    _fields_ = [("lmcs_min_bin_idx", u8), ("lmcs_delta_max_bin_idx", u8), ("lmcs_delta_cw_prec_minus1", u8),
                ("lmcs_delta_abs_crs", u8), ("lmcs_delta_sign_crs_flag", u8), ("pad0", u8), ("pad1", u8),
                ("pad2", u8), ("lmcs_delta_abs_cw", u32 * 16), ("lmcs_delta_sign_cw_flag", u8 * 16)]


class LmcsDerived(C.Structure):
    """LmcsDerived, src/lmcs.rs:127-153."""
    _fields_ = [("bit_depth", u32), ("org_cw", u32), ("lmcs_cw", u32 * 16), ("input_pivot", u32 * 16),
                ("lmcs_pivot", u32 * 17), ("scale_coeff", u32 * 16), ("inv_scale_coeff", u32 * 16),
                ("lmcs_delta_crs", i32), ("chroma_scale_coeff", u32 * 16)]


class LmcsBinding(C.Structure):
    """LmcsBinding, src/ctu.rs:712-725."""
    _fields_ = [("derived", C.POINTER(LmcsDerived)), ("min_bin_idx", u8), ("max_bin_idx", u8), ("pad0", u8),
                ("pad1", u8)]


class Rect(C.Structure):
    _fields_ = [("x", u32), ("y", u32), ("w", u32), ("h", u32)]


class CrsBlock(C.Structure):
    """One chroma TB with LMCS chroma residual scaling (src/ctu.rs:1752-1815)."""
    _fields_ = [("c_idx", u32), ("x", u32), ("y", u32), ("w", u32), ("h", u32), ("x_cu", u32), ("y_cu", u32),
                ("avail_l", u8), ("avail_t", u8), ("pad0", u8), ("pad1", u8), ("size_y", u32), ("pad2", u32),
                ("res_offset", u64)]


class Crop(C.Structure):
    _fields_ = [("x", u32), ("y", u32), ("w", u32), ("h", u32)]


class CtbSyntax(C.Structure):
    """Per-CTB syntax record (16 bytes, bit layout in include/vvclf.h)."""
    _fields_ = [("bytes", u8 * 16)]


class CtbSyntaxConfig(C.Structure):
    _fields_ = [("sao_luma_used", u8), ("sao_chroma_used", u8), ("alf_enabled", u8), ("alf_cb_enabled", u8),
                ("alf_cr_enabled", u8), ("alf_cc_cb_enabled", u8), ("alf_cc_cr_enabled", u8), ("pad0", u8),
                ("bit_depth", u32), ("chroma_format_idc", u32), ("num_alf_aps_ids_luma", u32),
                ("alf_chroma_num_alt_filters_minus1", u32), ("alf_cc_cb_filters_signalled_minus1", u32),
                ("alf_cc_cr_filters_signalled_minus1", u32)]


class Frame(C.Structure):
    """Flattened walker state of apply_in_loop_filters_with_alf, src/ctu.rs:11822-12066."""
    _fields_ = [("pic", Picture), ("chroma_format_idc", u32), ("pad0", u32), ("lmcs", LmcsBinding),
                ("cus", C.POINTER(DeblockCu)), ("n_cus", u32), ("has_chroma_cus", u32),
                ("chroma_cus", C.POINTER(DeblockCu)), ("n_chroma_cus", u32), ("pad1", u32),
                ("mv_grid", C.POINTER(MvGrid)), ("deblock", DeblockParams),
                ("no_filter_cols", C.POINTER(u32)), ("n_no_filter_cols", u32), ("n_no_filter_rows", u32),
                ("no_filter_rows", C.POINTER(u32)),
                ("sao", SaoPicture), ("sao_cfg", SaoConfig), ("alf", AlfPicture), ("alf_cfg", AlfConfig),
                ("alf_binding", AlfApsBinding), ("clip_bounds", C.POINTER(Bounds))]


class Tune(C.Structure):
    _fields_ = [("tile_w", u32), ("tile_h", u32), ("threads", u32), ("wgs_per_cu", u32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", u64), ("max_inflight", u32), ("max_batch", u32), ("prep_threads", u32),
                ("pad", u32), ("tune", Tune)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("failed", u64), ("rejected", u64), ("luma_pixels", u64),
                ("launches", u64), ("bytes_admitted_peak", u64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_ms", C.c_double), ("wall_s", C.c_double),
                ("stage_p50_ms", C.c_double * 5), ("stage_p99_ms", C.c_double * 5)]


class SynthSpec(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("bit_depth", u32), ("ctb_log2", u32), ("seed", u64),
                ("min_cu_w", u32), ("min_cu_h", u32), ("split_scale", f32), ("p_intra", f32), ("p_mv_change", f32),
                ("p_coded", f32), ("qp_min", u32), ("qp_max", u32), ("rich", u32), ("lmcs", u32), ("sao", u32),
                ("alf", u32), ("p_alf_luma", f32), ("n_luma_aps", u32), ("p_fixed_set", f32), ("chroma_alf", u32),
                ("cc_alf", u32), ("p_cc", f32), ("tiles", u32), ("aps_seed", u64)]


STRUCTS = {"vvclf_plane": Plane, "vvclf_picture": Picture, "vvclf_deblock_cu": DeblockCu,
           "vvclf_mv_cell": MvCell, "vvclf_mv_grid": MvGrid, "vvclf_deblock_params": DeblockParams,
           "vvclf_sao_ctb": SaoCtb, "vvclf_sao_ctb_params": SaoCtbParams, "vvclf_sao_picture": SaoPicture,
           "vvclf_sao_config": SaoConfig, "vvclf_bounds": Bounds, "vvclf_alf_ctb": AlfCtb,
           "vvclf_alf_picture": AlfPicture, "vvclf_alf_config": AlfConfig, "vvclf_alf_aps": AlfAps, "vvclf_alf_clip_rdo": AlfClipRdo,
           "vvclf_alf_aps_binding": AlfApsBinding, "vvclf_lmcs_data": LmcsData, "vvclf_lmcs_derived": LmcsDerived,
           "vvclf_lmcs_binding": LmcsBinding, "vvclf_frame": Frame, "vvclf_rect": Rect, "vvclf_crs_block": CrsBlock,
           "vvclf_crop": Crop, "vvclf_ctb_syntax": CtbSyntax, "vvclf_ctb_syntax_config": CtbSyntaxConfig, "vvclf_tune": Tune,
           "vvclf_server_config": ServerConfig, "vvclf_server_stats": ServerStats}

P = C.POINTER
VP = C.c_void_p
# exported entry points and their signatures (restype, argtypes)
SIGNATURES = {
    "vvclf_apply_deblocking": (C.c_int, [P(Picture), P(DeblockCu), u32, P(DeblockParams), u32]),
    "vvclf_apply_deblocking_clipped": (C.c_int, [P(Picture), P(DeblockCu), u32, P(DeblockCu), u32, P(MvGrid),
                                                 P(DeblockParams), u32, P(u32), u32, P(u32), u32]),
    "vvclf_apply_sao": (C.c_int, [P(Picture), P(SaoPicture), P(SaoConfig)]),
    "vvclf_apply_sao_clipped": (C.c_int, [P(Picture), P(SaoPicture), P(SaoConfig), P(Bounds)]),
    "vvclf_apply_alf": (C.c_int, [P(Picture), P(AlfPicture), P(AlfConfig), P(AlfApsBinding)]),
    "vvclf_apply_alf_clipped": (C.c_int, [P(Picture), P(AlfPicture), P(AlfConfig), P(AlfApsBinding), P(Bounds)]),
    "vvclf_lmcs_inverse_luma": (C.c_int, [P(Plane), P(LmcsBinding)]),
    "vvclf_apply_in_loop_filters": (C.c_int, [P(Frame)]),
    "vvclf_apply_in_loop_filters_noalloc": (C.c_int, [P(Frame)]),
    "vvclf_derive_lmcs": (C.c_int, [P(LmcsData), u32, P(LmcsDerived)]),
    "vvclf_lmcs_forward_luma": (C.c_int, [P(Plane), P(LmcsBinding), P(Rect), u32]),
    "vvclf_lmcs_chroma_residual": (C.c_int, [P(Picture), P(LmcsBinding), P(CrsBlock), u32, P(i32), u64]),
    "vvclf_conformance_crop": (C.c_int, [u32, u32, u32, P(u32), P(Crop)]),
    "vvclf_picture_md5": (C.c_int, [P(Picture), u32, u32, P(Crop), P((u8 * 16) * 3)]),
    "vvclf_job_output_md5": (C.c_int, [VP, u32, P(Crop), P((u8 * 16) * 3)]),
    "vvclf_alf_decide_luma": (C.c_int, [P(Picture), P(Picture), u32, u32, u32, P(P(AlfAps)), u32, P(Bounds),
                                         P(AlfCtb), P(u64)]),
    "vvclf_alf_decide_chroma": (C.c_int, [P(Picture), P(Picture), P(AlfCtb), P(AlfAps), u32, u32, u32, u32,
                                           P(Bounds), P(u64)]),
    "vvclf_cc_alf_decide": (C.c_int, [P(Picture), P(Picture), P(Plane), P(AlfCtb), P(AlfAps), u32, u32, u32, u32,
                                       P(Bounds), P(u64)]),
    "vvclf_sao_decide_picture": (C.c_int, [P(Picture), P(Picture), u32, u32, u32, u32, P(SaoCtbParams)]),
    "vvclf_sao_chroma_merge": (C.c_int, [P(Picture), P(Picture), P(SaoCtbParams), u32, u32, P(SaoCtbParams),
                                          P(u8)]),
    "vvclf_alf_clip_rdo_luma": (C.c_int, [P(Picture), P(Picture), P(i32), u32, u32, u32, P(AlfClipRdo)]),
    "vvclf_jobs_graph_create": (C.c_int, [VP, P(VP), u32, u32, P(VP)]),
    "vvclf_jobs_graph_launch": (C.c_int, [VP, VP]),
    "vvclf_jobs_graph_destroy": (C.c_int, [VP]),
    "vvclf_ctb_syntax_decode": (C.c_int, [P(CtbSyntax), u32, u32, P(CtbSyntaxConfig), P(SaoCtbParams), P(AlfCtb)]),
    "vvclf_ctb_syntax_encode": (C.c_int, [P(SaoCtbParams), P(AlfCtb), u32, u32, P(CtbSyntaxConfig), P(CtbSyntax)]),
    "vvclf_server_submit_syntax": (C.c_int, [VP, u32, P(Frame), P(CtbSyntax), P(CtbSyntaxConfig), P(u64)]),
    "vvclf_frame_check": (C.c_int, [P(Frame), P(u64)]),
    "vvclf_last_error": (C.c_char_p, []),
    "vvclf_abi_sizeof": (C.c_size_t, [C.c_char_p]),
    "vvclf_ctx_create": (C.c_int, [C.c_int, P(VP)]),
    "vvclf_ctx_destroy": (C.c_int, [VP]),
    "vvclf_ctx_stream": (VP, [VP]),
    "vvclf_ctx_synchronize": (C.c_int, [VP]),
    "vvclf_job_create": (C.c_int, [VP, P(Frame), P(VP)]),
    "vvclf_job_create_flags": (C.c_int, [VP, P(Frame), C.c_uint32, P(VP)]),
    "vvclf_job_deblock_table": (C.c_int, [VP, C.c_uint32, VP, C.c_size_t, P(C.c_size_t)]),
    "vvclf_job_destroy": (C.c_int, [VP]),
    "vvclf_job_set_input": (C.c_int, [VP, P(Picture)]),
    "vvclf_job_run": (C.c_int, [VP, VP]),
    "vvclf_job_graph_launch": (C.c_int, [VP, VP]),
    "vvclf_job_download": (C.c_int, [VP, P(Picture)]),
    "vvclf_job_output": (C.c_int, [VP, P(VP), P(VP), P(VP), P(u32), P(u32)]),
    "vvclf_job_device_bytes": (u64, [VP]),
    "vvclf_prof_create": (C.c_int, [VP, u32, P(VP)]),
    "vvclf_prof_destroy": (C.c_int, [VP]),
    "vvclf_prof_reset": (C.c_int, [VP]),
    "vvclf_job_run_prof": (C.c_int, [VP, VP, VP]),
    "vvclf_prof_stage_ms": (C.c_int, [VP, P(f32), P(u32)]),
    "vvclf_job_luma_pixels": (u64, [VP]),
    "vvclf_batch_create": (C.c_int, [VP, P(VP), u32, P(Tune), P(VP)]),
    "vvclf_batch_destroy": (C.c_int, [VP]),
    "vvclf_batch_run": (C.c_int, [VP, VP]),
    "vvclf_batch_graph_launch": (C.c_int, [VP, VP]),
    "vvclf_batch_get_tune": (C.c_int, [VP, P(Tune)]),
    "vvclf_autotune": (C.c_int, [VP, P(VP), u32, P(Tune)]),
    "vvclf_batch_time": (C.c_int, [VP, VP, u32, u32, C.c_int, P(C.c_double)]),
    "vvclf_server_create": (C.c_int, [VP, P(ServerConfig), P(VP)]),
    "vvclf_server_destroy": (C.c_int, [VP]),
    "vvclf_server_open_stream": (C.c_int, [VP, P(u32)]),
    "vvclf_server_close_stream": (C.c_int, [VP, u32]),
    "vvclf_server_submit": (C.c_int, [VP, u32, P(Frame), P(u64)]),
    "vvclf_server_poll": (C.c_int, [VP, u64]),
    "vvclf_server_wait": (C.c_int, [VP, u64, u32]),
    "vvclf_server_drain": (C.c_int, [VP, u32]),
    "vvclf_server_get_stats": (C.c_int, [VP, P(ServerStats)]),
    "vvclf_server_reset_stats": (C.c_int, [VP]),
    "vvclf_server_inject_failures": (C.c_int, [VP, u32]),
}


class VvclfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vvclf error {code}: {msg}")
        self.code = code


def _load(path, what):
    if not os.path.exists(path):
        raise ImportError(f"{what} not built: {path} is missing (run __graft_entry__.build())")
    return C.CDLL(path)


lib = _load(LIB_PATH, "libvvclf.so")
for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args

synth = _load(SYNTH_PATH, "libvvclf_synth.so")
synth.vvclf_synth_default_spec.argtypes = [u32, u32, u32, u64, P(SynthSpec)]
synth.vvclf_synth_default_spec.restype = None
synth.vvclf_synth_make.argtypes = [P(SynthSpec), P(VP)]
synth.vvclf_synth_make.restype = C.c_int
synth.vvclf_synth_free.argtypes = [VP]
synth.vvclf_synth_free.restype = None
synth.vvclf_synth_desc.argtypes = [VP]
synth.vvclf_synth_desc.restype = P(Frame)
synth.vvclf_synth_reset.argtypes = [VP]
synth.vvclf_synth_reset.restype = None
synth.vvclf_synth_input.argtypes = [VP]
synth.vvclf_synth_input.restype = P(Picture)
synth.vvclf_synth_stats.argtypes = [VP, P(u64)]
synth.vvclf_synth_stats.restype = None


def check(rc):
    if rc != OK:
        raise VvclfError(rc, lib.vvclf_last_error().decode(errors="replace"))
    return rc


def plane_array(pl):
    """numpy view (height, stride) of a plane's host samples."""
    if not pl.samples or pl.height == 0:
        return np.zeros((0, 0), np.uint16)
    buf = np.ctypeslib.as_array(pl.samples, shape=(pl.height * pl.stride,))
    return buf.reshape(pl.height, pl.stride)[:, :pl.width]


def picture_arrays(pic):
    return plane_array(pic.luma), plane_array(pic.cb), plane_array(pic.cr)


class Owned:
    """Keeps numpy buffers alive next to the ctypes structs that point at them."""

    def __init__(self):
        self.keep = []

    def hold(self, obj):
        self.keep.append(obj)
        return obj


def make_plane(arr, bit_depth, owner):
    arr = owner.hold(np.ascontiguousarray(arr, dtype=np.uint16))
    h, w = arr.shape
    return Plane(arr.ctypes.data_as(P(u16)), w, w, h, bit_depth)


def make_picture(y, cb, cr, bit_depth, owner):
    return Picture(make_plane(y, bit_depth, owner), make_plane(cb, bit_depth, owner), make_plane(cr, bit_depth, owner))


class SynthFrame:
    """A generated workload (synth.h): descriptor + owned planes."""

    def __init__(self, config=0, width=416, height=240, seed=0, **overrides):
        spec = SynthSpec()
        synth.vvclf_synth_default_spec(config, width, height, seed, C.byref(spec))
        for k, v in overrides.items():
            setattr(spec, k, v)
        self.spec = spec
        h = VP()
        rc = synth.vvclf_synth_make(C.byref(spec), C.byref(h))
        if rc != OK:
            raise VvclfError(rc, "synth_make failed")
        self.h = h

    @property
    def desc(self):
        return synth.vvclf_synth_desc(self.h)

    def reset(self):
        synth.vvclf_synth_reset(self.h)

    def input_arrays(self):
        return [a.copy() for a in picture_arrays(synth.vvclf_synth_input(self.h).contents)]

    def output_arrays(self):
        return [a.copy() for a in picture_arrays(self.desc.contents.pic)]

    def stats(self):
        out = (u64 * 8)()
        synth.vvclf_synth_stats(self.h, out)
        return list(out)

    def close(self):
        if self.h:
            synth.vvclf_synth_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        self.h = VP()
        check(lib.vvclf_ctx_create(device, C.byref(self.h)))

    @property
    def stream(self):
        return C.c_void_p(lib.vvclf_ctx_stream(self.h))

    def synchronize(self):
        check(lib.vvclf_ctx_synchronize(self.h))

    def close(self):
        if self.h:
            lib.vvclf_ctx_destroy(self.h)
            self.h = None


class Job:
    """One picture resident in HBM (vvclf_job_*)."""

    GPU_DERIVE = 1  # VVCLF_JOB_GPU_DERIVE

    def __init__(self, ctx, frame_ptr, flags=0):
        self.h = VP()
        check(lib.vvclf_job_create_flags(ctx.h, frame_ptr, flags, C.byref(self.h)))

    def deblock_table(self, which):
        """One deblocking table (0..5 descriptor maps, 6..9 work lists) as a
        numpy uint32 array."""
        import numpy as np
        size = C.c_size_t()
        check(lib.vvclf_job_deblock_table(self.h, which, None, 0, C.byref(size)))
        out = np.zeros(size.value // 4, dtype=np.uint32)
        if size.value:
            check(lib.vvclf_job_deblock_table(self.h, which, out.ctypes.data_as(VP), size.value, C.byref(size)))
        return out

    def run(self, stream=None):
        check(lib.vvclf_job_run(self.h, stream))

    def graph_launch(self, stream=None):
        check(lib.vvclf_job_graph_launch(self.h, stream))

    def download(self, pic_ptr):
        check(lib.vvclf_job_download(self.h, pic_ptr))

    @property
    def device_bytes(self):
        return lib.vvclf_job_device_bytes(self.h)

    @property
    def luma_pixels(self):
        return lib.vvclf_job_luma_pixels(self.h)

    def close(self):
        if self.h:
            lib.vvclf_job_destroy(self.h)
            self.h = None


class JobsGraph:
    """Many jobs' plain sequences in one HIP graph of `lanes` parallel
    branches (vvclf_jobs_graph_*); the jobs must outlive it."""

    def __init__(self, ctx, jobs, lanes=2):
        self.h = VP()
        self._jobs = list(jobs)
        arr = (VP * max(1, len(jobs)))(*[j.h for j in jobs])
        check(lib.vvclf_jobs_graph_create(ctx.h, arr, len(jobs), lanes, C.byref(self.h)))

    def launch(self, stream=None):
        check(lib.vvclf_jobs_graph_launch(self.h, stream))

    def close(self):
        if self.h:
            lib.vvclf_jobs_graph_destroy(self.h)
            self.h = None
