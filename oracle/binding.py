"""ctypes binding of the CPU oracle (oracle/_build/libvvclf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker.  Uses the struct mirrors of the
package binding, since the oracle runs on the same C ABI types.
"""
import ctypes as C
import importlib.util
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libvvclf_oracle.so")


def load_package():
    spec = importlib.util.spec_from_file_location(
        "vvclf_amd", os.path.join(ROOT, "oxideav-h266_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(vv=None, path=None):
    """Loads the oracle library (or, for tests of the tests, another build of
    the same source at `path`) and declares its signatures."""
    if path is None and not os.path.exists(LIB):
        build()
    vv = vv or load_package()
    lib = C.CDLL(path or LIB)
    P = C.POINTER
    u32, i32, u8 = C.c_uint32, C.c_int32, C.c_uint8
    sig = {
        "oracle_derive_lmcs": (C.c_int, [P(vv.LmcsData), u32, P(vv.LmcsDerived)]),
        "oracle_lmcs_idx_y_inv": (u32, [P(vv.LmcsDerived), u32, u32, u32]),
        "oracle_lmcs_inverse_map": (u32, [P(vv.LmcsDerived), u32, u32]),
        "oracle_lmcs_forward_map": (C.c_int64, [P(vv.LmcsDerived), u32]),
        "oracle_lmcs_inverse_luma": (None, [P(vv.Plane), P(vv.LmcsBinding)]),
        "oracle_lmcs_forward_luma_rect": (C.c_int, [P(vv.Plane), P(vv.LmcsBinding), u32, u32, u32, u32]),
        "oracle_lmcs_chroma_var_scale": (u32, [P(vv.Plane), P(vv.LmcsBinding), u32, u32, C.c_int, C.c_int, u32]),
        "oracle_lmcs_scale_chroma_residual_sample": (u32, [P(vv.LmcsDerived), C.c_int64, C.c_int64, u32]),
        "oracle_lmcs_chroma_residual": (C.c_int, [P(vv.Picture), P(vv.LmcsBinding), P(vv.CrsBlock), u32,
                                                  P(C.c_int32)]),
        "oracle_beta_prime": (i32, [i32]),
        "oracle_tc_prime": (i32, [i32]),
        "oracle_scale_tc": (i32, [i32, u32]),
        "oracle_scale_beta": (i32, [i32, u32]),
        "oracle_tu_coded_at": (C.c_int, [P(vv.DeblockCu), u32, i32, i32]),
        "oracle_luma_tb_len": (u32, [P(vv.DeblockCu), i32, i32, C.c_int]),
        "oracle_chroma_tb_len": (u32, [P(vv.DeblockCu), i32, i32, C.c_int, u32]),
        "oracle_is_luma_tb_boundary": (C.c_int, [P(vv.DeblockCu), u32, C.c_int]),
        "oracle_is_chroma_tb_boundary": (C.c_int, [P(vv.DeblockCu), u32, C.c_int]),
        "oracle_apply_deblocking_clipped": (None, [P(vv.Picture), P(vv.DeblockCu), u32, P(vv.DeblockCu), u32,
                                                   P(vv.MvGrid), P(vv.DeblockParams), u32, P(u32), u32, P(u32),
                                                   u32]),
        "oracle_apply_sao_clipped": (None, [P(vv.Picture), P(vv.SaoPicture), P(vv.SaoConfig), P(vv.Bounds)]),
        "oracle_sao_derive_offset_val": (None, [P(u32), P(u32), u32, P(i32)]),
        "oracle_alf_clip_value": (i32, [u32, u8]),
        "oracle_alf_table_45": (None, [i32, i32, C.c_int, P(i32)]),
        "oracle_alf_table_46": (None, [i32, i32, C.c_int, P(i32)]),
        "oracle_alf_table_47": (None, [i32, i32, C.c_int, P(i32)]),
        "oracle_alf_transpose_idx": (None, [u8, P(u8)]),
        "oracle_alf_fixed_coeff_row": (C.c_int, [u32, u32, P(i32)]),
        "oracle_alf_classify_ctb": (None, [P(C.c_uint16), u32, i32, i32, P(i32), u32, u32, u32, u32,
                                           P(C.c_uint16)]),
        "oracle_apply_alf_clipped": (C.c_int, [P(vv.Picture), P(vv.AlfPicture), P(vv.AlfConfig),
                                               P(vv.AlfApsBinding), P(vv.Bounds)]),
        "oracle_apply_in_loop_filters": (C.c_int, [P(vv.Frame)]),
        "oracle_alf_decide_luma": (C.c_int, [P(vv.Picture), P(vv.Picture), u32, u32, u32, P(P(vv.AlfAps)), u32,
                                             P(vv.Bounds), P(vv.AlfCtb), P(C.c_uint64)]),
        "oracle_alf_decide_chroma": (C.c_int, [P(vv.Picture), P(vv.Picture), P(vv.AlfCtb), P(vv.AlfAps), u32, u32,
                                               u32, u32, P(vv.Bounds), P(C.c_uint64)]),
        "oracle_cc_alf_decide": (C.c_int, [P(vv.Picture), P(vv.Picture), P(vv.Plane), P(vv.AlfCtb), P(vv.AlfAps),
                                           u32, u32, u32, u32, P(vv.Bounds), P(C.c_uint64)]),
        "oracle_sao_decide_picture": (C.c_int, [P(vv.Picture), P(vv.Picture), u32, u32, u32, u32,
                                                P(vv.SaoCtbParams)]),
        "oracle_sao_chroma_merge": (C.c_int, [P(vv.Picture), P(vv.Picture), P(vv.SaoCtbParams), u32, u32,
                                              P(vv.SaoCtbParams), P(C.c_uint8)]),
        "oracle_alf_clip_rdo_luma": (C.c_int, [P(vv.Picture), P(vv.Picture), P(C.c_int32), u32, u32, u32,
                                               P(vv.AlfClipRdo)]),
        "oracle_apply_in_loop_filters_mt": (C.c_int, [P(P(vv.Frame)), u32, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
