"""Small builders for hand-made test pictures and side information."""
import ctypes as C

import numpy as np


class Pic:
    """A 4:2:0 picture owned by numpy arrays, with a ctypes Picture view."""

    def __init__(self, vv, w, h, seed, bd=8, chroma=None):
        self.vv = vv
        mid = (1 << (bd - 1)) if chroma is None else chroma
        self.y = np.full((h, w), seed, np.uint16)
        self.cb = np.full((h // 2, w // 2), mid, np.uint16)
        self.cr = np.full((h // 2, w // 2), mid, np.uint16)
        self.bd = bd

    def ptr(self):
        vv = self.vv
        P = C.POINTER(C.c_uint16)

        def pl(a):
            h, w = a.shape
            return vv.Plane(a.ctypes.data_as(P), w, w, h, self.bd)

        self._pic = vv.Picture(pl(self.y), pl(self.cb), pl(self.cr))
        return C.byref(self._pic)

    def copy(self):
        o = Pic.__new__(Pic)
        o.vv, o.bd = self.vv, self.bd
        o.y, o.cb, o.cr = self.y.copy(), self.cb.copy(), self.cr.copy()
        return o


def cu(vv, x, y, w, h, qp, intra=True, coded=True):
    c = vv.DeblockCu()
    c.x, c.y, c.w, c.h, c.qp_y = x, y, w, h, qp
    c.intra = int(intra)
    c.tu_y_coded = int(coded)
    c.qp_c[0] = c.qp_c[1] = c.qp_c[2] = vv.QP_C_LEGACY
    return c


def cu_array(vv, cus):
    arr = (vv.DeblockCu * len(cus))(*cus)
    return arr


def dparams(vv, bd=8, ctb_log2=7, disabled=False):
    p = vv.DeblockParams()
    p.disabled = int(disabled)
    p.bit_depth = bd
    p.ctb_log2_size_y = ctb_log2
    return p


def sao_band(vv, pos, offs_abs, signs, bd=8):
    s = vv.SaoCtb()
    s.sao_type_idx = 1
    s.band_position = pos & 31
    shift = bd - min(10, bd)
    for i in range(4):
        s.offset_val[i + 1] = (1 - 2 * (signs[i] & 1)) * (offs_abs[i] << shift)
    return s


def sao_edge(vv, cls, offs_abs, bd=8):
    s = vv.SaoCtb()
    s.sao_type_idx = 2
    s.eo_class = cls
    shift = bd - min(10, bd)
    sign = [0, 0, 1, 1]
    for i in range(4):
        s.offset_val[i + 1] = (1 - 2 * sign[i]) * (offs_abs[i] << shift)
    return s


def sao_picture(vv, wc, hc, entries):
    arr = (vv.SaoCtbParams * (wc * hc))()
    for (rx, ry), p in entries.items():
        arr[ry * wc + rx] = p
    sp = vv.SaoPicture(wc, hc, C.cast(arr, C.POINTER(vv.SaoCtbParams)))
    return sp, arr


def sao_cfg(vv, luma=True, chroma=False, bd=8, ctb_log2=5):
    c = vv.SaoConfig()
    c.luma_used, c.chroma_used, c.bit_depth, c.ctb_log2_size_y, c.chroma_format_idc = int(luma), int(chroma), bd, ctb_log2, 1
    return c


def alf_picture(vv, wc, hc, entries):
    arr = (vv.AlfCtb * (wc * hc))()
    for (rx, ry), p in entries.items():
        arr[ry * wc + rx] = p
    return vv.AlfPicture(wc, hc, C.cast(arr, C.POINTER(vv.AlfCtb))), arr


def alf_cfg(vv, bd=8, ctb_log2=5, cb=False, cr=False, cc_cb=False, cc_cr=False):
    c = vv.AlfConfig()
    c.alf_enabled = 1
    c.cb_enabled, c.cr_enabled, c.cc_cb_enabled, c.cc_cr_enabled = int(cb), int(cr), int(cc_cb), int(cc_cr)
    c.bit_depth, c.ctb_log2_size_y, c.chroma_format_idc = bd, ctb_log2, 1
    return c


def binding_of(vv, luma=(), chroma=None, cc_cb=None, cc_cr=None):
    b = vv.AlfApsBinding()
    ptrs = (C.POINTER(vv.AlfAps) * max(1, len(luma)))()
    for i, a in enumerate(luma):
        ptrs[i] = C.pointer(a) if a is not None else C.POINTER(vv.AlfAps)()
    b.luma_apses = C.cast(ptrs, C.POINTER(C.POINTER(vv.AlfAps)))
    b.n_luma_apses = len(luma)
    b._keep = (ptrs, luma, chroma, cc_cb, cc_cr)
    if chroma is not None:
        b.chroma_aps = C.pointer(chroma)
    if cc_cb is not None:
        b.cc_cb_aps = C.pointer(cc_cb)
    if cc_cr is not None:
        b.cc_cr_aps = C.pointer(cc_cr)
    return b


def put_wide_samples(vv, frames, seed, share=200):
    """Sets about 1/share of every plane's samples, at the same places and to
    the same values in each frame, to values of 2^15 and more (beyond every
    bit depth; a decoder never produces them, the reference filters them all
    the same) and to 4095."""
    rng = np.random.default_rng(seed)
    planes = [vv.picture_arrays(f.desc.contents.pic) for f in frames]
    for p in range(3):
        a0 = planes[0][p]
        if a0.size == 0:
            continue
        n = max(1, a0.size // share)
        ys = rng.integers(0, a0.shape[0], n)
        xs = rng.integers(0, a0.shape[1], n)
        v = rng.choice([32768, 40000, 65535, 4095], n).astype(np.uint16)
        for pl in planes:
            pl[p][ys, xs] = v
