"""Pictures and APS sets for the encoder's per-CTB luma ALF decision
(src/alf_enc.rs:185-316, :352-470), shared by the oracle checks and the GPU
parity tests, plus one call wrapper for both implementations."""
import ctypes as C

import numpy as np

from helpers import Pic


def smooth_gradient(vv, w, h, bd=8):
    """The reference's smooth_gradient_picture (src/alf_enc.rs:1148-1157),
    scaled to the bit depth: a horizontal ramp from 32 over 191 steps."""
    p = Pic(vv, w, h, 0, bd=bd)
    x = np.arange(w, dtype=np.float32)
    ramp = (32 + (x / max(w, 1) * 191.0).astype(np.int32)).astype(np.uint16) << (bd - 8)
    p.y[:] = ramp[None, :]
    p.cb[:] = 128 << (bd - 8)
    p.cr[:] = 128 << (bd - 8)
    return p


def hash_noise(n, mult, mask):
    """(i * mult) & mask over the raster index, as the reference's tests
    distort their reconstructions (src/alf_enc.rs:1218-1221, :1344-1348)."""
    i = np.arange(n, dtype=np.uint64)
    return ((i * np.uint64(mult)) & np.uint64(mask)).astype(np.int64)


def noisy(src, mult, mask, bias=0):
    r = src.copy()
    h, w = r.y.shape
    v = r.y.astype(np.int64).ravel() + hash_noise(w * h, mult, mask) - bias
    r.y[:] = np.clip(v, 0, (1 << src.bd) - 1).reshape(h, w)
    return r


def luma_aps(vv, rows=None, clip=None, signal=True):
    """A luma APS whose 25 filters all carry `rows` (dict tap -> coeff) and
    the clip indices `clip` (dict tap -> idx)."""
    a = vv.AlfAps()
    a.alf_luma_filter_signal_flag = int(signal)
    a.n_luma_coeff = 25
    for k in range(25):
        for t, c in (rows or {}).items():
            a.luma_coeff[k][t] = c
        for t, c in (clip or {}).items():
            a.luma_clip_idx[k][t] = c
    if clip:
        a.alf_luma_clip_flag = 1
    return a


def random_aps(vv, rng, scale=24):
    a = vv.AlfAps()
    a.alf_luma_filter_signal_flag = 1
    a.alf_luma_clip_flag = 1
    a.n_luma_coeff = 25
    for k in range(25):
        for t in range(12):
            a.luma_coeff[k][t] = int(rng.integers(-scale, scale + 1))
            a.luma_clip_idx[k][t] = int(rng.integers(0, 4))
    return a


def tile_bounds(vv, cols, rows):
    ca = (C.c_uint32 * len(cols))(*cols)
    ra = (C.c_uint32 * len(rows))(*rows)
    b = vv.Bounds(C.cast(ca, C.POINTER(C.c_uint32)), len(cols), C.cast(ra, C.POINTER(C.c_uint32)), len(rows))
    b._keep = (ca, ra)
    return b


def decide(vv, fn, src, rec, ctb_log2, bd, apses=(), tb=None):
    """Runs fn (oracle_alf_decide_luma or vvclf_alf_decide_luma) on a copy of
    rec; returns (rc, choice, sse, luma) with choice[i] = -1 (off) or the set
    index, sse shaped (n_ctbs, 1 + 16 + n_aps)."""
    out = rec.copy()
    h, w = out.y.shape
    ctb = 1 << ctb_log2
    n = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    ctbs = (vv.AlfCtb * n)()
    sse = np.zeros((n, 17 + len(apses)), np.uint64)
    ptrs = (C.POINTER(vv.AlfAps) * max(1, len(apses)))()
    for i, a in enumerate(apses):
        ptrs[i] = C.pointer(a) if a is not None else C.POINTER(vv.AlfAps)()
    rc = fn(src.ptr(), out.ptr(), ctb_log2, bd, 1, C.cast(ptrs, C.POINTER(C.POINTER(vv.AlfAps))), len(apses),
            C.byref(tb) if tb is not None else None, ctbs, sse.ctypes.data_as(C.POINTER(C.c_uint64)))
    choice = np.array([ctbs[i].luma_filt_set_idx if ctbs[i].luma_on else -1 for i in range(n)], np.int64)
    extra = [(ctbs[i].cb_on, ctbs[i].cr_on, ctbs[i].cc_cb_idc, ctbs[i].cc_cr_idc, ctbs[i].cb_alt_idx,
              ctbs[i].cr_alt_idx) for i in range(n)]
    if rc == 0:
        assert all(e == (0, 0, 0, 0, 0, 0) for e in extra)
    return rc, choice, sse, out


def chroma_aps(vv, rng, n_alts, scale=24, clip=True):
    """A chroma APS with n_alts random alternatives."""
    a = vv.AlfAps()
    a.alf_chroma_filter_signal_flag = 1
    a.alf_chroma_clip_flag = int(clip)
    a.n_chroma_coeff = n_alts
    a.alf_chroma_num_alt_filters_minus1 = max(0, n_alts - 1)
    for k in range(n_alts):
        for t in range(6):
            a.chroma_coeff[k][t] = int(rng.integers(-scale, scale + 1))
            a.chroma_clip_idx[k][t] = int(rng.integers(0, 4)) if clip else 0
    return a


def cc_aps(vv, rng, c_idx, n, scale=32, rows=None):
    """A CC-ALF APS with n random filters for component c_idx (or `rows`)."""
    a = vv.AlfAps()
    rows = rows if rows is not None else [[int(rng.integers(-scale, scale + 1)) for _ in range(7)] for _ in range(n)]
    coeff = a.cc_cb_coeff if c_idx == 1 else a.cc_cr_coeff
    for k, r in enumerate(rows):
        for t in range(7):
            coeff[k][t] = r[t]
    if c_idx == 1:
        a.alf_cc_cb_filter_signal_flag, a.n_cc_cb_coeff = 1, len(rows)
    else:
        a.alf_cc_cr_filter_signal_flag, a.n_cc_cr_coeff = 1, len(rows)
    return a


def ctb_fields(c):
    return (c.luma_on, c.luma_filt_set_idx, c.cb_on, c.cb_alt_idx, c.cr_on, c.cr_alt_idx, c.cc_cb_idc, c.cc_cr_idc)


def decide_chroma(vv, fn, src, rec, aps, c_idx, ctb_log2, bd, tb=None, pre_luma=None, init=None, cfi=1):
    """Runs a chroma decision (oracle_alf_decide_chroma / vvclf_alf_decide_chroma,
    or with pre_luma the CC-ALF forms) on a copy of rec and of the CTB records
    `init` (list of field tuples, default all zero); returns (rc, records,
    sse, picture)."""
    out = rec.copy()
    h, w = out.y.shape
    ctb = 1 << ctb_log2
    n = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    apply = (vv.AlfCtb * n)()
    for i, f in enumerate(init or []):
        (apply[i].luma_on, apply[i].luma_filt_set_idx, apply[i].cb_on, apply[i].cb_alt_idx, apply[i].cr_on,
         apply[i].cr_alt_idx, apply[i].cc_cb_idc, apply[i].cc_cr_idc) = f
    if pre_luma is None:
        ncand = aps.n_chroma_coeff
    else:
        ncand = aps.n_cc_cb_coeff if c_idx == 1 else aps.n_cc_cr_coeff
    sse = np.zeros((n, 1 + ncand), np.uint64)
    args = [src.ptr(), out.ptr()]
    if pre_luma is not None:
        pl = vv.Plane(pre_luma.ctypes.data_as(C.POINTER(C.c_uint16)), pre_luma.shape[1], pre_luma.shape[1],
                      pre_luma.shape[0], bd)
        args.append(C.byref(pl))
    rc = fn(*args, apply, C.byref(aps), c_idx, ctb_log2, bd, cfi, C.byref(tb) if tb is not None else None,
            sse.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, [ctb_fields(apply[i]) for i in range(n)], sse, out
