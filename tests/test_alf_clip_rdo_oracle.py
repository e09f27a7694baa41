"""The oracle's clip-index search of a luma APS
(design_clip_rdo_for_luma_aps, src/alf_aps_design.rs:697-768) against the
closed forms of tests/alf_clip_rdo_cases.py, the reference's own tests
(src/alf_aps_design.rs:1024-1170) and an independent form of the search:
per class, since a 4x4 block's ALF output depends only on its class's row
(the picture's SSE is the sum of the classes'), the 25 rows searched
together from per-class SSEs of 1 + 12 x 3 whole-picture ALF passes, the
classes taken from the oracle's classification."""
import ctypes as C

import numpy as np
import pytest

import alf_clip_rdo_cases as cases
from helpers import Pic, alf_cfg, alf_picture, binding_of


@pytest.mark.parametrize("name", sorted(cases.ALL_CASES))
def test_case(orc, vv, name):
    cases.ALL_CASES[name](vv, orc, orc.oracle_alf_clip_rdo_luma)


def block_classes(orc, pic, ctb_log2, bd):
    """Per luma sample the class (0..24) of its 4x4 block."""
    h, w = pic.y.shape
    ctb = 1 << ctb_log2
    out = np.zeros((h, w), np.int64)
    clip = (C.c_int32 * 4)(0, 0, w, h)
    cls = (C.c_uint16 * ((ctb // 4) ** 2))()
    for ry in range((h + ctb - 1) // ctb):
        for rx in range((w + ctb - 1) // ctb):
            orc.oracle_alf_classify_ctb(pic.y.ctypes.data_as(C.POINTER(C.c_uint16)), w, w, h, clip, rx, ry, ctb,
                                        bd, cls)
            m = np.array(cls, np.int64).reshape(ctb // 4, ctb // 4) & 0xFF
            blk = np.kron(m, np.ones((4, 4), np.int64))
            y0, x0 = ry * ctb, rx * ctb
            hh, ww = min(ctb, h - y0), min(ctb, w - x0)
            out[y0:y0 + hh, x0:x0 + ww] = blk[:hh, :ww]
    return out


def per_class_sse(vv, orc, src, rec, coeff, clip, classes, ctb_log2, bd):
    a = vv.AlfAps()
    a.alf_luma_filter_signal_flag = a.alf_luma_clip_flag = 1
    a.n_luma_coeff = 25
    for k in range(25):
        for j in range(12):
            a.luma_coeff[k][j], a.luma_clip_idx[k][j] = int(coeff[k][j]), int(clip[k][j])
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    wc, hc = (w + ctb - 1) // ctb, (h + ctb - 1) // ctb
    on = vv.AlfCtb()
    on.luma_on, on.luma_filt_set_idx = 1, 16
    ap, _keep = alf_picture(vv, wc, hc, {(x, y): on for x in range(wc) for y in range(hc)})
    o = rec.copy()
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                        C.byref(binding_of(vv, [a])), None) == 0
    d = (src.y.astype(np.int64) - o.y.astype(np.int64)) ** 2
    return np.bincount(classes.ravel(), weights=d.ravel(), minlength=25).astype(np.int64)[:25]


def per_class_search(vv, orc, src, rec, coeff, ctb_log2, bd):
    classes = block_classes(orc, rec, ctb_log2, bd)
    clip = np.zeros((25, 12), np.int64)
    cur = per_class_sse(vv, orc, src, rec, coeff, clip, classes, ctb_log2, bd)
    base = int(cur.sum())
    for j in range(12):
        trials = []
        for c in (1, 2, 3):
            t = clip.copy()
            t[:, j] = c
            trials.append(per_class_sse(vv, orc, src, rec, coeff, t, classes, ctb_log2, bd))
        for k in range(25):
            for c in (1, 2, 3):
                if trials[c - 1][k] < cur[k]:
                    cur[k], clip[k][j] = trials[c - 1][k], c
    return clip, base, int(cur.sum())


def noisy_pair(vv, w, h, bd, seed):
    rng = np.random.default_rng(seed)
    src = Pic(vv, w, h, 0, bd=bd)
    # smooth ramps with sharp steps, so that both clipping and smoothing pay
    y, x = np.mgrid[0:h, 0:w]
    base = (x * 3 + y * 2) % 97 + ((x // 7 + y // 5) % 2) * 120
    src.y[:] = (base << (bd - 8)).astype(np.uint16)
    rec = src.copy()
    rec.y[:] = np.clip(src.y.astype(np.int64) + rng.integers(-12, 13, (h, w)) * (1 << (bd - 8)), 0,
                       (1 << bd) - 1)
    return src, rec, rng.integers(-24, 25, (25, 12))


@pytest.mark.parametrize("w,h,ctb_log2,bd,seed", [(64, 64, 7, 8, 1), (48, 40, 4, 10, 2), (72, 56, 5, 12, 3)])
def test_matches_per_class_search(orc, vv, w, h, ctb_log2, bd, seed):
    src, rec, coeff = noisy_pair(vv, w, h, bd, seed)
    rc, clip, flag, base, post = cases.search(vv, orc.oracle_alf_clip_rdo_luma, src, rec, coeff, ctb_log2, bd)
    want_clip, want_base, want_post = per_class_search(vv, orc, src, rec, coeff, ctb_log2, bd)
    assert rc == 0 and base == want_base and post == want_post
    assert (clip == want_clip).all()
    assert clip.any() and flag == 1 and post < base  # the fixture exercises the search


def test_reference_never_increases_sse(orc, vv):
    # src/alf_aps_design.rs:1044-1077 with a fixed low-pass row in place of
    # the Wiener-designed rows: 8-wide stripes of 32 / 224, rec a 3-tap
    # horizontal blur of src
    w = h = 64
    src = cases.stripes(vv, w, h, period=16, lo=32, hi=224)
    rec = src.copy()
    s = src.y.astype(np.int64)
    rec.y[:, 1:-1] = ((s[:, :-2] + s[:, 1:-1] + s[:, 2:]) // 3).astype(np.uint16)
    row = np.array([1, 2, 3, 2, 1, 4, 8, 4, 1, 2, 6, 10])
    rc, clip, flag, base, post = cases.search(vv, orc.oracle_alf_clip_rdo_luma, src, rec, np.tile(row, (25, 1)))
    assert rc == 0 and post <= base
