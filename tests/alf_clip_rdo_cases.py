"""Pictures and closed forms for the luma ALF clip-index search
(design_clip_rdo_for_luma_aps, src/alf_aps_design.rs:697-768), shared by the
oracle checks and the GPU parity tests.  Each case takes the search to test
(the oracle's or the library's) and the oracle's ALF stage (pinned on its
own) for the facts a case needs about the picture."""
import ctypes as C

import numpy as np

from helpers import Pic, alf_cfg, alf_picture, binding_of

# AlfClip exponents at 8 bits (Table 8, src/alf.rs:627-644): clip index
# 0..3 bounds a neighbour difference by 256, 32, 8, 2
# The 4x4 blocks of vertical stripes 4 samples wide fall in class 24 with a
# transposeIdx of 1 or 3, which reads APS tap 2 (the (0, +-2) pair of the
# untransposed diamond) at the (+-2, 0) position (eqs. 1442-1445): across
# the stripes, every sample has exactly one of that pair in the other stripe
ACROSS = 2


def search(vv, fn, src, rec, coeff, ctb_log2=7, bd=8):
    """Runs fn on (src, rec, coeff); returns (rc, clip_idx (25, 12), flag,
    baseline, post)."""
    flat = np.ascontiguousarray(np.asarray(coeff, np.int32).reshape(25, 12))
    out = vv.AlfClipRdo()
    rc = fn(src.ptr(), rec.ptr(), flat.ctypes.data_as(C.POINTER(C.c_int32)), ctb_log2, bd, 1, C.byref(out))
    clip = np.array([[out.clip_idx[k][j] for j in range(12)] for k in range(25)], np.int64)
    coeff_back = np.array([[out.coeff[k][j] for j in range(12)] for k in range(25)], np.int64)
    if rc == 0:
        assert (coeff_back == flat).all()
    return rc, clip, out.alf_luma_clip_flag, out.baseline_sse_y, out.post_clip_sse_y


def alf_sse(vv, orc, src, rec, coeff, clip, ctb_log2=7, bd=8):
    """SSE of src against the oracle's ALF of rec with every CTB on an APS of
    rows `coeff` and clip indices `clip` (replay_aps_full_picture_sse)."""
    a = vv.AlfAps()
    a.alf_luma_filter_signal_flag = 1
    a.alf_luma_clip_flag = 1
    a.n_luma_coeff = 25
    for k in range(25):
        for j in range(12):
            a.luma_coeff[k][j] = int(coeff[k][j])
            a.luma_clip_idx[k][j] = int(clip[k][j])
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    wc, hc = (w + ctb - 1) // ctb, (h + ctb - 1) // ctb
    on = vv.AlfCtb()
    on.luma_on, on.luma_filt_set_idx = 1, 16
    ap, _keep = alf_picture(vv, wc, hc, {(x, y): on for x in range(wc) for y in range(hc)})
    o = rec.copy()
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                        C.byref(binding_of(vv, [a])), None) == 0
    d = src.y.astype(np.int64) - o.y.astype(np.int64)
    return int((d * d).sum())


def stripes(vv, w, h, period=8, lo=16, hi=240, bd=8):
    """Vertical stripes period / 2 samples wide."""
    p = Pic(vv, w, h, 0, bd=bd)
    x = np.arange(w)
    p.y[:] = np.where((x // (period // 2)) % 2 == 0, lo, hi).astype(np.uint16)[None, :]
    return p


def one_tap_rows(tap, c, rows=range(25)):
    coeff = np.zeros((25, 12), np.int64)
    for k in rows:
        coeff[k][tap] = c
    return coeff


def effective_rows(vv, orc, pic, tap, c):
    """The rows k for which a filter with only coefficient c at `tap` in row
    k changes the picture: the classes present whose transposes put the tap
    across the stripes."""
    zero = np.zeros((25, 12), np.int64)
    return [k for k in range(25) if alf_sse(vv, orc, pic, pic, one_tap_rows(tap, c, [k]), zero) > 0]


# --------------------------------------------------------------------------
# cases: each asserts its closed form on fn
# --------------------------------------------------------------------------

def case_flat_picture_keeps_no_clip(vv, orc, fn):
    # src/alf_aps_design.rs:1024-1040: rec == src flat, zero rows
    src = Pic(vv, 64, 64, 128)
    rc, clip, flag, base, post = search(vv, fn, src, src.copy(), np.zeros((25, 12)))
    assert rc == 0 and base == 0 and post == 0 and flag == 0 and not clip.any()


def case_smallest_clip_restores_edges(vv, orc, fn):
    # rec == src, 4-wide stripes of 16 / 240, every row 16 at tap ACROSS:
    # the pair sums 16 * clip(+-224) on every sample, the output
    # moves by (16 * 224 + 64) >> 7 = 28, (16 * 32 + 64) >> 7 = 4,
    # (16 * 8 + 64) >> 7 = 1 and (16 * 2 + 64) >> 7 = 0 for indices 0..3
    # (-28, -4, -1, 0 below): each index strictly improves on the one before,
    # so a row that reaches a stripe edge takes 3, and its other taps (zero
    # coefficients, no change) stay 0, as does every row that changes nothing
    src = stripes(vv, 64, 64)
    coeff = one_tap_rows(ACROSS, 16)
    eff = effective_rows(vv, orc, src, ACROSS, 16)
    assert eff, "the stripes reach no row"
    rc, clip, flag, base, post = search(vv, fn, src, src.copy(), coeff)
    want = np.zeros((25, 12), np.int64)
    want[eff, ACROSS] = 3
    assert rc == 0 and (clip == want).all(), (clip[:, ACROSS], eff)
    assert flag == 1 and post == 0 and base == alf_sse(vv, orc, src, src, coeff, np.zeros((25, 12))) > 0


def case_equal_trials_keep_first_index(vv, orc, fn):
    # as above with coefficient 1: indices 1, 2 and 3 all give
    # (clip + 64) >> 7 = 0 (SSE 0) against 2 / -2 for index 0: the first
    # strict improvement, index 1, is kept and the equal later ones are not
    src = stripes(vv, 64, 64)
    coeff = one_tap_rows(ACROSS, 1)
    eff = effective_rows(vv, orc, src, ACROSS, 16)
    rc, clip, flag, base, post = search(vv, fn, src, src.copy(), coeff)
    want = np.zeros((25, 12), np.int64)
    want[eff, ACROSS] = 1
    assert rc == 0 and (clip == want).all(), (clip[:, ACROSS], eff)
    assert flag == 1 and post == 0 < base


def case_no_change_keeps_zero(vv, orc, fn):
    # every neighbour difference within 2 (samples 100..102), so no clip
    # index changes any output: each trial equals the current SSE and no
    # tap leaves 0, although the rows filter (base > 0)
    rng = np.random.default_rng(5)
    src = Pic(vv, 48, 40, 0)
    src.y[:] = rng.integers(100, 103, src.y.shape)
    rec = src.copy()
    rec.y[:] = rng.integers(100, 103, src.y.shape)
    coeff = rng.integers(-30, 31, (25, 12))
    rc, clip, flag, base, post = search(vv, fn, src, rec, coeff, ctb_log2=5)
    assert rc == 0 and not clip.any() and flag == 0 and post == base > 0
    assert base == alf_sse(vv, orc, src, rec, coeff, np.zeros((25, 12)), ctb_log2=5)


def case_result_replays(vv, orc, fn):
    # the reference's own high-contrast fixture
    # (src/alf_aps_design.rs:1122-1170): 4-wide stripes of 16 / 240 and an
    # LCG-perturbed reconstruction, every row 16 at tap 6, 8 at taps 2 and
    # 10; the search never raises the SSE, and baseline / post are the SSE
    # of ALF with clip 0 / with the chosen indices
    src, rec = high_contrast_fixture(vv)
    row = np.zeros(12, np.int64)
    row[6], row[2], row[10] = 16, 8, 8
    coeff = np.tile(row, (25, 1))
    rc, clip, flag, base, post = search(vv, fn, src, rec, coeff)
    assert rc == 0 and post <= base
    assert base == alf_sse(vv, orc, src, rec, coeff, np.zeros((25, 12)))
    assert post == alf_sse(vv, orc, src, rec, coeff, clip)
    assert flag == int(clip.any())


def high_contrast_fixture(vv, w=128, h=128):
    src = stripes(vv, w, h)
    y, x = np.mgrid[0:h, 0:w].astype(np.uint64)
    with np.errstate(over="ignore"):
        seed = y * np.uint64(2654435761) + x
        bits = seed * np.uint64(6364136223846793005) + np.uint64(1442695040888963407)
    noise = ((bits >> np.uint64(56)) & np.uint64(0x1F)).astype(np.int64) - 16
    rec = src.copy()
    rec.y[:] = np.clip(src.y.astype(np.int64) + noise, 0, 255)
    return src, rec


ALL_CASES = {f.__name__: f for f in [case_flat_picture_keeps_no_clip, case_smallest_clip_restores_edges,
                                     case_equal_trials_keep_first_index, case_no_change_keeps_zero,
                                     case_result_replays]}
