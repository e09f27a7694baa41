"""GPU parity of the luma ALF clip-index search (vvclf_alf_clip_rdo_luma;
design_clip_rdo_for_luma_aps, src/alf_aps_design.rs:697-768): the closed
forms of tests/alf_clip_rdo_cases.py, the oracle's literal search (901
picture replays) bit for bit on small pictures, the oracle's per-class form
at 1920x1080, and at 3840x2160 the reported SSEs against the ALF stage run
with the found indices."""
import ctypes as C
import sys

import numpy as np
import pytest

import alf_clip_rdo_cases as cases
from helpers import alf_cfg, alf_picture, binding_of
from test_alf_clip_rdo_oracle import noisy_pair, per_class_search

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(cases.ALL_CASES))
def test_case(vv, orc, name):
    say(f"clip search case {name}")
    cases.ALL_CASES[name](vv, orc, vv.lib.vvclf_alf_clip_rdo_luma)


CASES = [
    # w, h, ctb_log2, bd, samples of 2^15 and above
    (64, 64, 7, 8, False),
    (96, 80, 5, 10, False),
    (72, 56, 3, 8, False),
    (136, 104, 6, 12, False),
    (128, 96, 7, 14, False),
    (120, 88, 4, 16, False),
    (128, 128, 6, 10, True),
    (256, 256, 8, 8, False),
]


@pytest.mark.parametrize("w,h,ctb_log2,bd,wide", CASES)
def test_matches_oracle(vv, orc, w, h, ctb_log2, bd, wide):
    say(f"clip search {w}x{h} ctb_log2={ctb_log2} bd={bd} wide={wide}")
    src, rec, coeff = noisy_pair(vv, w, h, bd, w + h + bd)
    if wide:
        rng = np.random.default_rng(9)
        idx = rng.choice(w * h, 200, replace=False)
        rec.y.ravel()[idx] = rng.integers(1 << 15, 1 << 16, 200)
    got = cases.search(vv, vv.lib.vvclf_alf_clip_rdo_luma, src, rec, coeff, ctb_log2, bd)
    assert got[0] == 0, vv.lib.vvclf_last_error()
    want = cases.search(vv, orc.oracle_alf_clip_rdo_luma, src, rec, coeff, ctb_log2, bd)
    assert want[0] == 0
    assert (got[1] == want[1]).all() and got[2:] == want[2:]
    assert got[1].any()


def test_1080p_matches_per_class_oracle(vv, orc):
    say("clip search 1920x1080 against the per-class oracle form")
    src, rec, coeff = noisy_pair(vv, 1920, 1080, 10, 7)
    rc, clip, flag, base, post = cases.search(vv, vv.lib.vvclf_alf_clip_rdo_luma, src, rec, coeff, 7, 10)
    assert rc == 0, vv.lib.vvclf_last_error()
    want_clip, want_base, want_post = per_class_search(vv, orc, src, rec, coeff, 7, 10)
    assert (clip == want_clip).all() and base == want_base and post == want_post


def gpu_alf_sse(vv, src, rec, coeff, clip, ctb_log2, bd):
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
    rc = vv.lib.vvclf_apply_alf(o.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                C.byref(binding_of(vv, [a])))
    assert rc == 0, vv.lib.vvclf_last_error()
    d = src.y.astype(np.int64) - o.y.astype(np.int64)
    return int((d * d).sum())


def test_4k_sses_replay(vv):
    say("clip search 3840x2160: reported SSEs against the ALF stage")
    src, rec, coeff = noisy_pair(vv, 3840, 2160, 10, 11)
    rc, clip, flag, base, post = cases.search(vv, vv.lib.vvclf_alf_clip_rdo_luma, src, rec, coeff, 7, 10)
    assert rc == 0, vv.lib.vvclf_last_error()
    assert post < base and flag == 1 and clip.any()
    assert base == gpu_alf_sse(vv, src, rec, coeff, np.zeros((25, 12), np.int64), 7, 10)
    assert post == gpu_alf_sse(vv, src, rec, coeff, clip, 7, 10)
