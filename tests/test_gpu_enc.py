"""GPU parity of the encoder's per-CTB luma ALF decision
(vvclf_alf_decide_luma; src/alf_enc.rs:185-316, :352-470) against the
oracle, bit for bit: the chosen set of every CTB, the whole table of per-CTB
squared errors (off, 16 fixed sets, each APS set) and the committed luma."""
import sys

import numpy as np
import pytest

from alf_decide_cases import decide, luma_aps, noisy, random_aps, smooth_gradient, tile_bounds

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # w, h, ctb_log2, bd, n random APS, unsignalled slot, tile bounds, wide samples
    (200, 136, 6, 10, 2, True, None, False),
    (256, 128, 7, 8, 0, False, None, False),
    (160, 96, 5, 12, 1, False, ([0, 96, 160], [0, 64, 96]), False),
    (96, 72, 4, 8, 1, True, ([0, 48, 96], [0, 32, 72]), False),
    (72, 40, 3, 8, 1, False, None, False),
    (416, 240, 8, 10, 1, False, None, False),
    (200, 136, 6, 16, 1, False, None, False),
    (256, 192, 6, 10, 1, False, None, True),
    (1920, 1080, 7, 10, 2, False, ([0, 640, 1280, 1920], [0, 512, 1080]), False),
]


@pytest.mark.parametrize("w,h,ctb_log2,bd,n_aps,unsig,tbs,wide", CASES)
def test_decide_matches_oracle(vv, orc, w, h, ctb_log2, bd, n_aps, unsig, tbs, wide):
    say(f"alf decide {w}x{h} ctb_log2={ctb_log2} bd={bd} aps={n_aps} unsig={unsig} tiles={tbs is not None} "
        f"wide={wide}")
    rng = np.random.default_rng(w + 3 * h + bd)
    src = smooth_gradient(vv, w, h, bd)
    src.y[:] = np.clip(src.y.astype(np.int64) + rng.integers(-3, 4, src.y.shape) * (1 << (bd - 8)), 0,
                       (1 << bd) - 1)
    rec = noisy(src, 0x9E3779B1, (1 << (bd - 5)) - 1, 1 << (bd - 7))
    if wide:
        idx = rng.choice(w * h, 300, replace=False)
        rec.y.ravel()[idx] = rng.integers(1 << 15, 1 << 16, 300)
    apses = [random_aps(vv, rng, 16) for _ in range(n_aps)]
    if unsig:
        apses.append(luma_aps(vv, {6: 40}, signal=False))
    tb = tile_bounds(vv, *tbs) if tbs else None
    rc, choice, sse, out = decide(vv, vv.lib.vvclf_alf_decide_luma, src, rec, ctb_log2, bd, apses, tb)
    assert rc == 0, vv.lib.vvclf_last_error()
    rc_o, choice_o, sse_o, out_o = decide(vv, orc.oracle_alf_decide_luma, src, rec, ctb_log2, bd, apses, tb)
    assert rc_o == 0
    assert np.array_equal(sse, sse_o)
    assert np.array_equal(choice, choice_o)
    assert np.array_equal(out.y, out_o.y)
    assert np.array_equal(out.cb, rec.cb) and np.array_equal(out.cr, rec.cr)
    assert (choice >= 0).any()


def test_decide_tie_rules(vv, orc):
    say("alf decide: identity and duplicate APS sets")
    bd, w, h = 10, 256, 192
    src = smooth_gradient(vv, w, h, bd)
    src.y[:] = 1 << (bd - 1)
    rec = noisy(src, 0x2545F491, 63, 32)
    smooth = {k: 5 for k in range(12)}
    apses = [luma_aps(vv), luma_aps(vv, smooth), luma_aps(vv, smooth)]
    rc, choice, sse, out = decide(vv, vv.lib.vvclf_alf_decide_luma, src, rec, 6, bd, apses)
    assert rc == 0, vv.lib.vvclf_last_error()
    _rc, choice_o, sse_o, out_o = decide(vv, orc.oracle_alf_decide_luma, src, rec, 6, bd, apses)
    assert np.array_equal(sse, sse_o) and np.array_equal(choice, choice_o) and np.array_equal(out.y, out_o.y)
    assert (choice == 17).any() and (choice != 16).all() and (choice != 18).all()


def test_decide_rejects_short_aps(vv):
    say("alf decide: APS with fewer than 25 luma filters")
    src = smooth_gradient(vv, 64, 64, 8)
    short = luma_aps(vv, {6: 8})
    short.n_luma_coeff = 5
    rc, *_ = decide(vv, vv.lib.vvclf_alf_decide_luma, src, src, 6, 8, [short])
    assert rc == vv.ERR_INVALID
