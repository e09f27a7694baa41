"""GPU parity of the encoder's per-CTB luma ALF decision
(vvclf_alf_decide_luma; src/alf_enc.rs:185-316, :352-470) against the
oracle, bit for bit: the chosen set of every CTB, the whole table of per-CTB
squared errors (off, 16 fixed sets, each APS set) and the committed luma."""
import sys

import numpy as np
import pytest

from alf_decide_cases import decide, decide_chroma, luma_aps, noisy, random_aps, smooth_gradient, tile_bounds
from test_alf_decide_oracle import cc_fixture, chroma_fixture, lowpass_chroma_aps

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


CHROMA_CASES = [
    # w, h, ctb_log2, bd, c_idx, n candidates, tile bounds, wide samples
    (200, 136, 6, 10, 1, 4, None, False),
    (256, 128, 7, 8, 2, 8, None, False),
    (160, 96, 5, 12, 1, 2, ([0, 96, 160], [0, 64, 96]), False),
    (96, 72, 4, 10, 2, 3, ([0, 48, 96], [0, 32, 72]), False),
    (72, 40, 3, 8, 1, 2, None, False),
    (416, 240, 8, 16, 2, 3, None, False),
    (256, 192, 6, 10, 1, 3, None, True),
    (1920, 1080, 7, 10, 2, 4, ([0, 640, 1280, 1920], [0, 512, 1080]), False),
]


@pytest.mark.parametrize("cc", [False, True])
@pytest.mark.parametrize("w,h,ctb_log2,bd,c_idx,n,tbs,wide", CHROMA_CASES)
def test_chroma_decisions_match_oracle(vv, orc, w, h, ctb_log2, bd, c_idx, n, tbs, wide, cc):
    say(f"{'cc-alf' if cc else 'chroma alf'} decide {w}x{h} ctb_log2={ctb_log2} bd={bd} c={c_idx} n={n} "
        f"tiles={tbs is not None} wide={wide}")
    rng, src, rec, pre = chroma_fixture(vv, w, h, bd, w + 7 * h + c_idx)
    tb = tile_bounds(vv, *tbs) if tbs else None
    if cc:
        aps, rec = cc_fixture(orc, vv, rng, src, rec, pre, c_idx, ctb_log2, bd, min(n, 4), tb)
    else:
        aps = lowpass_chroma_aps(vv, rng, n)
    if wide:  # samples of 2^15 and more in the right half
        for pl in (rec.cb, rec.cr, pre):
            r = pl[:, pl.shape[1] // 2:]
            idx = rng.choice(r.size, r.size // 50, replace=False)
            r.reshape(-1)[idx] = rng.integers(1 << 15, 1 << 16, idx.size)
            pl[:, pl.shape[1] // 2:] = r
    ctb = 1 << ctb_log2
    nct = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    init = [(1, 2, 0, 0, 0, 0, 0, 0)] * nct  # luma fields the decision keeps
    fg, fo = (vv.lib.vvclf_cc_alf_decide, orc.oracle_cc_alf_decide) if cc else \
        (vv.lib.vvclf_alf_decide_chroma, orc.oracle_alf_decide_chroma)
    p = pre if cc else None
    rc, recs, sse, out = decide_chroma(vv, fg, src, rec, aps, c_idx, ctb_log2, bd, tb, p, init)
    assert rc == 0, vv.lib.vvclf_last_error()
    rc_o, recs_o, sse_o, out_o = decide_chroma(vv, fo, src, rec, aps, c_idx, ctb_log2, bd, tb, p, init)
    assert rc_o == 0
    assert np.array_equal(sse, sse_o)
    assert recs == recs_o
    assert np.array_equal(out.cb, out_o.cb) and np.array_equal(out.cr, out_o.cr) and np.array_equal(out.y, rec.y)
    assert recs != init  # some CTBs take a filter
    if 2 * ctb <= w:  # and the exact left half stays off
        assert any(r == init[0] for r in recs)


def test_chroma_decisions_reject_on_the_host(vv):
    say("chroma decisions: refusals")
    rng, src, rec, pre = chroma_fixture(vv, 64, 64, 8, 3)
    aps = lowpass_chroma_aps(vv, rng, 2)
    fn = vv.lib.vvclf_alf_decide_chroma
    assert decide_chroma(vv, fn, src, rec, aps, 1, 6, 8, cfi=2)[0] == vv.ERR_UNSUPPORTED
    assert decide_chroma(vv, fn, src, rec, aps, 3, 6, 8)[0] == vv.ERR_INVALID
    aps.n_chroma_coeff = 9
    assert decide_chroma(vv, fn, src, rec, aps, 1, 6, 8)[0] == vv.ERR_INVALID
