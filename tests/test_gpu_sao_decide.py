"""GPU parity of the SAO encoder decision (vvclf_sao_decide_picture;
sao_decide_picture, src/sao_enc.rs:385-462) with the oracle, field for
field, and the hand-worked cases of tests/sao_decide_cases.py through the
library."""
import sys

import numpy as np
import pytest

import sao_decide_cases as cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(cases.ALL_CASES))
def test_case_through_the_gpu(vv, name):
    say(f"sao decide case {name}")
    cases.ALL_CASES[name](vv, vv.lib.vvclf_sao_decide_picture)


def deblocked_like(vv, rng, w, h, bd, mono=False):
    """A smooth source with structure and a reconstruction carrying
    quantisation-like errors: ringing near edges and a band bias."""
    src = cases.pic(vv, w, h, bd)
    yy, xx = np.mgrid[0:h, 0:w]
    base = (xx * 3 + yy * 2) % 512 + ((xx // 16 + yy // 16) % 2) * 200
    src.y[:] = (base * (1 << bd) // 1024).astype(np.uint16)
    cy, cx = np.mgrid[0:h // 2, 0:w // 2]
    src.cb[:] = (((cx * 5 + cy) % 256) * (1 << bd) // 512 + (1 << (bd - 2))).astype(np.uint16)
    src.cr[:] = (((cx + cy * 7) % 256) * (1 << bd) // 512 + (1 << (bd - 2))).astype(np.uint16)
    rec = src.copy()
    q = 1 << (bd - 8)
    for p in (rec.y, rec.cb, rec.cr):
        noise = rng.integers(-4, 5, p.shape) * q + (((p.astype(np.int64) >> (bd - 5)) % 3) - 1) * 2 * q
        p[:] = np.clip(p.astype(np.int64) + noise, 0, (1 << bd) - 1)
    # a flat corner carrying only a level shift: no edge category, so band
    # offset is what the decision takes there
    fy, fx = min(h, 320), min(w, 320)  # beyond a 256 CTB, so that its border neighbours are flat too
    for a, b in ((src.y, rec.y), (src.cb, rec.cb), (src.cr, rec.cr)):
        sy, sx = (fy, fx) if a is src.y else (fy // 2, fx // 2)
        b[:sy, :sx] = 1 << (bd - 1)
        a[:sy, :sx] = (1 << (bd - 1)) - 5 * q
    if mono:
        for p in (src, rec):
            p.cb = np.zeros((0, 0), np.uint16)
            p.cr = np.zeros((0, 0), np.uint16)
    return src, rec


CASES = [
    # w, h, bd, ctb_log2, luma, chroma, mono
    (416, 240, 10, 7, True, True, False),
    (200, 136, 8, 6, True, True, False),
    (333, 199, 12, 5, True, True, False),
    (96, 72, 10, 4, True, False, False),
    (72, 40, 8, 3, False, True, False),
    (256, 256, 16, 8, True, True, False),
    (160, 96, 10, 6, True, True, True),
    (1920, 1080, 10, 7, True, True, False),
    (3840, 2160, 10, 7, True, True, False),
]


@pytest.mark.parametrize("w,h,bd,ctb_log2,luma,chroma,mono", CASES)
def test_decision_matches_oracle(vv, orc, w, h, bd, ctb_log2, luma, chroma, mono):
    say(f"sao decide {w}x{h} bd={bd} ctb_log2={ctb_log2} luma={luma} chroma={chroma} mono={mono}")
    rng = np.random.default_rng(w * 3 + h + bd)
    src, rec = deblocked_like(vv, rng, w, h, bd, mono)
    rc, got = cases.run(vv, vv.lib.vvclf_sao_decide_picture, src, rec, ctb_log2, bd, luma, chroma)
    assert rc == 0, vv.lib.vvclf_last_error()
    rc_o, want = cases.run(vv, orc.oracle_sao_decide_picture, src, rec, ctb_log2, bd, luma, chroma)
    assert rc_o == 0
    for i, (g, o) in enumerate(zip(got, want)):
        for comp in ("luma", "cb", "cr"):
            assert cases.fields(getattr(g, comp)) == cases.fields(getattr(o, comp)), (i, comp)
    kinds = {cases.fields(getattr(o, c))[0] for o in want for c in ("luma", "cb", "cr")}
    if w * h >= 100000:
        assert {1, 2} <= kinds, kinds  # band and edge offsets both chosen somewhere


@pytest.mark.parametrize("name", sorted(cases.ALL_MERGE_CASES))
def test_merge_case_through_the_gpu(vv, name):
    say(f"sao merge case {name}")
    cases.ALL_MERGE_CASES[name](vv, vv.lib.vvclf_sao_chroma_merge, vv.lib.vvclf_sao_decide_picture)


@pytest.mark.parametrize("w,h,bd,ctb_log2", [(416, 240, 10, 6), (333, 199, 8, 5), (1920, 1080, 10, 7),
                                            (256, 256, 12, 4), (96, 72, 10, 3)])
def test_merge_matches_oracle(vv, orc, w, h, bd, ctb_log2):
    say(f"sao chroma merge {w}x{h} bd={bd} ctb_log2={ctb_log2}")
    rng = np.random.default_rng(w + h * 5 + bd)
    src, rec = deblocked_like(vv, rng, w, h, bd)
    rc, indep = cases.run(vv, orc.oracle_sao_decide_picture, src, rec, ctb_log2, bd)
    assert rc == 0
    rc, got, gm = cases.merge(vv, vv.lib.vvclf_sao_chroma_merge, src, rec, indep, ctb_log2, bd)
    assert rc == 0, vv.lib.vvclf_last_error()
    rc, want, wm = cases.merge(vv, orc.oracle_sao_chroma_merge, src, rec, indep, ctb_log2, bd)
    assert rc == 0 and gm == wm
    for g, o in zip(got, want):
        for comp in ("luma", "cb", "cr"):
            assert cases.fields(getattr(g, comp)) == cases.fields(getattr(o, comp))
    assert set(wm) >= {0, 1}, set(wm)  # independent and merged CTBs both occur
