"""Closed-form cases for the SAO encoder decision (sao_decide_picture,
src/sao_enc.rs:385-462), each worked by hand from the reference's rules;
`fn` is the oracle's or the library's entry point, so the GPU tests run the
same cases."""
import ctypes as C

import numpy as np

from helpers import Pic


def run(vv, fn, src, rec, ctb_log2, bd, luma=True, chroma=True):
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    n = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    out = (vv.SaoCtbParams * n)()
    rc = fn(src.ptr(), rec.ptr(), ctb_log2, bd, int(luma), int(chroma), out)
    return rc, [out[i] for i in range(n)]


def fields(c):
    return (c.sao_type_idx, c.eo_class, c.band_position, tuple(c.offset_val))


def pic(vv, w, h, bd, y=100, c=None):
    p = Pic(vv, w, h, y, bd=bd)
    if c is not None:
        p.cb[:] = c
        p.cr[:] = c
    return p


NA = (0, 0, 0, (0, 0, 0, 0, 0))


def case_identical_is_not_applied(vv, fn):
    # src/sao_enc.rs:767-772: src == rec, nothing improves
    rng = np.random.default_rng(1)
    p = pic(vv, 64, 48, 10)
    p.y[:] = rng.integers(0, 1024, p.y.shape)
    p.cb[:] = rng.integers(0, 1024, p.cb.shape)
    rc, out = run(vv, fn, p, p.copy(), 5, 10)
    assert rc == 0 and all(fields(o.luma) == NA and fields(o.cb) == NA and fields(o.cr) == NA for o in out)


def case_band_offset_closed_form(vv, fn):
    # flat rec in one band, src c below it: edge offset sees only category 0;
    # the first window holding band b starts at b - 3; over n (even) samples
    # the offset is (-c n + n / 2) / n = -c + 1/2, truncated toward zero to
    # -(c - 1), then clipped to cMax = 2^(Min(BitDepth, 10) - 5) - 1
    for bd, rv, c, want in [(8, 100, 20, -7), (10, 400, 20, -19), (12, 1600, 100, -31), (8, 40, 3, -2)]:
        src, rec = pic(vv, 32, 32, bd, rv - c, rv - c), pic(vv, 32, 32, bd, rv, rv)
        rc, out = run(vv, fn, src, rec, 5, bd)
        band = rv >> (bd - 5)
        start = max(0, band - 3)
        off = [0, 0, 0, 0]
        off[band - start] = want
        sh = bd - min(bd, 10)
        exp = (1, 0, start, (0, *[o << sh if o >= 0 else -((-o) << sh) for o in off]))
        assert rc == 0 and fields(out[0].luma) == exp, (bd, fields(out[0].luma), exp)
        assert fields(out[0].cb) == exp and fields(out[0].cr) == exp


def case_band_31_is_never_in_a_window(vv, fn):
    # the windows start at 0..27 (src/sao_enc.rs:117), so band 31 is never
    # offset: flat rec in band 31 is left alone, in band 30 it is offset
    src, rec = pic(vv, 32, 32, 8, 240, 240), pic(vv, 32, 32, 8, 250, 250)
    rc, out = run(vv, fn, src, rec, 5, 8)
    assert rc == 0 and fields(out[0].luma) == NA
    src, rec = pic(vv, 32, 32, 8, 236, 236), pic(vv, 32, 32, 8, 246, 246)
    rc, out = run(vv, fn, src, rec, 5, 8)
    assert rc == 0 and fields(out[0].luma) == (1, 0, 27, (0, 0, 0, 0, -7))


def case_offset_rounds_toward_zero(vv, fn):
    # (sum + n / 2) / n in integers truncates toward zero: 64 samples of band
    # 12 with sum(src - rec) = -100 give (-100 + 32) / 64 = -1 (floor: -2)
    src, rec = pic(vv, 8, 8, 8, 100, 100), pic(vv, 8, 8, 8, 100, 100)
    src.y.ravel()[:50] = 98
    rc, out = run(vv, fn, src, rec, 3, 8, chroma=False)
    assert rc == 0 and fields(out[0].luma) == (1, 0, 9, (0, 0, 0, 0, -1))


def case_edge_offset_class_and_counts(vv, fn):
    # columns alternating 96 / 104 over a flat 100: every inner sample is a
    # local minimum (category 1, src - rec = +4) or maximum (4, -4) along the
    # horizontal class; offsets round((4 n + n/2) / n) = 4; the diagonal
    # classes see the same pattern on fewer samples (their neighbour rows
    # leave the picture), the vertical class only category 0, so horizontal
    # wins; edge offset is taken over the (also improving) band offset
    src, rec = pic(vv, 16, 16, 8, 100, 100), pic(vv, 16, 16, 8, 100, 100)
    rec.y[:, 0::2] = 96
    rec.y[:, 1::2] = 104
    rc, out = run(vv, fn, src, rec, 4, 8, chroma=False)
    assert rc == 0 and fields(out[0].luma) == (2, 0, 0, (0, 4, 0, 0, -4))
    # a checkerboard in a 8x16 picture: horizontal and vertical neighbours
    # are both opposite; the vertical class leaves out 2 rows of 8 samples,
    # the horizontal one 2 columns of 16, so vertical has more samples
    src, rec = pic(vv, 8, 16, 8, 100, 100), pic(vv, 8, 16, 8, 100, 100)
    yy, xx = np.mgrid[0:16, 0:8]
    rec.y[:] = np.where((xx + yy) % 2 == 0, 96, 104)
    rc, out = run(vv, fn, src, rec, 4, 8, chroma=False)
    assert rc == 0 and fields(out[0].luma) == (2, 1, 0, (0, 4, 0, 0, -4))


def case_cr_follows_cb(vv, fn):
    # Cr keeps Cb's type and class (src/sao_enc.rs:286-347)
    src, rec = pic(vv, 32, 32, 8, 100, 100), pic(vv, 32, 32, 8, 100, 100)
    rec.cb[:, 0::2] = 96
    rec.cb[:, 1::2] = 104
    rec.cr[:] = 100  # nothing to gain: edge offset of Cb's class, zero offsets
    rc, out = run(vv, fn, src, rec, 5, 8)
    assert rc == 0 and fields(out[0].cb) == (2, 0, 0, (0, 4, 0, 0, -4))
    assert fields(out[0].cr) == (2, 0, 0, (0, 0, 0, 0, 0))
    src, rec = pic(vv, 32, 32, 8, 100, 100), pic(vv, 32, 32, 8, 100, 120)
    rec.cr[:] = 100  # Cb band offset, Cr nothing: band offset with zeros
    rc, out = run(vv, fn, src, rec, 5, 8)
    assert rc == 0 and fields(out[0].cb)[0] == 1 and fields(out[0].cr) == (1, 0, 0, (0, 0, 0, 0, 0))
    src, rec = pic(vv, 32, 32, 8, 100, 100), pic(vv, 32, 32, 8, 100, 100)
    rec.cr[:] = 120  # Cb not applied: Cr not applied though it could gain
    rc, out = run(vv, fn, src, rec, 5, 8)
    assert rc == 0 and fields(out[0].cb) == NA and fields(out[0].cr) == NA


def case_chroma_region_rounds_down(vv, fn):
    # width 73, CTB 64: the second CTB is 9 luma columns wide, its chroma
    # region floor(9 / 2) = 4 columns (32..35); chroma column 36 is in no
    # CTB's statistics
    src, rec = pic(vv, 73, 16, 8, 100, 100), pic(vv, 73, 16, 8, 100, 100)
    for p in (src, rec):  # chroma planes of ceil(73 / 2) = 37 columns
        p.cb = np.full((8, 37), 100, np.uint16)
        p.cr = np.full((8, 37), 100, np.uint16)
    rec.cb[:, 36] = 130
    rc, out = run(vv, fn, src, rec, 6, 8)
    assert rc == 0 and fields(out[1].cb) == NA
    rec.cb[:, 35] = 130
    rc, out = run(vv, fn, src, rec, 6, 8)
    assert rc == 0 and fields(out[1].cb) != NA


def case_samples_beyond_the_bit_depth(vv, fn):
    src, rec = pic(vv, 32, 32, 8, 100, 100), pic(vv, 32, 32, 8, 100, 100)
    rec.y[3, 3] = 300  # band 37 at 8 bits: the reference indexes past its bands
    rc, _out = run(vv, fn, src, rec, 5, 8)
    assert rc == vv.ERR_INVALID
    rc, _out = run(vv, fn, src, rec, 5, 8, luma=False)
    assert rc == 0


ALL_CASES = {k: v for k, v in globals().items() if k.startswith("case_")}


# ------------------------------------------------ chroma SAO merge decision
def merge(vv, fn, src, rec, indep, ctb_log2, bd):
    n = len(indep)
    arr = (vv.SaoCtbParams * n)(*indep)
    out = (vv.SaoCtbParams * n)()
    m = (C.c_uint8 * n)()
    rc = fn(src.ptr(), rec.ptr(), arr, bd, ctb_log2, out, m)
    return rc, [out[i] for i in range(n)], list(m)


def merge_case_single_ctb_keeps_independent(vv, fn, decide_fn):
    # src/sao_enc.rs:845-859
    src = pic(vv, 64, 64, 8, 100, 128)
    rec = src.copy()
    rec.cb.ravel()[0] ^= 0x10
    rc, indep = run(vv, decide_fn, src, rec, 7, 8)
    assert rc == 0
    rc, out, m = merge(vv, fn, src, rec, indep, 7, 8)
    assert rc == 0 and m == [0] and fields(out[0].cb) == fields(indep[0].cb)


def merge_case_flat_picture_merges_everywhere(vv, fn, decide_fn):
    # src/sao_enc.rs:863-897, worked out: identical CTBs give every candidate
    # the same SSE, so a neighbour's parameters cost 8 (left) or 16 (above)
    # against 8 * 33 for two independent band offsets: every CTB after the
    # first merges left, the first of each other row merges up
    src = pic(vv, 128, 128, 8, 100, 100)
    rec = src.copy()
    rec.cb[:] = 110
    rec.cr[:] = 90
    rc, indep = run(vv, decide_fn, src, rec, 5, 8, luma=False)
    assert rc == 0 and all(fields(o.cb)[0] == 1 for o in indep)
    rc, out, m = merge(vv, fn, src, rec, indep, 5, 8)
    assert rc == 0
    assert m == [0 if i == 0 else (1 if i % 4 else 2) for i in range(16)]


def merge_case_rate_weight(vv, fn, decide_fn):
    # picture 24x16 (chroma 12x8), CTB 16: CTB 0 exact (both chroma planes
    # not applied), CTB 1 (chroma columns 8..11, 32 samples per plane) two
    # levels high in one band: its independent pick is band offset -1 on Cb
    # and Cr, SSE 2 * 32 * 1 = 64, bits 1 + 16 + 16; merging left (not
    # applied) leaves SSE 2 * 32 * 4 = 256: 256 + 8 < 64 + 8 * 33, merge.
    # (A weight of 4, or 10 bits for a band offset, would keep it
    # independent.)  The same pair stacked merges up: 256 + 16 < 328.
    for w, h, idx in ((24, 16, 1), (16, 24, 1)):
        src, rec = pic(vv, w, h, 8, 100, 102), pic(vv, w, h, 8, 100, 102)
        if w > h:
            src.cb[:, 8:] = 100
            src.cr[:, 8:] = 100
        else:
            src.cb[8:, :] = 100
            src.cr[8:, :] = 100
        rc, indep = run(vv, decide_fn, src, rec, 4, 8, luma=False)
        assert rc == 0 and fields(indep[0].cb) == NA and fields(indep[idx].cb) == (1, 0, 9, (0, 0, 0, 0, -1))
        assert fields(indep[idx].cr) == (1, 0, 9, (0, 0, 0, 0, -1))
        rc, out, m = merge(vv, fn, src, rec, indep, 4, 8)
        assert rc == 0 and m == ([0, 1] if w > h else [0, 2])
        assert fields(out[idx].cb) == NA and fields(out[idx].cr) == NA
        # with a source three levels low the band offset (-2) gains more
        # than the merge saves: 2 * 32 * 1 + 264 against 2 * 32 * 9 + 8
        src.cb[src.cb == 100] = 99
        src.cr[src.cr == 100] = 99
        rc, indep = run(vv, decide_fn, src, rec, 4, 8, luma=False)
        rc, out, m = merge(vv, fn, src, rec, indep, 4, 8)
        assert rc == 0 and m == [0, 0]


def merge_case_above_costs_two_bits(vv, fn, decide_fn):
    # picture 12x30 (chroma 6x15), CTB 16: the lower CTB's chroma region is
    # 6 x 7 = 42 samples per plane, two levels high: independent band offset
    # -1 gives 2 * 42 * 1 + 8 * 33 = 348, merging up (not applied)
    # 2 * 42 * 4 + 8 * 2 = 352: independent by 4 (a 1-bit merge-up would
    # merge: 344)
    src, rec = pic(vv, 12, 30, 8, 100, 102), pic(vv, 12, 30, 8, 100, 102)
    src.cb[8:, :] = 100
    src.cr[8:, :] = 100
    rc, indep = run(vv, decide_fn, src, rec, 4, 8, luma=False)
    assert rc == 0 and fields(indep[1].cb) == (1, 0, 9, (0, 0, 0, 0, -1)) and fields(indep[0].cb) == NA
    rc, out, m = merge(vv, fn, src, rec, indep, 4, 8)
    assert rc == 0 and m == [0, 0]


ALL_MERGE_CASES = {k: v for k, v in globals().items() if k.startswith("merge_case_")}
