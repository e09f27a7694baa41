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
