"""Hand-worked and closed-form pins of the oracle arms that the reference's own
unit tests leave open.

Every expected value here comes from the H.266 text or from arithmetic worked
by hand in the comments, never from the oracle's code:

- deblocking: the beta'/tC' tables (spec Table 43, typed here independently)
  and the chroma QP picks of compute_thresholds_chroma
  (src/deblock.rs:1144-1174) observed through tC-limited weak chroma filters;
- ALF luma classification (src/alf.rs:822-1010) on contents whose four
  Laplacian sums are known in closed form, including the virtual-boundary
  window, the activity scale ac = 3 and the virtual-boundary row padding;
- the ALF luma filter (src/alf.rs:652-724) tap by tap under every
  transposeIdx, with the tap positions taken from the spec's diamond and the
  transposes applied as geometric transforms (not as the permutation tables
  the oracle uses), at clip indices 0 and 2, with a class-dependent
  coefficient so that the class row is pinned too;
- Table 45's rows near the virtual boundary (src/alf.rs:1117-1130) and
  Table 46's (:1133-1143) through symmetric tap sets on random content;
- CC-ALF (src/alf.rs:447-539): each of the 7 taps on Cb and Cr, Table 47 with
  and without applyAlfLineBufBoundary, and the +-2^(BD-1) clip;
- the SAO and ALF tile-clip rectangles (src/sao.rs:426-450, src/alf.rs:
  1071-1096) in closed form and as the invariance "filtering a picture with
  tile bounds equals filtering each tile as a picture of its own".

Each check is a plain function of the loaded oracle so that
tests/test_oracle_mutants.py can run it against one-token mutants of the
oracle and show that it fails there.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import Pic, alf_cfg, alf_picture, binding_of, cu, cu_array, dparams, sao_cfg, sao_edge, sao_picture

# --------------------------------------------------------------------------- spec data, typed here
# H.266 Table 43.
BETA_PRIME = [0] * 16 + list(range(6, 19)) + list(range(20, 89, 2))
TC_PRIME = [0] * 18 + [3, 4, 4, 4, 4, 5, 5, 5, 5, 7, 7, 8, 9, 10, 10, 11, 13, 14, 15, 17, 19, 21, 24, 25, 29, 33,
                       36, 41, 45, 51, 57, 64, 71, 80, 89, 100, 112, 125, 141, 157, 177, 198, 222, 250, 280, 314,
                       352, 395]
# H.266 Table 8 at BitDepth 10: AlfClip[clipIdx].
ALF_CLIP_10 = [1024, 128, 32, 8]
# The 7x7 luma diamond (eq. 1449): tap k reads the sample pair +-(dx, dy).
LUMA_TAPS = [(0, 3), (1, 2), (0, 2), (-1, 2), (2, 1), (1, 1), (0, 1), (-1, 1), (-2, 1), (3, 0), (2, 0), (1, 0)]
# transposeIdx as geometry: 1 transposes the diamond (x <-> y), 2 mirrors it
# left-right, 3 rotates it by 90 degrees.  Eqs. 1442-1444 write these out as
# index permutations; the pairs are point-symmetric, so (dx, dy) and
# (-dx, -dy) name the same pair.
TRANSFORM = {0: lambda dx, dy: (dx, dy), 1: lambda dx, dy: (dy, dx), 2: lambda dx, dy: (-dx, dy),
             3: lambda dx, dy: (-dy, dx)}
# The 5x5 chroma diamond (eq. 1490).
CHROMA_TAPS = [(0, 2), (1, 1), (0, 1), (-1, 1), (2, 0), (1, 0)]


def table_45(y, ctb):
    """(alfShiftY, y1, y2, y3) of H.266 Table 45 with applyAlfLineBufBoundary = 1."""
    if y in (ctb - 5, ctb - 4):
        return 10, 0, 0, 0
    if y in (ctb - 6, ctb - 3):
        return 7, 1, 1, 1
    if y in (ctb - 7, ctb - 2):
        return 7, 1, 2, 2
    return 7, 1, 2, 3


def table_46(y, ctb_h_c):
    """(alfShiftC, y1, y2) of H.266 Table 46 with applyAlfLineBufBoundary = 1."""
    if y in (ctb_h_c - 2, ctb_h_c - 3):
        return 10, 0, 0
    if y in (ctb_h_c - 1, ctb_h_c - 4):
        return 7, 1, 1
    return 7, 1, 2


def table_47(y_line, ctb, lbb):
    """(yP1, yP2) of H.266 Table 47 as src/alf.rs:528-539 states it."""
    if lbb:
        if y_line in (ctb - 5, ctb - 4):
            return 0, 0
        if y_line in (ctb - 6, ctb - 3):
            return 1, 1
    return 1, 2


def at(a, x, y):
    """a[y][x] with coordinates clamped to the array (picture-edge replication)."""
    h, w = a.shape
    return a[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)]


def rshift(v, s):
    return np.right_shift(v, s)  # arithmetic on int64


# --------------------------------------------------------------------------- deblocking tables and chroma QP
def pin_deblock_tables(orc, vv):
    assert len(BETA_PRIME) == 64 and len(TC_PRIME) == 66
    got_b = [orc.oracle_beta_prime(q) for q in range(64)]
    got_t = [orc.oracle_tc_prime(q) for q in range(66)]
    assert got_b == BETA_PRIME
    assert got_t == TC_PRIME


def test_deblock_beta_tc_tables(orc, vv):
    pin_deblock_tables(orc, vv)


def _chroma_edge(orc, vv, p_cu, q_cu, params):
    """Two intra 16x16 CUs meeting at luma x = 16 (chroma x = 8, on the
    chroma 8-grid) over a chroma 100 | 900 step at 10 bits.  Both chroma
    blocks are 8 samples wide, so the strong chroma filter is eligible, but
    its decision fails on this step and the weak filter runs with
    raw delta = ((800 << 2) + 100 - 900 + 4) >> 3 = 300, which tC (< 300)
    clips: p0 = 100 + tC, q0 = 900 - tC.  Returns (tC of Cb, tC of Cr)."""
    pic = Pic(vv, 32, 16, 512, bd=10)
    for plane in (pic.cb, pic.cr):
        plane[:, :8], plane[:, 8:] = 100, 900
    arr = cu_array(vv, [p_cu, q_cu])
    orc.oracle_apply_deblocking_clipped(pic.ptr(), arr, 2, None, 0, None, C.byref(params), 1, None, 0, None, 0)
    out = []
    for plane in (pic.cb, pic.cr):
        tcs = set(int(v) for v in plane[:, 7] - 100) | set(int(v) for v in 900 - plane[:, 8].astype(np.int64))
        assert len(tcs) == 1, tcs
        assert (plane[:, :7] == 100).all() and (plane[:, 9:] == 900).all()
        out.append(tcs.pop())
    assert (pic.y == 512).all()
    return tuple(out)


def _ccu(vv, x, qp_y, qp_c=None, joint=False):
    c = cu(vv, x, 0, 16, 16, qp_y)
    if qp_c is not None:
        for i in range(3):
            c.qp_c[i] = qp_c[i]
    c.joint_cbcr2 = int(joint)
    return c


def pin_chroma_qp(orc, vv):
    # tC = tC'(QpC + 2 (bS - 1) + 2 tc_offset) at 10 bits (no scaling), bS = 2,
    # QpC = (QpP + QpQ + 1) >> 1 with each side's pick (src/deblock.rs:1155-1163).
    p = dparams(vv, bd=10, ctb_log2=7)
    p.luma_tc_offset_div2 = 5  # luma offsets never reach a chroma edge
    # explicit per-component QPs: Cb (30 + 32 + 1) >> 1 = 31 -> tC'(33) = 11;
    # Cr (40 + 42 + 1) >> 1 = 41 -> tC'(43) = 33
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 37, [30, 40, 50]), _ccu(vv, 16, 37, [32, 42, 52]), p) == (11, 33)
    # joint Cb-Cr (TuCResMode 2) on both sides: qp_c[2] for both components,
    # (50 + 52 + 1) >> 1 = 51 -> tC'(53) = 100
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 37, [30, 40, 50], True), _ccu(vv, 16, 37, [32, 42, 52], True),
                        p) == (100, 100)
    # joint on P only: Cb (50 + 32 + 1) >> 1 = 41 -> 33; Cr (50 + 42 + 1) >> 1 = 46 -> tC'(48) = 57
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 37, [30, 40, 50], True), _ccu(vv, 16, 37, [32, 42, 52]),
                        p) == (33, 57)
    # legacy arm on P (QpY 37 + PPS offset: Cb +5 -> 42, Cr -3 -> 34) against
    # explicit Q [20, 24, 28]: Cb (42 + 20 + 1) >> 1 = 31 -> 11;
    # Cr (34 + 24 + 1) >> 1 = 29 -> tC'(31) = 10
    p2 = dparams(vv, bd=10, ctb_log2=7)
    p2.chroma_qp_offset_cb, p2.chroma_qp_offset_cr = 5, -3
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 37), _ccu(vv, 16, 37, [20, 24, 28]), p2) == (11, 10)
    # legacy clamp to 63: QpY 60 + 6 = 66 -> 63 on both sides; Cb with
    # tc_offset_div2 -6: 63 + 2 - 12 = 53 -> 100 (unclamped 66 would give
    # tC'(56) = 141); Cr (offset 0): 60 + 2 = 62 -> 280
    p3 = dparams(vv, bd=10, ctb_log2=7)
    p3.chroma_qp_offset_cb = 6
    p3.cb_tc_offset_div2 = -6
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 60), _ccu(vv, 16, 60), p3) == (100, 280)
    # per-component tC offsets: QpC 30; Cb +2 -> 30 + 2 + 4 = 36 -> 15; Cr -1 -> 30 -> 9
    p4 = dparams(vv, bd=10, ctb_log2=7)
    p4.cb_tc_offset_div2, p4.cr_tc_offset_div2 = 2, -1
    assert _chroma_edge(orc, vv, _ccu(vv, 0, 37, [30, 30, 30]), _ccu(vv, 16, 37, [30, 30, 30]), p4) == (15, 9)


def test_deblock_chroma_qp_picks(orc, vv):
    pin_chroma_qp(orc, vv)


# --------------------------------------------------------------------------- ALF luma classification
# Content of one 4x4 block's reach (x4-3 .. x4+6, y4-3 .. y4+6):
#   P = K + alpha u^2 + beta v^2 + gamma u v + a (-1)^x + b (-1)^y,  u = x - x4 - 1, v = y - y4 - 1.
# Its 1-D Laplacians (2c - n1 - n2) are  H: -2 alpha + 4a (-1)^x,  V: -2 beta + 4b (-1)^y,
# D0: -2 (alpha + beta + gamma) + 4a (-1)^x + 4b (-1)^y,  D1: the same with -gamma.
# The checkerboard positions of the 8x8 window are 16 with x, y even and 16
# with both odd, and |p - q| + |p + q| = 2 max(|p|, |q|), so with
#   mH = max(|alpha|, 2|a|), mV = max(|beta|, 2|b|),
#   m0 = max(|alpha + beta + gamma|, 2|a + b|), m1 = max(|alpha + beta - gamma|, 2|a + b|)
# the sums are sumH = 64 mH, sumV = 64 mV, sumD0 = 64 m0, sumD1 = 64 m1, and at
# 10 bits the activity index is (2 (sumH + sumV)) >> 9 = (mH + mV) >> 2.
# Each row: (alpha, beta, gamma, a, b) -> (filtIdx, transposeIdx), worked below.
LADDER = [
    # flat: all sums 0; sumV > sumH fails -> dirHV 3; sumD0 > sumD1 fails -> dirD 2;
    # dirS 0; transposeIdx table[3*2 + 1] = 3
    ((0, 0, 0, 0, 0), (0, 3)),
    # mH 1, mV 3, m0 6, m1 2: act 1 -> 1; dirHV 1, dirD 0; d1 hv0 = 6 = hv1 d0
    # (not greater: H/V wins); hvd 3/1: 3 > 2 -> dirS 1; filtIdx 1 + (2 + 1) 5 = 16; t 0
    ((1, 3, 2, 0, 0), (16, 0)),
    # mirror of the above in H/V: dirHV 3, dirD 0 -> t 1
    ((3, 1, 2, 0, 0), (16, 1)),
    # m0 2, m1 6: dirD 2, dirHV 1 -> t 2
    ((1, 3, -2, 0, 0), (16, 2)),
    ((3, 1, -2, 0, 0), (16, 3)),
    # mH = mV = 2 (tie -> dirHV 3), m0 7, m1 1 (dirD 0): 7*2 > 2*1 -> diagonal
    # wins; hvd 7/1: 14 > 9 -> dirS 2; dir1 = dirD = 0 -> filtIdx 1 + (0 + 2) 5 = 11;
    # t = table[0*2 + (3 >> 1)] = 1
    ((2, 2, 3, 0, 0), (11, 1)),
    # mH 2, mV 3 (dirHV 1), m0 1, m1 9 (dirD 2): 9*2 > 3*1; dirS 2; act 5 >> 2 = 1;
    # filtIdx 11; t = table[2*2 + 0] = 2
    ((2, 3, -4, 0, 0), (11, 2)),
    # mH 2, mV 3, m0 9, m1 1: diagonal wins with dirD 0 -> t = table[0] = 0
    ((2, 3, 4, 0, 0), (11, 0)),
    # mH = mV = 2, m0 5, m1 3: 5*2 > 2*3 -> diagonal; hvd 5/3: 10 > 27 no, 5 > 6 no
    # -> dirS 0; filtIdx 1; t = table[0*2 + 1] = 1
    ((2, 2, 1, 0, 0), (1, 1)),
    # a = 1: mH max(1, 2) = 2, mV 1, m0 max(5, 2) = 5, m1 max(1, 2) = 2: act 3 >> 2 = 0;
    # dirHV 3 (hv 2/1), dirD 0 (d 5/2): 5*1 > 2*2 -> diagonal; hvd 5/2: 10 > 18 no,
    # 5 > 4 -> dirS 1; filtIdx 0 + (0 + 1) 5 = 5; t = table[1] = 1
    ((1, 1, 3, 1, 0), (5, 1)),
    # mH 9, mV 2, m0 = m1 = 11: act 11 >> 2 = 2 -> 2; dirHV 3, dirD 2 (tie);
    # 11*2 > 9*11 no; hvd 9/2: 18 > 18 no (boundary), 9 > 4 -> dirS 1;
    # filtIdx 2 + (2 + 1) 5 = 17; t 3
    ((9, 2, 0, 0, 0), (17, 3)),
    # mH 10, mV 2, m0 = m1 = 12: act 3 -> 2; hvd 10/2: 20 > 18 -> dirS 2; filtIdx 2 + 20 = 22
    ((10, 2, 0, 0, 0), (22, 3)),
    # mH 4, mV 2: hvd 4/2: 4 > 4 no (boundary) -> dirS 0; act 6 >> 2 = 1; filtIdx 1
    ((4, 2, 0, 0, 0), (1, 3)),
    # mH 5, mV 2: 5 > 4 -> dirS 1; act 7 >> 2 = 1; filtIdx 1 + 15 = 16
    ((5, 2, 0, 0, 0), (16, 3)),
    # mH 1, mV 2 (dirHV 1), m0 = m1 = 3 (dirD 2): 3*1 > 2*3 no; hvd 2/1: 2 > 2 no
    # -> dirS 0; act 3 >> 2 = 0; filtIdx 0; t = table[1*2 + 1] = 2
    ((1, 2, 0, 0, 0), (0, 2)),
    # a = b = n: mH = mV = 2n, m0 = m1 = 4n; ties on both axes (dirHV 3, dirD 2),
    # 4n*2n > 2n*4n no, hvd 2n/2n -> dirS 0; act n -> varTab[n]; t 3
    ((0, 0, 0, 1, 1), (1, 3)),
    ((0, 0, 0, 2, 2), (2, 3)),
    ((0, 0, 0, 6, 6), (2, 3)),
    ((0, 0, 0, 7, 7), (3, 3)),
    ((0, 0, 0, 14, 14), (3, 3)),
    ((0, 0, 0, 15, 15), (4, 3)),
    ((0, 0, 0, 40, 40), (4, 3)),  # act 40 clamps to 15
    # a = n, b = 0: mH 2n, mV 0, m0 = m1 = 2n: dirHV 3 (hv 2n/0), dirD 2 (tie),
    # 2n*0 > 2n*2n no; hvd 2n/0 -> dirS 2; filtIdx varTab[2n >> 2] + (2 + 2) 5; t 3
    ((0, 0, 0, 30, 0), (24, 3)),  # act 15 -> 4 -> 24
    ((0, 0, 0, 8, 0), (22, 3)),   # act 4 -> 2 -> 22
    # a = 0, b = n: dirHV 1 (hv 2n/0): t = table[1*2 + 1] = 2; filtIdx as above
    ((0, 0, 0, 0, 30), (24, 2)),
]
# transposeIdx of each ladder row is fixed by (dirHV, dirD) alone:
# (1, 0) -> 0, (3, 0) -> 1, (1, 2) -> 2, (3, 2) -> 3, on either branch.
SLOTS = [4, 16, 28, 40, 52]  # block origins whose reach (o-3 .. o+6) stays inside a 64-sample CTB, away
                             # from the picture edge and from the virtual-boundary rows (>= 56)


def block_pattern(params, x4, y4, bd=10):
    alpha, beta, gamma, a, b = params
    ys, xs = np.mgrid[y4 - 3:y4 + 7, x4 - 3:x4 + 7]
    u, v = xs - x4 - 1, ys - y4 - 1
    px, py = np.where(xs % 2 == 0, 1, -1), np.where(ys % 2 == 0, 1, -1)
    p = alpha * u * u + beta * v * v + gamma * u * v + a * px + b * py
    k = (1 << (bd - 1)) - (int(p.min()) + int(p.max())) // 2
    p = p + k
    assert p.min() >= 0 and p.max() < (1 << bd), params
    return p


def ladder_picture(cases, bd=10):
    """A 128x64 picture (two 64x64 CTBs) holding one ladder case per block
    slot; returns (luma, [(x4, y4, expected filtIdx, transposeIdx)])."""
    W, H = 128, 64
    y = np.full((H, W), 1 << (bd - 1), np.int64)
    placed = []
    slots = [(cx * 64 + sx, sy) for cx in range(2) for sy in SLOTS for sx in SLOTS]
    assert len(cases) <= len(slots)
    for (params, want), (x4, y4) in zip(cases, slots):
        y[y4 - 3:y4 + 7, x4 - 3:x4 + 7] = block_pattern(params, x4, y4, bd)
        placed.append((x4, y4) + tuple(want))
    return y, placed


def classify(orc, plane, ctb, bd, rx=0, ry=0):
    h, w = plane.shape
    cls = (C.c_uint16 * ((ctb // 4) ** 2))()
    clip = (C.c_int32 * 4)(0, 0, w, h)
    arr = np.ascontiguousarray(plane, np.uint16)
    orc.oracle_alf_classify_ctb(arr.ctypes.data_as(C.POINTER(C.c_uint16)), w, w, h, clip, rx, ry, ctb, bd, cls)
    g = np.array(cls).reshape(ctb // 4, ctb // 4)
    return g & 0xff, g >> 8


def pin_alf_classification_ladder(orc, vv):
    y, placed = ladder_picture(LADDER)
    grids = [classify(orc, y, 64, 10, rx=rx) for rx in range(2)]
    for x4, y4, f, t in placed:
        fg, tg = grids[x4 // 64]
        got = (int(fg[y4 // 4, (x4 % 64) // 4]), int(tg[y4 // 4, (x4 % 64) // 4]))
        assert got == (f, t), (x4, y4, got, (f, t))


def test_alf_classification_ladder(orc, vv):
    pin_alf_classification_ladder(orc, vv)


def pin_alf_classification_bit_depth(orc, vv):
    # The activity shift is BitDepth - 1: a = b = n parity content gives
    # sumH = sumV = 128 n and index (2 * 256 n) >> (BD - 1).
    # 8 bits, n = 1: 512 >> 7 = 4 -> varTab 2;  12 bits, n = 4: 2048 >> 11 = 1 -> 1;
    # 12 bits, n = 8: 4096 >> 11 = 2 -> 2.  Directions as in the ladder (t 3).
    for bd, n, want in ((8, 1, 2), (12, 4, 1), (12, 8, 2), (10, 1, 1)):
        y, placed = ladder_picture([((0, 0, 0, n, n), (want, 3))], bd=bd)
        f, t = classify(orc, y, 64, bd)
        x4, y4 = placed[0][:2]
        assert (int(f[y4 // 4, x4 // 4]), int(t[y4 // 4, x4 // 4])) == (want, 3), (bd, n)


def test_alf_classification_bit_depth(orc, vv):
    pin_alf_classification_bit_depth(orc, vv)


def pin_alf_classification_virtual_boundary(orc, vv):
    # Content constant along y, P = 512 + a (-1)^x: sumV = 0 and every other
    # Laplacian is 4|a| per checkerboard position, so sumH = sumD0 = sumD1 = 4|a| n
    # over a window of n positions: dirHV 3, dirD 2, dirS 2, t 3, filtIdx
    # varTab[idx] + 20.  Interior blocks: n = 32, ac = 2: idx = 256|a| >> 9.
    # Blocks at y4 = CtbSizeY - 8 and - 4 (eq. 1458-1462 carve-outs): six
    # window rows, n = 24, ac = 3: idx = 288|a| >> 9.
    #   a = 10: interior 2560 >> 9 = 5 -> 22; boundary 2880 >> 9 = 5 -> 22
    #   a = 13: interior 3328 >> 9 = 6 -> 22; boundary 3744 >> 9 = 7 -> 23
    # (ac = 2 at the boundary would give 4 -> 22 for a = 13; a full window with
    # ac = 3 would give 7 -> 23 for a = 10.)
    for a, want_vb in ((10, 22), (13, 23)):
        xs = np.arange(64)
        y = np.tile(512 + a * np.where(xs % 2 == 0, 1, -1), (64, 1))
        f, t = classify(orc, y, 64, 10)
        inner = f[:, 1:15]
        assert (inner[:14] == 22).all() and (t[:, 1:15] == 3).all(), a
        assert (inner[14:] == want_vb).all(), (a, inner[14:])
    # The last CTB row of a picture whose height past yCtb is <= CtbSizeY - 4
    # has no carve-out: a 64 + 58 row picture, block y4 = 56 of CTB row 1 uses
    # the full window (reads past the picture's last row replicate it, and the
    # content is constant along y) with ac = 2: a = 13 -> 22.
    xs = np.arange(64)
    y = np.tile(512 + 13 * np.where(xs % 2 == 0, 1, -1), (122, 1))
    f, _ = classify(orc, y, 64, 10, ry=1)
    assert (f[14, 1:15] == 22).all(), f[14]


def test_alf_classification_virtual_boundary(orc, vv):
    pin_alf_classification_virtual_boundary(orc, vv)


def pin_alf_classification_vb_padding(orc, vv):
    # Section 8.8.5.5: a block at y4 = CtbSizeY - 8 reads no row at or below the
    # virtual boundary (yCtb + CtbSizeY - 4; those rows are replaced by the row
    # above it), and a block at y4 = CtbSizeY - 4 reads no row above it.  With
    # noise only on the far side of the boundary, both blocks see a flat
    # window: filtIdx 0, transposeIdx 3 (as the flat ladder row).
    rng = np.random.default_rng(5)
    y = np.full((64, 64), 300)
    y[60:] = rng.integers(0, 1024, (4, 64))
    f, t = classify(orc, y, 64, 10)
    assert (f[14, 1:15] == 0).all() and (t[14, 1:15] == 3).all(), f[14]
    assert (f[13, 1:15] == 0).all()  # y4 = 52 reads rows 49 .. 58 only
    y = np.full((64, 64), 300)
    y[:60] = rng.integers(0, 1024, (60, 64))
    f, t = classify(orc, y, 64, 10)
    assert (f[15, 1:15] == 0).all() and (t[15, 1:15] == 3).all(), f[15]


def test_alf_classification_vb_padding(orc, vv):
    pin_alf_classification_vb_padding(orc, vv)


# --------------------------------------------------------------------------- ALF luma filter
def class_coeff(c):
    """A coefficient per class (nonzero, distinct, both signs)."""
    return (20 + 3 * c) * (1 if c % 2 == 0 else -1)


def luma_aps(vv, taps, clip_idx, coeff_of_class=class_coeff):
    aps = vv.AlfAps()
    aps.alf_luma_filter_signal_flag = 1
    aps.alf_luma_clip_flag = 1
    aps.n_luma_coeff = 25
    for c in range(25):
        for k in taps:
            aps.luma_coeff[c][k] = coeff_of_class(c)
            aps.luma_clip_idx[c][k] = clip_idx
    return aps


def run_luma_alf(orc, vv, luma, aps, ctb_log2=6, bd=10, tiles=None):
    H, W = luma.shape
    pic = Pic(vv, W, H, 0, bd=bd)
    pic.y[:] = luma
    ctb = 1 << ctb_log2
    wc, hc = -(-W // ctb), -(-H // ctb)
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, 16
    ap, _k = alf_picture(vv, wc, hc, {(x, y): c for x in range(wc) for y in range(hc)})
    b = binding_of(vv, luma=[aps])
    assert orc.oracle_apply_alf_clipped(pic.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                        C.byref(b), tiles) == 0
    return pic.y.astype(np.int64)


def single_tap_expect(P, x, y, f, d, lim, shift=7, bd=10):
    s = P[y, x]
    dx, dy = d
    acc = np.clip(at(P, x + dx, y + dy) - s, -lim, lim) + np.clip(at(P, x - dx, y - dy) - s, -lim, lim)
    return np.clip(s + rshift(f * acc + (1 << (shift - 1)), shift), 0, (1 << bd) - 1)


def pin_alf_luma_taps_by_transpose(orc, vv):
    # One nonzero tap k (the same in all 25 classes, scaled per class by
    # class_coeff) reduces eq. 1449 to out = s + ((f (clip(a - s) + clip(b - s))
    # + 64) >> 7) for the pair a, b at +-T_t(dx_k, dy_k), where t is the block's
    # transposeIdx; the clip index moves with its coefficient.  The ladder
    # blocks fix (filtIdx, t) per block by hand.
    y, placed = ladder_picture(LADDER)
    yy, xx = np.mgrid[0:4, 0:4]
    for k in range(12):
        for clip_idx in (0, 2):
            out = run_luma_alf(orc, vv, y, luma_aps(vv, [k], clip_idx))
            discriminating = False
            for x4, y4, f, t in placed:
                X, Y = xx + x4, yy + y4
                d = TRANSFORM[t](*LUMA_TAPS[k])
                want = single_tap_expect(y, X, Y, class_coeff(f), d, ALF_CLIP_10[clip_idx])
                assert (out[Y, X] == want).all(), (k, clip_idx, x4, y4, f, t)
                for t2 in range(4):  # the pin must tell transposes apart
                    d2 = TRANSFORM[t2](*LUMA_TAPS[k])
                    if (d2 != d and d2 != (-d[0], -d[1]) and
                            (single_tap_expect(y, X, Y, class_coeff(f), d2, ALF_CLIP_10[clip_idx]) != want).any()):
                        discriminating = True
            assert discriminating, k


def test_alf_luma_taps_by_transpose(orc, vv):
    pin_alf_luma_taps_by_transpose(orc, vv)


# Tap sets that every transposeIdx maps onto themselves: with one coefficient
# and clip index on the whole set the filter no longer depends on the class.
TAP_ORBITS = [(0, 9), (2, 10), (6, 11), (5, 7), (1, 3, 4, 8)]


def pin_alf_luma_table45_rows(orc, vv):
    # Random content over two CTB rows (CTB 64) at 10 bits: every row of each
    # CTB, including Table 45's rows CtbSizeY - 7 .. CtbSizeY - 2, against
    # out = s + ((f sum_k (clip(a_k - s) + clip(b_k - s)) + 2^(sh-1)) >> sh),
    # where the pair offsets (dx, dy) take |dy| = 3, 2, 1 -> y3, y2, y1 and sh
    # from Table 45, and reads replicate at the picture edges.
    rng = np.random.default_rng(45)
    H, W = 128, 48
    y = rng.integers(0, 1024, (H, W)).astype(np.int64)
    ys, xs = np.mgrid[0:H, 0:W]
    for orbit in TAP_ORBITS:
        for clip_idx, f in ((0, 23), (2, -41)):
            out = run_luma_alf(orc, vv, y, luma_aps(vv, orbit, clip_idx, lambda c: f))
            for row in range(H):
                sh, y1, y2, y3 = table_45(row % 64, 64)
                ymap = {0: 0, 1: y1, 2: y2, 3: y3}
                s = y[row]
                X = xs[row]
                acc = np.zeros(W, np.int64)
                lim = ALF_CLIP_10[clip_idx]
                for k in orbit:
                    dx, dy = LUMA_TAPS[k]
                    dyr = ymap[abs(dy)] * (1 if dy >= 0 else -1)
                    acc += np.clip(at(y, X + dx, row + dyr) - s, -lim, lim)
                    acc += np.clip(at(y, X - dx, row - dyr) - s, -lim, lim)
                want = np.clip(s + rshift(f * acc + (1 << (sh - 1)), sh), 0, 1023)
                assert (out[row] == want).all(), (orbit, clip_idx, row)


def test_alf_luma_table45_rows(orc, vv):
    pin_alf_luma_table45_rows(orc, vv)


# --------------------------------------------------------------------------- chroma ALF, all rows
def pin_alf_chroma_table46_rows(orc, vv):
    # One chroma tap per component (Cb alternative 0, Cr alternative 1) on
    # random content, every row of two chroma CTB rows (CTB 64 -> 32 chroma
    # rows), with Table 46's offsets and shift: out = s + ((f (clip(a - s) +
    # clip(b - s)) + 2^(sh-1)) >> sh), |dy| = 2, 1 -> y2, y1.
    rng = np.random.default_rng(46)
    W, H = 96, 128
    for k in range(6):
        pic = Pic(vv, W, H, 0, bd=10, chroma=0)
        pic.cb[:] = rng.integers(0, 1024, pic.cb.shape)
        pic.cr[:] = rng.integers(0, 1024, pic.cr.shape)
        pre = {"cb": pic.cb.astype(np.int64), "cr": pic.cr.astype(np.int64)}
        aps = vv.AlfAps()
        aps.alf_chroma_filter_signal_flag = 1
        aps.alf_chroma_clip_flag = 1
        aps.alf_chroma_num_alt_filters_minus1 = 1
        aps.n_chroma_coeff = 2
        plan = {"cb": (0, k, 45, 2), "cr": (1, (k + 3) % 6, -38, 1)}
        for alt, kk, f, ci in plan.values():
            aps.chroma_coeff[alt][kk] = f
            aps.chroma_clip_idx[alt][kk] = ci
        ctb = vv.AlfCtb()
        ctb.cb_on = ctb.cr_on = 1
        ctb.cb_alt_idx, ctb.cr_alt_idx = 0, 1
        ap, _k = alf_picture(vv, 2, 2, {(x, y): ctb for x in range(2) for y in range(2)})
        b = binding_of(vv, chroma=aps)
        orc.oracle_apply_alf_clipped(pic.ptr(), C.byref(ap),
                                     C.byref(alf_cfg(vv, bd=10, ctb_log2=6, cb=True, cr=True)), C.byref(b), None)
        for name, (alt, kk, f, ci) in plan.items():
            s_all = pre[name]
            h, w = s_all.shape
            lim = ALF_CLIP_10[ci]
            xs = np.arange(w)
            got = getattr(pic, name).astype(np.int64)
            for row in range(h):
                sh, y1, y2 = table_46(row % 32, 32)
                dx, dy = CHROMA_TAPS[kk]
                dyr = {0: 0, 1: y1, 2: y2}[dy]
                s = s_all[row]
                acc = np.clip(at(s_all, xs + dx, row + dyr) - s, -lim, lim) + \
                    np.clip(at(s_all, xs - dx, row - dyr) - s, -lim, lim)
                want = np.clip(s + rshift(f * acc + (1 << (sh - 1)), sh), 0, 1023)
                assert (got[row] == want).all(), (name, kk, row)


def test_alf_chroma_table46_rows(orc, vv):
    pin_alf_chroma_table46_rows(orc, vv)


# --------------------------------------------------------------------------- CC-ALF
# Eq. 1515: the 7 taps around the co-located luma sample (2x, 2y).
CC_TAPS = [lambda p1, p2: (0, -p1), lambda p1, p2: (-1, 0), lambda p1, p2: (1, 0), lambda p1, p2: (-1, p1),
           lambda p1, p2: (0, p1), lambda p1, p2: (1, p1), lambda p1, p2: (0, p2)]


def cc_expect(L, ch_pre, coeffs, ph, ctb=64, bd=10):
    """CC-ALF of a whole chroma plane in closed form (4:2:0)."""
    h, w = ch_pre.shape
    out = ch_pre.copy()
    half = 1 << (bd - 1)
    xs = np.arange(w)
    for y in range(h):
        y_ctb_c = (y // (ctb // 2)) * (ctb // 2)
        yl = 2 * y
        lbb = not (yl + ctb >= ph and ph - 2 * y_ctb_c <= ctb - 4)
        p1, p2 = table_47(2 * (y - y_ctb_c), ctb, lbb)
        cen = at(L, 2 * xs, yl)
        acc = np.zeros(w, np.int64)
        for k, f in enumerate(coeffs):
            if f:
                dx, dy = CC_TAPS[k](p1, p2)
                acc += f * (at(L, 2 * xs + dx, yl + dy) - cen)
        out[y] = np.clip(ch_pre[y] + np.clip(rshift(acc + 64, 7), -half, half - 1), 0, (1 << bd) - 1)
    return out


def run_cc(orc, vv, L, cb, cr, cb_coeffs, cr_coeffs):
    H, W = L.shape
    pic = Pic(vv, W, H, 0, bd=10, chroma=0)
    pic.y[:], pic.cb[:], pic.cr[:] = L, cb, cr
    aps_cb, aps_cr = vv.AlfAps(), vv.AlfAps()
    aps_cb.n_cc_cb_coeff = 2
    aps_cr.n_cc_cr_coeff = 3
    for k in range(7):
        aps_cb.cc_cb_coeff[1][k] = cb_coeffs[k]
        aps_cr.cc_cr_coeff[2][k] = cr_coeffs[k]
    ctb = vv.AlfCtb()
    ctb.cc_cb_idc, ctb.cc_cr_idc = 2, 3
    wc, hc = -(-W // 64), -(-H // 64)
    ap, _k = alf_picture(vv, wc, hc, {(x, y): ctb for x in range(wc) for y in range(hc)})
    b = binding_of(vv, cc_cb=aps_cb, cc_cr=aps_cr)
    assert orc.oracle_apply_alf_clipped(pic.ptr(), C.byref(ap),
                                        C.byref(alf_cfg(vv, bd=10, ctb_log2=6, cc_cb=True, cc_cr=True)),
                                        C.byref(b), None) == 0
    assert (pic.y == L).all()
    return pic.cb.astype(np.int64), pic.cr.astype(np.int64)


def pin_cc_alf_taps(orc, vv):
    # Each tap alone on Cb (filter 1 of its APS) and, shifted by three, on Cr
    # (filter 2), over pictures 128 rows high (Table 47 with
    # applyAlfLineBufBoundary in both CTB rows), 88 rows (last CTB row 24 rows
    # high: no line-buffer boundary there) and 124 rows (last CTB row exactly
    # CtbSizeY - 4 high: none either).
    rng = np.random.default_rng(47)
    for H in (128, 88, 124):
        W = 80
        L = rng.integers(0, 1024, (H, W)).astype(np.int64)
        cb = rng.integers(0, 1024, (H // 2, W // 2)).astype(np.int64)
        cr = rng.integers(0, 1024, (H // 2, W // 2)).astype(np.int64)
        for k in range(7):
            fcb = [0] * 7
            fcr = [0] * 7
            fcb[k] = 37 if k % 2 else -53
            fcr[(k + 3) % 7] = -29 if k % 2 else 61
            got_cb, got_cr = run_cc(orc, vv, L, cb, cr, fcb, fcr)
            assert (got_cb == cc_expect(L, cb, fcb, H)).all(), (H, k, "cb")
            assert (got_cr == cc_expect(L, cr, fcr, H)).all(), (H, k, "cr")


def test_cc_alf_taps(orc, vv):
    pin_cc_alf_taps(orc, vv)


def pin_cc_alf_clip(orc, vv):
    # All seven taps at +64 (Cb) and -64 (Cr) on a luma checkerboard of 0 and
    # 1023: around a 0 centre the sum reaches 64 * 1023 * (number of 1023
    # neighbours), so (sum + 64) >> 7 exceeds 511 and is clipped to 511 (Cb) /
    # -512 (Cr) at 10 bits; a 1023 centre gives the mirror image.
    H, W = 128, 64
    yy, xx = np.mgrid[0:H, 0:W]
    L = np.where((xx + yy) % 2 == 0, 0, 1023).astype(np.int64)
    cb = np.full((H // 2, W // 2), 512, np.int64)
    cr = np.full((H // 2, W // 2), 512, np.int64)
    got_cb, got_cr = run_cc(orc, vv, L, cb, cr, [64] * 7, [-64] * 7)
    want_cb, want_cr = cc_expect(L, cb, [64] * 7, H), cc_expect(L, cr, [-64] * 7, H)
    assert (got_cb == want_cb).all() and (got_cr == want_cr).all()
    # the co-located luma of every chroma sample is 0 here ((2x + 2y) even)
    assert (want_cb == 1023).sum() > 0 and (want_cr == 0).sum() > 0  # 512 + 511, 512 - 512
    assert set(np.unique(want_cb)) <= {512, 512 + 511, 1023} and 1023 in set(np.unique(want_cb))


def test_cc_alf_clip(orc, vv):
    pin_cc_alf_clip(orc, vv)


# --------------------------------------------------------------------------- tile clips
def bounds(vv, cols, rows):
    ca = (C.c_uint32 * len(cols))(*cols)
    ra = (C.c_uint32 * len(rows))(*rows)
    b = vv.Bounds(C.cast(ca, C.POINTER(C.c_uint32)), len(cols), C.cast(ra, C.POINTER(C.c_uint32)), len(rows))
    b._keep = (ca, ra)
    return b


def pin_sao_tile_edges(orc, vv):
    # Edge offset on an alternating pattern: every sample is a strict local
    # peak (p = 520) or valley (480) along the class's direction, so
    # edgeIdx = 4 (offset_val[4] = -6) or 0 -> 1 (offset_val[1] = +7), except
    # where a neighbour lies outside the picture or (with tile bounds) outside
    # the CTB's tile: there the sample is copied (eq. 1429 and its tile rule).
    # Picture 64x64, CTB 16, tile columns at 0 | 32 | 64, rows at 0 | 16 | 64;
    # chroma edges at half those positions.
    for cls in range(4):
        for tiled in (False, True):
            pic = Pic(vv, 64, 64, 0, bd=10, chroma=0)
            for plane in (pic.y, pic.cb, pic.cr):
                h, w = plane.shape
                yy, xx = np.mgrid[0:h, 0:w]
                alt = yy if cls == 1 else xx
                plane[:] = np.where(alt % 2 == 0, 520, 480)
            e = sao_edge(vv, cls, [7, 5, 4, 6], bd=10)
            prm = vv.SaoCtbParams(e, e, e)
            sp, _k = sao_picture(vv, 4, 4, {(x, y): prm for x in range(4) for y in range(4)})
            cfg = sao_cfg(vv, luma=True, chroma=True, bd=10, ctb_log2=4)
            tb = bounds(vv, [0, 32, 64], [0, 16, 64]) if tiled else None
            pre = [p.copy() for p in (pic.y, pic.cb, pic.cr)]
            orc.oracle_apply_sao_clipped(pic.ptr(), C.byref(sp), C.byref(cfg), C.byref(tb) if tb else None)
            for plane, before, sub in zip((pic.y, pic.cb, pic.cr), pre, (1, 2, 2)):
                h, w = plane.shape
                ex_x = {0, w - 1} | ({32 // sub - 1, 32 // sub} if tiled else set())
                ex_y = {0, h - 1} | ({16 // sub - 1, 16 // sub} if tiled else set())
                moved = np.where(before == 520, 514, 487)
                for yv in range(h):
                    for xv in range(w):
                        copy = (cls != 1 and xv in ex_x) or (cls != 0 and yv in ex_y)
                        want = before[yv, xv] if copy else moved[yv, xv]
                        assert plane[yv, xv] == want, (cls, tiled, sub, xv, yv)


def test_sao_tile_edges(orc, vv):
    pin_sao_tile_edges(orc, vv)


def random_alf_setup(vv, rng, wc, hc):
    aps = vv.AlfAps()
    aps.alf_luma_filter_signal_flag = aps.alf_chroma_filter_signal_flag = 1
    aps.alf_luma_clip_flag = aps.alf_chroma_clip_flag = 1
    aps.n_luma_coeff = 25
    aps.alf_chroma_num_alt_filters_minus1 = 1
    aps.n_chroma_coeff = 2
    aps.n_cc_cb_coeff = aps.n_cc_cr_coeff = 2
    for c in range(25):
        for k in range(12):
            aps.luma_coeff[c][k] = int(rng.integers(-64, 65))
            aps.luma_clip_idx[c][k] = int(rng.integers(0, 4))
    for a in range(2):
        for k in range(6):
            aps.chroma_coeff[a][k] = int(rng.integers(-64, 65))
            aps.chroma_clip_idx[a][k] = int(rng.integers(0, 4))
        for k in range(7):
            aps.cc_cb_coeff[a][k] = int(rng.choice([0, 1, -1, 2, -2, 4, -4, 8, -8, 16, -16, 32, -32, 64, -64]))
            aps.cc_cr_coeff[a][k] = int(rng.choice([0, 1, -1, 2, -2, 4, -4, 8, -8, 16, -16, 32, -32, 64, -64]))
    ctbs = {}
    for y in range(hc):
        for x in range(wc):
            c = vv.AlfCtb()
            c.luma_on, c.cb_on, c.cr_on = 1, 1, 1
            c.luma_filt_set_idx = int(rng.choice([16, 3]))
            c.cb_alt_idx, c.cr_alt_idx = int(rng.integers(0, 2)), int(rng.integers(0, 2))
            c.cc_cb_idc, c.cc_cr_idc = int(rng.integers(0, 3)), int(rng.integers(0, 3))
            ctbs[(x, y)] = c
    return aps, ctbs


def pin_tile_clip_equals_separate_pictures(orc, vv):
    # With loop filtering across tiles off, ALF reads replicate at the tile's
    # edges exactly as at the picture's (src/alf.rs:1071-1109), and SAO edge
    # offset treats a neighbour in another tile like one outside the picture
    # (src/sao.rs:426-450): filtering a picture with tile bounds must equal
    # filtering each tile as a picture of its own.  Tiles: columns 0 | 128 | 192,
    # rows 0 | 64 | 128, CTB 64, random samples and side information.
    rng = np.random.default_rng(1071)
    W, H, ctb = 192, 128, 64
    wc, hc = W // ctb, H // ctb
    cols, rows = [0, 128, 192], [0, 64, 128]
    aps, ctbs = random_alf_setup(vv, rng, wc, hc)
    sao = {}
    for y in range(hc):
        for x in range(wc):
            comp = []
            for _ in range(3):
                if rng.integers(0, 2):
                    comp.append(sao_edge(vv, int(rng.integers(0, 4)), [int(v) for v in rng.integers(0, 16, 4)], 10))
                else:
                    s = vv.SaoCtb()
                    s.sao_type_idx = 1
                    s.band_position = int(rng.integers(0, 32))
                    for i in range(4):
                        s.offset_val[i + 1] = int(rng.integers(-15, 16))
                    comp.append(s)
            sao[(x, y)] = vv.SaoCtbParams(*comp)
    full = Pic(vv, W, H, 0, bd=10)
    full.y[:] = rng.integers(0, 1024, (H, W))
    full.cb[:] = rng.integers(0, 1024, (H // 2, W // 2))
    full.cr[:] = rng.integers(0, 1024, (H // 2, W // 2))
    src = full.copy()

    def filt(pic, wcc, hcc, ctb_map, tb):
        sp, _k1 = sao_picture(vv, wcc, hcc, {k: sao[v] for k, v in ctb_map.items()})
        orc.oracle_apply_sao_clipped(pic.ptr(), C.byref(sp),
                                     C.byref(sao_cfg(vv, luma=True, chroma=True, bd=10, ctb_log2=6)),
                                     C.byref(tb) if tb else None)
        ap, _k2 = alf_picture(vv, wcc, hcc, {k: ctbs[v] for k, v in ctb_map.items()})
        b = binding_of(vv, luma=[aps], chroma=aps, cc_cb=aps, cc_cr=aps)
        cfg = alf_cfg(vv, bd=10, ctb_log2=6, cb=True, cr=True, cc_cb=True, cc_cr=True)
        assert orc.oracle_apply_alf_clipped(pic.ptr(), C.byref(ap), C.byref(cfg), C.byref(b),
                                            C.byref(tb) if tb else None) == 0

    filt(full, wc, hc, {(x, y): (x, y) for x in range(wc) for y in range(hc)}, bounds(vv, cols, rows))
    untiled = src.copy()
    filt(untiled, wc, hc, {(x, y): (x, y) for x in range(wc) for y in range(hc)}, None)
    assert (untiled.y != full.y).any()  # the tile edges matter for this content
    for x0, x1 in zip(cols, cols[1:]):
        for y0, y1 in zip(rows, rows[1:]):
            t = Pic(vv, x1 - x0, y1 - y0, 0, bd=10)
            t.y[:] = src.y[y0:y1, x0:x1]
            t.cb[:] = src.cb[y0 // 2:y1 // 2, x0 // 2:x1 // 2]
            t.cr[:] = src.cr[y0 // 2:y1 // 2, x0 // 2:x1 // 2]
            twc, thc = (x1 - x0) // ctb, (y1 - y0) // ctb
            filt(t, twc, thc, {(x, y): (x + x0 // ctb, y + y0 // ctb) for x in range(twc) for y in range(thc)},
                 None)
            assert (t.y == full.y[y0:y1, x0:x1]).all(), (x0, y0, "luma")
            assert (t.cb == full.cb[y0 // 2:y1 // 2, x0 // 2:x1 // 2]).all(), (x0, y0, "cb")
            assert (t.cr == full.cr[y0 // 2:y1 // 2, x0 // 2:x1 // 2]).all(), (x0, y0, "cr")


def test_tile_clip_equals_separate_pictures(orc, vv):
    pin_tile_clip_equals_separate_pictures(orc, vv)


ALL_PINS = {name: fn for name, fn in globals().items() if name.startswith("pin_")}
