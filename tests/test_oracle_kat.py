"""Pins the CPU oracle against the reference's own known-answer unit tests.

Each test cites the reference test it ports (file:line).  Values come from
those tests; where a reference test only checks a direction of change, the
port does the same and a closed-form invariant test below adds an exact pin.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import (Pic, alf_cfg, alf_picture, binding_of, cu, cu_array, dparams, sao_band, sao_cfg, sao_edge,
                     sao_picture)


def deblock(orc, vv, pic, cus, params, cfi=1):
    arr = cu_array(vv, cus)
    orc.oracle_apply_deblocking_clipped(pic.ptr(), arr, len(cus), None, 0, None, C.byref(params), cfi, None, 0, None, 0)


# --------------------------------------------------------------------------- deblock
def test_beta_prime_table_anchors(orc):
    # src/deblock.rs:2083-2092
    assert [orc.oracle_beta_prime(q) for q in (15, 16, 32, 50, 63, -1, 99)] == [0, 6, 26, 62, 88, 0, 88]


def test_tc_prime_table_anchors(orc):
    # src/deblock.rs:2096-2102
    assert [orc.oracle_tc_prime(q) for q in (17, 18, 32, 50, 65)] == [0, 3, 10, 71, 395]


def test_bit_depth_scaling(orc):
    # src/deblock.rs:2107-2124
    assert orc.oracle_scale_beta(26, 8) == 26
    assert orc.oracle_scale_tc(10, 8) == (10 + 2) >> 2
    assert orc.oracle_scale_beta(26, 10) == 26 << 2
    assert orc.oracle_scale_tc(10, 10) == 10
    assert orc.oracle_scale_beta(26, 12) == 26 << 4
    assert orc.oracle_scale_tc(10, 12) == 10 << 2
    assert orc.oracle_scale_tc(9, 9) == (9 + 1) >> 1


def test_disabled_does_not_modify_picture(orc, vv):
    # src/deblock.rs:2128-2189
    pic = Pic(vv, 16, 16, 64)
    pic.y[:, 8:] = 200
    before = pic.y.copy()
    deblock(orc, vv, pic, [cu(vv, 0, 0, 8, 16, 32), cu(vv, 8, 0, 8, 16, 32)], dparams(vv, disabled=True))
    assert (pic.y == before).all()


def test_vertical_edge_smooths_weak_filter(orc, vv):
    # src/deblock.rs:2199-2260 (direction of change only)
    pic = Pic(vv, 16, 16, 100)
    pic.y[:, 8:] = 110
    deblock(orc, vv, pic, [cu(vv, 0, 0, 8, 16, 37), cu(vv, 8, 0, 8, 16, 37)], dparams(vv))
    assert pic.y[8, 7] > 100 and pic.y[8, 8] < 110
    assert pic.y[8, 0] == 100 and pic.y[8, 15] == 110


def test_tb_split_geometry_lookups(orc, vv):
    # src/deblock.rs:2342-2413
    base = cu(vv, 32, 16, 32, 16, 32)
    assert orc.oracle_luma_tb_len(C.byref(base), 40, 20, 1) == 32
    assert orc.oracle_luma_tb_len(C.byref(base), 40, 20, 0) == 16
    assert not orc.oracle_is_luma_tb_boundary(C.byref(base), 8, 1)
    isp = cu(vv, 32, 16, 32, 16, 32)
    isp.has_tb_split, isp.tb_vertical, isp.tb_n_bounds, isp.tb_luma_only = 1, 1, 3, 1
    isp.tb_bounds[0], isp.tb_bounds[1], isp.tb_bounds[2] = 8, 16, 24
    for i in range(4):
        isp.tb_y_coded[i] = 1
    assert orc.oracle_luma_tb_len(C.byref(isp), 40, 20, 1) == 8
    assert orc.oracle_luma_tb_len(C.byref(isp), 63, 20, 1) == 8
    assert orc.oracle_luma_tb_len(C.byref(isp), 40, 20, 0) == 16
    assert orc.oracle_is_luma_tb_boundary(C.byref(isp), 8, 1)
    assert orc.oracle_is_luma_tb_boundary(C.byref(isp), 16, 1)
    assert not orc.oracle_is_luma_tb_boundary(C.byref(isp), 4, 1)
    assert not orc.oracle_is_luma_tb_boundary(C.byref(isp), 8, 0)
    assert not orc.oracle_is_chroma_tb_boundary(C.byref(isp), 16, 1)
    assert orc.oracle_chroma_tb_len(C.byref(isp), 40, 20, 1, 2) == 16
    sbt = cu(vv, 32, 16, 32, 16, 32, intra=False)
    sbt.has_tb_split, sbt.tb_vertical, sbt.tb_n_bounds, sbt.tb_luma_only = 1, 1, 1, 0
    sbt.tb_bounds[0] = 16
    sbt.tb_y_coded[0] = 1
    sbt.tb_cb_coded[0] = 1
    assert orc.oracle_is_chroma_tb_boundary(C.byref(sbt), 16, 1)
    assert orc.oracle_chroma_tb_len(C.byref(sbt), 34, 20, 1, 2) == 8
    assert orc.oracle_tu_coded_at(C.byref(sbt), 0, 40, 20)
    assert not orc.oracle_tu_coded_at(C.byref(sbt), 0, 50, 20)
    assert orc.oracle_tu_coded_at(C.byref(sbt), 1, 40, 20)
    assert not orc.oracle_tu_coded_at(C.byref(sbt), 1, 50, 20)


def test_long_luma_filter_activates_for_32x16(orc, vv):
    # src/deblock.rs:2420-2488
    pic = Pic(vv, 64, 16, 100)
    pic.y[:, 32:] = 110
    deblock(orc, vv, pic, [cu(vv, 0, 0, 32, 16, 32), cu(vv, 32, 0, 32, 16, 32)], dparams(vv))
    assert pic.y[8, 31] > 100 and pic.y[8, 32] < 110
    assert pic.y[8, 0] == 100 and pic.y[8, 63] == 110


def test_weak_filter_p1_clip_bound_is_neg_of_half_tc(orc, vv):
    # src/deblock.rs:2523-2547: beta 52, tC 13, delta -13 -> 127 / 113 / 134 / 106
    pic = Pic(vv, 32, 8, 128)
    pic.y[:, :16] = 140
    pic.y[:, 16:] = 100
    deblock(orc, vv, pic, [cu(vv, 0, 0, 16, 8, 45), cu(vv, 16, 0, 16, 8, 45)], dparams(vv, ctb_log2=7))
    assert pic.y[2, 15] == 127
    assert pic.y[2, 16] == 113
    assert pic.y[2, 14] == 134
    assert pic.y[2, 17] == 106


def test_asymmetric_long_filter_7_3_vertical(orc, vv):
    # src/deblock.rs:2556-2581
    pic = Pic(vv, 48, 8, 128)
    pic.y[:, :32] = 100
    pic.y[:, 32:] = 104
    deblock(orc, vv, pic, [cu(vv, 0, 0, 32, 8, 45), cu(vv, 32, 0, 16, 8, 45)], dparams(vv, ctb_log2=7))
    row = pic.y[3]
    assert list(row[25:32]) == [100, 100, 101, 101, 101, 102, 102]
    assert list(row[32:36]) == [102, 103, 104, 104]
    assert row[24] == 100


# Long-filter arms on a flat 100 | 104 step at x = 32 (intra, qp 45, 8-bit),
# worked by hand from src/deblock.rs:1552-1672: every refMiddle arm
# (eqs. 1389-1392) gives 102 here ((1640 + 8) >> 4, or (820 + 4) >> 3 for
# 5/3), refP = 100, refQ = 104, so p'_i = 100 + ((2 f_i + 32) >> 6) and
# q'_j = (6688 - 2 g_j) >> 6 with F7 = 59 50 41 32 23 14 5, F5 = 58 45 32 19 6
# and F3 = 53 32 11 (no clipping: tC is large).  A side is 3 when its CU is
# 16 wide, and 5 through the sub-block cap at the CU edge (src/deblock.rs:
# 857-866): P when the P CU has sub-blocks, Q when the Q CU has more than one
# along the edge normal; the interior sub-block edges sit in flat samples.
P7 = [100, 100, 101, 101, 101, 102, 102]  # x = 25 .. 31
P5 = [100, 101, 101, 101, 102]  # x = 27 .. 31
Q7 = [102, 102, 103, 103, 103, 104, 104]  # x = 32 .. 38
Q5 = [102, 103, 103, 103, 104]  # x = 32 .. 36
P3 = [100, 101, 102]  # x = 29 .. 31
Q3 = [102, 103, 104]  # x = 32 .. 34


@pytest.mark.parametrize("p_w,p_sb,q_w,q_sb,want_p,want_q",
                         [(32, False, 32, False, P7, Q7), (32, True, 32, True, P5, Q5),
                          (32, False, 32, True, P7, Q5), (32, True, 32, False, P5, Q7),
                          (32, True, 16, False, P5, Q3), (16, False, 32, True, P3, Q5)],
                         ids=["7-7", "5-5", "7-5", "5-7", "5-3", "3-5"])
def test_long_filter_arms_on_a_step(orc, vv, p_w, p_sb, q_w, q_sb, want_p, want_q):
    pic = Pic(vv, 64, 8, 0)
    pic.y[:, :32] = 100
    pic.y[:, 32:] = 104
    p_cu, q_cu = cu(vv, 32 - p_w, 0, p_w, 8, 45), cu(vv, 32, 0, q_w, 8, 45)
    for c, sb in ((p_cu, p_sb), (q_cu, q_sb)):
        if sb:
            c.has_num_sb, c.num_sb_x, c.num_sb_y = 1, 2, 1
    cus = ([cu(vv, 0, 0, 16, 8, 45)] if p_w == 16 else []) + [p_cu, q_cu] + \
        ([cu(vv, 48, 0, 16, 8, 45)] if q_w == 16 else [])  # the extra CUs' edges lie in flat samples
    deblock(orc, vv, pic, cus, dparams(vv, ctb_log2=7))
    for y in range(8):
        row = [int(v) for v in pic.y[y]]
        assert row[32 - len(want_p):32] == want_p
        assert row[32:32 + len(want_q)] == want_q
        assert row[31 - len(want_p)] == 100 and row[32 + len(want_q)] == 104  # beyond the arm: untouched
        assert set(row[:31 - len(want_p)]) == {100} and set(row[33 + len(want_q):]) == {104}


@pytest.mark.parametrize("vertical", [True, False], ids=["ver", "hor"])
def test_chroma_strong_filter_on_a_step(orc, vv, vertical):
    # A flat 100 | 108 chroma step between two intra CUs whose chroma blocks
    # are 16 samples across the edge (maxFilterLength 3 on both sides), away
    # from a CTB row: the strong-filter decision holds (flat sides, |p0 - q0|
    # = 8 below (5 tC + 1) >> 1), and eqs. 1411-1416 (src/deblock.rs:
    # 1986-2023) worked by hand give p2..p0 = (7P+Q+4)>>3, (6P+2Q+4)>>3,
    # (5P+3Q+4)>>3 = 101, 102, 103 and q0..q2 = 105, 106, 107.  The weak
    # filter would move p0 and q0 only.
    W, H = (64, 16) if vertical else (16, 64)
    pic = Pic(vv, W, H, 128)
    for plane in (pic.cb, pic.cr):
        if vertical:
            plane[:, :16], plane[:, 16:] = 100, 108
        else:
            plane[:16, :], plane[16:, :] = 100, 108
    cus = [cu(vv, 0, 0, 32, 16, 45), cu(vv, 32, 0, 32, 16, 45)] if vertical else \
        [cu(vv, 0, 0, 16, 32, 45), cu(vv, 0, 32, 16, 32, 45)]
    deblock(orc, vv, pic, cus, dparams(vv, ctb_log2=7))
    want = [100, 101, 102, 103, 105, 106, 107, 108]  # positions 12 .. 19 across the edge
    for plane in (pic.cb, pic.cr):
        lines = plane if vertical else plane.T
        for line in lines:
            assert [int(v) for v in line[12:20]] == want
            assert set(int(v) for v in line[:13]) == {100} and set(int(v) for v in line[19:]) == {108}


def test_chroma_strong_filter_at_a_ctb_row(orc, vv):
    # The same 100 | 108 step on a horizontal chroma edge that lies on a luma
    # CTB row (CTB 64, edge at luma y = 64): maxFilterLengthP is 1 and the
    # asymmetric arm of eqs. 1417-1420 (src/deblock.rs:2041-2055) runs.  By
    # hand: p0 = (5P+3Q+4)>>3 = 103, q0..q2 = 105, 106, 107 as in the full arm,
    # but p1 and p2 keep their input (100).
    pic = Pic(vv, 16, 128, 128)
    for plane in (pic.cb, pic.cr):
        plane[:32, :], plane[32:, :] = 100, 108
    deblock(orc, vv, pic, [cu(vv, 0, 0, 16, 64, 45), cu(vv, 0, 64, 16, 64, 45)], dparams(vv, ctb_log2=6))
    for plane in (pic.cb, pic.cr):
        for col in plane.T:
            assert [int(v) for v in col[28:36]] == [100, 100, 100, 103, 105, 106, 107, 108]
            assert set(int(v) for v in col[:31]) == {100} and set(int(v) for v in col[35:]) == {108}


def test_ctb_row_boundary_caps_luma_p_side_at_3(orc, vv):
    # src/deblock.rs:2589-2638
    pic = Pic(vv, 8, 128, 128)
    pic.y[:64] = 100
    pic.y[64:] = 104
    deblock(orc, vv, pic, [cu(vv, 0, 0, 8, 64, 45), cu(vv, 0, 64, 8, 64, 45)], dparams(vv, ctb_log2=6))
    col = pic.y[:, 2]
    assert list(col[61:64]) == [100, 101, 102]
    assert list(col[64:71]) == [102, 102, 103, 103, 103, 104, 104]
    assert (col[57:61] == 100).all()


def test_no_neighbour_skips_filtering(orc, vv):
    # src/deblock.rs:2642-2673
    pic = Pic(vv, 8, 8, 100)
    before = pic.y.copy()
    deblock(orc, vv, pic, [cu(vv, 0, 0, 8, 8, 32)], dparams(vv))
    assert (pic.y == before).all()


@pytest.mark.parametrize("seed", [3, 7, 11])
def test_deblock_flat_plane_is_identity(orc, vv, seed):
    # closed form: every filter arm moves a sample by a clipped weighted
    # difference of its neighbours, which is zero on a flat plane.
    f = vv.SynthFrame(0, 256, 128, seed=seed, rich=1, sao=0, alf=0)
    d = f.desc.contents
    y, cb, cr = vv.picture_arrays(d.pic)
    y[:] = 500
    cb[:] = 300
    cr[:] = 700
    orc.oracle_apply_deblocking_clipped(C.byref(d.pic), d.cus, d.n_cus, d.chroma_cus if d.has_chroma_cus else None,
                                        d.n_chroma_cus, d.mv_grid, C.byref(d.deblock), 1, None, 0, None, 0)
    y, cb, cr = vv.picture_arrays(d.pic)
    assert (y == 500).all() and (cb == 300).all() and (cr == 700).all()


def _motion_edge(orc, vv, p_cell, q_cell, p_kind="inter", q_kind="inter"):
    """Two uncoded 16x8 CUs meeting at x = 16 over a 100 | 130 step, with a
    4x4 motion grid whose cells left of the edge hold `p_cell` and right of it
    `q_cell` (src/deblock.rs:351-363).  Returns the deblocked luma."""
    pic = Pic(vv, 32, 8, 0)
    pic.y[:, :16] = 100
    pic.y[:, 16:] = 130
    cus = []
    for x, kind in ((0, p_kind), (16, q_kind)):
        c = cu(vv, x, 0, 16, 8, 37, intra=(kind == "intra"), coded=False)
        c.ibc = int(kind == "ibc")
        c.ciip = int(kind == "ciip")
        cus.append(c)
    cells = (vv.MvCell * 16)()
    for cy in range(2):
        for cx in range(8):
            src = p_cell if cx < 4 else q_cell
            for name, val in src.items():
                if name in ("mv_l0", "mv_l1"):
                    getattr(cells[cy * 8 + cx], name)[:] = val
                else:
                    setattr(cells[cy * 8 + cx], name, val)
    grid = vv.MvGrid(C.cast(cells, C.POINTER(vv.MvCell)), 8, 2)
    arr = cu_array(vv, cus)
    orc.oracle_apply_deblocking_clipped(pic.ptr(), arr, 2, None, 0, C.byref(grid), C.byref(dparams(vv)), 1,
                                        None, 0, None, 0)
    return pic.y.copy()


def _uni(poc, mv, l1=False):
    return {"pred_l1": 1, "poc_l1": poc, "mv_l1": mv} if l1 else {"pred_l0": 1, "poc_l0": poc, "mv_l0": mv}


def _bi(poc0, mv0, poc1, mv1):
    return {"pred_l0": 1, "pred_l1": 1, "poc_l0": poc0, "poc_l1": poc1, "mv_l0": mv0, "mv_l1": mv1}


# Known answers of the section 8.8.3.5 motion arms (src/deblock.rs:984-1043):
# bS = 1 iff the number of MVs differs, the reference pictures differ (by
# identity, whichever list holds them), or an MV component differs by >= 8 in
# 1/16-sample units; for bi-prediction from one picture twice, only when both
# the list-aligned and the crossed pairing differ.
MOTION_CASES = [
    ("uni mv dx 7", _uni(4, (0, 0)), _uni(4, (7, 0)), False),
    ("uni mv dx 8", _uni(4, (0, 0)), _uni(4, (8, 0)), True),
    ("uni mv dy -8", _uni(4, (3, 5)), _uni(4, (3, -3)), True),
    ("uni mv dx -7 dy 7", _uni(4, (10, 10)), _uni(4, (3, 17)), False),
    ("uni same picture other list", _uni(4, (1, 2)), _uni(4, (1, 2), l1=True), False),
    ("uni other picture", _uni(4, (1, 2)), _uni(8, (1, 2)), True),
    ("uni vs bi", _uni(4, (0, 0)), _bi(4, (0, 0), 8, (0, 0)), True),
    ("no mv either side", {}, {}, False),
    ("bi lists swapped", _bi(4, (5, 5), 8, (-9, 2)), _bi(8, (-9, 2), 4, (5, 5)), False),
    ("bi lists swapped mv 8", _bi(4, (5, 5), 8, (-9, 2)), _bi(8, (-9, 2), 4, (13, 5)), True),
    ("bi other picture", _bi(4, (0, 0), 8, (0, 0)), _bi(4, (0, 0), 12, (0, 0)), True),
    ("bi same pic crossed equal", _bi(4, (0, 0), 4, (16, 0)), _bi(4, (16, 0), 4, (0, 0)), False),
    ("bi same pic both differ", _bi(4, (0, 0), 4, (0, 0)), _bi(4, (8, 0), 4, (8, 0)), True),
    ("bi same pic one pairing", _bi(4, (0, 0), 4, (8, 0)), _bi(4, (0, 0), 4, (0, 0)), True),
]


@pytest.mark.parametrize("name,p,q,filtered", MOTION_CASES, ids=[m[0] for m in MOTION_CASES])
def test_deblock_motion_bs_arms(orc, vv, name, p, q, filtered):
    # With both sides uncoded inter, the edge is filtered iff the motion arms
    # give bS = 1: the 100 | 130 step then moves; with bS = 0 it stays.
    y = _motion_edge(orc, vv, p, q)
    changed = (y[:, :16] != 100).any() or (y[:, 16:] != 130).any()
    assert changed == filtered


@pytest.mark.parametrize("bv_dx,filtered", [(7, False), (-7, False), (8, True), (-8, True)])
def test_deblock_ibc_bs_arms(orc, vv, bv_dx, filtered):
    # src/deblock.rs:1093-1106: IBC on both sides compares the block vectors
    # (held in the L0 slot) with the same >= 8 threshold; IBC against inter is
    # a prediction-mode difference and always gives bS = 1.
    p, q = _uni(0, (4, 4)), _uni(0, (4 + bv_dx, 4))
    y = _motion_edge(orc, vv, p, q, "ibc", "ibc")
    assert ((y[:, :16] != 100).any() or (y[:, 16:] != 130).any()) == filtered
    same = _uni(0, (4, 4))
    y = _motion_edge(orc, vv, same, same, "ibc", "inter")
    assert (y[:, :16] != 100).any()


def test_deblock_ciip_is_bs_2(orc, vv):
    # src/deblock.rs:1072-1078: CIIP on either side takes the intra arm
    # (bS = 2, tC from Q + 2), so it filters exactly like an intra
    # neighbour, and differently from a bS = 1 motion edge on this step.
    same = _uni(4, (0, 0))
    intra = _motion_edge(orc, vv, same, same, "intra", "inter")
    ciip = _motion_edge(orc, vv, same, same, "inter", "ciip")
    bs1 = _motion_edge(orc, vv, _uni(4, (0, 0)), _uni(4, (8, 0)))
    assert (ciip == intra).all()
    assert (intra != bs1).any()


def test_deblock_tc_zero_is_identity(orc, vv):
    # closed form: tC' = 0 for Q < 18 (Table 43) -> the segment is untouched
    pic = Pic(vv, 32, 8, 0)
    rng = np.random.default_rng(1)
    pic.y[:] = rng.integers(0, 256, pic.y.shape)
    before = pic.y.copy()
    deblock(orc, vv, pic, [cu(vv, 0, 0, 16, 8, 10), cu(vv, 16, 0, 16, 8, 10)], dparams(vv))
    assert (pic.y == before).all()


# --------------------------------------------------------------------------- SAO
def run_sao(orc, vv, pic, sp, cfg, tb=None):
    orc.oracle_apply_sao_clipped(pic.ptr(), C.byref(sp), C.byref(cfg), tb)


def test_sao_offset_val(orc):
    # src/sao.rs:663-673
    out = (C.c_int32 * 5)()
    orc.oracle_sao_derive_offset_val((C.c_uint32 * 4)(1, 2, 3, 4), (C.c_uint32 * 4)(0, 1, 0, 1), 8, out)
    assert list(out) == [0, 1, -2, 3, -4]
    orc.oracle_sao_derive_offset_val((C.c_uint32 * 4)(1, 1, 1, 1), (C.c_uint32 * 4)(0, 0, 0, 0), 12, out)
    assert list(out) == [0, 4, 4, 4, 4]


def test_sao_no_op_when_all_not_applied(orc, vv):
    pic = Pic(vv, 32, 32, 100)
    sp, _k = sao_picture(vv, 1, 1, {})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert (pic.y == 100).all()


def test_sao_no_op_when_flags_off(orc, vv):
    pic = Pic(vv, 32, 32, 100)
    p = vv.SaoCtbParams()
    p.luma = sao_band(vv, 0, [10, 0, 0, 0], [0] * 4)
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv, luma=False, chroma=False))
    assert (pic.y == 100).all()


@pytest.mark.parametrize("pos,seed,offs,want", [(12, 100, [10, 0, 0, 0], 110),   # src/sao.rs:724-740
                                                (0, 100, [50, 50, 50, 50], 100),  # :743-758
                                                (31, 250, [50, 0, 0, 0], 255)])   # :761-775
def test_sao_band_offset(orc, vv, pos, seed, offs, want):
    pic = Pic(vv, 32, 32, seed)
    p = vv.SaoCtbParams()
    p.luma = sao_band(vv, pos, offs, [0] * 4)
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert (pic.y == want).all()


def test_sao_edge_flat_no_op(orc, vv):
    pic = Pic(vv, 32, 32, 100)
    p = vv.SaoCtbParams()
    p.luma = sao_edge(vv, 0, [10, 10, 10, 10])
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert (pic.y == 100).all()


def test_sao_edge_horizontal_pattern(orc, vv):
    # src/sao.rs:805-833
    pic = Pic(vv, 32, 1, 0)
    pic.y[0, :] = [100 if x % 2 == 0 else 110 for x in range(32)]
    p = vv.SaoCtbParams()
    p.luma = sao_edge(vv, 0, [5, 0, 0, 2])
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert pic.y[0, 0] == 100 and pic.y[0, 31] == 110
    assert list(pic.y[0, 1:31]) == [105 if x % 2 == 0 else 108 for x in range(1, 31)]


def test_sao_edge_vertical_pattern(orc, vv):
    # src/sao.rs:945-971
    pic = Pic(vv, 2, 32, 0)
    pic.y[:, 0] = [100 if y % 2 == 0 else 110 for y in range(32)]
    pic.y[:, 1] = pic.y[:, 0]
    p = vv.SaoCtbParams()
    p.luma = sao_edge(vv, 1, [5, 0, 0, 2])
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert pic.y[0, 0] == 100 and pic.y[31, 0] == 110
    assert list(pic.y[1:31, 0]) == [105 if y % 2 == 0 else 108 for y in range(1, 31)]


def test_sao_edge_135_peak(orc, vv):
    # src/sao.rs:880-900
    pic = Pic(vv, 4, 4, 100)
    pic.y[1, 1] = 110
    p = vv.SaoCtbParams()
    p.luma = sao_edge(vv, 2, [5, 0, 0, 2])
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv))
    assert pic.y[1, 1] == 108 and pic.y[0, 0] == 100


def test_sao_chroma_band_offset(orc, vv):
    # src/sao.rs:906-927
    pic = Pic(vv, 32, 32, 100)
    p = vv.SaoCtbParams()
    p.cb = sao_band(vv, 16, [20, 0, 0, 0], [0] * 4)
    p.cr = sao_band(vv, 16, [20, 0, 0, 0], [0] * 4)
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv, luma=False, chroma=True))
    assert (pic.cb == 148).all() and (pic.cr == 148).all() and (pic.y == 100).all()


def test_sao_band_is_per_sample_offset(orc, vv):
    # closed form: band offset = a per-sample lookup of (sample >> (bd-5)) & 31
    rng = np.random.default_rng(5)
    pic = Pic(vv, 64, 64, 0, bd=10)
    pic.y[:] = rng.integers(0, 1024, pic.y.shape)
    before = pic.y.astype(np.int64)
    p = vv.SaoCtbParams()
    p.luma = sao_band(vv, 30, [7, 3, 5, 9], [1, 0, 1, 0], bd=10)
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    run_sao(orc, vv, pic, sp, sao_cfg(vv, bd=10, ctb_log2=6))
    band = before >> 5
    table = {30: -7, 31: 3, 0: -5, 1: 9}
    want = np.clip(before + np.vectorize(lambda b: table.get(b, 0))(band), 0, 1023)
    assert (pic.y == want).all()


def test_sao_edge_matches_independent_numpy(orc, vv):
    # closed form: EO classification by neighbour signs, restated in numpy
    rng = np.random.default_rng(9)
    for cls in range(4):
        pic = Pic(vv, 48, 40, 0, bd=10)
        pic.y[:] = rng.integers(0, 1024, pic.y.shape)
        pre = pic.y.astype(np.int64)
        p = vv.SaoCtbParams()
        p.luma = sao_edge(vv, cls, [9, 4, 6, 2], bd=10)
        sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
        run_sao(orc, vv, pic, sp, sao_cfg(vv, bd=10, ctb_log2=6))
        (h0, v0), (h1, v1) = [((-1, 0), (1, 0)), ((0, -1), (0, 1)), ((-1, -1), (1, 1)), ((1, -1), (-1, 1))][cls]
        want = pre.copy()
        H, W = pre.shape
        for y in range(1, H - 1):
            for x in range(1, W - 1):
                c, a, b = pre[y, x], pre[y + v0, x + h0], pre[y + v1, x + h1]
                local_min = c < a and c < b
                local_max = c > a and c > b
                corner_lo = (c < a and c == b) or (c == a and c < b)
                corner_hi = (c > a and c == b) or (c == a and c > b)
                off = 9 if local_min else 4 if corner_lo else -6 if corner_hi else -2 if local_max else 0
                want[y, x] = min(max(c + off, 0), 1023)
        assert (pic.y[1:-1, 1:-1] == want[1:-1, 1:-1]).all()
        # picture-border samples whose neighbour leaves the picture are copied
        if cls == 0:
            assert (pic.y[:, 0] == pre[:, 0]).all()


# --------------------------------------------------------------------------- ALF
def test_alf_clip_values(orc):
    # src/alf.rs:1411-1429 and Table 8 at 10-bit
    for bd in range(8, 17):
        vals = [orc.oracle_alf_clip_value(bd, i) for i in range(4)]
        assert vals[0] == 1 << bd
        assert vals[0] > vals[1] > vals[2] > vals[3]
    assert [orc.oracle_alf_clip_value(10, i) for i in range(4)] == [1024, 128, 32, 8]


def test_alf_tables_45_46(orc):
    # src/alf.rs:1480-1516
    o = (C.c_int32 * 4)()
    orc.oracle_alf_table_45(10, 32, 1, o)
    assert list(o) == [7, 1, 2, 3]
    for y, want in [(27, [10, 0, 0, 0]), (28, [10, 0, 0, 0]), (26, [7, 1, 1, 1]), (29, [7, 1, 1, 1]),
                    (25, [7, 1, 2, 2]), (30, [7, 1, 2, 2])]:
        orc.oracle_alf_table_45(y, 32, 1, o)
        assert list(o) == want, y
    for y in range(32):
        orc.oracle_alf_table_45(y, 32, 0, o)
        assert list(o) == [7, 1, 2, 3]
    o3 = (C.c_int32 * 3)()
    for y, want in [(0, [7, 1, 2]), (13, [10, 0, 0]), (14, [10, 0, 0]), (12, [7, 1, 1]), (15, [7, 1, 1])]:
        orc.oracle_alf_table_46(y, 16, 1, o3)
        assert list(o3) == want


def test_alf_transpose_involutions(orc):
    # src/alf.rs:1532-1554
    for t in range(4):
        p = (C.c_uint8 * 12)()
        orc.oracle_alf_transpose_idx(t, p)
        assert [p[p[k]] for k in range(12)] == list(range(12))
    orc.oracle_alf_transpose_idx(0, p)
    assert list(p) == list(range(12))


def test_alf_fixed_filter_rows(orc):
    # src/alf_fixed.rs:228-246
    row = (C.c_int32 * 12)()
    assert orc.oracle_alf_fixed_coeff_row(0, 0, row) == 0
    assert list(row) == [0, 0, 8, -8, -2, -7, 4, 4, 2, 1, -1, 25]
    assert orc.oracle_alf_fixed_coeff_row(15, 24, row) == 0
    assert list(row) == [0, 1, 1, -2, 2, -10, 10, 0, -2, 1, -7, 23]
    assert orc.oracle_alf_fixed_coeff_row(16, 0, row) == -1
    assert orc.oracle_alf_fixed_coeff_row(0, 25, row) == -1


def classify(orc, plane, ctb=32, bd=8):
    h, w = plane.shape
    cls = (C.c_uint16 * ((ctb // 4) ** 2))()
    clip = (C.c_int32 * 4)(0, 0, w, h)
    arr = np.ascontiguousarray(plane, np.uint16)
    orc.oracle_alf_classify_ctb(arr.ctypes.data_as(C.POINTER(C.c_uint16)), w, w, h, clip, 0, 0, ctb, bd, cls)
    return np.array(cls).reshape(ctb // 4, ctb // 4)


def test_alf_classification_flat_is_class_zero(orc):
    # src/alf.rs:1564-1573
    cls = classify(orc, np.full((32, 32), 100))
    assert ((cls & 0xff) == 0).all()


def test_alf_classification_linear_gradient_interior(orc):
    # src/alf.rs:1579-1592
    plane = np.tile(np.arange(32) * 7, (32, 1))
    assert classify(orc, plane)[3, 3] & 0xff == 0


def test_alf_classification_directional(orc):
    # closed form: a horizontal step has only vertical gradients in a block
    # straddling it: dir1 = vertical-dominant (dirHV = 1), dirS = 2 (strong),
    # so filtIdx = activity + ((1 & 1) << 1 + 2) * 5 = activity + 20.
    plane = np.zeros((32, 32))
    plane[16:] = 200
    cls = classify(orc, plane)
    f = cls[3, 2] & 0xff  # block rows 12..15 window -2..5 crosses row 16
    assert f >= 20


def run_alf(orc, vv, pic, ap, cfg, b):
    return orc.oracle_apply_alf_clipped(pic.ptr(), C.byref(ap), C.byref(cfg), C.byref(b), None)


def test_alf_fixed_set_identity_on_flat(orc, vv):
    # src/alf.rs:1240-1261
    pic = Pic(vv, 32, 32, 100)
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, 0
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): c})
    assert run_alf(orc, vv, pic, ap, alf_cfg(vv), vv.AlfApsBinding()) == 0
    assert (pic.y == 100).all()


def test_alf_fixed_set_modifies_spike(orc, vv):
    # src/alf.rs:1263-1292
    pic = Pic(vv, 32, 32, 100)
    pic.y[15, 15] = 250
    before = pic.y.copy()
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, 5
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): c})
    run_alf(orc, vv, pic, ap, alf_cfg(vv), vv.AlfApsBinding())
    assert (pic.y != before).any()


def test_alf_disabled_slice_is_no_op(orc, vv):
    pic = Pic(vv, 32, 32, 100)
    pic.y[3, 3] = 7
    before = pic.y.copy()
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, 0
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): c})
    cfg = alf_cfg(vv)
    cfg.alf_enabled = 0
    run_alf(orc, vv, pic, ap, cfg, vv.AlfApsBinding())
    assert (pic.y == before).all()


def test_alf_zero_coeff_identity(orc, vv):
    # src/alf.rs:1296-1324
    pic = Pic(vv, 32, 32, 100)
    aps = vv.AlfAps()
    aps.alf_luma_filter_signal_flag = 1
    aps.n_luma_coeff = 25
    b = binding_of(vv, luma=[aps])
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, 16
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): c})
    run_alf(orc, vv, pic, ap, alf_cfg(vv), b)
    assert (pic.y == 100).all()


def test_alf_chroma_zero_coeff_identity(orc, vv):
    # src/alf.rs:1436-1465
    pic = Pic(vv, 32, 32, 100)
    aps = vv.AlfAps()
    aps.alf_chroma_filter_signal_flag = 1
    aps.n_chroma_coeff = 1
    b = binding_of(vv, chroma=aps)
    c = vv.AlfCtb()
    c.cb_on = c.cr_on = 1
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): c})
    run_alf(orc, vv, pic, ap, alf_cfg(vv, cb=True, cr=True), b)
    assert (pic.cb == 128).all() and (pic.cr == 128).all() and (pic.y == 100).all()


def test_alf_luma_five_point_stencil(orc, vv):
    # closed form: an APS whose 25 classes all carry f = c only on the
    # horizontal (k = 11) and vertical (k = 6) nearest-neighbour taps, with no
    # clipping, is invariant under every transposeIdx (eqs. 1442-1445 swap 6
    # and 11 or keep both) and reduces eq. 1449-1451 to the 5-point Laplacian
    # stencil out = s + ((c * (l + r + u + d - 4 s) + 64) >> 7).  Checked away
    # from the virtual-boundary rows, with picture-edge replication.
    rng = np.random.default_rng(2)
    H = W = 64
    pic = Pic(vv, W, H, 0, bd=10)
    pic.y[:] = rng.integers(0, 1024, (H, W))
    pre = pic.y.astype(np.int64)
    aps = vv.AlfAps()
    aps.alf_luma_filter_signal_flag = 1
    aps.n_luma_coeff = 25
    c = 21
    for k in range(25):
        aps.luma_coeff[k][6] = c
        aps.luma_coeff[k][11] = c
    b = binding_of(vv, luma=[aps])
    ctb = vv.AlfCtb()
    ctb.luma_on, ctb.luma_filt_set_idx = 1, 16
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): ctb})
    run_alf(orc, vv, pic, ap, alf_cfg(vv, bd=10, ctb_log2=6), b)
    pad = np.pad(pre, 1, mode="edge")
    lap = pad[1:-1, :-2] + pad[1:-1, 2:] + pad[:-2, 1:-1] + pad[2:, 1:-1] - 4 * pre
    want = np.clip(pre + ((c * lap + 64) >> 7), 0, 1023)
    rows = [y for y in range(H) if y < 64 - 7]  # rows unaffected by Table 45
    assert (pic.y[rows] == want[rows]).all()


# eq. 1490 pairs (src/alf.rs:1051-1056) with y1 = 1, y2 = 2 (Table 46 away
# from the virtual boundary): each chroma tap k adds f[k] times the sum of two
# clipped differences at these offsets (dx, dy).
CHROMA_TAPS = [((0, 2), (0, -2)), ((1, 1), (-1, -1)), ((0, 1), (0, -1)),
               ((-1, 1), (1, -1)), ((2, 0), (-2, 0)), ((1, 0), (-1, 0))]


@pytest.mark.parametrize("clip_idx", [0, 2])
@pytest.mark.parametrize("tap", range(6))
def test_alf_chroma_single_tap_closed_form(orc, vv, tap, clip_idx):
    # closed form: with a single nonzero chroma coefficient the 5x5 diamond
    # reduces to out = s + ((f * (clip(a - s) + clip(b - s)) + 64) >> 7) for
    # the tap's two mirrored neighbours a, b (eqs. 1490-1492), with picture
    # edge replication.  Cb takes alternative 0 with f = +37 on `tap`, Cr takes
    # alternative 1 with f = -29 on the next tap, so the alternative index and
    # the coefficient sign are pinned too.  AlfClip[10][clip_idx] is 2^10 or
    # 2^5 (Table 8).  Two CTBs side by side check neighbours across the CTB
    # edge; rows 28..31 of each 32-row chroma CTB use Table 46's modified
    # offsets and are left out.
    rng = np.random.default_rng(10 + tap)
    W, H = 128, 64
    pic = Pic(vv, W, H, 0, bd=10, chroma=0)
    pic.cb[:] = rng.integers(0, 1024, pic.cb.shape)
    pic.cr[:] = rng.integers(0, 1024, pic.cr.shape)
    pre = {"cb": pic.cb.astype(np.int64), "cr": pic.cr.astype(np.int64)}
    aps = vv.AlfAps()
    aps.alf_chroma_filter_signal_flag = 1
    aps.alf_chroma_clip_flag = 1
    aps.alf_chroma_num_alt_filters_minus1 = 1
    aps.n_chroma_coeff = 2
    plan = {"cb": (0, tap, 37), "cr": (1, (tap + 1) % 6, -29)}
    for alt, k, f in plan.values():
        aps.chroma_coeff[alt][k] = f
        aps.chroma_clip_idx[alt][k] = clip_idx
    b = binding_of(vv, chroma=aps)
    ctb = vv.AlfCtb()
    ctb.cb_on = ctb.cr_on = 1
    ctb.cb_alt_idx, ctb.cr_alt_idx = 0, 1
    ap, _k = alf_picture(vv, 2, 1, {(0, 0): ctb, (1, 0): ctb})
    run_alf(orc, vv, pic, ap, alf_cfg(vv, bd=10, ctb_log2=6, cb=True, cr=True), b)
    lim = 1 << (10 if clip_idx == 0 else 5)
    rows = [y for y in range(32) if y < 28]
    for name, (alt, k, f) in plan.items():
        s = pre[name]
        pad = np.pad(s, 2, mode="edge")
        h, w = s.shape
        acc = np.zeros_like(s)
        for dx, dy in CHROMA_TAPS[k]:
            nb = pad[2 + dy:2 + dy + h, 2 + dx:2 + dx + w]
            acc += np.clip(nb - s, -lim, lim)
        want = np.clip(s + ((f * acc + 64) >> 7), 0, 1023)
        got = getattr(pic, name)
        assert (got[rows] == want[rows]).all(), (name, k)
    assert (pic.y == 0).all()


def test_cc_alf_single_tap(orc, vv):
    # closed form: CC-ALF with only the right-neighbour tap (k = 2) set to 64
    # adds clip((64 * (L[2y][2x+1] - L[2y][2x]) + 64) >> 7) to the chroma
    # sample (eqs. 1513-1518), i.e. half the horizontal luma difference.
    rng = np.random.default_rng(4)
    pic = Pic(vv, 64, 64, 0, bd=10, chroma=512)
    pic.y[:] = rng.integers(0, 1024, (64, 64))
    L = pic.y.astype(np.int64)
    cc = vv.AlfAps()
    cc.n_cc_cb_coeff = 1
    cc.cc_cb_coeff[0][2] = 64
    b = binding_of(vv, cc_cb=cc)
    ctb = vv.AlfCtb()
    ctb.cc_cb_idc = 1
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): ctb})
    run_alf(orc, vv, pic, ap, alf_cfg(vv, bd=10, ctb_log2=6, cc_cb=True), b)
    xs = np.arange(32)
    right = np.minimum(2 * xs + 1, 63)
    want = np.clip(512 + ((64 * (L[0::2][:, right] - L[0::2][:, 2 * xs]) + 64) >> 7), 0, 1023)
    assert (pic.cb == want).all()
    assert (pic.cr == 512).all()


# --------------------------------------------------------------------------- LMCS
def lmcs_data(vv, **kw):
    d = vv.LmcsData()
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def near_identity_bd8(vv):
    d = lmcs_data(vv)
    d.lmcs_delta_abs_cw[0] = 1
    d.lmcs_delta_sign_cw_flag[0] = 1
    return d


def derive(orc, vv, d, bd):
    out = vv.LmcsDerived()
    rc = orc.oracle_derive_lmcs(C.byref(d), bd, C.byref(out))
    return rc, out


def test_lmcs_org_cw_across_bit_depths(orc, vv):
    # src/lmcs.rs:730-760
    d = lmcs_data(vv, lmcs_delta_max_bin_idx=15)
    for bd, want in [(8, 16), (9, 32), (10, 64), (11, 128), (12, 256), (13, 512), (14, 1024), (15, 2048), (16, 4096)]:
        rc, o = derive(orc, vv, d, bd)
        assert rc == 0 and o.org_cw == want and o.lmcs_cw[0] == want
        assert o.scale_coeff[0] == 2048 and o.inv_scale_coeff[0] == 2048


def test_lmcs_rejects_bit_depth(orc, vv):
    d = near_identity_bd8(vv)
    assert derive(orc, vv, d, 7)[0] != 0 and derive(orc, vv, d, 17)[0] != 0


def test_lmcs_near_identity_full_window_bd8(orc, vv):
    # src/lmcs.rs:770-805
    rc, o = derive(orc, vv, near_identity_bd8(vv), 8)
    assert rc == 0 and o.org_cw == 16 and o.lmcs_cw[0] == 15
    assert all(o.lmcs_cw[i] == 16 for i in range(1, 16)) and sum(o.lmcs_cw) == 255
    assert list(o.input_pivot) == [16 * i for i in range(16)]
    assert o.lmcs_pivot[0] == 0 and all(o.lmcs_pivot[i] == 15 + 16 * (i - 1) for i in range(1, 17))
    assert o.scale_coeff[0] == 1920 and o.inv_scale_coeff[0] == 2184
    assert all(o.scale_coeff[i] == 2048 and o.inv_scale_coeff[i] == 2048 for i in range(1, 16))
    assert o.chroma_scale_coeff[0] == 2184


def test_lmcs_full_window_identity_budget_rejected(orc, vv):
    # src/lmcs.rs:808-820
    for bd in range(8, 17):
        assert derive(orc, vv, lmcs_data(vv), bd)[0] != 0


def test_lmcs_narrowed_window_bd10(orc, vv):
    # src/lmcs.rs:823-850
    rc, o = derive(orc, vv, lmcs_data(vv, lmcs_min_bin_idx=1, lmcs_delta_max_bin_idx=1), 10)
    assert rc == 0 and o.lmcs_cw[0] == 0 and o.lmcs_cw[15] == 0
    assert all(o.lmcs_cw[i] == 64 and o.scale_coeff[i] == 2048 for i in range(1, 15))
    for i in (0, 15):
        assert o.scale_coeff[i] == 0 and o.inv_scale_coeff[i] == 0 and o.chroma_scale_coeff[i] == 2048
    assert o.lmcs_pivot[1] == 0 and o.lmcs_pivot[15] == 896 and o.lmcs_pivot[16] == 896


def test_lmcs_band_boundaries(orc, vv):
    # src/lmcs.rs:853-880 and :883-897
    d = lmcs_data(vv, lmcs_delta_max_bin_idx=15, lmcs_delta_cw_prec_minus1=6)
    d.lmcs_delta_abs_cw[0], d.lmcs_delta_sign_cw_flag[0] = 14, 1
    rc, o = derive(orc, vv, d, 8)
    assert rc == 0 and o.lmcs_cw[0] == 2 and o.scale_coeff[0] == 256 and o.inv_scale_coeff[0] == 16384
    d.lmcs_delta_abs_cw[0] = 15
    assert derive(orc, vv, d, 8)[0] != 0
    d.lmcs_delta_abs_cw[0], d.lmcs_delta_sign_cw_flag[0] = 111, 0
    rc, o = derive(orc, vv, d, 8)
    assert rc == 0 and o.lmcs_cw[0] == 127 and o.scale_coeff[0] == 16256 and o.inv_scale_coeff[0] == 258
    d.lmcs_delta_abs_cw[0] = 112
    assert derive(orc, vv, d, 8)[0] != 0
    d = lmcs_data(vv, lmcs_delta_max_bin_idx=15, lmcs_delta_cw_prec_minus1=14)
    d.lmcs_delta_abs_cw[0] = 28671
    rc, o = derive(orc, vv, d, 16)
    assert rc == 0 and o.lmcs_cw[0] == 32767 and o.scale_coeff[0] == 16384 and o.inv_scale_coeff[0] == 256
    d.lmcs_delta_abs_cw[0] = 28672
    assert derive(orc, vv, d, 16)[0] != 0


def test_lmcs_pivot_crossing_clause(orc, vv):
    # src/lmcs.rs:900-925
    d = lmcs_data(vv, lmcs_delta_max_bin_idx=14, lmcs_delta_cw_prec_minus1=6)
    for i in (0, 1):
        d.lmcs_delta_abs_cw[i], d.lmcs_delta_sign_cw_flag[i] = 14, 1
    assert derive(orc, vv, d, 8)[0] != 0
    d.lmcs_delta_abs_cw[1], d.lmcs_delta_sign_cw_flag[1] = 111, 0
    rc, o = derive(orc, vv, d, 8)
    assert rc == 0 and o.lmcs_pivot[1] == 2


def test_lmcs_forward_idx_inverse(orc, vv):
    # src/lmcs.rs:1019-1075
    rc, o = derive(orc, vv, near_identity_bd8(vv), 8)
    assert [orc.oracle_lmcs_forward_map(C.byref(o), s) for s in (80, 85, 8)] == [79, 84, 8]
    assert [orc.oracle_lmcs_idx_y_inv(C.byref(o), s, 0, 15) for s in (0, 15, 30, 31, 255, 1000)] == \
        [0, 1, 1, 2, 15, 15]
    for s in range(16, 256):
        m = orc.oracle_lmcs_forward_map(C.byref(o), s)
        idx = orc.oracle_lmcs_idx_y_inv(C.byref(o), m, 0, 15)
        assert orc.oracle_lmcs_inverse_map(C.byref(o), m, idx) == s
    assert orc.oracle_lmcs_inverse_map(C.byref(o), 255, 15) <= 255


def test_lmcs_round_trip_bd10_narrowed(orc, vv):
    # src/lmcs.rs:1114-1132
    rc, o = derive(orc, vv, lmcs_data(vv, lmcs_min_bin_idx=1, lmcs_delta_max_bin_idx=1), 10)
    for s in (64, 100, 500, 895):
        m = orc.oracle_lmcs_forward_map(C.byref(o), s)
        idx = orc.oracle_lmcs_idx_y_inv(C.byref(o), m, 1, 14)
        assert orc.oracle_lmcs_inverse_map(C.byref(o), m, idx) == s


def test_lmcs_inverse_pass_bd10(orc, vv):
    # src/ctu.rs:16583-16625: lmcsCW[0] = 56 at 10-bit; 40 -> 46; plane of 10 -> 11
    d = lmcs_data(vv, lmcs_delta_cw_prec_minus1=3)
    d.lmcs_delta_abs_cw[0], d.lmcs_delta_sign_cw_flag[0] = 8, 1
    rc, o = derive(orc, vv, d, 10)
    assert rc == 0
    assert orc.oracle_lmcs_idx_y_inv(C.byref(o), 40, 0, 15) == 0
    assert orc.oracle_lmcs_inverse_map(C.byref(o), 40, 0) == 46
    pic = Pic(vv, 64, 64, 10, bd=10)
    b = vv.LmcsBinding(C.pointer(o), 0, 15)
    orc.oracle_lmcs_inverse_luma(C.byref(pic.ptr()._obj.luma), C.byref(b))
    assert (pic.y == 11).all()


# --------------------------------------------------------------------------- driver
def test_driver_is_stage_composition(orc, vv):
    # the driver runs LMCS -> deblock -> SAO -> ALF on one picture
    # (src/ctu.rs:11837-12064): same result as calling the stages in order.
    f1 = vv.SynthFrame(1, 256, 192, seed=21)
    f2 = vv.SynthFrame(1, 256, 192, seed=21)
    assert orc.oracle_apply_in_loop_filters(f1.desc) == 0
    d = f2.desc.contents
    orc.oracle_lmcs_inverse_luma(C.byref(d.pic.luma), C.byref(d.lmcs))
    orc.oracle_apply_deblocking_clipped(C.byref(d.pic), d.cus, d.n_cus, None, 0, d.mv_grid, C.byref(d.deblock), 1,
                                        None, 0, None, 0)
    orc.oracle_apply_sao_clipped(C.byref(d.pic), C.byref(d.sao), C.byref(d.sao_cfg), None)
    orc.oracle_apply_alf_clipped(C.byref(d.pic), C.byref(d.alf), C.byref(d.alf_cfg), C.byref(d.alf_binding), None)
    for a, b in zip(f1.output_arrays(), f2.output_arrays()):
        assert (a == b).all()
