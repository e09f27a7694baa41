"""Host-side parts of the reconstruction-side steps: the plane MD5 of the
conformance convention (RFC 1321 test vectors, and hashlib on random crops),
the conformance crop, host-side rejections, and the oracle's LMCS
forward-map and chroma-scaling arms pinned by the reference's own
known answers."""
import ctypes as C
import hashlib

import numpy as np
import pytest

from helpers import Pic
from recon_cases import lmcs_binding


def md5_of(vv, pic, cfi, bd, crop):
    out = ((C.c_uint8 * 16) * 3)()
    rc = vv.lib.vvclf_picture_md5(pic.ptr(), cfi, bd, C.byref(crop), C.byref(out))
    return rc, [bytes(out[k]).hex() for k in range(3)]


@pytest.mark.parametrize("text,want", [
    (b"", "d41d8cd98f00b204e9800998ecf8427e"),
    (b"a", "0cc175b9c0f1b6a831c399e269772661"),
    (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
    (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
    (b"abcdefghijklmnopqrstuvwxyz", "c3fcd3d76192e4007dfb496cca67e13b"),
    (b"12345678901234567890123456789012345678901234567890123456789012345678901234567890",
     "57edf4a22be3c955ac49da2e2107b67a"),
])
def test_md5_rfc1321_vectors(vv, text, want):
    # RFC 1321 appendix A.5: an 8-bit luma row holding the message bytes
    n = max(1, len(text))
    pic = Pic(vv, n, 2, 0, bd=8)
    pic.y[0, :len(text)] = np.frombuffer(text, np.uint8)
    rc, h = md5_of(vv, pic, 0, 8, vv.Crop(0, 0, len(text), 1))
    assert rc == 0 and h[0] == want


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_md5_matches_hashlib_on_cropped_planes(vv, bd):
    rng = np.random.default_rng(bd)
    W, H = 136, 72
    pic = Pic(vv, W, H, 0, bd=bd)
    for a in (pic.y, pic.cb, pic.cr):
        a[:] = rng.integers(0, 1 << 16 if bd == 8 else 1 << bd, a.shape)  # 8-bit: the high byte is dropped
    crop = vv.Crop(6, 4, 120, 60)
    rc, h = md5_of(vv, pic, 1, bd, crop)
    assert rc == 0

    def ref(a, x, y, w, hh):
        sub = a[y:y + hh, x:x + w]
        data = sub.astype("<u2").tobytes() if bd > 8 else (sub & 0xff).astype(np.uint8).tobytes()
        return hashlib.md5(data).hexdigest()
    assert h[0] == ref(pic.y, 6, 4, 120, 60)
    assert h[1] == ref(pic.cb, 3, 2, 60, 30)
    assert h[2] == ref(pic.cr, 3, 2, 60, 30)


def test_md5_rejects_a_crop_outside_the_picture(vv):
    pic = Pic(vv, 16, 16, 0, bd=10)
    rc, _ = md5_of(vv, pic, 1, 10, vv.Crop(8, 0, 16, 16))
    assert rc == vv.ERR_INVALID


def test_conformance_crop(vv):
    # src/stream.rs:816-842: window offsets in chroma units times SubWidthC /
    # SubHeightC (2 / 2 for 4:2:0, 2 / 1 for 4:2:2, 1 / 1 otherwise)
    out = vv.Crop()
    win = (C.c_uint32 * 4)(1, 2, 3, 4)  # left, right, top, bottom
    assert vv.lib.vvclf_conformance_crop(1920, 1088, 1, win, C.byref(out)) == 0
    assert (out.x, out.y, out.w, out.h) == (2, 6, 1920 - 2 - 4, 1088 - 6 - 8)
    assert vv.lib.vvclf_conformance_crop(1920, 1088, 2, win, C.byref(out)) == 0
    assert (out.x, out.y, out.w, out.h) == (2, 3, 1914, 1081)
    assert vv.lib.vvclf_conformance_crop(64, 64, 0, win, C.byref(out)) == 0
    assert (out.x, out.y, out.w, out.h) == (1, 3, 61, 57)
    assert vv.lib.vvclf_conformance_crop(64, 64, 1, None, C.byref(out)) == 0
    assert (out.x, out.y, out.w, out.h) == (0, 0, 64, 64)
    big = (C.c_uint32 * 4)(20, 20, 0, 0)
    assert vv.lib.vvclf_conformance_crop(64, 64, 1, big, C.byref(out)) == vv.ERR_INVALID


def test_forward_map_rejects_overlapping_rects_on_the_host(vv):
    b, _k = lmcs_binding(vv, 1, 10)
    pic = Pic(vv, 32, 32, 300, bd=10)
    rects = (vv.Rect * 2)(vv.Rect(0, 0, 16, 16), vv.Rect(8, 8, 16, 16))
    before = pic.y.copy()
    assert vv.lib.vvclf_lmcs_forward_luma(C.byref((pic.ptr(), pic._pic.luma)[1]), C.byref(b), rects, 2) == vv.ERR_INVALID
    assert b"overlaps" in vv.lib.vvclf_last_error()
    assert (pic.y == before).all()


def lmcs_payload(vv, abs_cw, sign_cw, crs):
    d = vv.LmcsData()
    d.lmcs_min_bin_idx = 0
    d.lmcs_delta_max_bin_idx = 0
    d.lmcs_delta_cw_prec_minus1 = 3
    for i in range(16):
        d.lmcs_delta_abs_cw[i] = abs_cw[i]
        d.lmcs_delta_sign_cw_flag[i] = sign_cw[i]
    d.lmcs_delta_abs_crs = crs
    return d


def bound(vv, orc, d, bd=8):
    der = vv.LmcsDerived()
    assert orc.oracle_derive_lmcs(C.byref(d), bd, C.byref(der)) == 0
    b = vv.LmcsBinding()
    b.derived = C.pointer(der)
    b.min_bin_idx, b.max_bin_idx = 0, 15
    return b, der


def test_oracle_chroma_var_scale_reference_answers(vv, orc):
    # src/ctu.rs:16258-16357: lmcsCW[0] = 8, others 16 (8 bits), crs +4; no
    # neighbour -> invAvgLuma 128 (eq. 1217) -> bin 8 -> 16 * 2048 / 20 = 1638
    b, der = bound(vv, orc, lmcs_payload(vv, [8] + [0] * 15, [1] + [0] * 15, 4))
    pic = Pic(vv, 64, 64, 100, bd=8)
    lum = (pic.ptr(), pic._pic.luma)[1]
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum), C.byref(b), 0, 0, 0, 0, 64) == 1638
    # src/ctu.rs:16359-16460: lmcsCW 16 on bins 0..7 and 12 on 8..15, crs +4:
    # a left column of luma 60 -> bin 3 -> 32768 / 20 = 1638; mid-grey 128 ->
    # bin 8 -> 32768 / 16 = 2048 (the CU at (64, 0): availL only)
    b2, der2 = bound(vv, orc, lmcs_payload(vv, [0] * 8 + [4] * 8, [0] * 8 + [1] * 8, 4))
    pic2 = Pic(vv, 128, 64, 60, bd=8)
    lum2 = (pic2.ptr(), pic2._pic.luma)[1]
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum2), C.byref(b2), 64, 0, 1, 0, 64) == 1638
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum2), C.byref(b2), 64, 0, 0, 0, 64) == 2048
    # src/ctu.rs:16626-: the CU at (0, 64), top row of luma 60 (availT only)
    pic3 = Pic(vv, 64, 128, 60, bd=8)
    lum3 = (pic3.ptr(), pic3._pic.luma)[1]
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum3), C.byref(b2), 0, 64, 0, 1, 64) == 1638
    # the rounded mean over both sides (eq. 1216): 64 samples of 60 and 64 of
    # 200 -> (64*60 + 64*200 + 64) >> 7 = 130 -> bin 8 -> 2048; of 60 and 100 ->
    # 80 -> bin 5 -> 1638
    pic4 = Pic(vv, 128, 128, 60, bd=8)
    pic4.y[63, :] = 200
    lum4 = (pic4.ptr(), pic4._pic.luma)[1]
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum4), C.byref(b2), 64, 64, 1, 1, 64) == 2048
    pic4.y[63, :] = 100
    assert orc.oracle_lmcs_chroma_var_scale(C.byref(lum4), C.byref(b2), 64, 64, 1, 1, 64) == 1638


def test_oracle_chroma_fold_reference_answers(vv, orc):
    # eq. 1220 as worked in src/ctu.rs:16330-16350: Clip1(128 + Sign(r) *
    # ((|r| * 1638 + 1024) >> 11)) at 8 bits, and the eq. 1219 clamp of r to
    # [-256, 255]
    b, der = bound(vv, orc, lmcs_payload(vv, [8] + [0] * 15, [1] + [0] * 15, 4))
    for r, want in [(0, 128), (1, 129), (-1, 127), (3, 130), (-3, 126), (10, 136), (-10, 120), (200, 255),
                    (-200, 0), (1000, 255), (-1000, 0)]:
        mag = (min(abs(r), 256 if r < 0 else 255) * 1638 + 1024) >> 11
        assert want == max(0, min(255, 128 + (mag if r >= 0 else -mag)))
        assert orc.oracle_lmcs_scale_chroma_residual_sample(C.byref(der), 128, r, 1638) == want


def test_oracle_forward_rect_reference_answers(vv, orc):
    # src/lmcs.rs:1019-1035 (the near-identity payload of :719-727, lmcsCW[0]
    # = 15 at 8 bits): 80 -> 79, 85 -> 84, 8 -> 8, through the rectangle pass
    # (src/ctu.rs:1701-1725), which leaves samples outside the rectangle alone
    b, der = bound(vv, orc, lmcs_payload(vv, [1] + [0] * 15, [1] + [0] * 15, 0))
    pic = Pic(vv, 8, 4, 0, bd=8)
    pic.y[0, :4] = [80, 85, 8, 80]
    pic.y[1, :4] = [80, 85, 8, 80]
    lum = (pic.ptr(), pic._pic.luma)[1]
    assert orc.oracle_lmcs_forward_luma_rect(C.byref(lum), C.byref(b), 0, 0, 3, 1) == 0
    assert list(pic.y[0, :4]) == [79, 84, 8, 80] and list(pic.y[1, :4]) == [80, 85, 8, 80]
    pic.y[2, 2] = 256  # beyond 8 bits: the reference indexes past its pieces
    assert orc.oracle_lmcs_forward_luma_rect(C.byref(lum), C.byref(b), 0, 2, 4, 1) == -1
