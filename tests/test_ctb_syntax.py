"""The per-CTB syntax record (the compact side-information format of the
serving API): hand-built records against hand-worked SAO / ALF parameters
(the inference of decode_sao_ctb, src/sao_syntax.rs:154-284, and of
decode_alf_ctu, src/alf_syntax.rs:170-290, with the reference's own known
answers where it has them), invalid records, and round trips."""
import ctypes as C

import numpy as np
import pytest

from ctb_records import CB, CR, MERGE_L, MERGE_U, Y, Rec, config, decode, random_records


def recs_of(vv, *rs):
    return (vv.CtbSyntax * len(rs))(*[r.struct(vv) for r in rs])


def offs(s):
    return list(s.offset_val)


def test_band_offset_and_eq_153(vv):
    # src/sao_syntax.rs:722-731: abs [1, 2, 3, 4], signs [0, 1, 0, 1] at 8 bits
    # -> SaoOffsetVal [0, 1, -2, 3, -4]; abs [1, 1, 1, 1] at 12 bits -> 4 each
    cfg = config(vv, bd=8, alf=False)
    rc, sao, _ = decode(vv, recs_of(vv, Rec().sao(Y, 1, [1, 2, 3, 4], [0, 1, 0, 1], 5)), 1, 1, cfg)
    assert rc == 0
    s = sao[0].luma
    assert (s.sao_type_idx, s.band_position, s.eo_class) == (1, 5, 0) and offs(s) == [0, 1, -2, 3, -4]
    assert sao[0].cb.sao_type_idx == 0 and sao[0].cr.sao_type_idx == 0
    rc, sao, _ = decode(vv, recs_of(vv, Rec().sao(Y, 1, [1, 1, 1, 1], [0, 0, 0, 0], 0)), 1, 1, config(vv, bd=12, alf=False))
    assert rc == 0 and offs(sao[0].luma) == [0, 4, 4, 4, 4]


def test_edge_offset_signs_and_cr_inheritance(vv):
    # Cb edge offset, class 2, abs [3, 1, 2, 4] -> (+, +, -, -): [0, 3, 1, -2, -4];
    # Cr takes Cb's type and class with its own magnitudes [0, 5, 6, 7]
    rec = Rec().sao(CB, 2, [3, 1, 2, 4], [0] * 4, 2).sao_cr([0, 5, 6, 7], [0] * 4, 0)
    rc, sao, _ = decode(vv, recs_of(vv, rec), 1, 1, config(vv, alf=False))
    assert rc == 0
    cb, cr = sao[0].cb, sao[0].cr
    assert (cb.sao_type_idx, cb.eo_class, cb.band_position) == (2, 2, 0) and offs(cb) == [0, 3, 1, -2, -4]
    assert (cr.sao_type_idx, cr.eo_class) == (2, 2) and offs(cr) == [0, 0, 5, -6, -7]
    assert sao[0].luma.sao_type_idx == 0
    # Cr band offset keeps its own band position
    rec = Rec().sao(CB, 1, [1, 0, 0, 0], [1, 0, 0, 0], 9).sao_cr([2, 0, 0, 3], [0, 0, 0, 1], 30)
    rc, sao, _ = decode(vv, recs_of(vv, rec), 1, 1, config(vv, alf=False))
    assert rc == 0 and sao[0].cb.band_position == 9 and sao[0].cr.band_position == 30
    assert offs(sao[0].cb) == [0, -1, 0, 0, 0] and offs(sao[0].cr) == [0, 2, 0, 0, -3]


def test_merge_flags_copy_the_decoded_neighbour(vv):
    # 3 x 2 grid: (0,0) explicit, (1,0) and (2,0) merge left (a chain), (0,1)
    # merges up, (1,1) explicit, (2,1) merges up (from (2,0), i.e. (0,0))
    a = Rec().sao(Y, 1, [1, 2, 3, 4], [0, 0, 1, 1], 7)
    b = Rec().sao(Y, 2, [5, 0, 0, 5], [0] * 4, 1)
    grid = [a, Rec().put(MERGE_L, 1, 1), Rec().put(MERGE_L, 1, 1), Rec().put(MERGE_U, 1, 1), b,
            Rec().put(MERGE_U, 1, 1)]
    rc, sao, _ = decode(vv, recs_of(vv, *grid), 3, 2, config(vv, alf=False))
    assert rc == 0
    raw = [bytes(memoryview(sao[i])) for i in range(6)]
    assert raw[1] == raw[0] and raw[2] == raw[0] and raw[3] == raw[0] and raw[5] == raw[0] and raw[4] != raw[0]
    assert sao[4].luma.sao_type_idx == 2 and sao[4].luma.eo_class == 1


@pytest.mark.parametrize("rec,w,h,cfg_kw", [
    (Rec().put(MERGE_L, 1, 1), 1, 1, {}),                       # merge left at rx = 0
    (Rec().put(MERGE_U, 1, 1), 1, 1, {}),                       # merge up at ry = 0
    (Rec().sao(Y, 1, [8, 0, 0, 0], [0] * 4, 0), 1, 1, {"bd": 8}),  # abs above cMax 7 (Table 127)
    (Rec().sao(Y, 3, [0] * 4, [0] * 4, 0), 1, 1, {}),            # type 3
    (Rec().sao(Y, 2, [1, 1, 1, 1], [0, 0, 1, 0], 0), 1, 1, {}),  # signs on an edge offset
    (Rec().sao(Y, 1, [0, 1, 1, 1], [1, 0, 0, 0], 0), 1, 1, {}),  # sign without magnitude
    (Rec().sao(Y, 2, [1] * 4, [0] * 4, 4), 1, 1, {}),            # EO class 4
    (Rec().put(120, 1, 1), 1, 1, {}),                            # reserved bit
    (Rec().alf(luma_on=1, use_aps=1, idx=4), 1, 1, {}),          # prev idx beyond 4 APS ids
    (Rec().alf(luma_on=0, idx=3), 1, 1, {}),                     # filter index without the flag
    (Rec().alf(cb_on=1, cb_alt=2), 1, 1, {}),                    # alt idx beyond num_alt_filters_minus1
    (Rec().alf(cc_cb=3), 1, 1, {}),                              # CC idc beyond filters_signalled_minus1 + 1
    (Rec().alf(luma_on=1), 1, 1, {"alf": False}),                # ALF fields with ALF off
])
def test_invalid_records(vv, rec, w, h, cfg_kw):
    rc, _, _ = decode(vv, recs_of(vv, rec), w, h, config(vv, **cfg_kw))
    assert rc == vv.ERR_INVALID


def test_alf_fields(vv):
    cfg = config(vv, n_aps=4, alt_m1=1, cc_m1=(1, 2))
    rs = [Rec().alf(luma_on=1, use_aps=1, idx=2, cb_on=1, cb_alt=1, cc_cb=2, cc_cr=3),
          Rec().alf(luma_on=1, idx=7, cr_on=1, cr_alt=0),
          Rec().alf()]
    rc, _, alf = decode(vv, recs_of(vv, *rs), 3, 1, cfg)
    assert rc == 0
    got = [(a.luma_on, a.luma_filt_set_idx, a.cb_on, a.cb_alt_idx, a.cr_on, a.cr_alt_idx, a.cc_cb_idc, a.cc_cr_idc)
           for a in alf]
    # AlfCtbFiltSetIdxY = 16 + prev_idx (eq. 1439), fixed index as is
    assert got == [(1, 18, 1, 1, 0, 0, 2, 3), (1, 7, 0, 0, 1, 0, 0, 0), (0, 0, 0, 0, 0, 0, 0, 0)]


@pytest.mark.parametrize("seed,bd,cfi", [(1, 10, 1), (2, 8, 1), (3, 12, 1), (4, 10, 0)])
def test_round_trips(vv, seed, bd, cfi):
    rng = np.random.default_rng(seed)
    w, h = 15, 9
    cfg = config(vv, bd=bd, cfi=cfi, n_aps=int(rng.integers(0, 9)), alt_m1=int(rng.integers(0, 8)),
                 cc_m1=(int(rng.integers(0, 4)), int(rng.integers(0, 4))))
    recs = random_records(vv, rng, w, h, cfg)
    rc, sao, alf = decode(vv, recs, w, h, cfg)
    assert rc == 0, vv.lib.vvclf_last_error()
    out = (vv.CtbSyntax * (w * h))()
    assert vv.lib.vvclf_ctb_syntax_encode(sao, alf, w, h, C.byref(cfg), out) == 0, vv.lib.vvclf_last_error()
    rc, sao2, alf2 = decode(vv, out, w, h, cfg)
    assert rc == 0
    assert bytes(sao2) == bytes(sao) and bytes(alf2) == bytes(alf)
    # records without merges re-encode to themselves
    assert all(bytes(out[i]) == bytes(recs[i]) for i in range(w * h) if not (recs[i].bytes[0] & 3) and
               not (out[i].bytes[0] & 3))


def test_encode_rejects_parameters_without_syntax(vv):
    cfg = config(vv, alf=False)
    sao = (vv.SaoCtbParams * 1)()
    alf = (vv.AlfCtb * 1)()
    out = (vv.CtbSyntax * 1)()
    sao[0].cb.sao_type_idx, sao[0].cr.sao_type_idx = 1, 2  # Cr type differs from Cb's
    assert vv.lib.vvclf_ctb_syntax_encode(sao, alf, 1, 1, C.byref(cfg), out) == vv.ERR_INVALID
    sao[0].cr.sao_type_idx = 0
    sao[0].cb.sao_type_idx = 0
    sao[0].luma.sao_type_idx = 1
    sao[0].luma.offset_val[1] = 40  # above cMax 31 at 10 bits
    assert vv.lib.vvclf_ctb_syntax_encode(sao, alf, 1, 1, C.byref(cfg), out) == vv.ERR_INVALID
