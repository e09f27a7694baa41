"""Serving with per-CTB syntax records (vvclf_server_submit_syntax): the
records are expanded on the server's prep threads and the result matches the
oracle run on the expanded parameters, bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

from ctb_records import config, decode, random_records

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def syntax_config(vv, d):
    b = d.alf_binding
    chroma_aps = b.chroma_aps.contents if b.chroma_aps else None
    cc_cb = b.cc_cb_aps.contents.n_cc_cb_coeff if b.cc_cb_aps else 1
    cc_cr = b.cc_cr_aps.contents.n_cc_cr_coeff if b.cc_cr_aps else 1
    c = config(vv, bd=d.sao_cfg.bit_depth, cfi=1, n_aps=b.n_luma_apses,
               alt_m1=max(0, chroma_aps.n_chroma_coeff - 1) if chroma_aps else 0,
               cc_m1=(max(0, cc_cb - 1), max(0, cc_cr - 1)))
    c.alf_cb_enabled, c.alf_cr_enabled = d.alf_cfg.cb_enabled, d.alf_cfg.cr_enabled
    c.alf_cc_cb_enabled, c.alf_cc_cr_enabled = d.alf_cfg.cc_cb_enabled, d.alf_cfg.cc_cr_enabled
    c.alf_enabled = d.alf_cfg.alf_enabled
    c.sao_luma_used, c.sao_chroma_used = d.sao_cfg.luma_used, d.sao_cfg.chroma_used
    return c


def test_served_syntax_records_match_oracle(vv, orc):
    say("server: syntax records")
    ctx = vv.Context()
    scfg_s = vv.ServerConfig()
    scfg_s.hbm_budget_bytes = 64 << 30
    scfg_s.max_inflight = 16
    scfg_s.max_batch = 8
    scfg_s.prep_threads = 3
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, C.byref(scfg_s), C.byref(s)))
    rng = np.random.default_rng(9)
    work = []
    for k, (w, h) in enumerate([(832, 480), (1920, 1080), (416, 240), (832, 480)]):
        sid = C.c_uint32()
        vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
        fg = vv.SynthFrame(4, w, h, seed=300 + k)
        fo = vv.SynthFrame(4, w, h, seed=300 + k)
        d = fg.desc.contents
        cfg = syntax_config(vv, d)
        wc, hc = d.sao.pic_width_in_ctbs_y, d.sao.pic_height_in_ctbs_y
        recs = random_records(vv, rng, wc, hc, cfg)
        rc, sao, alf = decode(vv, recs, wc, hc, cfg)
        assert rc == 0, vv.lib.vvclf_last_error()
        od = fo.desc.contents
        od.sao.ctbs = C.cast(sao, C.POINTER(vv.SaoCtbParams))
        od.alf.ctbs = C.cast(alf, C.POINTER(vv.AlfCtb))
        assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit_syntax(s, sid.value, fg.desc, recs, C.byref(cfg), C.byref(t)))
        work.append((t.value, fg, fo, recs, sao, alf, (w, h)))
    for t, fg, fo, *_rest, wh in work:
        assert vv.lib.vvclf_server_wait(s, t, 60000) == 0, vv.lib.vvclf_last_error()
        for name, a, b in zip(("luma", "cb", "cr"), fo.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, b), f"{wh} {name} differs at {(a != b).sum()} samples"
        assert any((a != b).any() for a, b in zip(fg.input_arrays(), fg.output_arrays()))
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


def test_served_syntax_rejects_a_bad_record(vv):
    say("server: invalid syntax record")
    ctx = vv.Context()
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, None, C.byref(s)))
    sid = C.c_uint32()
    vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
    f = vv.SynthFrame(4, 416, 240, seed=3)
    d = f.desc.contents
    cfg = syntax_config(vv, d)
    n = d.sao.pic_width_in_ctbs_y * d.sao.pic_height_in_ctbs_y
    recs = (vv.CtbSyntax * n)()
    recs[0].bytes[0] = 1  # merge left at the picture's left edge
    before = f.output_arrays()
    t = C.c_uint64()
    vv.check(vv.lib.vvclf_server_submit_syntax(s, sid.value, f.desc, recs, C.byref(cfg), C.byref(t)))
    assert vv.lib.vvclf_server_wait(s, t.value, 60000) == vv.ERR_INVALID
    for a, b in zip(before, f.output_arrays()):
        assert np.array_equal(a, b)
    vv.lib.vvclf_server_destroy(s)
    ctx.close()
