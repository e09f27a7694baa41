"""Serving loop (continuous batching) on the GPU: every result bit-exact
against the oracle whether it ran alone or in a mix, HBM-budget admission,
and one failed job leaving the others intact."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def server(vv, ctx, budget=64 << 30, inflight=16, batch=8):
    cfg = vv.ServerConfig()
    cfg.hbm_budget_bytes = budget
    cfg.max_inflight = inflight
    cfg.max_batch = batch
    cfg.prep_threads = 3
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, C.byref(cfg), C.byref(s)))
    return s


def stats(vv, s):
    st = vv.ServerStats()
    vv.check(vv.lib.vvclf_server_get_stats(s, C.byref(st)))
    return st


STREAMS = [(0, 416, 240), (1, 832, 480), (1, 208, 120), (3, 416, 240), (0, 256, 128), (1, 200, 136)]


def test_served_mix_is_exact(vv, orc):
    say("server mix")
    ctx = vv.Context()
    s = server(vv, ctx)
    ids = []
    for _ in STREAMS:
        i = C.c_uint32()
        vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(i)))
        ids.append(i.value)
    frames, refs, tickets = [], [], []
    for rep in range(3):  # three pictures per stream, APS content changes each time
        for sid, (cfg, w, h) in zip(ids, STREAMS):
            seed = 100 * rep + sid
            fg = vv.SynthFrame(cfg, w, h, seed=seed, rich=1 if sid % 2 else 0)
            fo = vv.SynthFrame(cfg, w, h, seed=seed, rich=1 if sid % 2 else 0)
            assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
            t = C.c_uint64()
            vv.check(vv.lib.vvclf_server_submit(s, sid, fg.desc, C.byref(t)))
            frames.append(fg)
            refs.append(fo)
            tickets.append(t.value)
    for t in tickets:
        assert vv.lib.vvclf_server_wait(s, t, 60000) == 0, vv.lib.vvclf_last_error()
    for fg, fo in zip(frames, refs):
        for a, b in zip(fo.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, b)
    st = stats(vv, s)
    assert st.completed == len(frames) and st.failed == 0
    assert st.launches >= 1 and st.p99_ms >= st.p50_ms > 0
    # the same pictures alone (one job per drop-in call) give the same samples
    for fg, (cfg, w, h), sid in list(zip(frames, STREAMS, ids))[:4]:
        alone = vv.SynthFrame(cfg, w, h, seed=sid, rich=1 if sid % 2 else 0)
        assert vv.lib.vvclf_apply_in_loop_filters(alone.desc) == 0
        for a, b in zip(alone.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, b)
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


def test_budget_admission_and_rejection(vv, orc):
    say("server budget")
    ctx = vv.Context()
    one = vv.SynthFrame(1, 832, 480, seed=1)
    st0 = (vv.u64 * 8)()
    vv.check(vv.lib.vvclf_frame_check(one.desc, st0))
    budget = int(st0[6] * 1.5) + (4 << 20)  # about one 832x480 job at a time
    s = server(vv, ctx, budget=budget, inflight=8, batch=4)
    sid = C.c_uint32()
    vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
    frames, refs, tickets = [], [], []
    for k in range(6):
        fg, fo = vv.SynthFrame(1, 832, 480, seed=k), vv.SynthFrame(1, 832, 480, seed=k)
        orc.oracle_apply_in_loop_filters(fo.desc)
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sid.value, fg.desc, C.byref(t)))
        frames.append(fg)
        refs.append(fo)
        tickets.append(t.value)
    big = vv.SynthFrame(1, 3840, 2160, seed=9)
    tb = C.c_uint64()
    vv.check(vv.lib.vvclf_server_submit(s, sid.value, big.desc, C.byref(tb)))
    for t in tickets:
        assert vv.lib.vvclf_server_wait(s, t, 60000) == 0
    assert vv.lib.vvclf_server_wait(s, tb.value, 60000) == vv.ERR_BUDGET
    for fg, fo in zip(frames, refs):
        for a, b in zip(fo.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, b)
    st = stats(vv, s)
    assert st.bytes_admitted_peak <= budget and st.rejected == 1 and st.completed == 6
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


def test_injected_failure_spares_other_jobs(vv, orc):
    say("server injected failure")
    ctx = vv.Context()
    s = server(vv, ctx)
    sid = C.c_uint32()
    vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
    vv.check(vv.lib.vvclf_server_inject_failures(s, 2))
    frames, refs, tickets = [], [], []
    for k in range(6):
        fg, fo = vv.SynthFrame(0, 416, 240, seed=k), vv.SynthFrame(0, 416, 240, seed=k)
        orc.oracle_apply_in_loop_filters(fo.desc)
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sid.value, fg.desc, C.byref(t)))
        frames.append(fg)
        refs.append(fo)
        tickets.append(t.value)
    rcs = [vv.lib.vvclf_server_wait(s, t, 60000) for t in tickets]
    assert sorted(rcs).count(vv.ERR_HIP) == 2 and rcs.count(0) == 4
    for rc, fg, fo in zip(rcs, frames, refs):
        if rc == 0:
            for a, b in zip(fo.output_arrays(), fg.output_arrays()):
                assert np.array_equal(a, b)
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


def test_small_ctbs_pick_a_fitting_tile(vv, orc):
    """A launch holding pictures with 32-sample CTBs runs 32x32 tiles (the
    configured 64x64 tiles would straddle their ALF CTBs); all stay exact."""
    say("server small ctbs")
    ctx = vv.Context()
    s = server(vv, ctx, batch=8)
    cases = [(1, 416, 240, 31, {"ctb_log2": 5}), (0, 256, 128, 32, {}), (1, 208, 120, 33, {"ctb_log2": 5, "rich": 1}),
             (3, 416, 240, 34, {"ctb_log2": 6})]
    sid = C.c_uint32()
    vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
    frames, refs, tickets = [], [], []
    for cfg, w, h, seed, kw in cases:
        fg = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        fo = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sid.value, fg.desc, C.byref(t)))
        frames.append(fg)
        refs.append(fo)
        tickets.append(t.value)
    for t in tickets:
        assert vv.lib.vvclf_server_wait(s, t, 60000) == 0, vv.lib.vvclf_last_error()
    for fg, fo, case in zip(frames, refs, cases):
        for a, b in zip(fo.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, b), case
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


def test_server_rejects_unknown_tile_shape(vv):
    ctx = vv.Context()
    cfg = vv.ServerConfig()
    cfg.tune = vv.Tune(48, 48, 256, 0)
    s = C.c_void_p()
    assert vv.lib.vvclf_server_create(ctx.h, C.byref(cfg), C.byref(s)) == -1  # VVCLF_ERR_INVALID
    assert b"48x48" in vv.lib.vvclf_last_error()
    ctx.close()


def test_reset_stats_starts_afresh(vv, orc):
    """Statistics after a reset count only the jobs submitted after it, and
    the results of those jobs still match the oracle."""
    say("server reset stats")
    ctx = vv.Context()
    s = server(vv, ctx)
    sid = C.c_uint32()
    vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(sid)))
    for seed in (1, 2, 3):
        f = vv.SynthFrame(1, 416, 240, seed=seed)
        vv.check(vv.lib.vvclf_server_submit(s, sid.value, f.desc, C.byref(C.c_uint64())))
        vv.check(vv.lib.vvclf_server_drain(s, 60000))
        f.close()
    assert stats(vv, s).completed == 3
    vv.check(vv.lib.vvclf_server_reset_stats(s))
    st = stats(vv, s)
    assert st.completed == 0 and st.submitted == 0 and st.launches == 0 and st.p99_ms == 0.0
    fg = vv.SynthFrame(0, 832, 480, seed=9)
    fo = vv.SynthFrame(0, 832, 480, seed=9)
    assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
    t = C.c_uint64()
    vv.check(vv.lib.vvclf_server_submit(s, sid.value, fg.desc, C.byref(t)))
    assert vv.lib.vvclf_server_wait(s, t.value, 60000) == 0, vv.lib.vvclf_last_error()
    st = stats(vv, s)
    assert st.completed == 1 and st.submitted == 1 and st.luma_pixels == 832 * 480 and st.p50_ms > 0.0
    for a, b in zip(fo.output_arrays(), fg.output_arrays()):
        assert np.array_equal(a, b)
    vv.lib.vvclf_server_destroy(s)
    ctx.close()
