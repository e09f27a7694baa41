"""GPU parity at the benchmark sizes: the fused batched form on 1920x1080 and
3840x2160 pictures, the server on 1080p and 4K streams, and the 7680x4320
ALF-heavy dense-edge workload (every CTB with luma ALF and CC-ALF, 8 APS
sets, minimum CU 4x8) through the drop-in call and a fused batch of four.
Every output is compared with the CPU oracle bit for bit."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_outputs(vv, orc, cases):
    """Oracle results of `cases` ((workload, w, h, seed, overrides)), run on
    several host threads; returns the output arrays of each."""
    frames = [vv.SynthFrame(cfg, w, h, seed=seed, **kw) for cfg, w, h, seed, kw in cases]
    arr = (C.POINTER(vv.Frame) * len(frames))(*[f.desc for f in frames])
    assert orc.oracle_apply_in_loop_filters_mt(arr, len(frames), min(len(frames), os.cpu_count() or 1)) == 0
    outs = [f.output_arrays() for f in frames]
    for f in frames:
        f.close()
    return outs


def assert_same(ref, got, what):
    for name, a, b in zip(("luma", "cb", "cr"), ref, got):
        if not np.array_equal(a, b):
            yx = np.argwhere(a != b)[0]
            raise AssertionError(f"{what}: {name} differs at {(a != b).sum()} samples, first y,x={yx}")


def run_batch(vv, ctx, frames, tile):
    jobs = [vv.Job(ctx, f.desc) for f in frames]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    b = C.c_void_p()
    tune = vv.Tune(tile[0], tile[1], 256, 0) if tile else None
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune) if tune else None, C.byref(b)))
    got_tune = vv.Tune()
    vv.check(vv.lib.vvclf_batch_get_tune(b, C.byref(got_tune)))
    vv.check(vv.lib.vvclf_batch_graph_launch(b, ctx.stream))
    ctx.synchronize()
    outs = []
    for j, f in zip(jobs, frames):
        j.download(C.byref(f.desc.contents.pic))
        outs.append(f.output_arrays())
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    return outs, (got_tune.tile_w, got_tune.tile_h)


LARGE_MIX = [(0, 1920, 1080, 4000, {}), (2, 1920, 1080, 4001, {}), (1, 3840, 2160, 4002, {}),
             (1, 3840, 2160, 4003, {"rich": 1}), (1, 832, 480, 4004, {}), (0, 416, 240, 4005, {})]


@pytest.mark.parametrize("tile", [(128, 64), (64, 64)], ids=["128x64", "64x64"])
def test_batch_1080p_and_4k(vv, orc, tile):
    say(f"fused batch 2x1080p + 2x4K + 2 small, tile {tile}")
    ref = oracle_outputs(vv, orc, LARGE_MIX)
    ctx = vv.Context()
    frames = [vv.SynthFrame(cfg, w, h, seed=seed, **kw) for cfg, w, h, seed, kw in LARGE_MIX]
    outs, used = run_batch(vv, ctx, frames, tile)
    assert used == tile
    for r, g, case in zip(ref, outs, LARGE_MIX):
        assert_same(r, g, f"batch {tile} {case}")
    ctx.close()


def test_server_1080p_and_4k_streams(vv, orc):
    say("server: 2 streams of 1080p, 2 of 4K, two pictures each")
    cases = [(1, 1920, 1080), (0, 1920, 1080), (1, 3840, 2160), (1, 3840, 2160)]
    jobs = [(k, cfg, w, h, 4100 + 10 * k + i) for i in range(2) for k, (cfg, w, h) in enumerate(cases)]
    ref = oracle_outputs(vv, orc, [(cfg, w, h, seed, {"aps_seed": 77 + k}) for k, cfg, w, h, seed in jobs])
    ctx = vv.Context()
    cfgs = vv.ServerConfig()
    cfgs.hbm_budget_bytes = 64 << 30
    cfgs.max_inflight = 16
    cfgs.max_batch = 8
    cfgs.prep_threads = 4
    s = C.c_void_p()
    vv.check(vv.lib.vvclf_server_create(ctx.h, C.byref(cfgs), C.byref(s)))
    sids = []
    for _ in cases:
        i = C.c_uint32()
        vv.check(vv.lib.vvclf_server_open_stream(s, C.byref(i)))
        sids.append(i.value)
    frames, tickets = [], []
    for k, cfg, w, h, seed in jobs:
        f = vv.SynthFrame(cfg, w, h, seed=seed, aps_seed=77 + k)
        t = C.c_uint64()
        vv.check(vv.lib.vvclf_server_submit(s, sids[k], f.desc, C.byref(t)))
        frames.append(f)
        tickets.append(t.value)
    for t in tickets:
        assert vv.lib.vvclf_server_wait(s, t, 120000) == 0, vv.lib.vvclf_last_error()
    for f, r, j in zip(frames, ref, jobs):
        assert_same(r, f.output_arrays(), f"served {j}")
    st = vv.ServerStats()
    vv.check(vv.lib.vvclf_server_get_stats(s, C.byref(st)))
    assert st.completed == len(jobs) and st.failed == 0
    vv.lib.vvclf_server_destroy(s)
    ctx.close()


EIGHT_K = [(3, 7680, 4320, 6000 + i, {}) for i in range(4)]


def test_8k_alf_heavy_dropin(vv, orc):
    say("7680x4320 ALF-heavy dense-edge picture, drop-in call")
    ref = oracle_outputs(vv, orc, EIGHT_K[:1])
    f = vv.SynthFrame(*EIGHT_K[0][:3], seed=EIGHT_K[0][3])
    assert vv.lib.vvclf_apply_in_loop_filters(f.desc) == 0, vv.lib.vvclf_last_error()
    assert_same(ref[0], f.output_arrays(), "8K drop-in")


def test_8k_alf_heavy_batch_of_four(vv, orc):
    say("4x 7680x4320 ALF-heavy dense-edge pictures, one fused launch (autotuned tile)")
    ref = oracle_outputs(vv, orc, EIGHT_K)
    ctx = vv.Context()
    frames = [vv.SynthFrame(cfg, w, h, seed=seed, **kw) for cfg, w, h, seed, kw in EIGHT_K]
    jobs = [vv.Job(ctx, f.desc) for f in frames]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune()
    vv.check(vv.lib.vvclf_autotune(ctx.h, arr, len(jobs), C.byref(tune)))
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    vv.check(vv.lib.vvclf_batch_graph_launch(b, ctx.stream))
    ctx.synchronize()
    for j, f, r, case in zip(jobs, frames, ref, EIGHT_K):
        j.download(C.byref(f.desc.contents.pic))
        assert_same(r, f.output_arrays(), f"8K batch tile {tune.tile_w}x{tune.tile_h} {case}")
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    ctx.close()
