"""GPU parity of the batched fused form: many pictures of uneven sizes in one
persistent launch (and in a replayed HIP graph), each bit-exact against the
CPU oracle, and each identical alone and in the mix."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [(0, 416, 240, 1, {}), (1, 832, 480, 2, {}), (3, 416, 240, 3, {}), (1, 208, 120, 4, {"rich": 1}),
       (1, 832, 480, 5, {"rich": 1, "tiles": 1}), (0, 200, 136, 6, {}), (1, 256, 128, 7, {"rich": 1}),
       (1, 416, 240, 8, {"bit_depth": 8}), (1, 416, 240, 9, {"sao": 0}), (1, 416, 240, 10, {"alf": 0}),
       (3, 416, 240, 11, {"tiles": 1, "rich": 1}), (0, 64, 64, 12, {}), (1, 416, 240, 13, {"bit_depth": 12})]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(vv, orc, cases):
    gpu, ref = [], []
    for cfg, w, h, seed, kw in cases:
        fg = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        fo = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
        gpu.append(fg)
        ref.append(fo)
    return gpu, ref


def check_all(vv, jobs, gpu, ref, what, cases=MIX):
    for job, fg, fo, case in zip(jobs, gpu, ref, cases):
        job.download(C.byref(fg.desc.contents.pic))
        for name, a, b in zip(("luma", "cb", "cr"), fo.output_arrays(), fg.output_arrays()):
            if not np.array_equal(a, b):
                yx = np.argwhere(a != b)[0]
                raise AssertionError(f"{what} {case}: {name} differs at {(a != b).sum()} samples, first y,x={yx}")


@pytest.mark.parametrize("tw,th", [(64, 64), (128, 64), (64, 32), (32, 32)])
def test_batch_matches_oracle(vv, orc, tw, th):
    say(f"batch tile {tw}x{th}")
    gpu, ref = make(vv, orc, MIX)
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune(tw, th, 256, 0)
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    vv.check(vv.lib.vvclf_batch_run(b, ctx.stream))
    ctx.synchronize()
    check_all(vv, jobs, gpu, ref, f"batch {tw}x{th}")
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    ctx.close()


# A batch holding a job above 12 bits runs the fused kernel's 32-bit ALF
# variant for all of its jobs.
HBD_MIX = [(1, 416, 240, 21, {"bit_depth": 16}), (0, 256, 128, 22, {"bit_depth": 14, "rich": 1}),
           (1, 208, 120, 23, {}), (3, 416, 240, 24, {"tiles": 1, "bit_depth": 16})]


WIDE = [(1, 416, 240, 21, {}), (0, 256, 128, 22, {"rich": 1}), (1, 416, 240, 23, {"bit_depth": 12, "rich": 1}),
        (1, 208, 120, 24, {"lmcs": 0})]


@pytest.mark.parametrize("tw,th", [(128, 64), (64, 64)])
def test_batch_samples_beyond_the_bit_depth(vv, orc, tw, th):
    """Samples of 2^15 and more in every plane: tiles holding one filter in
    32-bit lanes, and every picture still matches the oracle."""
    from helpers import put_wide_samples
    say(f"batch wide samples tile {tw}x{th}")
    gpu, ref = [], []
    for cfg, w, h, seed, kw in WIDE:
        fg = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        fo = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        put_wide_samples(vv, [fg, fo], seed)
        assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
        gpu.append(fg)
        ref.append(fo)
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune(tw, th, 256, 0)
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    vv.check(vv.lib.vvclf_batch_run(b, ctx.stream))
    ctx.synchronize()
    check_all(vv, jobs, gpu, ref, f"batch wide samples {tw}x{th}", cases=WIDE)
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    ctx.close()


@pytest.mark.parametrize("tw,th", [(64, 64), (32, 32)])
def test_batch_high_bit_depth(vv, orc, tw, th):
    say(f"batch >12-bit tile {tw}x{th}")
    gpu, ref = make(vv, orc, HBD_MIX)
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune(tw, th, 256, 0)
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    vv.check(vv.lib.vvclf_batch_run(b, ctx.stream))
    ctx.synchronize()
    check_all(vv, jobs, gpu, ref, f"batch >12-bit {tw}x{th}", HBD_MIX)
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    ctx.close()


def test_graph_replays_and_alone_vs_mix(vv, orc):
    say("graph replay + alone vs mix")
    gpu, ref = make(vv, orc, MIX)
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    tune = vv.Tune(64, 64, 256, 0)
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), C.byref(tune), C.byref(b)))
    for rep in range(3):
        vv.check(vv.lib.vvclf_batch_graph_launch(b, ctx.stream))
        ctx.synchronize()
        check_all(vv, jobs, gpu, ref, f"graph replay {rep}")
        # eager plain runs between replays must not disturb the graph's result
        jobs[rep].run(ctx.stream)
        ctx.synchronize()
    vv.lib.vvclf_batch_destroy(b)
    # each picture alone in its own batch gives the same samples as in the mix
    for i, job in enumerate(jobs[:5]):
        one = (C.c_void_p * 1)(job.h)
        b1 = C.c_void_p()
        vv.check(vv.lib.vvclf_batch_create(ctx.h, one, 1, C.byref(tune), C.byref(b1)))
        vv.check(vv.lib.vvclf_batch_run(b1, ctx.stream))
        ctx.synchronize()
        job.download(C.byref(gpu[i].desc.contents.pic))
        for a, c in zip(ref[i].output_arrays(), gpu[i].output_arrays()):
            assert np.array_equal(a, c)
        vv.lib.vvclf_batch_destroy(b1)
    for j in jobs:
        j.close()
    ctx.close()


def test_autotune_picks_a_valid_shape(vv, orc):
    say("autotune")
    gpu, ref = make(vv, orc, MIX[:4])
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    t = vv.Tune()
    vv.check(vv.lib.vvclf_autotune(ctx.h, arr, len(jobs), C.byref(t)))
    assert (t.tile_w, t.tile_h) in [(64, 64), (128, 64), (64, 32), (32, 32)] and t.wgs_per_cu >= 1
    b = C.c_void_p()
    vv.check(vv.lib.vvclf_batch_create(ctx.h, arr, len(jobs), None, C.byref(b)))
    vv.check(vv.lib.vvclf_batch_run(b, ctx.stream))
    ctx.synchronize()
    for job, fg, fo in zip(jobs, gpu, ref):
        job.download(C.byref(fg.desc.contents.pic))
        for a, c in zip(fo.output_arrays(), fg.output_arrays()):
            assert np.array_equal(a, c)
    vv.lib.vvclf_batch_destroy(b)
    for j in jobs:
        j.close()
    ctx.close()


@pytest.mark.parametrize("lanes", [1, 2, 3])
def test_jobs_graph_lanes_match_oracle(vv, orc, lanes):
    """The plain sequences of many jobs in one graph of parallel lanes: every
    picture equals the oracle on each of three replays, with eager plain runs
    of one job between replays."""
    say(f"jobs graph lanes={lanes}")
    gpu, ref = make(vv, orc, MIX)
    ctx = vv.Context()
    jobs = [vv.Job(ctx, f.desc) for f in gpu]
    g = vv.JobsGraph(ctx, jobs, lanes)
    for rep in range(3):
        g.launch(ctx.stream)
        ctx.synchronize()
        check_all(vv, jobs, gpu, ref, f"jobs graph lanes {lanes} replay {rep}")
        jobs[rep].run(ctx.stream)
        ctx.synchronize()
    g.close()
    for j in jobs:
        j.close()
    ctx.close()


def test_jobs_graph_rejects_bad_arguments(vv):
    say("jobs graph: bad arguments")
    ctx = vv.Context()
    h = C.c_void_p()
    assert vv.lib.vvclf_jobs_graph_create(ctx.h, None, 1, 2, C.byref(h)) == vv.ERR_INVALID
    assert vv.lib.vvclf_jobs_graph_create(ctx.h, (C.c_void_p * 1)(None), 1, 2, C.byref(h)) == vv.ERR_INVALID
    assert vv.lib.vvclf_jobs_graph_create(ctx.h, (C.c_void_p * 1)(None), 0, 0, C.byref(h)) == vv.ERR_INVALID
    ctx.close()
