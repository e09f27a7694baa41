"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs, bit-exact (tolerance 0: all arithmetic is integer)."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import Pic, alf_cfg, put_wide_samples, alf_picture, binding_of, cu, cu_array, dparams, sao_band, sao_cfg, sao_edge, sao_picture

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_same(a_planes, b_planes, what):
    for name, a, b in zip(("luma", "cb", "cr"), a_planes, b_planes):
        if not np.array_equal(a, b):
            diff = np.argwhere(a != b)
            y, x = diff[0]
            raise AssertionError(f"{what}: {name} differs at {len(diff)} samples, first (x={x}, y={y}): "
                                 f"gpu {b[y, x]} oracle {a[y, x]}")


def full_chain(vv, orc, cfg, w, h, seed, **kw):
    say(f"case cfg={cfg} {w}x{h} seed={seed} {kw}")
    fo = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
    fg = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
    assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
    rc = vv.lib.vvclf_apply_in_loop_filters(fg.desc)
    assert rc == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), f"cfg {cfg} {w}x{h} seed {seed} {kw}")
    # the chain must actually have changed the picture
    assert any((a != b).any() for a, b in zip(fg.input_arrays(), fg.output_arrays()))


SMALL = [(0, 416, 240, s, {}) for s in range(3)] + \
        [(1, 832, 480, s, {}) for s in range(3)] + \
        [(3, 416, 240, s, {}) for s in range(2)] + \
        [(0, 256, 128, s, {"rich": 1}) for s in range(6)] + \
        [(1, 832, 480, s, {"rich": 1, "tiles": 1}) for s in range(3)] + \
        [(1, 200, 136, 1, {}), (1, 208, 120, 2, {"rich": 1}), (0, 64, 64, 3, {}), (0, 8, 8, 4, {})] + \
        [(1, 416, 240, 5, {"ctb_log2": 6}), (1, 416, 240, 6, {"ctb_log2": 5, "rich": 1}),
         (1, 416, 240, 7, {"bit_depth": 8}), (1, 416, 240, 8, {"bit_depth": 12, "rich": 1}),
         (1, 416, 240, 9, {"sao": 0}), (1, 416, 240, 10, {"alf": 0}), (1, 416, 240, 11, {"lmcs": 0, "sao": 0}),
         (3, 416, 240, 12, {"tiles": 1, "rich": 1}),
         # above 12 bits the ALF runs in 32-bit lanes instead of packed 16-bit pairs
         (1, 416, 240, 14, {"bit_depth": 14, "rich": 1}), (0, 256, 128, 15, {"bit_depth": 16, "tiles": 1})]


@pytest.mark.parametrize("cfg,w,h,seed,kw", SMALL)
def test_full_chain_small(vv, orc, cfg, w, h, seed, kw):
    full_chain(vv, orc, cfg, w, h, seed, **kw)


@pytest.mark.parametrize("cfg,w,h,seed", [(0, 1920, 1080, 0), (1, 3840, 2160, 1000), (3, 1920, 1080, 3000)])
def test_full_chain_full_size(vv, orc, cfg, w, h, seed):
    full_chain(vv, orc, cfg, w, h, seed)


@pytest.mark.parametrize("cfg,w,h,seed,kw", [(1, 416, 240, 31, {}), (0, 256, 128, 32, {"rich": 1}),
                                           (1, 416, 240, 33, {"bit_depth": 12, "rich": 1}),
                                           (1, 208, 120, 34, {"lmcs": 0})])
def test_samples_beyond_the_bit_depth(vv, orc, cfg, w, h, seed, kw):
    """Samples of 2^15 and more (which the 16-bit ALF lanes would wrap) in
    every plane: the drop-in call still matches the oracle."""
    say(f"wide samples cfg={cfg} {w}x{h} seed={seed} {kw}")
    fo = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
    fg = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
    put_wide_samples(vv, [fo, fg], seed)
    assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
    assert vv.lib.vvclf_apply_in_loop_filters(fg.desc) == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), f"wide samples cfg {cfg} {w}x{h} {kw}")


def test_stages_standalone(vv, orc):
    """Each drop-in stage entry point against the oracle's same stage."""
    for seed in range(3):
        say(f"standalone stages seed={seed}")
        fo = vv.SynthFrame(1, 416, 240, seed=seed, rich=1, tiles=1)
        fg = vv.SynthFrame(1, 416, 240, seed=seed, rich=1, tiles=1)
        do, dg = fo.desc.contents, fg.desc.contents
        # LMCS
        orc.oracle_lmcs_inverse_luma(C.byref(do.pic.luma), C.byref(do.lmcs))
        assert vv.lib.vvclf_lmcs_inverse_luma(C.byref(dg.pic.luma), C.byref(dg.lmcs)) == 0
        assert_same(fo.output_arrays(), fg.output_arrays(), "lmcs")
        # deblock
        orc.oracle_apply_deblocking_clipped(C.byref(do.pic), do.cus, do.n_cus, do.chroma_cus, do.n_chroma_cus,
                                            do.mv_grid, C.byref(do.deblock), 1, do.no_filter_cols,
                                            do.n_no_filter_cols, do.no_filter_rows, do.n_no_filter_rows)
        rc = vv.lib.vvclf_apply_deblocking_clipped(C.byref(dg.pic), dg.cus, dg.n_cus, dg.chroma_cus, dg.n_chroma_cus,
                                                   dg.mv_grid, C.byref(dg.deblock), 1, dg.no_filter_cols,
                                                   dg.n_no_filter_cols, dg.no_filter_rows, dg.n_no_filter_rows)
        assert rc == 0, vv.lib.vvclf_last_error()
        assert_same(fo.output_arrays(), fg.output_arrays(), "deblock")
        # SAO
        orc.oracle_apply_sao_clipped(C.byref(do.pic), C.byref(do.sao), C.byref(do.sao_cfg), do.clip_bounds)
        assert vv.lib.vvclf_apply_sao_clipped(C.byref(dg.pic), C.byref(dg.sao), C.byref(dg.sao_cfg),
                                              dg.clip_bounds) == 0
        assert_same(fo.output_arrays(), fg.output_arrays(), "sao")
        # ALF
        assert orc.oracle_apply_alf_clipped(C.byref(do.pic), C.byref(do.alf), C.byref(do.alf_cfg),
                                            C.byref(do.alf_binding), do.clip_bounds) == 0
        assert vv.lib.vvclf_apply_alf_clipped(C.byref(dg.pic), C.byref(dg.alf), C.byref(dg.alf_cfg),
                                              C.byref(dg.alf_binding), dg.clip_bounds) == 0
        assert_same(fo.output_arrays(), fg.output_arrays(), "alf")


@pytest.mark.parametrize("w,h", [(3840, 2160), (3848, 2168)])
def test_stages_standalone_banded(vv, orc, w, h):
    """The drop-in stage calls at sizes whose sample sets (16 MB and more) and
    side tables move in per-thread bands: the LMCS call's luma-only set, and
    a width whose rows are padded to the pitch, so that band borders fall
    inside and between the planes."""
    say(f"standalone stages {w}x{h}, banded transfers")
    fo = vv.SynthFrame(1, w, h, seed=21, rich=1, tiles=1)
    fg = vv.SynthFrame(1, w, h, seed=21, rich=1, tiles=1)
    do, dg = fo.desc.contents, fg.desc.contents
    orc.oracle_lmcs_inverse_luma(C.byref(do.pic.luma), C.byref(do.lmcs))
    assert vv.lib.vvclf_lmcs_inverse_luma(C.byref(dg.pic.luma), C.byref(dg.lmcs)) == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), "lmcs")
    orc.oracle_apply_deblocking_clipped(C.byref(do.pic), do.cus, do.n_cus, do.chroma_cus, do.n_chroma_cus,
                                        do.mv_grid, C.byref(do.deblock), 1, do.no_filter_cols,
                                        do.n_no_filter_cols, do.no_filter_rows, do.n_no_filter_rows)
    rc = vv.lib.vvclf_apply_deblocking_clipped(C.byref(dg.pic), dg.cus, dg.n_cus, dg.chroma_cus, dg.n_chroma_cus,
                                               dg.mv_grid, C.byref(dg.deblock), 1, dg.no_filter_cols,
                                               dg.n_no_filter_cols, dg.no_filter_rows, dg.n_no_filter_rows)
    assert rc == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), "deblock")
    orc.oracle_apply_sao_clipped(C.byref(do.pic), C.byref(do.sao), C.byref(do.sao_cfg), do.clip_bounds)
    assert vv.lib.vvclf_apply_sao_clipped(C.byref(dg.pic), C.byref(dg.sao), C.byref(dg.sao_cfg),
                                          dg.clip_bounds) == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), "sao")
    assert orc.oracle_apply_alf_clipped(C.byref(do.pic), C.byref(do.alf), C.byref(do.alf_cfg),
                                        C.byref(do.alf_binding), do.clip_bounds) == 0
    assert vv.lib.vvclf_apply_alf_clipped(C.byref(dg.pic), C.byref(dg.alf), C.byref(dg.alf_cfg),
                                          C.byref(dg.alf_binding), dg.clip_bounds) == 0, vv.lib.vvclf_last_error()
    assert_same(fo.output_arrays(), fg.output_arrays(), "alf")
    # and the whole chain in one call on a fresh copy
    fo.close()
    fg.close()
    full_chain(vv, orc, 1, w, h, 22, rich=1)


@pytest.mark.parametrize("bd,above", [(10, False), (10, True), (8, True), (12, True)])
def test_alf_unclipped_taps(vv, orc, bd, above):
    """ALF luma where the filter sets clip nothing (every fixed set; an APS
    without clipping): taps whose clip value exceeds the bit depth take the
    unclipped form only while the window's samples stay within the bit depth.
    `above` puts samples between 2^BitDepth and 2^15 into the picture, where
    the clip still acts."""
    say(f"alf unclipped taps bd={bd} above={above}")
    rng = np.random.default_rng(bd * 2 + above)
    W, H, ctb_log2 = 416, 240, 6
    pic = Pic(vv, W, H, 0, bd=bd)
    pic.y[:] = rng.integers(0, 1 << bd, (H, W))
    if above:
        n = W * H // 100
        pic.y.ravel()[rng.choice(W * H, n, replace=False)] = rng.integers(1 << bd, 1 << 15, n)
    aps = vv.AlfAps()
    aps.alf_luma_filter_signal_flag = 1
    aps.n_luma_coeff = 25
    for k in range(25):
        for t in range(12):
            aps.luma_coeff[k][t] = int(rng.integers(-20, 21))
    b = binding_of(vv, luma=[aps])
    ctb = 1 << ctb_log2
    wc, hc = (W + ctb - 1) // ctb, (H + ctb - 1) // ctb
    ent = {}
    for y in range(hc):
        for x in range(wc):
            c = vv.AlfCtb()
            c.luma_on, c.luma_filt_set_idx = 1, int(rng.integers(0, 17))
            ent[(x, y)] = c
    ap, _k = alf_picture(vv, wc, hc, ent)
    cfg = alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)
    o, g = pic.copy(), pic.copy()
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(cfg), C.byref(b), None) == 0
    assert vv.lib.vvclf_apply_alf_clipped(g.ptr(), C.byref(ap), C.byref(cfg), C.byref(b), None) == 0, \
        vv.lib.vvclf_last_error()
    assert np.array_equal(o.y, g.y)
    assert (g.y != pic.y).any()


def test_reference_kats_through_the_gpu(vv):
    """The reference's exact-value deblocking tests, run on the GPU path."""
    say("kat: weak p1 clip")
    pic = Pic(vv, 32, 8, 128)
    pic.y[:, :16] = 140
    pic.y[:, 16:] = 100
    arr = cu_array(vv, [cu(vv, 0, 0, 16, 8, 45), cu(vv, 16, 0, 16, 8, 45)])
    assert vv.lib.vvclf_apply_deblocking(pic.ptr(), arr, 2, C.byref(dparams(vv)), 1) == 0
    assert [pic.y[2, 15], pic.y[2, 16], pic.y[2, 14], pic.y[2, 17]] == [127, 113, 134, 106]
    say("kat: long 7/3")
    pic = Pic(vv, 48, 8, 128)
    pic.y[:, :32] = 100
    pic.y[:, 32:] = 104
    arr = cu_array(vv, [cu(vv, 0, 0, 32, 8, 45), cu(vv, 32, 0, 16, 8, 45)])
    assert vv.lib.vvclf_apply_deblocking(pic.ptr(), arr, 2, C.byref(dparams(vv)), 1) == 0
    assert list(pic.y[3, 24:36]) == [100, 100, 100, 101, 101, 101, 102, 102, 102, 103, 104, 104]
    say("kat: ctb row 3/7")
    pic = Pic(vv, 8, 128, 128)
    pic.y[:64] = 100
    pic.y[64:] = 104
    arr = cu_array(vv, [cu(vv, 0, 0, 8, 64, 45), cu(vv, 0, 64, 8, 64, 45)])
    assert vv.lib.vvclf_apply_deblocking(pic.ptr(), arr, 2, C.byref(dparams(vv, ctb_log2=6)), 1) == 0
    assert list(pic.y[57:71, 2]) == [100, 100, 100, 100, 100, 101, 102, 102, 102, 103, 103, 103, 104, 104]
    say("kat: sao edge pattern")
    pic = Pic(vv, 32, 2, 0)
    pic.y[:, :] = [100 if x % 2 == 0 else 110 for x in range(32)]
    p = vv.SaoCtbParams()
    p.luma = sao_edge(vv, 0, [5, 0, 0, 2])
    sp, _k = sao_picture(vv, 1, 1, {(0, 0): p})
    assert vv.lib.vvclf_apply_sao(pic.ptr(), C.byref(sp), C.byref(sao_cfg(vv))) == 0
    assert list(pic.y[0, :6]) == [100, 108, 105, 108, 105, 108] and pic.y[0, 31] == 110


def test_job_runs_are_repeatable(vv, orc):
    """A job keeps its input: running it again gives the same output, which
    equals the drop-in result."""
    say("job repeatability")
    ctx = vv.Context()
    f = vv.SynthFrame(1, 832, 480, seed=77)
    fo = vv.SynthFrame(1, 832, 480, seed=77)
    orc.oracle_apply_in_loop_filters(fo.desc)
    job = vv.Job(ctx, f.desc)
    outs = []
    for _ in range(3):
        job.run()
        job.download(C.byref(f.desc.contents.pic))
        outs.append(f.output_arrays())
    # the captured-graph form of the same sequence, replayed
    for _ in range(3):
        job.graph_launch()
        job.download(C.byref(f.desc.contents.pic))
        outs.append(f.output_arrays())
    for o in outs:
        assert_same(fo.output_arrays(), o, "job run")
    job.close()
    ctx.close()
