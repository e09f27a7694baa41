"""GPU parity of the reconstruction-side steps against the oracle, bit for
bit: LMCS forward mapping over inter-CU rectangles (src/ctu.rs:1701-1725),
chroma residual scaling per chroma TB (src/ctu.rs:1752-1815, :734-742), and
the MD5 of a job's cropped output against hashlib."""
import ctypes as C
import hashlib
import sys

import numpy as np
import pytest

from helpers import Pic
from recon_cases import crs_blocks, lmcs_binding, quadtree_rects

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("bd,w,h,seed", [(10, 416, 240, 1), (8, 832, 480, 2), (12, 256, 128, 3), (10, 1920, 1080, 4),
                                         (16, 200, 136, 5)])
def test_forward_luma_matches_oracle(vv, orc, bd, w, h, seed):
    say(f"lmcs forward bd={bd} {w}x{h} seed={seed}")
    rng = np.random.default_rng(seed)
    b, _k = lmcs_binding(vv, seed, bd)
    tiles = quadtree_rects(rng, w, h)
    inter = [t for t in tiles if rng.random() < 0.6]
    rects = (vv.Rect * len(inter))(*[vv.Rect(*t) for t in inter])
    g = Pic(vv, w, h, 0, bd=bd)
    g.y[:] = rng.integers(0, 1 << bd, (h, w))
    inp = g.y.copy()
    o = g.copy()
    g.ptr()
    assert vv.lib.vvclf_lmcs_forward_luma(C.byref(g._pic.luma), C.byref(b), rects, len(inter)) == 0, \
        vv.lib.vvclf_last_error()
    o.ptr()
    for t in inter:
        assert orc.oracle_lmcs_forward_luma_rect(C.byref(o._pic.luma), C.byref(b), *t) == 0
    assert np.array_equal(g.y, o.y)
    assert (g.y != inp).any() and len(inter) > 0


def test_forward_luma_rejects_samples_beyond_the_bit_depth(vv):
    say("lmcs forward: sample beyond the bit depth")
    b, _k = lmcs_binding(vv, 7, 10)
    g = Pic(vv, 64, 64, 500, bd=10)
    g.y[40, 40] = 1024
    before = g.y.copy()
    rects = (vv.Rect * 2)(vv.Rect(0, 0, 32, 32), vv.Rect(32, 32, 32, 32))
    g.ptr()
    assert vv.lib.vvclf_lmcs_forward_luma(C.byref(g._pic.luma), C.byref(b), rects, 2) == vv.ERR_INVALID
    assert (g.y == before).all()


@pytest.mark.parametrize("bd,w,h,seed,size_y", [(10, 416, 240, 11, 64), (8, 256, 128, 12, 32),
                                                (12, 832, 480, 13, 64), (10, 1920, 1080, 14, 32)])
def test_chroma_residual_matches_oracle(vv, orc, bd, w, h, seed, size_y):
    say(f"chroma residual scaling bd={bd} {w}x{h} seed={seed}")
    rng = np.random.default_rng(seed)
    b, _k = lmcs_binding(vv, seed, bd)
    g = Pic(vv, w, h, 0, bd=bd)
    g.y[:] = rng.integers(0, 1 << bd, (h, w))
    g.cb[:] = rng.integers(0, 1 << bd, g.cb.shape)
    g.cr[:] = rng.integers(0, 1 << bd, g.cr.shape)
    o = g.copy()
    blocks, n_res = crs_blocks(vv, rng, w // 2, h // 2, w, h, size_y)
    # residuals across and beyond the eq. 1219 clamp
    res = rng.integers(-(2 << bd), 2 << bd, n_res).astype(np.int32)
    res[rng.random(n_res) < 0.7] //= 64
    rp = res.ctypes.data_as(C.POINTER(C.c_int32))
    rc = vv.lib.vvclf_lmcs_chroma_residual(g.ptr(), C.byref(b), blocks, len(blocks), rp, n_res)
    assert rc == 0, vv.lib.vvclf_last_error()
    assert orc.oracle_lmcs_chroma_residual(o.ptr(), C.byref(b), blocks, len(blocks), rp) == 0
    assert np.array_equal(g.y, o.y) and np.array_equal(g.cb, o.cb) and np.array_equal(g.cr, o.cr)


@pytest.mark.parametrize("bd", [8, 10])
def test_job_output_md5_matches_hashlib(vv, bd):
    say(f"job output md5 bd={bd}")
    f = vv.SynthFrame(1, 832, 480, seed=5, bit_depth=bd)
    ctx = vv.Context()
    job = vv.Job(ctx, f.desc)
    job.run(ctx.stream)
    ctx.synchronize()
    job.download(C.byref(f.desc.contents.pic))
    y, cb, cr = f.output_arrays()
    crop = vv.Crop(8, 4, 800, 470)
    out = ((C.c_uint8 * 16) * 3)()
    assert vv.lib.vvclf_job_output_md5(job.h, bd, C.byref(crop), C.byref(out)) == 0, vv.lib.vvclf_last_error()

    def ref(a, x, yy, ww, hh):
        sub = a[yy:yy + hh, x:x + ww]
        return hashlib.md5(sub.astype("<u2").tobytes() if bd > 8 else (sub & 0xff).astype(np.uint8).tobytes()).hexdigest()
    want = [ref(y, 8, 4, 800, 470), ref(cb, 4, 2, 400, 235), ref(cr, 4, 2, 400, 235)]
    assert [bytes(out[k]).hex() for k in range(3)] == want
    host = ((C.c_uint8 * 16) * 3)()
    assert vv.lib.vvclf_picture_md5(C.byref(f.desc.contents.pic), 1, bd, C.byref(crop), C.byref(host)) == 0
    assert [bytes(host[k]).hex() for k in range(3)] == want
    job.close()
    ctx.close()
