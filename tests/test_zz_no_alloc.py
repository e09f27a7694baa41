"""The allocation-free driver entry point (vvclf_apply_in_loop_filters_noalloc)
against the allocating one (vvclf_apply_in_loop_filters), through the C ABI,
bit for bit: two calls in a row, the second on a smaller picture (the scratch
then holds the first call's larger picture), and a third at the first size
again."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(1, 3840, 2160, 1000, {}), (1, 832, 480, 7, {"rich": 1, "tiles": 1}), (3, 1920, 1080, 3001, {})]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_noalloc_matches_the_allocating_entry_point(vv, orc):
    for cfg, w, h, seed, kw in CASES:
        say(f"noalloc case cfg={cfg} {w}x{h} seed={seed} {kw}")
        fa = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        fn = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
        assert vv.lib.vvclf_apply_in_loop_filters(fa.desc) == 0, vv.lib.vvclf_last_error()
        assert vv.lib.vvclf_apply_in_loop_filters_noalloc(fn.desc) == 0, vv.lib.vvclf_last_error()
        for name, a, b in zip(("luma", "cb", "cr"), fa.output_arrays(), fn.output_arrays()):
            assert np.array_equal(a, b), f"{name} differs at {(a != b).sum()} samples ({w}x{h})"
        assert any((a != b).any() for a, b in zip(fn.input_arrays(), fn.output_arrays()))
    # the last (smaller than the first) also against the oracle
    fo = vv.SynthFrame(*CASES[-1][:3], seed=CASES[-1][3])
    assert orc.oracle_apply_in_loop_filters(fo.desc) == 0
    for a, b in zip(fo.output_arrays(), fn.output_arrays()):
        assert np.array_equal(a, b)


def test_noalloc_rejects_pictures_beyond_the_scratch(vv):
    f = vv.SynthFrame(3, 7680, 4320, seed=1)
    before = f.output_arrays()
    assert vv.lib.vvclf_apply_in_loop_filters_noalloc(f.desc) == vv.ERR_UNSUPPORTED
    assert b"scratch" in vv.lib.vvclf_last_error()
    for a, b in zip(before, f.output_arrays()):
        assert np.array_equal(a, b)
