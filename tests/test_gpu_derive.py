"""The deblocking descriptors derived on the GPU (one thread per CU record and
pass, vvclf_derive.hip) against the host walk of the same records: every
descriptor map and work list identical; and the GPU path rejects records
that do not tile the picture.

Skipped: the first GPU run of this path ended in a device memory fault
("illegal memory access" during the derivation of the first job) whose cause
has not been found (the same kernels run as host code under AddressSanitizer
match the host walk on every frame of these cases), so the path is switched
off (VVCLF_JOB_GPU_DERIVE returns VVCLF_ERR_UNSUPPORTED) and these tests stay
off the GPU until a cause is shown."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="GPU descriptor derivation switched off after an unexplained device memory "
                                      "fault; see the module docstring")]

TABLES = ["luma V", "luma H", "Cb V", "Cr V", "Cb H", "Cr H", "list luma V", "list luma H", "list chroma V",
          "list chroma H"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [(c, 416, 240, s, {}) for c in range(5) for s in range(2)] + \
        [(0, 256, 128, s, {"rich": 1}) for s in range(8)] + \
        [(1, 832, 480, s, {"rich": 1, "tiles": 1}) for s in range(3)] + \
        [(3, 416, 240, 12, {"tiles": 1, "rich": 1}), (1, 200, 136, 1, {}), (1, 208, 120, 2, {"rich": 1}),
         (0, 64, 64, 3, {}), (0, 8, 8, 4, {}), (1, 416, 240, 5, {"ctb_log2": 6}),
         (1, 416, 240, 6, {"ctb_log2": 5, "rich": 1}), (1, 416, 240, 8, {"bit_depth": 12, "rich": 1}),
         (1, 3840, 2160, 1000, {}), (3, 1920, 1080, 3000, {})]


@pytest.mark.parametrize("cfg,w,h,seed,kw", CASES)
def test_gpu_derivation_matches_host_walk(vv, cfg, w, h, seed, kw):
    say(f"derive cfg={cfg} {w}x{h} seed={seed} {kw}")
    f = vv.SynthFrame(cfg, w, h, seed=seed, **kw)
    ctx = vv.Context()
    jh = vv.Job(ctx, f.desc)
    jg = vv.Job(ctx, f.desc, vv.Job.GPU_DERIVE)
    n_entries = 0
    for k, name in enumerate(TABLES):
        a, b = jh.deblock_table(k), jg.deblock_table(k)
        assert a.shape == b.shape, f"{name}: host {a.shape} gpu {b.shape}"
        if not np.array_equal(a, b):
            i = int(np.argmax(a != b))
            raise AssertionError(f"{name}: {int((a != b).sum())} entries differ, first at {i}: "
                                 f"host {a[i]:#x} gpu {b[i]:#x}")
        if k >= 6:
            n_entries += a.size
    assert n_entries > 0  # the picture has edges to filter
    jh.close()
    jg.close()
    ctx.close()


def test_gpu_derivation_rejects_overlapping_records(vv):
    from helpers import cu, cu_array
    say("derive: overlapping records")
    f = vv.SynthFrame(0, 64, 64, seed=1, sao=0, alf=0)
    d = f.desc.contents
    # two records claim the same vertical edge segments at x = 16
    arr = cu_array(vv, [cu(vv, 0, 0, 16, 64, 30), cu(vv, 16, 0, 48, 64, 30), cu(vv, 16, 0, 16, 64, 30)])
    d.cus = C.cast(arr, C.POINTER(vv.DeblockCu))
    d.n_cus = 3
    d.mv_grid = C.POINTER(vv.MvGrid)()
    ctx = vv.Context()
    h = C.c_void_p()
    assert vv.lib.vvclf_job_create_flags(ctx.h, f.desc, vv.Job.GPU_DERIVE, C.byref(h)) == vv.ERR_INVALID
    assert b"overlap" in vv.lib.vvclf_last_error()
    ctx.close()
