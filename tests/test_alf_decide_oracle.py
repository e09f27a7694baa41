"""The oracle's per-CTB luma ALF decision (src/alf_enc.rs:185-316 without an
APS, :352-470 with APS sets) checked against what fixes it independently of
its own code: per-CTB squared errors recomputed in numpy from pictures the
(separately pinned) ALF stage filters with one candidate everywhere, the
first-minimum rule (off first, strict <), replay of the recorded decision
through the ALF stage, and the reference's own encoder tests
(src/alf_enc.rs:1159-1545) restated on the oracle."""
import ctypes as C

import numpy as np
import pytest

from alf_decide_cases import (cc_aps, chroma_aps, decide, decide_chroma, luma_aps, noisy, random_aps,
                              smooth_gradient, tile_bounds)
from helpers import alf_cfg, alf_picture, binding_of


def ctb_sse(src, rec, ctb):
    """Per-CTB luma SSE over the CTB clipped to the picture (ctb_sse_y,
    src/alf_enc.rs:79-113), raster order."""
    h, w = src.y.shape
    d = (src.y.astype(np.int64) - rec.y.astype(np.int64)) ** 2
    out = []
    for y0 in range(0, h, ctb):
        for x0 in range(0, w, ctb):
            out.append(int(d[y0:y0 + ctb, x0:x0 + ctb].sum()))
    return np.array(out, np.uint64)


def filtered_everywhere(orc, vv, rec, ctb_log2, bd, idx, apses, tb):
    """rec with every CTB's luma filtered by candidate idx (0..15 fixed,
    16 + i APS slot i), through the oracle's ALF stage."""
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    wc, hc = (w + ctb - 1) // ctb, (h + ctb - 1) // ctb
    c = vv.AlfCtb()
    c.luma_on, c.luma_filt_set_idx = 1, idx
    ap, _k = alf_picture(vv, wc, hc, {(x, y): c for y in range(hc) for x in range(wc)})
    o = rec.copy()
    b = binding_of(vv, luma=list(apses)) if apses else vv.AlfApsBinding()
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                        C.byref(b), C.byref(tb) if tb is not None else None) == 0
    return o


def replay(orc, vv, rec, choice, ctb_log2, bd, apses, tb):
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    wc, hc = (w + ctb - 1) // ctb, (h + ctb - 1) // ctb
    ent = {}
    for i, t in enumerate(choice):
        c = vv.AlfCtb()
        if t >= 0:
            c.luma_on, c.luma_filt_set_idx = 1, int(t)
        ent[(i % wc, i // wc)] = c
    ap, _k = alf_picture(vv, wc, hc, ent)
    o = rec.copy()
    b = binding_of(vv, luma=list(apses)) if apses else vv.AlfApsBinding()
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(alf_cfg(vv, bd=bd, ctb_log2=ctb_log2)),
                                        C.byref(b), C.byref(tb) if tb is not None else None) == 0
    return o


def first_min(row):
    """Off (column 0) first, then the candidates in order; a candidate is
    taken only when strictly lower than everything before it."""
    k = int(np.argmin(row))  # numpy returns the first minimum
    return k - 1


CASES = [
    # w, h, ctb_log2, bd, n random APS, unsignalled slot, tile bounds
    (200, 136, 6, 10, 2, True, None),
    (256, 128, 7, 8, 0, False, None),
    (160, 96, 5, 12, 1, False, ([0, 96, 160], [0, 64, 96])),
    (96, 72, 4, 8, 1, True, ([0, 48, 96], [0, 32, 72])),
]


@pytest.mark.parametrize("w,h,ctb_log2,bd,n_aps,unsig,tbs", CASES)
def test_table_and_choice_from_independent_trials(orc, vv, w, h, ctb_log2, bd, n_aps, unsig, tbs):
    rng = np.random.default_rng(w * 7 + h)
    src = smooth_gradient(vv, w, h, bd)
    rec = noisy(src, 0x9E3779B1, (1 << (bd - 5)) - 1, 1 << (bd - 7))
    apses = [random_aps(vv, rng, 16) for _ in range(n_aps)]
    if unsig:
        apses.append(luma_aps(vv, {6: 40}, signal=False))
    tb = tile_bounds(vv, *tbs) if tbs else None
    rc, choice, sse, out = decide(vv, orc.oracle_alf_decide_luma, src, rec, ctb_log2, bd, apses, tb)
    assert rc == 0
    ctb = 1 << ctb_log2
    cols = [ctb_sse(src, rec, ctb)]
    trials = []
    for idx in range(16 + len(apses)):
        t = filtered_everywhere(orc, vv, rec, ctb_log2, bd, idx, apses, tb)
        trials.append(t)
        cols.append(ctb_sse(src, t, ctb))
    want = np.stack(cols, axis=1)
    assert np.array_equal(sse, want)
    if unsig:
        assert np.array_equal(sse[:, -1], sse[:, 0])
    assert [first_min(r) for r in want] == choice.tolist()
    assert (choice >= 0).any()
    # the committed samples: per CTB the chosen trial, else the input
    wc = (w + ctb - 1) // ctb
    for i, t in enumerate(choice):
        y0, x0 = (i // wc) * ctb, (i % wc) * ctb
        ref = (trials[t] if t >= 0 else rec).y[y0:y0 + ctb, x0:x0 + ctb]
        assert np.array_equal(out.y[y0:y0 + ctb, x0:x0 + ctb], ref)
    assert np.array_equal(out.cb, rec.cb) and np.array_equal(out.cr, rec.cr)


def test_identity_and_duplicate_aps_tie_rules(orc, vv):
    # a signalled all-zero APS filters to the input exactly, so its error
    # equals the off error and the strict comparison keeps the CTB off; two
    # identical winning APS sets tie and the first one is kept
    bd, w, h = 10, 256, 192
    src = flat_picture(vv, w, h, bd)
    rec = noisy(src, 0x2545F491, 63, 32)
    smooth = {k: 5 for k in range(12)}  # 24 neighbours at 5/128: close to their plain mean
    strong = luma_aps(vv, smooth)
    ident = luma_aps(vv)
    rc, choice, sse, _o = decide(vv, orc.oracle_alf_decide_luma, src, rec, 6, bd, [ident, strong, luma_aps(vv, smooth)])
    assert rc == 0
    assert np.array_equal(sse[:, 17], sse[:, 0])
    assert np.array_equal(sse[:, 18], sse[:, 19])
    assert (choice != 16).all() and (choice != 18).all()
    assert (choice == 17).any(), "the strong low-pass set should win on white noise"


def flat_picture(vv, w, h, bd):
    p = smooth_gradient(vv, w, h, bd)
    p.y[:] = 1 << (bd - 1)
    return p


def test_rejects_short_aps_and_bad_arguments(orc, vv):
    src = smooth_gradient(vv, 64, 64, 8)
    short = luma_aps(vv, {6: 8})
    short.n_luma_coeff = 5
    rc, *_ = decide(vv, orc.oracle_alf_decide_luma, src, src, 6, 8, [short])
    assert rc == vv.ERR_INVALID


# -------------------------------------------------- the reference's encoder tests
def test_geometry(orc, vv):
    # src/alf_enc.rs:1159-1170
    src = smooth_gradient(vv, 128, 128)
    rec = src.copy()
    rec.y[0, 0] ^= 0x10
    rc, choice, sse, _o = decide(vv, orc.oracle_alf_decide_luma, src, rec, 7, 8)
    assert rc == 0 and choice.shape == (1,) and sse.shape == (1, 17)


def test_flat_source_does_not_hurt(orc, vv):
    # src/alf_enc.rs:1172-1206
    src = flat_picture(vv, 128, 128, 8)
    rec = src.copy()
    rec.y.ravel()[10] ^= 0x40
    rc, _c, _s, out = decide(vv, orc.oracle_alf_decide_luma, src, rec, 7, 8)
    assert rc == 0
    assert ctb_sse(src, out, 128).sum() <= ctb_sse(src, rec, 128).sum()


def test_multi_set_dominates_set_zero(orc, vv):
    # src/alf_enc.rs:1214-1295
    src = smooth_gradient(vv, 128, 128)
    rec = src.copy()
    v = rec.y.astype(np.int64).ravel()
    n = ((np.arange(v.size, dtype=np.uint64) * np.uint64(2654435761)) & np.uint64(0x1f)).astype(np.int64)
    rec.y[:] = ((v + n - 0x10) & 0xffff).reshape(rec.y.shape)
    set0 = filtered_everywhere(orc, vv, rec, 7, 8, 0, [], None)
    pre, post0 = ctb_sse(src, rec, 128), ctb_sse(src, set0, 128)
    sse_a = int(min(pre[0], post0[0]))
    rc, _c, _s, out = decide(vv, orc.oracle_alf_decide_luma, src, rec, 7, 8)
    assert rc == 0 and int(ctb_sse(src, out, 128).sum()) <= sse_a


def test_picks_a_set_when_better(orc, vv):
    # src/alf_enc.rs:1296-1337
    src = smooth_gradient(vv, 256, 128)
    rec = src.copy()
    yy, xx = np.mgrid[0:128, 0:256]
    left = (xx < 128) & (((xx ^ yy) & 1) == 1)
    rec.y[left] += 2
    right = xx >= 128
    rec.y[right] += (((xx + yy) >> 4) & 7)[right].astype(np.uint16)
    rc, choice, _s, _o = decide(vv, orc.oracle_alf_decide_luma, src, rec, 7, 8)
    assert rc == 0 and choice.shape == (2,)
    assert (choice >= 0).any() and (choice < 16).all()


def test_with_aps_never_increases_sse(orc, vv):
    # src/alf_enc.rs:1389-1423
    src = smooth_gradient(vv, 128, 128)
    rec = noisy(src, 0xc6a4a793, 0x0f)
    rc, _c, _s, out = decide(vv, orc.oracle_alf_decide_luma, src, rec, 7, 8, [luma_aps(vv, {6: 8, 2: 4})])
    assert rc == 0 and ctb_sse(src, out, 128).sum() <= ctb_sse(src, rec, 128).sum()


@pytest.mark.parametrize("apses", [[], [{6: 4}]])
def test_decision_replays_to_same_samples(orc, vv, apses):
    # src/alf_enc.rs:1338-1372 and :1496-1539
    src = smooth_gradient(vv, 128, 128)
    pre = noisy(src, 0xdeadbeef, 0x07)
    sets = [luma_aps(vv, a) for a in apses]
    rc, choice, _s, out = decide(vv, orc.oracle_alf_decide_luma, src, pre, 7, 8, sets)
    assert rc == 0
    assert np.array_equal(out.y, replay(orc, vv, pre, choice, 7, 8, sets, None).y)


# ----------------------------------------------- chroma ALF and CC-ALF decisions
def chroma_ctb_sse(src, rec, c_idx, ctb):
    """Per-CTB SSE of component c_idx over the chroma CTB clipped to the plane
    (ctb_sse_chroma, src/alf_enc.rs:812-840), raster order of luma CTBs."""
    a, b = (src.cb, rec.cb) if c_idx == 1 else (src.cr, rec.cr)
    h, w = src.y.shape
    cw = ctb // 2
    d = (a.astype(np.int64) - b.astype(np.int64)) ** 2
    return np.array([int(d[y0 // 2:y0 // 2 + cw, x0 // 2:x0 // 2 + cw].sum())
                     for y0 in range(0, h, ctb) for x0 in range(0, w, ctb)], np.uint64)


def all_on_trial(orc, vv, rec, c_idx, k, ctb_log2, bd, aps, tb, cc, pre_luma=None):
    """rec with every CTB's component c_idx at chroma alternative k (or CC
    filter k + 1), through the oracle's ALF stage (the reference's trial)."""
    h, w = rec.y.shape
    ctb = 1 << ctb_log2
    wc, hc = (w + ctb - 1) // ctb, (h + ctb - 1) // ctb
    ent = {}
    for y in range(hc):
        for x in range(wc):
            c = vv.AlfCtb()
            if cc:
                if c_idx == 1:
                    c.cc_cb_idc = k + 1
                else:
                    c.cc_cr_idc = k + 1
            elif c_idx == 1:
                c.cb_on, c.cb_alt_idx = 1, k
            else:
                c.cr_on, c.cr_alt_idx = 1, k
            ent[(x, y)] = c
    ap, _k = alf_picture(vv, wc, hc, ent)
    o = rec.copy()
    if pre_luma is not None:
        o.y[:] = pre_luma
    cfg = alf_cfg(vv, bd=bd, ctb_log2=ctb_log2, cb=c_idx == 1, cr=c_idx == 2, cc_cb=c_idx == 1, cc_cr=c_idx == 2)
    if cc:
        b = binding_of(vv, cc_cb=aps if c_idx == 1 else None, cc_cr=aps if c_idx == 2 else None)
    else:
        b = binding_of(vv, chroma=aps)
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(cfg), C.byref(b),
                                        C.byref(tb) if tb is not None else None) == 0
    return o


CHROMA_CASES = [
    # w, h, ctb_log2, bd, c_idx, n candidates, tile bounds
    (200, 136, 6, 10, 1, 4, None),
    (256, 128, 7, 8, 2, 8, None),
    (160, 96, 5, 12, 1, 2, ([0, 96, 160], [0, 64, 96])),
    (96, 72, 4, 10, 2, 3, ([0, 48, 96], [0, 32, 72])),
]


def chroma_fixture(vv, w, h, bd, seed):
    """Smooth chroma source, reconstruction = source + noise, luma with
    texture; pre = the luma plane before luma ALF."""
    rng = np.random.default_rng(seed)
    src = smooth_gradient(vv, w, h, bd)
    yy, xx = np.mgrid[0:h // 2, 0:w // 2]
    span = max(1, w // 2 + h // 2)
    src.cb[:] = (1 << (bd - 1)) + ((xx * 3 + yy * 2) * 64 // (3 * span)) * (1 << (bd - 8))
    src.cr[:] = (1 << (bd - 1)) - ((xx + yy * 3) * 48 // (3 * span)) * (1 << (bd - 8))
    src.y[:] = np.clip(src.y.astype(np.int64) + rng.integers(-24, 25, src.y.shape) * (1 << (bd - 8)), 0,
                       (1 << bd) - 1)
    rec = src.copy()
    # noise on the right half: CTBs on the left stay exact, so that no
    # filter can beat leaving them off
    q = max(1, (w // 2) // 2)
    for pl in (rec.cb, rec.cr):
        pl[:, q:] = np.clip(pl[:, q:].astype(np.int64) + rng.integers(-10, 11, pl[:, q:].shape) * (1 << (bd - 8)),
                            0, (1 << bd) - 1)
    pre = src.y.copy()
    return rng, src, rec, pre


def lowpass_chroma_aps(vv, rng, n):
    """n random alternatives with a low-pass one (taps summing to ~1) first."""
    a = chroma_aps(vv, rng, n)
    for t, c in enumerate([8, 10, 16, 10, 12, 20]):
        a.chroma_coeff[0][t] = c
        a.chroma_clip_idx[0][t] = 0
    return a


def cc_fixture(orc, vv, rng, src, rec, pre, c_idx, ctb_log2, bd, n, tb):
    """A CC-ALF APS whose filter 0 undoes a correction taken out of rec's
    component (rec := rec with the negated filter applied), so that the
    decision has something to find."""
    rows = [[int(rng.integers(-40, 41)) for _ in range(7)] for _ in range(n)]
    neg = cc_aps(vv, rng, c_idx, 1, rows=[[-v for v in rows[0]]])
    bent = all_on_trial(orc, vv, rec, c_idx, 0, ctb_log2, bd, neg, tb, True, pre)
    out = rec.copy()
    q = max(1, (rec.y.shape[1] // 2) // 2)
    dst, bsrc, ssrc = (out.cb, bent.cb, src.cb) if c_idx == 1 else (out.cr, bent.cr, src.cr)
    dst[:, q:] = bsrc[:, q:]
    dst[:, :q] = ssrc[:, :q]
    return cc_aps(vv, rng, c_idx, n, rows=rows), out


@pytest.mark.parametrize("cc", [False, True])
@pytest.mark.parametrize("w,h,ctb_log2,bd,c_idx,n,tbs", CHROMA_CASES)
def test_chroma_decisions_from_independent_trials(orc, vv, w, h, ctb_log2, bd, c_idx, n, tbs, cc):
    rng, src, rec, pre = chroma_fixture(vv, w, h, bd, w + h + c_idx)
    tb = tile_bounds(vv, *tbs) if tbs else None
    if cc:
        aps, rec = cc_fixture(orc, vv, rng, src, rec, pre, c_idx, ctb_log2, bd, min(n, 4), tb)
    else:
        aps = lowpass_chroma_aps(vv, rng, n)
    ctb = 1 << ctb_log2
    nct = ((w + ctb - 1) // ctb) * ((h + ctb - 1) // ctb)
    init = [(1, 3, 0, 0, 0, 0, 0, 0)] * nct  # luma fields the decision must keep
    fn = orc.oracle_cc_alf_decide if cc else orc.oracle_alf_decide_chroma
    rc, recs, sse, out = decide_chroma(vv, fn, src, rec, aps, c_idx, ctb_log2, bd, tb, pre if cc else None, init)
    assert rc == 0
    ncand = sse.shape[1] - 1
    cols = [chroma_ctb_sse(src, rec, c_idx, ctb)]
    trials = []
    for k in range(ncand):
        t = all_on_trial(orc, vv, rec, c_idx, k, ctb_log2, bd, aps, tb, cc, pre if cc else None)
        trials.append(t)
        cols.append(chroma_ctb_sse(src, t, c_idx, ctb))
    want = np.stack(cols, axis=1)
    assert np.array_equal(sse, want)
    choice = [first_min(r) for r in want]
    assert any(c >= 0 for c in choice) and any(c < 0 for c in choice)
    wc = (w + ctb - 1) // ctb
    comp = (lambda p: p.cb) if c_idx == 1 else (lambda p: p.cr)
    for i, k in enumerate(choice):
        f = list(init[i])
        if k >= 0:
            if cc:
                f[6 if c_idx == 1 else 7] = k + 1
            else:
                f[2 if c_idx == 1 else 4], f[3 if c_idx == 1 else 5] = 1, k
        assert recs[i] == tuple(f)
        y0, x0, cw = (i // wc) * ctb // 2, (i % wc) * ctb // 2, ctb // 2
        ref = comp(trials[k] if k >= 0 else rec)[y0:y0 + cw, x0:x0 + cw]
        assert np.array_equal(comp(out)[y0:y0 + cw, x0:x0 + cw], ref)
    other = (lambda p: p.cr) if c_idx == 1 else (lambda p: p.cb)
    assert np.array_equal(other(out), other(rec)) and np.array_equal(out.y, rec.y)


def test_chroma_decisions_are_no_ops_where_the_reference_returns(orc, vv):
    rng, src, rec, pre = chroma_fixture(vv, 64, 64, 8, 5)
    aps = chroma_aps(vv, rng, 2)
    aps.alf_chroma_filter_signal_flag = 0
    rc, recs, _s, out = decide_chroma(vv, orc.oracle_alf_decide_chroma, src, rec, aps, 1, 6, 8)
    assert rc == 0 and np.array_equal(out.cb, rec.cb) and recs[0] == (0,) * 8
    rc, recs, _s, out = decide_chroma(vv, orc.oracle_cc_alf_decide, src, rec, cc_aps(vv, rng, 1, 0, rows=[]), 1, 6,
                                      8, pre_luma=pre)
    assert rc == 0 and np.array_equal(out.cb, rec.cb)
    rc, recs, _s, out = decide_chroma(vv, orc.oracle_alf_decide_chroma, src, rec, chroma_aps(vv, rng, 2), 1, 6, 8,
                                      cfi=0)
    assert rc == 0 and np.array_equal(out.cb, rec.cb)


def test_cc_zero_filter_is_never_chosen(orc, vv):
    # an all-zero CC filter adds nothing: its error equals the off error and
    # the strict comparison keeps every CTB off
    rng, src, rec, pre = chroma_fixture(vv, 128, 128, 10, 9)
    aps = cc_aps(vv, rng, 2, 0, rows=[[0] * 7])
    rc, recs, sse, out = decide_chroma(vv, orc.oracle_cc_alf_decide, src, rec, aps, 2, 6, 10, pre_luma=pre)
    assert rc == 0 and np.array_equal(sse[:, 1], sse[:, 0])
    assert all(r[7] == 0 for r in recs) and np.array_equal(out.cr, rec.cr)


# the reference's CC-ALF encoder tests, src/alf_enc.rs:1574-1760
def cc_two_cb_filters(vv):
    return cc_aps(vv, None, 1, 0, rows=[[0] * 7, [0, 8, -8, 0, 0, 0, 0]])


def test_cc_flat_luma_keeps_chroma(orc, vv):
    src = flat_picture(vv, 64, 64, 8)
    src.y[:] = src.cb[:] = src.cr[:] = 100
    rec = src.copy()
    rec.cb.ravel()[0] ^= 0x20
    rec.cb.ravel()[10] ^= 0x10
    rc, recs, _s, out = decide_chroma(vv, orc.oracle_cc_alf_decide, src, rec, cc_two_cb_filters(vv), 1, 7, 8,
                                      pre_luma=rec.y.copy())
    assert rc == 0 and np.array_equal(out.cb, rec.cb) and recs[0][6] == 0


def test_cc_never_increases_chroma_sse_and_replays(orc, vv):
    src = flat_picture(vv, 128, 128, 8)
    src.y[:] = src.cb[:] = src.cr[:] = 100
    src.y[:, 64:] = 220
    rec = src.copy()
    rec.cb[0:64, 32:64] += 15
    pre = rec.y.copy()
    rc, recs, _s, out = decide_chroma(vv, orc.oracle_cc_alf_decide, src, rec, cc_two_cb_filters(vv), 1, 7, 8,
                                      pre_luma=pre)
    assert rc == 0
    assert chroma_ctb_sse(src, out, 1, 128).sum() <= chroma_ctb_sse(src, rec, 1, 128).sum()
    # replay through the ALF stage from the pre-decision chroma and luma
    ap, _k = alf_picture(vv, 1, 1, {(0, 0): vv.AlfCtb(cc_cb_idc=recs[0][6])})
    o = rec.copy()
    cfg = alf_cfg(vv, bd=8, ctb_log2=7, cb=True, cc_cb=True)
    assert orc.oracle_apply_alf_clipped(o.ptr(), C.byref(ap), C.byref(cfg),
                                        C.byref(binding_of(vv, cc_cb=cc_two_cb_filters(vv))), None) == 0
    assert np.array_equal(o.cb, out.cb)
