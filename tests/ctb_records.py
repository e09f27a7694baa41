"""Per-CTB syntax records built field by field in Python, following the bit
layout documented in include/vvclf.h (independently of the library's own
packer), and random canonical records for round trips."""
import ctypes as C

import numpy as np

# field positions: merge flags, luma / Cb groups (type, abs x4, sign x4,
# band position or EO class), the Cr group (abs x4, sign x4, band position),
# then the ALF fields
MERGE_L, MERGE_U, Y, CB, CR, ALF = 0, 1, 2, 33, 64, 93


class Rec:
    def __init__(self):
        self.bits = 0

    def put(self, pos, n, v):
        assert 0 <= v < (1 << n)
        self.bits |= v << pos
        return self

    def sao(self, g, typ, abs4, sign4, pos):
        self.put(g, 2, typ)
        for i, a in enumerate(abs4):
            self.put(g + 2 + 5 * i, 5, a)
        for i, s in enumerate(sign4):
            self.put(g + 22 + i, 1, s)
        return self.put(g + 26, 5, pos)

    def sao_cr(self, abs4, sign4, pos):
        for i, a in enumerate(abs4):
            self.put(CR + 5 * i, 5, a)
        for i, s in enumerate(sign4):
            self.put(CR + 20 + i, 1, s)
        return self.put(CR + 24, 5, pos)

    def alf(self, luma_on=0, use_aps=0, idx=0, cb_on=0, cb_alt=0, cr_on=0, cr_alt=0, cc_cb=0, cc_cr=0):
        return (self.put(ALF, 1, luma_on).put(ALF + 1, 1, use_aps).put(ALF + 2, 4, idx).put(ALF + 6, 1, cb_on)
                .put(ALF + 7, 3, cb_alt).put(ALF + 10, 1, cr_on).put(ALF + 11, 3, cr_alt).put(ALF + 14, 3, cc_cb)
                .put(ALF + 17, 3, cc_cr))

    def struct(self, vv):
        r = vv.CtbSyntax()
        for i in range(16):
            r.bytes[i] = (self.bits >> (8 * i)) & 0xff
        return r


def config(vv, bd=10, cfi=1, n_aps=4, alt_m1=1, cc_m1=(1, 1), sao=True, alf=True):
    c = vv.CtbSyntaxConfig()
    c.sao_luma_used = c.sao_chroma_used = int(sao)
    c.alf_enabled = int(alf)
    c.alf_cb_enabled = c.alf_cr_enabled = c.alf_cc_cb_enabled = c.alf_cc_cr_enabled = int(alf)
    c.bit_depth, c.chroma_format_idc = bd, cfi
    c.num_alf_aps_ids_luma = n_aps
    c.alf_chroma_num_alt_filters_minus1 = alt_m1
    c.alf_cc_cb_filters_signalled_minus1, c.alf_cc_cr_filters_signalled_minus1 = cc_m1
    return c


def random_records(vv, rng, w, h, cfg):
    """Random canonical records for a w x h CTB grid (absent fields zero)."""
    cmax = (1 << (min(cfg.bit_depth, 10) - 5)) - 1
    out = (vv.CtbSyntax * (w * h))()

    def group(rec, g):
        typ = int(rng.integers(0, 3))
        if typ == 0:
            return 0, 0
        abs4 = [int(v) for v in rng.integers(0, cmax + 1, 4)]
        if typ == 1:
            sign4 = [int(rng.integers(0, 2)) if a else 0 for a in abs4]
            pos = int(rng.integers(0, 32))
        else:
            sign4, pos = [0] * 4, int(rng.integers(0, 4))
        rec.sao(g, typ, abs4, sign4, pos)
        return typ, pos

    for ry in range(h):
        for rx in range(w):
            rec = Rec()
            if cfg.sao_luma_used or cfg.sao_chroma_used:
                m = rng.random()
                if rx > 0 and m < 0.25:
                    rec.put(MERGE_L, 1, 1)
                elif ry > 0 and m < 0.45:
                    rec.put(MERGE_U, 1, 1)
                else:
                    if cfg.sao_luma_used:
                        group(rec, Y)
                    if cfg.sao_chroma_used and cfg.chroma_format_idc:
                        typ, _ = group(rec, CB)
                        if typ:
                            abs4 = [int(v) for v in rng.integers(0, cmax + 1, 4)]
                            sign4 = [int(rng.integers(0, 2)) if a and typ == 1 else 0 for a in abs4]
                            rec.sao_cr(abs4, sign4, int(rng.integers(0, 32)) if typ == 1 else 0)
            if cfg.alf_enabled:
                luma_on = int(rng.random() < 0.8)
                use_aps = int(luma_on and cfg.num_alf_aps_ids_luma > 0 and rng.random() < 0.7)
                idx = int(rng.integers(0, cfg.num_alf_aps_ids_luma)) if use_aps else (int(rng.integers(0, 16)) if luma_on else 0)
                kw = dict(luma_on=luma_on, use_aps=use_aps, idx=idx)
                if cfg.chroma_format_idc:
                    for c in ("cb", "cr"):
                        on = int(rng.random() < 0.6)
                        kw[c + "_on"] = on
                        kw[c + "_alt"] = int(rng.integers(0, cfg.alf_chroma_num_alt_filters_minus1 + 1)) if on else 0
                    kw["cc_cb"] = int(rng.integers(0, cfg.alf_cc_cb_filters_signalled_minus1 + 2))
                    kw["cc_cr"] = int(rng.integers(0, cfg.alf_cc_cr_filters_signalled_minus1 + 2))
                rec.alf(**kw)
            out[ry * w + rx] = rec.struct(vv)
    return out


def decode(vv, recs, w, h, cfg):
    sao = (vv.SaoCtbParams * (w * h))()
    alf = (vv.AlfCtb * (w * h))()
    rc = vv.lib.vvclf_ctb_syntax_decode(recs, w, h, C.byref(cfg), sao, alf)
    return rc, sao, alf
