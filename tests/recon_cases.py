"""Random inputs for the reconstruction-side steps (LMCS forward mapping of
inter-CU rectangles, chroma residual scaling per chroma TB)."""
import ctypes as C

import numpy as np


def quadtree_rects(rng, w, h, root=64, min_size=4):
    """A tiling of a w x h plane by a random quadtree / binary split of
    root-sized squares (clipped at the plane edge)."""
    out = []

    def split(x, y, s_w, s_h):
        if x >= w or y >= h:
            return
        r = rng.random()
        if s_w > min_size and s_h > min_size and r < 0.45:
            hw, hh = s_w // 2, s_h // 2
            for dx, dy in ((0, 0), (hw, 0), (0, hh), (hw, hh)):
                split(x + dx, y + dy, hw, hh)
        elif s_w > min_size and r < 0.6:
            split(x, y, s_w // 2, s_h)
            split(x + s_w // 2, y, s_w // 2, s_h)
        elif s_h > min_size and r < 0.75:
            split(x, y, s_w, s_h // 2)
            split(x, y + s_h // 2, s_w, s_h // 2)
        else:
            out.append((x, y, min(s_w, w - x), min(s_h, h - y)))

    for y in range(0, h, root):
        for x in range(0, w, root):
            split(x, y, root, root)
    return out


def lmcs_binding(vv, seed, bd):
    """A derived LMCS table of a random valid payload (budget and pivot
    clauses hold) and its binding; returns (binding, keepalive)."""
    rng = np.random.default_rng(seed)
    for _ in range(1000):
        d = vv.LmcsData()
        d.lmcs_min_bin_idx = int(rng.integers(0, 3))
        d.lmcs_delta_max_bin_idx = int(rng.integers(0, 3))
        d.lmcs_delta_cw_prec_minus1 = 5
        org = (1 << bd) // 16
        for i in range(16):
            d.lmcs_delta_abs_cw[i] = int(rng.integers(0, org // 2))
            d.lmcs_delta_sign_cw_flag[i] = int(rng.integers(0, 2))
        d.lmcs_delta_abs_crs = int(rng.integers(0, org // 4))
        d.lmcs_delta_sign_crs_flag = int(rng.integers(0, 2))
        der = vv.LmcsDerived()
        if vv.lib.vvclf_derive_lmcs(C.byref(d), bd, C.byref(der)) == 0:
            b = vv.LmcsBinding()
            b.derived = C.pointer(der)
            b.min_bin_idx = d.lmcs_min_bin_idx
            b.max_bin_idx = 15 - d.lmcs_delta_max_bin_idx
            return b, (der, d)
    raise RuntimeError("no valid LMCS payload drawn")


def crs_blocks(vv, rng, wc, hc, w, h, size_y):
    """Chroma TBs tiling both chroma planes, each with a random CU anchor and
    availability; returns (array of CrsBlock, residual count)."""
    blocks = []
    off = 0
    for c_idx in (1, 2):
        for x, y, bw, bh in quadtree_rects(rng, wc, hc, root=32, min_size=2):
            b = vv.CrsBlock()
            b.c_idx, b.x, b.y, b.w, b.h = c_idx, x, y, bw, bh
            b.x_cu = int(rng.integers(0, w // 4)) * 4
            b.y_cu = int(rng.integers(0, h // 4)) * 4
            b.avail_l = int(b.x_cu > 0 and rng.random() < 0.7)
            b.avail_t = int(b.y_cu > 0 and rng.random() < 0.7)
            b.size_y = size_y
            b.res_offset = off
            off += bw * bh
            blocks.append(b)
    arr = (vv.CrsBlock * len(blocks))(*blocks)
    return arr, off
