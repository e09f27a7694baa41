// vvclf_fused.hip — the batched form: one persistent launch filters every
// tile of many pictures (uneven sizes) through the whole chain in LDS:
//   load luma tile + 16-sample halo (chroma + 4) with the LMCS LUT applied
//   -> deblock V -> deblock H -> SAO into a second buffer -> ALF / CC-ALF
//   -> write the tile.
// Halo (luma, output tile [x0,x1) x [y0,y1)):
//   ALF reads SAO output on [x0-3, x1+3); SAO reads deblocked [x0-4, x1+4);
//   horizontal edges y in [y0-8, y1+8] change rows [y-7, y+7) and read the
//   vertical-pass output on [y-8, y+8) => rows [y0-16, y1+16); vertical edges
//   x in [x0-8, x1+8] likewise read columns [x0-16, x1+16).
// Chroma (4:2:0): 5x5 ALF reads SAO output +-2, SAO reads deblocked +-3,
//   chroma edges (every 8 samples) write <= 3 and read <= 4 => +-4.
// Edges of one direction run concurrently in place (extents never reach the
// samples another edge of that direction uses; the host checks this).
#include <hip/hip_runtime.h>

#include "vvclf_filters.h"
#include "vvclf_internal.h"

namespace vvclf {

template <int TW, int TH>
struct fused_lds {
  static constexpr int LW = TW + 32, LH = TH + 32;          // luma input region
  static constexpr int CW = TW / 2 + 8, CH = TH / 2 + 8;    // chroma input region
  static constexpr int SW = TW + 6, SH = TH + 6;            // SAO'd luma
  static constexpr int SCW = TW / 2 + 4, SCH = TH / 2 + 4;  // SAO'd chroma
  uint16_t L[LH * LW];
  uint16_t Cc[2][CH * CW];
  uint16_t S[SH * SW];
  uint16_t Sc[2][SCH * SCW];
  uint32_t bank[ALF_BANK_WORDS];
  uint16_t lut[1024];
};

// A rectangular LDS window holding picture samples [ox, ox+w) x [oy, oy+h)
// (the tile region intersected with the picture).  Reads clamp to the window,
// which reproduces the reference's picture clamp at picture edges; interior
// window edges are never crossed (halo sizes above).
struct lwin {
  uint16_t* p;
  int pitch, ox, oy, w, h;
  __device__ __forceinline__ int rd(int x, int y) const {
    x = clampi(x - ox, 0, w - 1);
    y = clampi(y - oy, 0, h - 1);
    return p[y * pitch + x];
  }
  __device__ __forceinline__ bool inside(int x, int y) const {
    return x >= ox && y >= oy && x < ox + w && y < oy + h;
  }
  __device__ __forceinline__ void wr(int x, int y, int v) { p[(y - oy) * pitch + (x - ox)] = (uint16_t)v; }
};

__device__ __forceinline__ int fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ int cdiv_i(int a, int b) { return -fdiv(-a, b); }

__device__ __forceinline__ luma_terms shfl_terms4(const luma_terms& t, int src) {
  luma_terms r;
  r.dp = __shfl(t.dp, src); r.dq = __shfl(t.dq, src); r.sp = __shfl(t.sp, src);
  r.sq = __shfl(t.sq, src); r.spq = __shfl(t.spq, src); r.dpl = __shfl(t.dpl, src);
  r.dql = __shfl(t.dql, src); r.sp7 = __shfl(t.sp7, src); r.sq7 = __shfl(t.sq7, src);
  r.spm = __shfl(t.spm, src); r.sqm = __shfl(t.sqm, src);
  return r;
}

// Luma deblocking of one direction over a grid of segments (in place in LDS).
// VERT: edge columns gx in [gx0, gx1], segment rows gy in [gy0, gy1].
// HOR : edge rows gy in [gy0, gy1], segment columns gx in [gx0, gx1].
template <bool VERT>
__device__ void lds_deblock_luma(const dev_job& j, lwin& L, int gx0, int gx1, int gy0, int gy1, int maxv) {
  const int tid = threadIdx.x;
  const int ncol = gx1 - gx0 + 1, nrow = gy1 - gy0 + 1;
  const int nseg = ncol * nrow;
  const uint32_t* map = VERT ? j.e_lv : j.e_lh;
  const int lane = tid & 63;
  for (int base = 0; base < nseg; base += blockDim.x / 4) {
    const int seg = base + (tid >> 2), line = tid & 3;
    uint32_t e = 0;
    int gx = 0, gy = 0;
    if (seg < nseg) {
      gx = gx0 + seg % ncol;
      gy = gy0 + seg / ncol;
      e = map[(size_t)gy * j.gw + gx];
    }
    if (e != 0) {  // uniform across the quad
      const int x0 = VERT ? 4 * gx : 4 * gx + line, y0 = VERT ? 4 * gy + line : 4 * gy;
      int p[8], q[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        p[i] = VERT ? L.rd(x0 - 1 - i, y0) : L.rd(x0, y0 - 1 - i);
        q[i] = VERT ? L.rd(x0 + i, y0) : L.rd(x0, y0 + i);
      }
      int beta, tc;
      edge_thresholds(e, j.bd_dbk, beta, tc);
      int mflp = (e >> E_MFLP_SHIFT) & 7, mflq = (e >> E_MFLQ_SHIFT) & 7;
      luma_terms mine = luma_line_terms(p, q, mflp, mflq);
      int qb = lane & ~3;
      luma_terms t0 = shfl_terms4(mine, qb), t3 = shfl_terms4(mine, qb + 3);
      luma_decision d = luma_decide(t0, t3, beta, tc, mflp, mflq, (e & E_CTBROW_BIT) != 0);
      int np[8], nq[8], nwp, nwq;
      luma_filter_line(p, q, d, tc, np, nq, nwp, nwq);
      if (e & E_PLTP_BIT) nwp = 0;
      if (e & E_PLTQ_BIT) nwq = 0;
#pragma unroll
      for (int i = 0; i < 7; i++) {
        if (i < nwp) {
          int x = VERT ? x0 - 1 - i : x0, y = VERT ? y0 : y0 - 1 - i;
          if (L.inside(x, y)) L.wr(x, y, clampi(np[i], 0, maxv));
        }
        if (i < nwq) {
          int x = VERT ? x0 + i : x0, y = VERT ? y0 : y0 + i;
          if (L.inside(x, y)) L.wr(x, y, clampi(nq[i], 0, maxv));
        }
      }
    }
  }
}

// Chroma deblocking of one direction (both components) in LDS.
// VERT: edge columns sx in [sx0, sx1] (x = 8 sx), segment rows sy in [sy0, sy1] (y = 2 sy).
// HOR : edge rows sy in [sy0, sy1] (y = 8 sy), segment columns sx in [sx0, sx1] (x = 2 sx).
template <bool VERT>
__device__ void lds_deblock_chroma(const dev_job& j, lwin cb, uint16_t* cr_p, int sx0, int sx1, int sy0, int sy1,
                                   int maxv) {
  const int tid = threadIdx.x;
  const int ncol = sx1 - sx0 + 1, nrow = sy1 - sy0 + 1;
  const int nseg = ncol * nrow;
  const int mw = VERT ? j.cvw : j.chw;
  for (int base = 0; base < 2 * nseg; base += blockDim.x / 2) {
    const int item = base + (tid >> 1), line = tid & 1;
    uint32_t e = 0;
    int sx = 0, sy = 0, comp = 0;
    if (item < 2 * nseg) {
      comp = item / nseg;
      int seg = item % nseg;
      sx = sx0 + seg % ncol;
      sy = sy0 + seg / ncol;
      e = (VERT ? j.e_cv[comp] : j.e_ch[comp])[(size_t)sy * mw + sx];
    }
    if (e != 0) {  // uniform across the pair
      lwin L = cb;  // Cb and Cr windows share their geometry
      if (comp) L.p = cr_p;
      int x0, y0;
      if (VERT) { x0 = 8 * sx; y0 = 2 * sy + line; } else { x0 = 2 * sx + line; y0 = 8 * sy; }
      int c[8];
#pragma unroll
      for (int i = 0; i < 8; i++) c[i] = VERT ? L.rd(x0 - 4 + i, y0) : L.rd(x0, y0 - 4 + i);
      int beta, tc;
      edge_thresholds(e, j.bd_dbk, beta, tc);
      int mflp = (e >> E_MFLP_SHIFT) & 7, mflq = (e >> E_MFLQ_SHIFT) & 7;
      bool eligible, short_p;
      if (VERT) { eligible = mflp == 3 && mflq == 3; short_p = false; }
      else { eligible = mflq == 3 && (mflp == 3 || mflp == 1); short_p = mflp == 1; }
      bool mine = eligible && chroma_strong_row(c, beta, tc, short_p);
      bool other = __shfl_xor((int)mine, 1) != 0;
      int n[8];
      uint32_t wm = chroma_filter_line(c, eligible && mine && other, short_p, tc, n);
      if (e & E_PLTP_BIT) wm &= 0xf0u;
      if (e & E_PLTQ_BIT) wm &= 0x0fu;
#pragma unroll
      for (int i = 1; i < 7; i++) {
        if (!((wm >> i) & 1)) continue;
        int x = VERT ? x0 - 4 + i : x0, y = VERT ? y0 : y0 - 4 + i;
        if (L.inside(x, y)) L.wr(x, y, clampi(n[i], 0, maxv));
      }
    }
  }
}

// SAO of one component from window `src` into window `dst` (dst's extent is
// the region produced).  sub = 1 for chroma.
__device__ void lds_sao(const dev_job& j, const lwin& src, lwin& dst, int comp, int pw, int ph) {
  const int sub = comp ? 1 : 0;
  const int log2c = j.sao_ctb_log2 - sub;
  const bool active = j.do_sao && (comp == 0 ? j.sao_luma : j.sao_chroma);
  const int maxv = (1 << j.sao_bd) - 1;
  // rows over the waves, columns over the lanes (no integer division)
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  for (int yy = threadIdx.x >> 6; yy < dst.h; yy += nw)
  for (int xx = lane; xx < dst.w; xx += 64) {
    const int x = dst.ox + xx, y = dst.oy + yy;
    const int pv = src.rd(x, y);
    int out = pv;
    const int rx = x >> log2c, ry = y >> log2c;
    if (active && rx < j.swctb && ry < j.shctb) {
      const size_t ci = (size_t)ry * j.swctb + rx;
      const uint64_t prm = j.sao[ci * 3 + comp];
      const int type = (int)(prm & 3);
      if (type == 2) {
        int h0, v0, h1, v1;
        eo_offsets((int)((prm >> 2) & 31), h0, v0, h1, v1);
        const int nx0 = x + h0, ny0 = y + v0, nx1 = x + h1, ny1 = y + v1;
        bool outside = nx0 < 0 || ny0 < 0 || nx1 < 0 || ny1 < 0 || nx0 >= pw || ny0 >= ph || nx1 >= pw || ny1 >= ph;
        if (j.sao_rect) {
          uint64_t r = j.sao_rect[ci];
          int cx0 = (int)(r & 0xffff), cy0 = (int)((r >> 16) & 0xffff), cx1 = (int)((r >> 32) & 0xffff),
              cy1 = (int)(r >> 48);
          if (sub) { cx0 >>= 1; cy0 >>= 1; cx1 = (cx1 + 1) >> 1; cy1 = (cy1 + 1) >> 1; }
          outside = outside || nx0 < cx0 || ny0 < cy0 || nx1 < cx0 || ny1 < cy0 || nx0 >= cx1 || ny0 >= cy1 ||
                    nx1 >= cx1 || ny1 >= cy1;
        }
        out = outside ? clampi(pv, 0, maxv) : sao_edge_sample(pv, src.rd(nx0, ny0), src.rd(nx1, ny1), prm, maxv);
      } else if (type == 1) {
        out = sao_band_sample(pv, prm, j.sao_bd, maxv);
      }
    }
    dst.wr(x, y, out);
  }
}

template <int TW, int TH>
__device__ void fused_tile(const dev_job& j, fused_lds<TW, TH>& s, int x0, int y0) {
  typedef fused_lds<TW, TH> LD;
  const int W = j.w, H = j.h;
  const int x1 = min(x0 + TW, W), y1 = min(y0 + TH, H);
  const int maxv = (1 << j.bd) - 1;
  const bool chroma = j.wc > 0;
  // ---- load (LMCS inverse folded into the load)
  lwin L{s.L, LD::LW, max(0, x0 - 16), max(0, y0 - 16), 0, 0};
  L.w = min(W, x1 + 16) - L.ox;
  L.h = min(H, y1 + 16) - L.oy;
  {
    const uint16_t* src = j.in.y;
    const bool lut_lds = j.do_lmcs && j.bd <= 10;
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int yy = threadIdx.x >> 6; yy < L.h; yy += nw)
    for (int xx = lane; xx < L.w; xx += 64) {
      uint16_t v = src[(size_t)(L.oy + yy) * j.py + L.ox + xx];
      if (j.do_lmcs) v = (lut_lds && v < 1024) ? s.lut[v] : j.lmcs_lut[v];
      s.L[yy * LD::LW + xx] = v;
    }
  }
  const int cx0 = x0 >> 1, cy0 = y0 >> 1, cx1 = min(x1 >> 1, j.wc), cy1 = min(y1 >> 1, j.hc);
  lwin Cw[2];
  if (chroma) {
    for (int c = 0; c < 2; c++) {
      Cw[c] = lwin{s.Cc[c], LD::CW, max(0, cx0 - 4), max(0, cy0 - 4), 0, 0};
      Cw[c].w = min(j.wc, cx1 + 4) - Cw[c].ox;
      Cw[c].h = min(j.hc, cy1 + 4) - Cw[c].oy;
      const uint16_t* src = c ? j.in.cr : j.in.cb;
      const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
      for (int yy = threadIdx.x >> 6; yy < Cw[c].h; yy += nw)
        for (int xx = lane; xx < Cw[c].w; xx += 64)
          s.Cc[c][yy * LD::CW + xx] = src[(size_t)(Cw[c].oy + yy) * j.pc + Cw[c].ox + xx];
    }
  }
  __syncthreads();
  // ---- deblocking: all vertical edges, then all horizontal edges
  if (j.do_dbk) {
    {
      int gx0 = max(0, cdiv_i(x0 - 10, 4)), gx1 = min(j.gw - 1, fdiv(x1 + 10, 4));  // x in [x0-10, x1+10]
      int gy0 = L.oy / 4, gy1 = min(j.gh - 1, (L.oy + L.h - 1) / 4);
      lds_deblock_luma<true>(j, L, gx0, gx1, gy0, gy1, maxv);
    }
    if (chroma && j.do_dbk_chroma) {
      int sx0 = max(0, cdiv_i(cx0 - 5, 8)), sx1 = min(j.cvw - 1, fdiv(cx1 + 5, 8));  // x in [cx0-5, cx1+5]
      int sy0 = Cw[0].oy / 2, sy1 = min(j.cvh - 1, (Cw[0].oy + Cw[0].h - 1) / 2);
      lds_deblock_chroma<true>(j, Cw[0], Cw[1].p, sx0, sx1, sy0, sy1, maxv);
    }
    __syncthreads();
    {
      int gy0 = max(0, cdiv_i(y0 - 10, 4)), gy1 = min(j.gh - 1, fdiv(y1 + 10, 4));  // y in [y0-10, y1+10]
      int gx0 = max(0, fdiv(x0 - 4, 4)), gx1 = min(j.gw - 1, fdiv(min(W, x1 + 4) - 1, 4));
      lds_deblock_luma<false>(j, L, gx0, gx1, gy0, gy1, maxv);
    }
    if (chroma && j.do_dbk_chroma) {
      int sy0 = max(0, cdiv_i(cy0 - 5, 8)), sy1 = min(j.chh - 1, fdiv(cy1 + 5, 8));  // y in [cy0-5, cy1+5]
      int sx0 = max(0, fdiv(cx0 - 3, 2)), sx1 = min(j.chw - 1, fdiv(min(j.wc, cx1 + 3) - 1, 2));
      lds_deblock_chroma<false>(j, Cw[0], Cw[1].p, sx0, sx1, sy0, sy1, maxv);
    }
    __syncthreads();
  }
  // ---- SAO into the second buffer
  lwin S{s.S, LD::SW, max(0, x0 - 3), max(0, y0 - 3), 0, 0};
  S.w = min(W, x1 + 3) - S.ox;
  S.h = min(H, y1 + 3) - S.oy;
  lds_sao(j, L, S, 0, W, H);
  lwin Sc[2];
  if (chroma) {
    for (int c = 0; c < 2; c++) {
      Sc[c] = lwin{s.Sc[c], LD::SCW, max(0, cx0 - 2), max(0, cy0 - 2), 0, 0};
      Sc[c].w = min(j.wc, cx1 + 2) - Sc[c].ox;
      Sc[c].h = min(j.hc, cy1 + 2) - Sc[c].oy;
      lds_sao(j, Cw[c], Sc[c], c + 1, j.wc, j.hc);
    }
  }
  // ---- ALF: this tile lies inside one CTB
  const int ctb = 1 << j.ctb_log2;
  const int rx = x0 >> j.ctb_log2, ry = y0 >> j.ctb_log2;
  const bool in_grid = j.do_alf && rx < j.wctb && ry < j.hctb;
  const size_t ci = in_grid ? (size_t)ry * j.wctb + rx : 0;
  const uint32_t cw = in_grid ? j.alf_ctb[ci] : 0xffffffffu;
  const int bank_idx = (int)(cw & 0xff);
  if (bank_idx != (int)ALF_OFF8) {
    for (int i = threadIdx.x; i < ALF_BANK_WORDS; i += blockDim.x)
      s.bank[i] = j.alf_bank[(size_t)bank_idx * ALF_BANK_WORDS + i];
  }
  __syncthreads();
  uint64_t r = in_grid ? j.alf_rect[ci] : (((uint64_t)W << 32) | ((uint64_t)H << 48));
  const int rx0 = (int)(r & 0xffff), ry0 = (int)((r >> 16) & 0xffff), rx1 = (int)((r >> 32) & 0xffff) - 1,
            ry1 = (int)(r >> 48) - 1;
  // luma: one 4x4 block per thread
  {
    const int bw = (x1 - x0 + 3) / 4, bh = (y1 - y0 + 3) / 4;
    const int amax = (1 << j.alf_bd) - 1;
    for (int b = threadIdx.x; b < bw * bh; b += blockDim.x) {
      const int x4 = x0 + 4 * (b % bw), y4 = y0 + 4 * (b / bw);
      uint16_t* dst = j.out.y;
      if (bank_idx == (int)ALF_OFF8) {
        for (int rr = 0; rr < 4; rr++)
          for (int c = 0; c < 4; c++) {
            int x = x4 + c, y = y4 + rr;
            if (x < x1 && y < y1) dst[(size_t)y * j.py + x] = (uint16_t)S.rd(x, y);
          }
        continue;
      }
      // W[r][c] = sample(x4-3+c, y4-3+r): clamp to the fetch rectangle, then
      // read the SAO'd window (the clamped position always lies inside it).
      int W[10][10];
#pragma unroll
      for (int rr = 0; rr < 10; rr++) {
        const uint16_t* row = S.p + (clampi(y4 - 3 + rr, ry0, ry1) - S.oy) * S.pitch - S.ox;
#pragma unroll
        for (int cc = 0; cc < 10; cc++) W[rr][cc] = row[clampi(x4 - 3 + cc, rx0, rx1)];
      }
      const int y_ctb = ry * ctb;
      const bool carve = !(y_ctb + ctb >= H) || (H - y_ctb) > ctb - 4;
      int outv[4][4];
      alf_block(W, s.bank, y4 - y_ctb, ctb, carve, j.alf_bd, amax, outv);
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (y4 + rr < y1 && x4 + c < x1) dst[(size_t)(y4 + rr) * j.py + x4 + c] = (uint16_t)outv[rr][c];
    }
  }
  // chroma ALF + CC-ALF: one sample per thread
  if (chroma) {
    const int cwid = cx1 - cx0, chei = cy1 - cy0, n = cwid * chei;
    const int ctbc = ctb >> 1;
    const int amax = (1 << j.alf_bd) - 1, half = 1 << (j.alf_bd - 1);
    const int ccx0 = rx0 >> 1, ccy0 = ry0 >> 1, ccx1 = ((rx1 + 2) >> 1) - 1, ccy1 = ((ry1 + 2) >> 1) - 1;
    for (int i = threadIdx.x; i < 2 * n; i += blockDim.x) {
      const int c = i / n, k = i % n;
      const int x = cx0 + k % cwid, y = cy0 + k / cwid;
      lwin SC = Sc[0];  // Cb and Cr windows share their geometry
      if (c) SC.p = Sc[1].p;
      int out = SC.rd(x, y);
      if (in_grid) {
        const int alt = (int)((cw >> (c == 0 ? 8 : 16)) & 0xff);
        if (alt != (int)ALF_OFF8) {
          const uint32_t* f = j.alf_cbank + alt * 6;
          auto B = [&](int xx, int yy) -> int { return SC.rd(clampi(xx, ccx0, ccx1), clampi(yy, ccy0, ccy1)); };
          int sft, d1, d2;
          alf_t46(y - ry * ctbc, ctbc, sft, d1, d2);
          const int c0 = B(x, y);
          int sum = alf_pair(f[0], B(x, y + d2), B(x, y - d2), c0);
          sum += alf_pair(f[1], B(x + 1, y + d1), B(x - 1, y - d1), c0);
          sum += alf_pair(f[2], B(x, y + d1), B(x, y - d1), c0);
          sum += alf_pair(f[3], B(x - 1, y + d1), B(x + 1, y - d1), c0);
          sum += alf_pair(f[4], B(x + 2, y), B(x - 2, y), c0);
          sum += alf_pair(f[5], B(x + 1, y), B(x - 1, y), c0);
          out = clampi(c0 + ((sum + (1 << (sft - 1))) >> sft), 0, amax);
        }
        const int ccf = (int)((cw >> (c == 0 ? 24 : 28)) & 0xf);
        if (ccf != (int)ALF_OFF4) {
          const int32_t* cf = j.alf_ccbank + c * ALF_CCBANK_WORDS + ccf * 7;
          const int xl = 2 * x, yl = 2 * y;
          const int y_ctb_c = ry * ctbc;
          const bool lbb = !(yl + ctb >= H && (H - y_ctb_c * 2) <= ctb - 4);
          int yp1, yp2;
          alf_t47((y - y_ctb_c) * 2, ctb, lbb, yp1, yp2);
          auto Lm = [&](int xx, int yy) -> int { return S.rd(clampi(xx, rx0, rx1), clampi(yy, ry0, ry1)); };
          const int cen = Lm(xl, yl);
          int sum = cf[0] * (Lm(xl, yl - yp1) - cen) + cf[1] * (Lm(xl - 1, yl) - cen) + cf[2] * (Lm(xl + 1, yl) - cen) +
                    cf[3] * (Lm(xl - 1, yl + yp1) - cen) + cf[4] * (Lm(xl, yl + yp1) - cen) +
                    cf[5] * (Lm(xl + 1, yl + yp1) - cen) + cf[6] * (Lm(xl, yl + yp2) - cen);
          out = clampi(out + clampi((sum + 64) >> 7, -half, half - 1), 0, amax);
        }
      }
      (c ? j.out.cr : j.out.cb)[(size_t)y * j.pc + x] = (uint16_t)out;
    }
  }
  __syncthreads();  // LDS is reused by the next tile
}

// Persistent launch: the tile list is cut into 8 contiguous chunks, one per
// XCD (workgroups are dispatched round-robin over the 8 XCDs), so spatially
// adjacent tiles - which share halo samples - hit the same L2.
template <int TW, int TH>
__global__ void __launch_bounds__(256) k_fused(const dev_job* __restrict__ jobs, const tile_ref* __restrict__ tiles,
                                               uint32_t n_tiles) {
  __shared__ fused_lds<TW, TH> s;
  const uint32_t xcd = blockIdx.x & 7, k = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const uint32_t chunk = (n_tiles + 7) >> 3;
  const uint32_t begin = xcd * chunk, end = min(n_tiles, begin + chunk);
  int cur_job = -1;
  for (uint32_t t = begin + k; t < end; t += per_xcd) {
    const tile_ref tr = tiles[t];
    const dev_job& j = jobs[tr.job];
    if ((int)tr.job != cur_job) {
      // LMCS inverse table in LDS for bit depths <= 10
      if (j.do_lmcs && j.bd <= 10)
        for (int i = threadIdx.x; i < 1024; i += blockDim.x) s.lut[i] = j.lmcs_lut[i];
      cur_job = (int)tr.job;
      __syncthreads();
    }
    fused_tile<TW, TH>(j, s, tr.x0, tr.y0);
  }
}

// Host-side launch helpers ---------------------------------------------------
size_t fused_lds_bytes(int tw, int th) {
  if (tw == 64 && th == 64) return sizeof(fused_lds<64, 64>);
  if (tw == 128 && th == 64) return sizeof(fused_lds<128, 64>);
  if (tw == 64 && th == 32) return sizeof(fused_lds<64, 32>);
  if (tw == 32 && th == 32) return sizeof(fused_lds<32, 32>);
  return 0;
}

int fused_max_blocks_per_cu(int tw, int th) {
  int n = 0;
  hipError_t e = hipErrorInvalidValue;
  if (tw == 64 && th == 64) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<64, 64>, 256, 0);
  else if (tw == 128 && th == 64) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<128, 64>, 256, 0);
  else if (tw == 64 && th == 32) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<64, 32>, 256, 0);
  else if (tw == 32 && th == 32) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<32, 32>, 256, 0);
  return e == hipSuccess ? n : 0;
}

hipError_t launch_fused(int tw, int th, const dev_job* jobs, const tile_ref* tiles, uint32_t n_tiles, uint32_t grid,
                        hipStream_t s) {
  grid = (grid + 7) & ~7u;
  if (tw == 64 && th == 64) hipLaunchKernelGGL((k_fused<64, 64>), dim3(grid), dim3(256), 0, s, jobs, tiles, n_tiles);
  else if (tw == 128 && th == 64)
    hipLaunchKernelGGL((k_fused<128, 64>), dim3(grid), dim3(256), 0, s, jobs, tiles, n_tiles);
  else if (tw == 64 && th == 32)
    hipLaunchKernelGGL((k_fused<64, 32>), dim3(grid), dim3(256), 0, s, jobs, tiles, n_tiles);
  else if (tw == 32 && th == 32)
    hipLaunchKernelGGL((k_fused<32, 32>), dim3(grid), dim3(256), 0, s, jobs, tiles, n_tiles);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace vvclf
