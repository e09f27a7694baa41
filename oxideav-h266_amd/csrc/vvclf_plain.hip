// vvclf_plain.hip — the plain per-picture form: one kernel per stage on one
// stream, picture resident in HBM.
//   prep      in -> A        LMCS inverse LUT (src/ctu.rs:11837-11849) + copy
//   dbk_lv    A in place     luma vertical edges   (all V before any H,
//   dbk_lh    A in place     luma horizontal edges  src/deblock.rs:528-555)
//   dbk_c     A in place     Cb/Cr vertical, then horizontal
//   sao       A -> B         src/sao.rs:389-626
//   alf_luma  B -> out       classification + 7x7 diamond per 4x4 block
//   alf_chroma B -> out      5x5 chroma ALF + CC-ALF (src/alf.rs:447-523, :1014-1065)
// Deblocking writes in place: edges of one direction never touch samples
// another edge of that direction reads (filter extents <= edge spacing), so
// all segments of a pass run concurrently with the sequential result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "vvclf_filters.h"
#include "vvclf_internal.h"

namespace vvclf {

struct plane_view {
  uint16_t* p;
  int pitch, w, h;
};

__device__ __forceinline__ plane_view pick(const dev_planes& d, const dev_job& j, int comp) {
  plane_view v;
  if (comp == 0) {
    v.p = d.y; v.pitch = j.py; v.w = j.w; v.h = j.h;
  } else {
    v.p = comp == 1 ? d.cb : d.cr; v.pitch = j.pc; v.w = j.wc; v.h = j.hc;
  }
  return v;
}

__device__ __forceinline__ int rdc(const plane_view& v, int x, int y) {
  x = clampi(x, 0, v.w - 1);
  y = clampi(y, 0, v.h - 1);
  return v.p[(size_t)y * v.pitch + x];
}

// ---------------------------------------------------------------------------
// prep: 8 samples per thread, rows of every plane (blockIdx.z = component).
// ---------------------------------------------------------------------------
// A workgroup covers 2048 samples of the rows blockIdx.y + k * gridDim.y (at
// most PREP_GRID_Y row workgroups per plane, so that for luma with LMCS at
// <= 10 bits the first 1024 entries of the inverse table are staged in LDS
// once per workgroup, not once per row).
constexpr int PREP_GRID_Y = 512;

__global__ void __launch_bounds__(256) k_prep(dev_job j) {
  __shared__ uint16_t lut[1024];
  const int comp = blockIdx.z;
  plane_view src = pick(j.in, j, comp), dst = pick(j.a, j, comp);
  const int x = (blockIdx.x * 256 + threadIdx.x) * 8;
  const bool map = comp == 0 && j.do_lmcs;
  const bool lut_lds = map && j.bd <= 10;
  if (lut_lds) {
    for (int i = threadIdx.x; i < 1024; i += 256) lut[i] = j.lmcs_lut[i];
    __syncthreads();
  }
  if (x >= src.w) return;
  for (int y = blockIdx.y; y < src.h; y += gridDim.y) {
    uint4 v = *reinterpret_cast<const uint4*>(src.p + (size_t)y * src.pitch + x);
    if (map) {
      // samples below 1024 through the LDS table with no branch per sample;
      // chunks with a larger sample through the global table
      if (lut_lds && ((v.x | v.y | v.z | v.w) & 0xfc00fc00u) == 0) {
        v = make_uint4(lut_pair(lut, v.x), lut_pair(lut, v.y), lut_pair(lut, v.z), lut_pair(lut, v.w));
      } else {
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++)
          w[i] = (uint32_t)gld(j.lmcs_lut + (w[i] & 0xffffu)) | ((uint32_t)gld(j.lmcs_lut + (w[i] >> 16)) << 16);
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    *reinterpret_cast<uint4*>(dst.p + (size_t)y * dst.pitch + x) = v;
  }
}

// plain copy between buffers (used when a stage is inactive)
__global__ void __launch_bounds__(256) k_copy(dev_job j, dev_planes from, dev_planes to) {
  int comp = blockIdx.z;
  plane_view src = pick(from, j, comp), dst = pick(to, j, comp);
  int y = blockIdx.y;
  int x = (blockIdx.x * 256 + threadIdx.x) * 8;
  if (y >= src.h || x >= src.w) return;
  *reinterpret_cast<uint4*>(dst.p + (size_t)y * dst.pitch + x) =
      *reinterpret_cast<const uint4*>(src.p + (size_t)y * src.pitch + x);
}

// ---------------------------------------------------------------------------
// Luma deblocking.  One lane per sample line; the four lines of a segment sit
// in four consecutive lanes, which exchange the decision terms of lines 0 and
// 3.  Block = 4 waves; each wave covers 16 segments.
// ---------------------------------------------------------------------------
template <bool VERT>
__device__ __forceinline__ void dbk_luma_lane(const dev_job& j, int t) {
  // one list entry (a segment with a nonzero descriptor) per quad of lanes
  const int line = t & 3, it = t >> 2;
  if (it >= (VERT ? j.n_lv : j.n_lh)) return;  // whole quad leaves together
  const uint32_t w = (VERT ? j.l_lv : j.l_lh)[it];
  const int gx = (int)(w & 0xffff), gy = (int)(w >> 16);
  const uint32_t e = (VERT ? j.e_lv : j.e_lh)[(size_t)gy * j.gw + gx];
  plane_view v = pick(j.a, j, 0);
  const int x0 = VERT ? 4 * gx : 4 * gx + line;  // q0 position of this line
  const int y0 = VERT ? 4 * gy + line : 4 * gy;
  int p[8], q[8];
  if (VERT && x0 - 8 >= 0 && x0 + 8 <= v.w) {
    // the line's 16 samples in four 8-byte reads (x0 - 8 is a multiple of 4;
    // a row below the picture reads its last row, as the clamped reads do)
    const uint16_t* rp = v.p + (size_t)clampi(y0, 0, v.h - 1) * v.pitch + (x0 - 8);
    const uint2* r2 = reinterpret_cast<const uint2*>(rp);
    const uint2 a0 = r2[0], a1 = r2[1], b0 = r2[2], b1 = r2[3];
    const uint32_t aw[4] = {a0.x, a0.y, a1.x, a1.y}, bw[4] = {b0.x, b0.y, b1.x, b1.y};
#pragma unroll
    for (int i = 0; i < 8; i++) {
      p[i] = (int)((aw[(7 - i) >> 1] >> (16 * ((7 - i) & 1))) & 0xffffu);
      q[i] = (int)((bw[i >> 1] >> (16 * (i & 1))) & 0xffffu);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      p[i] = VERT ? rdc(v, x0 - 1 - i, y0) : rdc(v, x0, y0 - 1 - i);
      q[i] = VERT ? rdc(v, x0 + i, y0) : rdc(v, x0, y0 + i);
    }
  }
  int beta, tc;
  edge_thresholds(e, j.bd_dbk, kTc, beta, tc);
  int mflp = (e >> E_MFLP_SHIFT) & 7, mflq = (e >> E_MFLQ_SHIFT) & 7;
  luma_terms mine = luma_line_terms(p, q, mflp, mflq);
  luma_terms t0 = quad_terms<0>(mine), t3 = quad_terms<3>(mine);
  luma_decision d = luma_decide(t0, t3, beta, tc, mflp, mflq, (e & E_CTBROW_BIT) != 0);
  int np[8], nq[8], nwp, nwq;
  luma_filter_line(p, q, d, tc, np, nq, nwp, nwq);
  if (e & E_PLTP_BIT) nwp = 0;
  if (e & E_PLTQ_BIT) nwq = 0;
  const int maxv = (1 << j.bd) - 1;
  if (VERT) {
    if (y0 >= v.h) return;
    uint16_t* row = v.p + (size_t)y0 * v.pitch;
#pragma unroll
    for (int i = 0; i < 7; i++) {
      if (i < nwp && x0 - 1 - i >= 0) row[x0 - 1 - i] = (uint16_t)clampi(np[i], 0, maxv);
      if (i < nwq && x0 + i < v.w) row[x0 + i] = (uint16_t)clampi(nq[i], 0, maxv);
    }
  } else {
    if (x0 >= v.w) return;
#pragma unroll
    for (int i = 0; i < 7; i++) {
      if (i < nwp && y0 - 1 - i >= 0) v.p[(size_t)(y0 - 1 - i) * v.pitch + x0] = (uint16_t)clampi(np[i], 0, maxv);
      if (i < nwq && y0 + i < v.h) v.p[(size_t)(y0 + i) * v.pitch + x0] = (uint16_t)clampi(nq[i], 0, maxv);
    }
  }
}

// ---------------------------------------------------------------------------
// Chroma deblocking: one lane per chroma line, the two lines of a segment in
// lanes 2k, 2k+1.  blockIdx.z = Cb / Cr.
// ---------------------------------------------------------------------------
template <bool VERT>
__device__ __forceinline__ void dbk_chroma_lane(const dev_job& j, int t) {
  // one list entry (a segment with a nonzero descriptor) per pair of lanes
  const int line = t & 1, it = t >> 1;
  if (it >= (VERT ? j.n_cv : j.n_ch)) return;  // whole pair leaves together
  const uint32_t w = (VERT ? j.l_cv : j.l_ch)[it];
  const int comp = (int)(w >> 31) + 1, sx = (int)(w & 0xffff), sy = (int)((w >> 16) & 0x7fff);
  const int mw = VERT ? j.cvw : j.chw;
  const uint32_t e = (VERT ? j.e_cv[comp - 1] : j.e_ch[comp - 1])[(size_t)sy * mw + sx];
  plane_view v = pick(j.a, j, comp);
  int x0, y0;  // q0 of this line
  if (VERT) { x0 = 8 * sx; y0 = 2 * sy + line; } else { x0 = 2 * sx + line; y0 = 8 * sy; }
  int c[8];
  if (VERT && x0 - 4 >= 0 && x0 + 4 <= v.w) {
    // the line's 8 samples in two 8-byte reads (x0 - 4 is a multiple of 4)
    const uint2* r2 = reinterpret_cast<const uint2*>(v.p + (size_t)clampi(y0, 0, v.h - 1) * v.pitch + (x0 - 4));
    const uint2 a0 = r2[0], a1 = r2[1];
    const uint32_t aw[4] = {a0.x, a0.y, a1.x, a1.y};
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = (int)((aw[i >> 1] >> (16 * (i & 1))) & 0xffffu);
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = VERT ? rdc(v, x0 - 4 + i, y0) : rdc(v, x0, y0 - 4 + i);
  }
  int beta, tc;
  edge_thresholds(e, j.bd_dbk, kTc, beta, tc);
  int mflp = (e >> E_MFLP_SHIFT) & 7, mflq = (e >> E_MFLQ_SHIFT) & 7;
  bool eligible, short_p;
  if (VERT) { eligible = mflp == 3 && mflq == 3; short_p = false; }
  else { eligible = mflq == 3 && (mflp == 3 || mflp == 1); short_p = mflp == 1; }
  bool mine = eligible && chroma_strong_row(c, beta, tc, short_p);
  bool other = __shfl_xor((int)mine, 1) != 0;
  bool strong = eligible && mine && other;
  int n[8];
  uint32_t wm = chroma_filter_line(c, strong, short_p, tc, n);
  const int maxv = (1 << j.bd) - 1;
  if (e & E_PLTP_BIT) wm &= 0xf0u;
  if (e & E_PLTQ_BIT) wm &= 0x0fu;
#pragma unroll
  for (int i = 1; i < 7; i++) {
    if (!((wm >> i) & 1)) continue;
    int x = VERT ? x0 - 4 + i : x0, y = VERT ? y0 : y0 - 4 + i;
    if (x < 0 || y < 0 || x >= v.w || y >= v.h) continue;
    v.p[(size_t)y * v.pitch + x] = (uint16_t)clampi(n[i], 0, maxv);
  }
}

// One launch per direction: workgroups [0, luma_blocks) filter luma
// segments, the rest chroma segments (the planes are independent).
template <bool VERT>
__global__ void __launch_bounds__(256) k_dbk(dev_job j, int luma_blocks) {
  if ((int)blockIdx.x < luma_blocks) dbk_luma_lane<VERT>(j, blockIdx.x * 256 + threadIdx.x);
  else dbk_chroma_lane<VERT>(j, (blockIdx.x - luma_blocks) * 256 + threadIdx.x);
}

// ---------------------------------------------------------------------------
// SAO: A -> B, a strip of 8 samples by SAO_ROWS rows per thread, inside one
// CTB (SAO CTBs are at least 16 luma / 8 chroma samples wide and high; strips
// are aligned to 8 samples and SAO_ROWS rows).  blockIdx.z = component.  The
// strip's own row is read with the CTB parameters, the rows above and below
// only by waves holding an edge class that reads them (vertical and diagonal
// classes: band offset, off and the horizontal class need the sample's row
// alone); rows are kept in registers; rows whose CTB has every offset in i8
// use the packed
// table form, the other rows (and edge-offset rows whose neighbourhoods leave
// the picture or the CTB's rectangle) go sample by sample.
// ---------------------------------------------------------------------------
constexpr int SAO_ROWS = 1;
static_assert(SAO_ROWS == 1, "k_sao gives each lane of a wave its own row");
// log2 of the lanes per row of a k_sao wave for a CTB 2^log2c samples wide
__host__ __device__ inline int sao_lanes_log2(int log2c) { return log2c < 3 ? 0 : (log2c > 9 ? 6 : log2c - 3); }

__global__ void __launch_bounds__(256) k_sao(dev_job j) {
  const int comp = blockIdx.z;
  plane_view src = pick(j.a, j, comp), dst = pick(j.b, j, comp);
  const int sub = comp ? 1 : 0;
  const int log2c = j.sao_ctb_log2 - sub;
  // A wave covers a CTB-wide strip: 2^l2l lanes of 8 samples per row and
  // 64 >> l2l rows, so that its lanes read one CTB's parameters (two when the
  // CTB holds fewer than 512 samples) and take one filter path.  The four
  // waves of a workgroup are stacked vertically.
  const int l2l = sao_lanes_log2(log2c);
  const int lane = threadIdx.x & 63, seg = lane & ((1 << l2l) - 1);
  const int ys = ((int)blockIdx.y * 4 + (int)(threadIdx.x >> 6)) * (64 >> l2l) + (lane >> l2l);
  const int x0 = (((int)blockIdx.x << l2l) + seg) * 8;
  if (ys >= src.h || x0 >= src.w) return;  // lanes of a row leave in x order: a lane's right neighbour
                                           // lane is present whenever x0 + 8 is inside the picture
  const int nr = min(SAO_ROWS, src.h - ys);
  const bool active = comp == 0 ? j.sao_luma : j.sao_chroma;
  const int maxv = (1 << j.sao_bd) - 1;
  // Rows ys-1 .. ys+SAO_ROWS (clamped) as 8-sample reads, and per row the
  // sample beyond the wave's end (x0-1 on lane 0, x0+8 elsewhere, clamped),
  // all issued with the parameter read; rows are padded to the pitch (a
  // multiple of 64), so the reads stay inside the plane past its width.
  const int ex = seg == 0 ? max(x0 - 1, 0) : min(x0 + 8, src.w - 1);
  uint4 q[SAO_ROWS + 2];
  uint32_t e[SAO_ROWS + 2];
  {
    const uint16_t* row = src.p + (size_t)ys * src.pitch;
    q[1] = *reinterpret_cast<const uint4*>(row + x0);
    e[1] = row[ex];
  }
  const int rx = x0 >> log2c, ry = ys >> log2c;
  uint64_t prm = 0;
  size_t ci = 0;
  if (active && rx < j.swctb && ry < j.shctb) {
    ci = (size_t)ry * j.swctb + rx;
    prm = j.sao[ci * 3 + comp];
  }
  // the rows above and below only where an edge class reads them (band
  // offset, off and the horizontal class read the sample's own row)
  if (__any((prm & 3) == 2 && ((prm >> 2) & 3) != 0)) {
#pragma unroll
    for (int r = 0; r < SAO_ROWS + 2; r += 2) {
      const uint16_t* row = src.p + (size_t)clampi(ys - 1 + r, 0, src.h - 1) * src.pitch;
      q[r] = *reinterpret_cast<const uint4*>(row + x0);
      e[r] = row[ex];
    }
  } else {
    q[0] = q[2] = q[1];
    e[0] = e[2] = e[1];
  }
  // samples x0-1 and x0+8 of each row: from the neighbouring lanes (which
  // hold the adjacent chunks), at the wave's ends from e[]
  sao_row8 Rw[SAO_ROWS + 2];
#pragma unroll
  for (int r = 0; r < SAO_ROWS + 2; r++) {
    const uint4& v = q[r];
    Rw[r].w[0] = v.x; Rw[r].w[1] = v.y; Rw[r].w[2] = v.z; Rw[r].w[3] = v.w;
    const uint32_t up = (uint32_t)__shfl_up((int)(v.w >> 16), 1), down = (uint32_t)__shfl_down((int)(v.x & 0xffffu), 1);
    Rw[r].l = seg == 0 ? (x0 > 0 ? e[r] : (v.x & 0xffffu)) : up;
    Rw[r].r = x0 + 8 >= src.w ? (v.w >> 16) : (seg == (1 << l2l) - 1 ? e[r] : down);
  }
  const sao_run run = sao_run_of(prm);
  const bool packed = run.fits && j.sao_bd <= 15;
  const uint32_t maxv2 = (uint32_t)maxv * 0x10001u;
  const int cls = run.aux & 3;
  // the strip's neighbourhoods inside the picture and the CTB rectangle?
  int cx0 = 0, cy0 = 0, cx1 = src.w, cy1 = src.h;
  if (run.type == 2 && j.sao_rect) {
    const uint64_t r = j.sao_rect[ci];
    cx0 = (int)(r & 0xffff); cy0 = (int)((r >> 16) & 0xffff); cx1 = (int)((r >> 32) & 0xffff); cy1 = (int)(r >> 48);
    if (sub) { cx0 >>= 1; cy0 >>= 1; cx1 = (cx1 + 1) >> 1; cy1 = (cy1 + 1) >> 1; }
    cx0 = max(cx0, 0); cy0 = max(cy0, 0); cx1 = min(cx1, src.w); cy1 = min(cy1, src.h);
  }
  const bool xin = x0 - 1 >= cx0 && x0 + 8 < cx1;
  // f(r, o) filters row ys + r into o
  auto each_row = [&](auto&& f) {
#pragma unroll
    for (int r = 0; r < SAO_ROWS; r++) {
      if (r >= nr) break;
      uint32_t o[4] = {Rw[r + 1].w[0], Rw[r + 1].w[1], Rw[r + 1].w[2], Rw[r + 1].w[3]};
      f(r, o);
      uint16_t* drow = dst.p + (size_t)(ys + r) * dst.pitch;
      if (x0 + 8 <= src.w) {
        *reinterpret_cast<uint4*>(drow + x0) = make_uint4(o[0], o[1], o[2], o[3]);
      } else {
        for (int i = 0; i < 8; i++)
          if (x0 + i < src.w) drow[x0 + i] = (uint16_t)(o[i >> 1] >> (16 * (i & 1)));
      }
    }
  };
  if (run.type == 1) {
    each_row([&](int r, uint32_t (&o)[4]) {
      if (packed) {
        sao_bo8(Rw[r + 1], run.aux, run.btab, j.sao_bd, maxv2, o);
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) sao_put(o, i, sao_band_sample(sao_at(Rw[r + 1], i), prm, j.sao_bd, maxv));
      }
    });
  } else if (run.type == 2) {
    int dx0, dy0, dx1, dy1;
    eo_offsets(cls, dx0, dy0, dx1, dy1);
    each_row([&](int r, uint32_t (&o)[4]) {
      const int y = ys + r;
      if (packed && xin && y - 1 >= cy0 && y + 1 < cy1)
        sao_eo8_any(cls, Rw[r], Rw[r + 1], Rw[r + 2], run.btab, maxv2, o);
      else  // per sample, with the picture / rectangle check of each neighbour
        sao_eo8_checked(dx0, dy0, dx1, dy1, Rw[r], Rw[r + 1], Rw[r + 2], x0, y, cx0, cy0, cx1, cy1, prm, maxv, o);
    });
  } else {  // off: copy
    each_row([&](int, uint32_t (&)[4]) {});
  }
}

// ---------------------------------------------------------------------------
// ALF luma: one thread per 4x4 block.  The block's 10x10 neighbourhood
// (-3..+6) feeds both the classification window and the 7x7 diamond.
// A workgroup covers one tile inside a CTB, 2^tw_log2 x 2^th_log2 samples;
// each block reads its class's 12 tap words of the CTB's permuted filter bank
// straight from memory (cached), so that no wave waits on a workgroup
// barrier.  The two block rows next to the virtual boundary (CtbSizeY - 8 and
// - 4) take their own classification and filter variants: with 128-wide
// tiles a wave holds two block rows, so that only the wave holding both of
// those rows runs two variants (with 64 x 64 tiles, four rows per wave, the
// wave holding them ran three: 37.8 -> 30.7 us per 4K picture).
// ---------------------------------------------------------------------------
template <bool HBD>
__global__ void __launch_bounds__(256) k_alf_luma(dev_job j, int tw_log2, int th_log2) {
  const int tb = 1 << (tw_log2 - 2), nby = 1 << (th_log2 - 2);
  const int bx = threadIdx.x % tb, by = threadIdx.x / tb;
  const int tx0 = blockIdx.x << tw_log2, ty0 = blockIdx.y << th_log2;
  const int ctb = 1 << j.ctb_log2;
  const int rx = tx0 >> j.ctb_log2, ry = ty0 >> j.ctb_log2;
  plane_view src = pick(j.b, j, 0), dst = pick(j.out, j, 0);
  // CTBs outside the ALF picture's grid are left unfiltered (copied).
  const bool in_grid = rx < j.wctb && ry < j.hctb;
  const size_t ci = in_grid ? (size_t)ry * j.wctb + rx : 0;
  uint32_t cw = (j.do_alf && in_grid) ? j.alf_ctb[ci] : 0xffffffffu;
  int bank_idx = (int)(cw & 0xff);
  const int x4 = tx0 + bx * 4, y4 = ty0 + by * 4;
  if (bank_idx == (int)ALF_OFF8) {
    if (by >= nby) return;
    for (int r = 0; r < 4; r++) {
      int y = y4 + r;
      if (y >= src.h) break;
      for (int c = 0; c < 4; c++) {
        int x = x4 + c;
        if (x < src.w) dst.p[(size_t)y * dst.pitch + x] = src.p[(size_t)y * src.pitch + x];
      }
    }
    return;
  }
  const uint32_t* bank = j.alf_bank + (size_t)bank_idx * ALF_BANK_WORDS;
  if (by >= nby) return;
  if (x4 >= src.w || y4 >= src.h) return;
  uint64_t r = j.alf_rect[ci];
  const int cx0 = (int)(r & 0xffff), cy0 = (int)((r >> 16) & 0xffff), cx1 = (int)((r >> 32) & 0xffff) - 1,
            cy1 = (int)(r >> 48) - 1;
  const int y_ctb = ry * ctb;
  const bool carve = !(y_ctb + ctb >= j.h) || (j.h - y_ctb) > ctb - 4;
  const int amax = (1 << j.alf_bd) - 1;
  uint32_t outp[4][2];
  if (HBD) {
    // 10x10 window, rect-clamped: W[r][c] = sample(x4-3+c, y4-3+r)
    int W[10][10];
#pragma unroll
    for (int rr = 0; rr < 10; rr++) {
      int yy = clampi(y4 - 3 + rr, cy0, cy1);
      const uint16_t* row = src.p + (size_t)yy * src.pitch;
#pragma unroll
      for (int cc = 0; cc < 10; cc++) W[rr][cc] = row[clampi(x4 - 3 + cc, cx0, cx1)];
    }
    int outv[4][4];
    alf_block(W, bank, y4 - y_ctb, ctb, carve, j.alf_bd, amax, outv);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      outp[rr][0] = (uint32_t)outv[rr][0] | ((uint32_t)outv[rr][1] << 16);
      outp[rr][1] = (uint32_t)outv[rr][2] | ((uint32_t)outv[rr][3] << 16);
    }
  } else {
    // packed window P[r][k] = samples (x4-4+2k, x4-3+2k) of row y4-3+r
    uint32_t P[10][6];
    if (x4 - 3 >= cx0 && x4 + 6 <= cx1 && y4 - 3 >= cy0 && y4 + 6 <= cy1) {
      // inside the rectangle: three aligned 8-byte reads per row (x4 - 4 is
      // a multiple of 4 and plane pitches are multiples of 64); rows as
      // 24-bit offsets from the window's first row
      const uint16_t* w0 = src.p + (size_t)(y4 - 3) * src.pitch + x4 - 4;
#pragma unroll
      for (int rr = 0; rr < 10; rr++) {
        const uint2* q = reinterpret_cast<const uint2*>(w0 + __mul24(rr, src.pitch));
        const uint2 a = q[0], b = q[1], c = q[2];
        P[rr][0] = a.x; P[rr][1] = a.y; P[rr][2] = b.x; P[rr][3] = b.y; P[rr][4] = c.x; P[rr][5] = c.y;
      }
    } else {
#pragma unroll
      for (int rr = 0; rr < 10; rr++) {
        const uint16_t* row = src.p + (size_t)clampi(y4 - 3 + rr, cy0, cy1) * src.pitch;
        uint32_t v[12];
        v[0] = 0;
        v[11] = 0;
#pragma unroll
        for (int q = 1; q <= 10; q++) v[q] = row[clampi(x4 - 4 + q, cx0, cx1)];
#pragma unroll
        for (int k = 0; k < 6; k++) P[rr][k] = v[2 * k] | (v[2 * k + 1] << 16);
      }
    }
    // 16-bit lanes are exact for samples below 2^15; a window holding a
    // larger sample (beyond the bit depth) takes the 32-bit form
    uint32_t any = 0;
#pragma unroll
    for (int rr = 0; rr < 10; rr++)
#pragma unroll
      for (int k = 0; k < 6; k++) any |= P[rr][k];
    if (any & 0x80008000u) {
      auto at = [&](int dx, int dy) -> int {
        return src.p[(size_t)clampi(y4 + dy, cy0, cy1) * src.pitch + clampi(x4 + dx, cx0, cx1)];
      };
      const int vb = y4 - y_ctb == ctb - 8 ? 1 : (y4 - y_ctb == ctb - 4 ? 2 : 0);
      const int cls = alf_class_at(at, vb, carve, j.alf_bd);
      alf_filter_at(at, bank + ((cls & 31) * 4 + (cls >> 5)) * 12, vb, amax, [&](int c, int r, int v) {
        if (x4 + c < src.w && y4 + r < src.h) dst.p[(size_t)(y4 + r) * dst.pitch + x4 + c] = (uint16_t)v;
      });
      return;
    } else {
      // every window sample within the bit depth: taps whose clip value
      // exceeds it take the unclipped form
      const bool in_range = (((any & 0xffffu) | (any >> 16)) >> j.alf_bd) == 0u;
      alf_block_pk(P, bank, y4 - y_ctb, ctb, carve, j.alf_bd, amax, outp, in_range);
    }
  }
  if (x4 + 4 <= src.w) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
      if (y4 + rr < src.h)
        *reinterpret_cast<uint2*>(dst.p + (size_t)(y4 + rr) * dst.pitch + x4) = make_uint2(outp[rr][0], outp[rr][1]);
  } else {
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (y4 + rr < src.h && x4 + c < src.w)
          dst.p[(size_t)(y4 + rr) * dst.pitch + x4 + c] = (uint16_t)(outp[rr][c >> 1] >> (16 * (c & 1)));
  }
}

// ---------------------------------------------------------------------------
// ALF chroma + CC-ALF: one thread per chroma sample, B -> out.
// ---------------------------------------------------------------------------
// log2 of the lanes per row of a k_alf_chroma wave for a chroma CTB 2^log2c
// samples wide (two samples per lane)
__host__ __device__ inline int alf_chroma_lanes_log2(int log2c) {
  return log2c < 1 ? 0 : (log2c > 7 ? 6 : log2c - 1);
}

__global__ void __launch_bounds__(256) k_alf_chroma(dev_job j) {
  // two horizontally adjacent samples (x, x+1), x even, per thread
  // a wave covers a chroma-CTB-wide strip (2^l2l lanes of two samples per
  // row, 64 >> l2l rows) so that its lanes share one CTB's filters and path;
  // the four waves of a workgroup are stacked vertically
  const int comp = blockIdx.z + 1;
  plane_view src = pick(j.b, j, comp), dst = pick(j.out, j, comp);
  const int l2l = alf_chroma_lanes_log2(j.ctb_log2 - 1);
  const int lane = threadIdx.x & 63;
  const int x = 2 * (((int)blockIdx.x << l2l) + (lane & ((1 << l2l) - 1)));
  const int y = ((int)blockIdx.y * 4 + (int)(threadIdx.x >> 6)) * (64 >> l2l) + (lane >> l2l);
  if (x >= src.w || y >= src.h) return;
  // rows are padded to the pitch (a multiple of 64): the pair read stays in
  // the plane even when x + 1 is past the width
  const uint16_t* srow = src.p + (size_t)y * src.pitch;
  uint32_t outp = *reinterpret_cast<const uint32_t*>(srow + x);
  const int ctb = 1 << j.ctb_log2, ctbc = ctb >> 1;
  const int rx = (2 * x) >> j.ctb_log2, ry = (2 * y) >> j.ctb_log2;  // one CTB for both samples
  if (j.do_alf && rx < j.wctb && ry < j.hctb) {
    const size_t ci = (size_t)ry * j.wctb + rx;
    const uint32_t cw = j.alf_ctb[ci];
    const uint64_t r = j.alf_rect[ci];
    const int lx0 = (int)(r & 0xffff), ly0 = (int)((r >> 16) & 0xffff), lx1 = (int)((r >> 32) & 0xffff),
              ly1 = (int)(r >> 48);
    const int maxv = (1 << j.alf_bd) - 1;
    const int alt = (int)((cw >> (comp == 1 ? 8 : 16)) & 0xff);
    if (alt != (int)ALF_OFF8) {
      const int cx0 = lx0 >> 1, cy0 = ly0 >> 1, cx1 = ((lx1 + 1) >> 1) - 1, cy1 = ((ly1 + 1) >> 1) - 1;
      const uint32_t* f = j.alf_cbank + alt * 6;
      int sft, y1, y2;
      alf_t46(y - ry * ctbc, ctbc, sft, y1, y2);
      bool done = false;
      if (j.alf_bd <= 12 && x - 2 >= cx0 && x + 3 <= cx1 && y - 2 >= cy0 && y + 2 <= cy1) {
        // neighbourhood inside the rectangle: packed pairs (see the fused form)
        const uint16_t* c0p = src.p + (size_t)y * src.pitch + x;  // rows as 24-bit offsets from here
        auto P2 = [&](int dx, int dy) -> uint32_t {
          return *reinterpret_cast<const uint32_t*>(c0p + (__mul24(dy, src.pitch) + dx));
        };
        const uint32_t m2 = P2(-2, 0), m0 = outp, p2 = P2(2, 0);
        const uint32_t am2 = P2(-2, y1), a0 = P2(0, y1), ap2 = P2(2, y1);
        const uint32_t bm2 = P2(-2, -y1), b0 = P2(0, -y1), bp2 = P2(2, -y1);
        const uint32_t u2 = P2(0, y2), l2 = P2(0, -y2);
        // 16-bit lanes are exact for samples below 2^15 (else the 32-bit form)
        done = ((m2 | m0 | p2 | am2 | a0 | ap2 | bm2 | b0 | bp2 | u2 | l2) & 0x80008000u) == 0;
        if (done) {
          const s16x2 cur = as_s16x2(m0);
          int a_0 = 1 << (sft - 1), a_1 = a_0;
          auto tap = [&](int t, uint32_t a, uint32_t b) {
            const uint32_t w = f[t], kk = w >> 16;
            alf_pk_tap(w & 0xffffu, w << 16, kk | ((0u - kk) << 16), cur, a, b, a_0, a_1);
          };
          tap(0, u2, l2);
          tap(1, __builtin_amdgcn_alignbit(ap2, a0, 16), __builtin_amdgcn_alignbit(b0, bm2, 16));
          tap(2, a0, b0);
          tap(3, __builtin_amdgcn_alignbit(a0, am2, 16), __builtin_amdgcn_alignbit(bp2, b0, 16));
          tap(4, p2, m2);
          tap(5, __builtin_amdgcn_alignbit(p2, m0, 16), __builtin_amdgcn_alignbit(m0, m2, 16));
          const s16x2 d = __builtin_amdgcn_cvt_pk_i16(a_0 >> sft, a_1 >> sft);
          const s16x2 zero2 = {0, 0}, vmax2 = {(short)maxv, (short)maxv};
          outp = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(__builtin_elementwise_add_sat(cur, d), zero2),
                                                   vmax2));
        }
      }
      if (!done) {
        uint32_t o2[2];
        for (int h = 0; h < 2; h++) {
          const int xs = x + h;
          auto S = [&](int dx, int dy) -> int {
            return src.p[(size_t)clampi(y + dy, cy0, cy1) * src.pitch + clampi(xs + dx, cx0, cx1)];
          };
          const int c0 = S(0, 0);
          int sum = 0;
          sum += alf_pair(f[0], S(0, y2), S(0, -y2), c0);
          sum += alf_pair(f[1], S(1, y1), S(-1, -y1), c0);
          sum += alf_pair(f[2], S(0, y1), S(0, -y1), c0);
          sum += alf_pair(f[3], S(-1, y1), S(1, -y1), c0);
          sum += alf_pair(f[4], S(2, 0), S(-2, 0), c0);
          sum += alf_pair(f[5], S(1, 0), S(-1, 0), c0);
          o2[h] = (uint32_t)clampi(c0 + ((sum + (1 << (sft - 1))) >> sft), 0, maxv);
        }
        outp = o2[0] | (o2[1] << 16);
      }
    }
    const int ccf = (int)((cw >> (comp == 1 ? 24 : 28)) & 0xf);
    if (ccf != (int)ALF_OFF4) {
      const int32_t* cf = j.alf_ccbank + (comp - 1) * ALF_CCBANK_WORDS + ccf * 7;
      plane_view lum = pick(j.b, j, 0);
      const int yl = 2 * y;
      const int y_ctb_c = ry * ctbc;
      const int ph = j.h;
      const bool lbb = !(yl + ctb >= ph && (ph - y_ctb_c * 2) <= ctb - 4);
      int yp1, yp2;
      alf_t47((y - y_ctb_c) * 2, ctb, lbb, yp1, yp2);
      const int half = 1 << (j.alf_bd - 1);
      // the four luma rows (rect-clamped), as 24-bit offsets from the centre row
      const int rc = clampi(yl, ly0, ly1 - 1);
      const uint16_t* lc = lum.p + (size_t)rc * lum.pitch;
      auto lrow = [&](int dy) { return lc + __mul24(clampi(yl + dy, ly0, ly1 - 1) - rc, lum.pitch); };
      const uint16_t *la = lrow(-yp1), *lb = lrow(yp1), *ld = lrow(yp2);
      const cc_taps k(cf);
      const int xl = 2 * x;  // a multiple of 4
      int sum[2];
      if (xl - 1 >= lx0 && xl + 3 <= lx1 - 1) {
        // luma columns xl-1 .. xl+3 inside the rectangle: aligned pair reads
        const uint2 A = *reinterpret_cast<const uint2*>(la + xl), C = *reinterpret_cast<const uint2*>(lc + xl);
        const uint2 B = *reinterpret_cast<const uint2*>(lb + xl), D = *reinterpret_cast<const uint2*>(ld + xl);
        const uint32_t Cm = *reinterpret_cast<const uint32_t*>(lc + xl - 2);
        const uint32_t Bm = *reinterpret_cast<const uint32_t*>(lb + xl - 2);
        sum[0] = k.sum(lo16(A.x), hi16(Cm), hi16(C.x), hi16(Bm), lo16(B.x), hi16(B.x), lo16(D.x), lo16(C.x));
        sum[1] = k.sum(lo16(A.y), hi16(C.x), hi16(C.y), hi16(B.x), lo16(B.y), hi16(B.y), lo16(D.y), lo16(C.y));
      } else {
        for (int h = 0; h < 2; h++) {
          auto cx = [&](int dx) { return clampi(xl + 2 * h + dx, lx0, lx1 - 1); };
          sum[h] = k.sum(la[cx(0)], lc[cx(-1)], lc[cx(1)], lb[cx(-1)], lb[cx(0)], lb[cx(1)], ld[cx(0)], lc[cx(0)]);
        }
      }
      uint32_t o2[2];
      for (int h = 0; h < 2; h++) {
        const int base = (int)((outp >> (16 * h)) & 0xffffu);
        o2[h] = (uint32_t)clampi(base + clampi((sum[h] + 64) >> 7, -half, half - 1), 0, maxv);
      }
      outp = o2[0] | (o2[1] << 16);
    }
  }
  uint16_t* drow = dst.p + (size_t)y * dst.pitch;
  if (x + 1 < src.w) *reinterpret_cast<uint32_t*>(drow + x) = outp;
  else drow[x] = (uint16_t)outp;
}

// ---------------------------------------------------------------------------
// Host launcher for the plain form.
// ---------------------------------------------------------------------------
static inline unsigned cdiv(unsigned a, unsigned b) { return (a + b - 1) / b; }

// Stage order of the plain form; `ev` (optional, PLAIN_STAGES + 1 events)
// brackets every stage so callers can time each kernel on this stream.
hipError_t launch_plain(const dev_job& j, hipStream_t s, hipEvent_t* ev) {
  const int ncomp = j.wc > 0 ? 3 : 1;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  mark(0);
  // prep: in -> A (rows of 2048 samples per block)
  hipLaunchKernelGGL(k_prep, dim3(cdiv(j.w, 2048), std::min(j.h, PREP_GRID_Y), ncomp), dim3(256), 0, s, j);
  mark(1);
  // deblocking: all vertical edges (luma and chroma in one launch), then all
  // horizontal edges
  const bool dc = j.do_dbk && j.do_dbk_chroma && ncomp == 3;
  {
    const unsigned lb = j.do_dbk ? cdiv(4u * j.n_lv, 256) : 0, cb = dc ? cdiv(2u * j.n_cv, 256) : 0;
    if (lb + cb) hipLaunchKernelGGL(k_dbk<true>, dim3(lb + cb), dim3(256), 0, s, j, (int)lb);
  }
  mark(2);
  {
    const unsigned lb = j.do_dbk ? cdiv(4u * j.n_lh, 256) : 0, cb = dc ? cdiv(2u * j.n_ch, 256) : 0;
    if (lb + cb) hipLaunchKernelGGL(k_dbk<false>, dim3(lb + cb), dim3(256), 0, s, j, (int)lb);
  }
  mark(3);
  if (j.do_sao) {
    // strips of 8 << l2l samples, 4 * (64 >> l2l) rows per workgroup; the
    // grid covers the largest plane in each direction
    unsigned gx = 0, gy = 0;
    for (int c = 0; c < ncomp; c++) {
      const int l = sao_lanes_log2(j.sao_ctb_log2 - (c ? 1 : 0));
      gx = std::max(gx, cdiv(c ? j.wc : j.w, 8u << l));
      gy = std::max(gy, cdiv(c ? j.hc : j.h, 4u * (64u >> l)));
    }
    hipLaunchKernelGGL(k_sao, dim3(gx, gy, ncomp), dim3(256), 0, s, j);
  }
  mark(4);
  if (j.do_alf) {
    const int twl = j.ctb_log2 < 6 ? j.ctb_log2 : (j.ctb_log2 >= 7 ? 7 : 6);
    const int thl = j.ctb_log2 < 6 ? j.ctb_log2 : (j.ctb_log2 >= 7 ? 5 : 6);
    const unsigned threads = (1u << (twl - 2)) << (thl - 2);
    const dim3 grid(cdiv(j.w, 1u << twl), cdiv(j.h, 1u << thl));
    if (j.alf_bd > 12)
      hipLaunchKernelGGL(k_alf_luma<true>, grid, dim3(threads < 64 ? 64 : threads), 0, s, j, twl, thl);
    else
      hipLaunchKernelGGL(k_alf_luma<false>, grid, dim3(threads < 64 ? 64 : threads), 0, s, j, twl, thl);
  }
  mark(5);
  if (j.do_alf && ncomp == 3) {
    const int l = alf_chroma_lanes_log2(j.ctb_log2 - 1);
    hipLaunchKernelGGL(k_alf_chroma, dim3(cdiv(j.wc, 2u << l), cdiv(j.hc, 4u * (64u >> l)), 2), dim3(256), 0, s, j);
  }
  mark(6);
  return hipGetLastError();
}

}  // namespace vvclf
