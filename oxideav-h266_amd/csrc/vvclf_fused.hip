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

#include <type_traits>

#include "vvclf_filters.h"
#include "vvclf_internal.h"

namespace vvclf {

template <int TW, int TH>
struct fused_lds {
  // threads per workgroup (four waves, one per SIMD; a workgroup of six
  // waves would place two on some SIMDs, which the register budget of a
  // second workgroup then cannot match)
  static constexpr int NT = 256;
  static constexpr int LW = TW + 32, LH = TH + 32;          // luma input region
  static constexpr int CW = TW / 2 + 8, CH = TH / 2 + 8;    // chroma input region
  static constexpr int SW = TW + 8, SH = TH + 6;            // SAO'd luma, origin (x0-4, y0-3)
  static constexpr int SCW = TW / 2 + 4, SCH = TH / 2 + 4;  // SAO'd chroma, origin (cx0-2, cy0-2)
  // bounds of the edge-descriptor rectangles a tile's deblocking visits
  static constexpr int MLV_W = TW / 4 + 6, MLV_H = TH / 4 + 9;  // luma vertical edges
  static constexpr int MLH_W = TW / 4 + 3, MLH_H = TH / 4 + 6;  // luma horizontal edges
  static constexpr int MCV_W = TW / 16 + 3, MCV_H = TH / 4 + 5; // chroma vertical edges
  static constexpr int MCH_W = TW / 4 + 5, MCH_H = TH / 16 + 3; // chroma horizontal edges
  struct maps_t {
    uint32_t lv[MLV_H * MLV_W], lh[MLH_H * MLH_W], cv[2][MCV_H * MCV_W], ch[2][MCH_H * MCH_W];
    // lists of the nonzero descriptors: (comp << 15) | (row << 6) | column
    uint16_t llv[MLV_H * MLV_W], llh[MLH_H * MLH_W], lcv[2 * MCV_H * MCV_W], lch[2 * MCH_H * MCH_W];
  };
  struct sao_out_t {
    uint16_t S[SH * SW];
    uint16_t Sc[2][SCH * SCW];
  };
  // L / Cc also hold the tile's output between the ALF and the final stores
  alignas(16) uint16_t L[LH * LW];
  alignas(16) uint16_t Cc[2][CH * CW];
  // the edge descriptors are dead once deblocking is done, before SAO writes
  union {
    maps_t m;
    sao_out_t o;
  };
  uint32_t bank[ALF_BANK_WORDS];
  uint32_t cbank[ALF_CBANK_WORDS];
  int32_t ccbank[2 * ALF_CCBANK_WORDS];
  uint64_t saop[9][3];  // SAO parameters of the (up to) 3x3 CTBs around the tile
  uint4 saorun[9][3];   // the same decoded for the run form (lds_sao)
  uint64_t saor[9];     // and their sample rectangles
  uint64_t safe[9][3];  // per CTB and component: [x0, x1) x [y0, y1) whose 8-neighbourhoods lie
                        // inside the picture and the CTB's rectangle (u16 fields)
  uint32_t nlist[4];    // entries in llv, llh, lcv, lch
  uint32_t wide;        // the tile's windows hold a sample of 2^15 or more
  uint16_t lut[1024];
  int16_t tct[68];  // the deblocking tC' table, copied once per workgroup
  static_assert(MLV_W <= 64 && MLH_W <= 64 && MCV_W <= 64 && MCH_W <= 64, "list column field");
  static_assert(sizeof(maps_t) <= sizeof(sao_out_t), "maps alias the SAO output");
  static_assert(TW <= LW && TH <= LH && TW / 2 <= CW && TH / 2 <= CH, "output buffers fit the input windows");
  static_assert(TW % 16 == 0, "16-byte output chunks");
};

// The thread index, opaque to the compiler: index arithmetic derived from it
// is then recomputed where it is used instead of being hoisted out of the
// tile loop, where it would hold dozens of registers for the whole kernel.
__device__ __forceinline__ int ltid() {
  int t;
  __asm__ volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
  return t;
}

// A rectangle of an edge-descriptor map (map units), staged in LDS with
// origin (x0, y0).
struct mrect {
  int x0, y0, nx, ny;
};

// Global -> register -> LDS staging of a map rectangle (at most MH x MW
// words, LDS row pitch MW); issue() for every rectangle first, then
// commit(), so that all loads of a tile are in flight together.  Lanes
// follow the rectangle's rows: a wave covers 64 / LPR rows per step, LPR
// lanes (a power of two >= MW) per row, so that a lane's column is its lane
// index within the row and a step's row is uniform across those lanes; no
// index needs a division, and each wave walks its own rows.
template <int MW, int MH, int NT>
struct rect_stage {
  static constexpr int LPR = MW <= 16 ? 16 : (MW <= 32 ? 32 : 64);  // lanes per row
  static constexpr int RPI = 64 / LPR;                              // rows per wave and step
  static constexpr int NW = NT / 64;                                // waves
  static constexpr int PER = (MH + NW * RPI - 1) / (NW * RPI);      // steps
  static_assert(MW <= 64 && NT % 64 == 0, "one row per LPR lanes");
  uint32_t v[PER];
  // row of step k for the calling lane; the column is col()
  __device__ __forceinline__ static int row(int k) {
    const int t = ltid();
    return (k * NW + (t >> 6)) * RPI + ((t & 63) / LPR);
  }
  __device__ __forceinline__ static int col() { return ltid() & (LPR - 1); }
  // map == nullptr: nothing to load (every element 0)
  __device__ __forceinline__ void issue(const uint32_t* map, int mw, const mrect& r) {
    // 64-bit base once; element offsets in 32 bits (rows < MH, map widths
    // below 2^24)
    const uint32_t* base = map ? map + (size_t)r.y0 * mw + r.x0 : nullptr;
    const int xx = col();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int yy = row(k);
      v[k] = (map && yy < r.ny && xx < r.nx) ? gld(base + (__mul24(yy, mw) + xx)) : 0u;
    }
  }
  // Stores the nonzero words (the only ones the deblocking reads) and lists
  // them as tag | (row << 6) | column, reserving the wave's list entries with
  // one LDS atomic.  Called by all lanes of the wave.
  __device__ __forceinline__ void commit(uint32_t* dst, const mrect& r, uint16_t* list, uint32_t* cnt,
                                         uint32_t tag) const {
    uint64_t m[PER];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      m[k] = __ballot(v[k] != 0);  // elements outside the rectangle hold 0
      tot += (uint32_t)__popcll(m[k]);
    }
    if (tot == 0) return;
    uint32_t base = 0;
    if ((ltid() & 63) == 0) base = atomicAdd(cnt, tot);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    const int xx = col();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m[k] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[k], 0u));
      if (v[k] != 0) {
        const int yy = row(k);
        dst[yy * MW + xx] = v[k];
        list[base + below] = (uint16_t)(tag | (uint32_t)(yy << 6) | (uint32_t)xx);
      }
      base += (uint32_t)__popcll(m[k]);
    }
  }
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

// Luma deblocking of one direction over a grid of segments (in place in LDS).
// VERT: edge columns gx in [gx0, gx1], segment rows gy in [gy0, gy1].
// HOR : edge rows gy in [gy0, gy1], segment columns gx in [gx0, gx1].
template <bool VERT, int PITCH>
__device__ __forceinline__ void lds_deblock_luma(const dev_job& j, lwin& L, const uint32_t* map, int mpitch, const mrect& mr,
                                 const uint16_t* list, int nlist, int maxv, const int16_t* tct) {
  const int tid = ltid();
  for (int base = 0; base < nlist; base += blockDim.x / 4) {
    const int it = base + (tid >> 2), line = tid & 3;
    uint32_t e = 0;
    int gx = 0, gy = 0;
    if (it < nlist) {
      const int w = list[it], ry = w >> 6, rx = w & 63;
      gx = mr.x0 + rx;
      gy = mr.y0 + ry;
      e = map[ry * mpitch + rx];
    }
    if (e != 0) {  // uniform across the quad
      const int x0 = VERT ? 4 * gx : 4 * gx + line, y0 = VERT ? 4 * gy + line : 4 * gy;
      // The line's 16 samples lie inside the window unless the edge is
      // within 8 samples of a picture border (where the window is clipped):
      // then reads clamp and writes outside the window are dropped.
      constexpr int STEP = VERT ? 1 : PITCH;
      const bool inwin = VERT ? (x0 - 8 >= L.ox && x0 + 8 <= L.ox + L.w && y0 < L.oy + L.h)
                              : (y0 - 8 >= L.oy && y0 + 8 <= L.oy + L.h && x0 < L.ox + L.w);
      uint16_t* c0 = L.p + (y0 - L.oy) * PITCH + (x0 - L.ox);  // q0
      int p[8], q[8];
      if (inwin) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          p[i] = c0[-(i + 1) * STEP];
          q[i] = c0[i * STEP];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          p[i] = VERT ? L.rd(x0 - 1 - i, y0) : L.rd(x0, y0 - 1 - i);
          q[i] = VERT ? L.rd(x0 + i, y0) : L.rd(x0, y0 + i);
        }
      }
      int beta, tc;
      edge_thresholds(e, j.bd_dbk, tct, beta, tc);
      int mflp = (e >> E_MFLP_SHIFT) & 7, mflq = (e >> E_MFLQ_SHIFT) & 7;
      luma_terms mine = luma_line_terms(p, q, mflp, mflq);
      luma_terms t0 = quad_terms<0>(mine), t3 = quad_terms<3>(mine);
      luma_decision d = luma_decide(t0, t3, beta, tc, mflp, mflq, (e & E_CTBROW_BIT) != 0);
      int np[8], nq[8], nwp, nwq;
      luma_filter_line(p, q, d, tc, np, nq, nwp, nwq);
      if (e & E_PLTP_BIT) nwp = 0;
      if (e & E_PLTQ_BIT) nwq = 0;
      if (inwin) {
#pragma unroll
        for (int i = 0; i < 7; i++) {
          if (i < nwp) c0[-(i + 1) * STEP] = (uint16_t)clampi(np[i], 0, maxv);
          if (i < nwq) c0[i * STEP] = (uint16_t)clampi(nq[i], 0, maxv);
        }
      } else {
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
}

// Chroma deblocking of one direction (both components) in LDS.
// VERT: edge columns sx in [sx0, sx1] (x = 8 sx), segment rows sy in [sy0, sy1] (y = 2 sy).
// HOR : edge rows sy in [sy0, sy1] (y = 8 sy), segment columns sx in [sx0, sx1] (x = 2 sx).
template <bool VERT, int PITCH>
__device__ __forceinline__ void lds_deblock_chroma(lwin cb, uint16_t* cr_p, const uint32_t* map0, const uint32_t* map1, int mpitch,
                                   const mrect& mr, const uint16_t* list, int nlist, int bd_dbk, int maxv,
                                   const int16_t* tct) {
  const int tid = ltid();
  for (int base = 0; base < nlist; base += blockDim.x / 2) {
    const int it = base + (tid >> 1), line = tid & 1;
    uint32_t e = 0;
    int sx = 0, sy = 0, comp = 0;
    if (it < nlist) {
      const int w = list[it], ry = (w >> 6) & 511, rx = w & 63;
      comp = w >> 15;
      sx = mr.x0 + rx;
      sy = mr.y0 + ry;
      e = (comp ? map1 : map0)[ry * mpitch + rx];
    }
    if (e != 0) {  // uniform across the pair
      lwin L = cb;  // Cb and Cr windows share their geometry
      if (comp) L.p = cr_p;
      int x0, y0;
      if (VERT) { x0 = 8 * sx; y0 = 2 * sy + line; } else { x0 = 2 * sx + line; y0 = 8 * sy; }
      constexpr int STEP = VERT ? 1 : PITCH;
      const bool inwin = VERT ? (x0 - 4 >= L.ox && x0 + 4 <= L.ox + L.w && y0 < L.oy + L.h)
                              : (y0 - 4 >= L.oy && y0 + 4 <= L.oy + L.h && x0 < L.ox + L.w);
      uint16_t* c0 = L.p + (y0 - L.oy) * PITCH + (x0 - L.ox);  // q0
      int c[8];
      if (inwin) {
#pragma unroll
        for (int i = 0; i < 8; i++) c[i] = c0[(i - 4) * STEP];
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) c[i] = VERT ? L.rd(x0 - 4 + i, y0) : L.rd(x0, y0 - 4 + i);
      }
      int beta, tc;
      edge_thresholds(e, bd_dbk, tct, beta, tc);
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
      if (inwin) {
#pragma unroll
        for (int i = 1; i < 7; i++)
          if ((wm >> i) & 1) c0[(i - 4) * STEP] = (uint16_t)clampi(n[i], 0, maxv);
      } else {
#pragma unroll
        for (int i = 1; i < 7; i++) {
          if (!((wm >> i) & 1)) continue;
          int x = VERT ? x0 - 4 + i : x0, y = VERT ? y0 : y0 - 4 + i;
          if (L.inside(x, y)) L.wr(x, y, clampi(n[i], 0, maxv));
        }
      }
    }
  }
}

// SAO parameters of the CTBs around a tile: CTB columns [bx, bx + 3) and
// rows [by, by + 3), staged in LDS.  Tile shapes are chosen so that these
// cover the tile's SAO reach (tile no larger than the SAO CTB); cells outside
// the SAO grid and components with SAO off hold type 0.
struct sao_tab {
  int bx, by;
};

// SAO of one component from window `src` into window `dst` (dst's extent is
// the region produced, at most RW x RH samples).  sub = 1 for chroma.  Work is
// split into strips of 8 samples at x = 8k by R rows (one CTB column per
// strip, SAO CTBs being at least 8 samples wide; at most two CTB rows, their
// height being at least 8 > R), one strip per thread (R is chosen so that
// one pass covers the region).  A strip reads the
// decoded parameters of its CTBs (runs[cell][comp]: type | class or band << 8
// | fits << 16, then the biased offset bytes) and its R + 2 sample rows up
// front, so that the LDS reads overlap and each row is read once.  Rows whose
// CTB has every offset in i8 use the packed table form (sao_eo8 / sao_bo8);
// the other rows, and edge-offset rows whose 8-neighbourhoods are not all
// inside the picture and the CTB's rectangle, are done afterwards sample by
// sample.  SRC16: src rows at x8 are 16-byte aligned (else 8); DST8: dst rows
// at x8 are 8-byte aligned (else 4).  NPL = 2 walks both chroma planes in one
// loop (component comp0 + plane; plane 1 at src.p + sd1 / dst.p + dd1, same
// geometry).
// The least strip height (at most 7) at which the strips of a RW x RH
// region, NPL planes, take one pass of NT threads.
constexpr int sao_strip_rows(int RW, int RH, int NPL, int NT) {
  int r = 1;
  while (r < 7 && NPL * ((RW + 14) / 8) * ((RH + r - 1) / r) > NT) r++;
  return r;
}

template <int RW, int RH, bool SRC16, bool DST8, int NPL, int NT, int R = sao_strip_rows(RW, RH, NPL, NT)>
__device__ __forceinline__ void lds_sao(const dev_job& j, const uint64_t (*saop)[3], const uint4 (*runs)[3],
                                        const uint64_t* saor, const uint64_t (*safe)[3], const sao_tab& tb,
                                        const lwin& src0, const lwin& dst0, int comp0, int sd1, int dd1, int pw,
                                        int ph) {
  static_assert(R + 1 <= 8, "a strip spans at most two CTB rows");
  const int sub = comp0 ? 1 : 0;
  const int log2c = j.sao_ctb_log2 - sub;
  const int maxv = (1 << j.sao_bd) - 1;
  const uint32_t maxv2 = (uint32_t)maxv * 0x10001u;
  const bool has_rect = j.sao_rect != nullptr;
  const int ka = dst0.ox >> 3, nk = ((dst0.ox + dst0.w - 1) >> 3) - ka + 1;
  constexpr int NK = (RW + 14) / 8;      // strip columns, at most
  constexpr int NS = (RH + R - 1) / R;  // strips per column, at most
  constexpr int per_plane = NK * NS;
  for (int i0 = ltid(); i0 < NPL * per_plane; i0 += NT) {
    const int pl = (NPL > 1 && i0 >= per_plane) ? 1 : 0, i = i0 - pl * per_plane;
    const int comp = comp0 + pl;
    lwin src = src0, dst = dst0;
    src.p += pl * sd1;
    dst.p += pl * dd1;
    // column-major: a wave's strips share few CTB columns, and so mostly one
    // SAO type and class
    const int kk = i / NS, st = i - __mul24(kk, NS);
    const int ys = dst.oy + st * R;
    if (kk >= nk || ys >= dst.oy + dst.h) continue;
    const int nr = min(R, dst.oy + dst.h - ys);
    const int x8 = (ka + kk) * 8;
    const int xa = max(x8, dst.ox), xb = min(x8 + 8, dst.ox + dst.w);  // samples produced
    // the CTBs of the strip's first and last rows
    const int cxr = (x8 >> log2c) - tb.bx;
    const int cell0 = ((ys >> log2c) - tb.by) * 3 + cxr, cell1 = (((ys + nr - 1) >> log2c) - tb.by) * 3 + cxr;
    const int ysplit = ((ys >> log2c) + 1) << log2c;  // first row in cell1
    const uint4 rp0 = runs[cell0][comp], rp1 = runs[cell1][comp];
    const uint64_t sr0 = safe[cell0][comp], sr1 = safe[cell1][comp];
    // rows ys-1 .. ys+R at x8-1 .. x8+8 (direct reads when inside the
    // window, else clamped)
    sao_row8 Rw[R + 2];
    const bool direct = x8 - 1 >= src.ox && x8 + 9 <= src.ox + src.w && ys - 1 >= src.oy && ys + R + 1 <= src.oy + src.h;
    if (direct) {
      const uint16_t* q = src.p + (ys - 1 - src.oy) * src.pitch + (x8 - src.ox);
#pragma unroll
      for (int r = 0; r < R + 2; r++, q += src.pitch) {
        if (SRC16) {
          const uint4 v = *reinterpret_cast<const uint4*>(q);
          Rw[r].w[0] = v.x; Rw[r].w[1] = v.y; Rw[r].w[2] = v.z; Rw[r].w[3] = v.w;
        } else {
          const uint2 a = *reinterpret_cast<const uint2*>(q), b = *reinterpret_cast<const uint2*>(q + 4);
          Rw[r].w[0] = a.x; Rw[r].w[1] = a.y; Rw[r].w[2] = b.x; Rw[r].w[3] = b.y;
        }
        Rw[r].l = q[-1];
        Rw[r].r = q[8];
      }
    } else {
#pragma unroll
      for (int r = 0; r < R + 2; r++) {
        const int yr = ys - 1 + r;
#pragma unroll
        for (int k = 0; k < 4; k++)
          Rw[r].w[k] = (uint32_t)src.rd(x8 + 2 * k, yr) | ((uint32_t)src.rd(x8 + 2 * k + 1, yr) << 16);
        Rw[r].l = src.rd(x8 - 1, yr);
        Rw[r].r = src.rd(x8 + 8, yr);
      }
    }
    uint32_t later = 0;  // rows left to the per-sample path
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (r < nr) {
        const int y = ys + r;
        const bool second = y >= ysplit;
        const uint4 rp = second ? rp1 : rp0;
        const uint64_t sr = second ? sr1 : sr0;
        const int type = (int)(rp.x & 3), aux = (int)((rp.x >> 8) & 31);
        const bool packed = ((rp.x >> 16) & 1) && j.sao_bd <= 15;
        const uint64_t btab = (uint64_t)rp.y | ((uint64_t)rp.z << 32);
        const sao_row8& P = Rw[r + 1];
        uint32_t o[4] = {P.w[0], P.w[1], P.w[2], P.w[3]};
        bool done = true;
        if (type == 1) {
          if (packed) sao_bo8(P, aux, btab, j.sao_bd, maxv2, o);
          else done = false;
        } else if (type == 2) {
          const bool inner = xa >= (int)(sr & 0xffff) && xb <= (int)((sr >> 16) & 0xffff) &&
                             y >= (int)((sr >> 32) & 0xffff) && y < (int)(sr >> 48);
          if (inner && packed) sao_eo8_any(aux & 3, Rw[r], P, Rw[r + 2], btab, maxv2, o);
          else done = false;
        }
        if (!done) {
          later |= 1u << r;
        } else if (xa == x8 && xb == x8 + 8) {
          uint16_t* d = dst.p + (y - dst.oy) * dst.pitch + (x8 - dst.ox);
          if (DST8) {
            reinterpret_cast<uint2*>(d)[0] = make_uint2(o[0], o[1]);
            reinterpret_cast<uint2*>(d)[1] = make_uint2(o[2], o[3]);
          } else {
#pragma unroll
            for (int k = 0; k < 4; k++) reinterpret_cast<uint32_t*>(d)[k] = o[k];
          }
        } else {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (x8 + k >= xa && x8 + k < xb) dst.wr(x8 + k, y, (o[k >> 1] >> (16 * (k & 1))) & 0xffffu);
        }
      }
    }
    // per sample, with the picture / rectangle check of each neighbour
    while (later) {
      const int r = __builtin_ctz(later);
      later &= later - 1;
      const int y = ys + r;
      const int cell = y >= ysplit ? cell1 : cell0;
      const uint64_t prm = saop[cell][comp];
      const int type = (int)(prm & 3);
      if (type == 1) {
        for (int x = xa; x < xb; x++) dst.wr(x, y, sao_band_sample(src.rd(x, y), prm, j.sao_bd, maxv));
        continue;
      }
      int h0, v0, h1, v1;
      eo_offsets((int)((prm >> 2) & 3), h0, v0, h1, v1);
      int cx0 = 0, cy0 = 0, cx1 = pw, cy1 = ph;
      if (has_rect) {
        const uint64_t rr = saor[cell];
        cx0 = (int)(rr & 0xffff); cy0 = (int)((rr >> 16) & 0xffff); cx1 = (int)((rr >> 32) & 0xffff); cy1 = (int)(rr >> 48);
        if (sub) { cx0 >>= 1; cy0 >>= 1; cx1 = (cx1 + 1) >> 1; cy1 = (cy1 + 1) >> 1; }
        cx0 = max(cx0, 0); cy0 = max(cy0, 0); cx1 = min(cx1, pw); cy1 = min(cy1, ph);
      }
      for (int x = xa; x < xb; x++) {
        const int pv = src.rd(x, y);
        const int nx0 = x + h0, ny0 = y + v0, nx1 = x + h1, ny1 = y + v1;
        const bool outside = nx0 < cx0 || ny0 < cy0 || nx1 < cx0 || ny1 < cy0 || nx0 >= cx1 || ny0 >= cy1 ||
                             nx1 >= cx1 || ny1 >= cy1;
        dst.wr(x, y, outside ? clampi(pv, 0, maxv) : sao_edge_sample(pv, src.rd(nx0, ny0), src.rd(nx1, ny1), prm, maxv));
      }
    }
  }
}

// Window copy (SAO off for the whole picture).
__device__ __forceinline__ void lds_copy(const lwin& src, lwin& dst) {
  const int lane = ltid() & 63, nw = blockDim.x >> 6;
  for (int yy = ltid() >> 6; yy < dst.h; yy += nw)
    for (int xx = lane; xx < dst.w; xx += 64) dst.wr(dst.ox + xx, dst.oy + yy, src.rd(dst.ox + xx, dst.oy + yy));
}

// Replaces every sample of buf (storage origin (ox, oy), pitch `pitch`) in
// [xa, xb) x [ya, yb) that lies outside the rectangle [rx0, rx1] x [ry0, ry1]
// by the sample at the clamped position, which lies inside the rectangle:
// reads touch only unchanged samples, so one pass suffices.  ALF reads of a
// tile (which lies inside one CTB and so has one rectangle) are then plain
// reads of the padded buffer.
__device__ __forceinline__ void pad_to_rect(uint16_t* buf, int pitch, int ox, int oy, int xa, int xb, int ya, int yb, int rx0, int rx1,
                            int ry0, int ry1) {
  const int w = xb - xa, n = w * (yb - ya);
  for (int i = ltid(); i < n; i += blockDim.x) {
    const int y = ya + i / w, x = xa + i % w;
    if (x >= rx0 && x <= rx1 && y >= ry0 && y <= ry1) continue;
    buf[(y - oy) * pitch + (x - ox)] = buf[(clampi(y, ry0, ry1) - oy) * pitch + (clampi(x, rx0, rx1) - ox)];
  }
}

// Geometry of one tile: the windows, edge-descriptor rectangles and SAO
// table origin that follow from the job and the tile origin.
struct tile_geom {
  int x0, y0, x1, y1, cx0, cy0, cx1, cy1;
  bool chroma, dbk, dbk_c, in_grid;
  lwin L, Cw;  // luma window (s.L), chroma window (Cb in s.Cc[0]; Cr alike in s.Cc[1])
  mrect mlv, mlh, mcv, mch;
  sao_tab tb;
};

template <int TW, int TH>
__device__ __forceinline__ tile_geom tile_geometry(const dev_job& j, fused_lds<TW, TH>& s, const tile_ref& tr) {
  typedef fused_lds<TW, TH> LD;
  tile_geom g;
  const int W = j.w, H = j.h;
  g.x0 = tr.x0;
  g.y0 = tr.y0;
  g.x1 = min(g.x0 + TW, W);
  g.y1 = min(g.y0 + TH, H);
  g.chroma = j.wc > 0;
  g.cx0 = g.x0 >> 1;
  g.cy0 = g.y0 >> 1;
  g.cx1 = min(g.x1 >> 1, j.wc);
  g.cy1 = min(g.y1 >> 1, j.hc);
  g.L = lwin{s.L, LD::LW, max(0, g.x0 - 16), max(0, g.y0 - 16), 0, 0};
  g.L.w = min(W, g.x1 + 16) - g.L.ox;
  g.L.h = min(H, g.y1 + 16) - g.L.oy;
  g.Cw = lwin{s.Cc[0], LD::CW, max(0, g.cx0 - 4), max(0, g.cy0 - 4), 0, 0};
  g.Cw.w = min(j.wc, g.cx1 + 4) - g.Cw.ox;
  g.Cw.h = min(j.hc, g.cy1 + 4) - g.Cw.oy;
  g.dbk = j.do_dbk;
  g.dbk_c = g.chroma && j.do_dbk && j.do_dbk_chroma;
  g.mlv = g.mlh = g.mcv = g.mch = mrect{0, 0, 0, 0};
  if (g.dbk) {
    // vertical edges x in [x0-10, x1+10] over the window rows
    g.mlv.x0 = max(0, cdiv_i(g.x0 - 10, 4));
    g.mlv.nx = min(j.gw - 1, fdiv(g.x1 + 10, 4)) - g.mlv.x0 + 1;
    g.mlv.y0 = g.L.oy / 4;
    g.mlv.ny = min(j.gh - 1, (g.L.oy + g.L.h - 1) / 4) - g.mlv.y0 + 1;
    // horizontal edges y in [y0-10, y1+10] over columns [x0-4, x1+4)
    g.mlh.y0 = max(0, cdiv_i(g.y0 - 10, 4));
    g.mlh.ny = min(j.gh - 1, fdiv(g.y1 + 10, 4)) - g.mlh.y0 + 1;
    g.mlh.x0 = max(0, fdiv(g.x0 - 4, 4));
    g.mlh.nx = min(j.gw - 1, fdiv(min(W, g.x1 + 4) - 1, 4)) - g.mlh.x0 + 1;
  }
  if (g.dbk_c) {
    // chroma vertical edges x in [cx0-5, cx1+5], horizontal y in [cy0-5, cy1+5]
    g.mcv.x0 = max(0, cdiv_i(g.cx0 - 5, 8));
    g.mcv.nx = min(j.cvw - 1, fdiv(g.cx1 + 5, 8)) - g.mcv.x0 + 1;
    g.mcv.y0 = g.Cw.oy / 2;
    g.mcv.ny = min(j.cvh - 1, (g.Cw.oy + g.Cw.h - 1) / 2) - g.mcv.y0 + 1;
    g.mch.y0 = max(0, cdiv_i(g.cy0 - 5, 8));
    g.mch.ny = min(j.chh - 1, fdiv(g.cy1 + 5, 8)) - g.mch.y0 + 1;
    g.mch.x0 = max(0, fdiv(g.cx0 - 3, 2));
    g.mch.nx = min(j.chw - 1, fdiv(min(j.wc, g.cx1 + 3) - 1, 2)) - g.mch.x0 + 1;
  }
  // SAO parameter table: CTBs covering luma x in [x0-4, x1+3] (chroma reach
  // included) and likewise y
  g.tb = sao_tab{0, 0};
  if (j.do_sao) {
    g.tb.bx = max(0, (g.x0 - 4) >> j.sao_ctb_log2);
    g.tb.by = max(0, (g.y0 - 4) >> j.sao_ctb_log2);
  }
  const int rx = g.x0 >> j.ctb_log2, ry = g.y0 >> j.ctb_log2;
  g.in_grid = j.do_alf && rx < j.wctb && ry < j.hctb;
  return g;
}

// Everything one tile reads from HBM, held in registers between the issue of
// the loads and their commit to LDS: samples (8 luma / 4 chroma per access),
// edge descriptors, the ALF banks and the SAO cells.  Window origins are
// multiples of 16 (luma) / 4 (chroma) samples and plane pitches multiples of
// 64, so the vector accesses are aligned; they may read up to 7 samples past
// the window's last column, inside the pitch.  Every element is written by
// issue() (0 where nothing is loaded), so none stays live from one tile to
// the next.
template <int TW, int TH>
struct tile_loads {
  typedef fused_lds<TW, TH> LD;
  static constexpr int NT = LD::NT;
  static constexpr int NCH = LD::LW / 8, PER = (LD::LH * NCH + NT - 1) / NT;
  static constexpr int NC4 = LD::CW / 4, PERC = (LD::CH * NC4 + NT - 1) / NT;  // per chroma plane
  static constexpr int PBK = (ALF_BANK_WORDS + NT - 1) / NT;
  uint4 v[PER];
  uint2 vc[2][PERC];
  rect_stage<LD::MLV_W, LD::MLV_H, NT> slv;
  rect_stage<LD::MLH_W, LD::MLH_H, NT> slh;
  rect_stage<LD::MCV_W, LD::MCV_H, NT> scv0, scv1;
  rect_stage<LD::MCH_W, LD::MCH_H, NT> sch0, sch1;
  uint32_t bk[PBK];
  uint64_t small;  // chroma ALF / CC-ALF bank words (threads 0..103)
  uint64_t sv;     // SAO cell parameters (threads 128..163)

  __device__ __forceinline__ void issue(const dev_job& j, const tile_geom& g, const tile_ref& tr) {
    const int tid = ltid();
    const lwin& L = g.L;
    const int nch = (L.w + 7) >> 3, total = L.h * NCH;
    const uint16_t* src = j.in.y + (size_t)L.oy * j.py + L.ox;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = tid + k * NT, yy = i / NCH, ch = i - __mul24(yy, NCH);
      v[k] = (i < total && ch < nch) ? gld(reinterpret_cast<const uint4*>(src + (__mul24(yy, j.py) + 8 * ch)))
                                     : make_uint4(0, 0, 0, 0);
    }
    const int nchc = (g.Cw.w + 3) >> 2, per_plane = g.Cw.h * NC4;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const uint16_t* csrc = (c ? j.in.cr : j.in.cb) + (size_t)g.Cw.oy * j.pc + g.Cw.ox;
#pragma unroll
      for (int k = 0; k < PERC; k++) {
        const int i = tid + k * NT, yy = i / NC4, ch = i - __mul24(yy, NC4);
        vc[c][k] = (g.chroma && i < per_plane && ch < nchc)
                       ? gld(reinterpret_cast<const uint2*>(csrc + (__mul24(yy, j.pc) + 4 * ch)))
                       : make_uint2(0, 0);
      }
    }
    slv.issue(g.dbk ? j.e_lv : nullptr, j.gw, g.mlv);
    slh.issue(g.dbk ? j.e_lh : nullptr, j.gw, g.mlh);
    scv0.issue(g.dbk_c ? j.e_cv[0] : nullptr, j.cvw, g.mcv);
    scv1.issue(g.dbk_c ? j.e_cv[1] : nullptr, j.cvw, g.mcv);
    sch0.issue(g.dbk_c ? j.e_ch[0] : nullptr, j.chw, g.mch);
    sch1.issue(g.dbk_c ? j.e_ch[1] : nullptr, j.chw, g.mch);
    const int bank_idx = (int)(tr.alf_cw & 0xff);
    const bool alf_on = bank_idx != (int)ALF_OFF8;
    const uint32_t* bsrc = j.alf_bank + (size_t)(alf_on ? bank_idx : 0) * ALF_BANK_WORDS;
#pragma unroll
    for (int k = 0; k < PBK; k++) bk[k] = (alf_on && tid + k * NT < ALF_BANK_WORDS) ? gld(bsrc + tid + k * NT) : 0u;
    small = 0;
    if (g.in_grid && tid < ALF_CBANK_WORDS) small = gld(j.alf_cbank + tid);
    else if (g.in_grid && tid < ALF_CBANK_WORDS + 2 * ALF_CCBANK_WORDS)
      small = (uint32_t)gld(j.alf_ccbank + tid - ALF_CBANK_WORDS);
    // SAO cells: thread 128 + 4 * cell + k loads component k (k < 3) or the
    // rectangle (k = 3); off components and cells outside the grid read 0
    sv = 0;
    const int scell = tid - 128, scx = g.tb.bx + (scell >> 2) % 3, scy = g.tb.by + (scell >> 2) / 3, sk = scell & 3;
    if (j.do_sao && scell >= 0 && scell < 36 && scx < j.swctb && scy < j.shctb) {
      const size_t sci = (size_t)scy * j.swctb + scx;
      const bool on = sk == 3 || (sk == 0 ? j.sao_luma : j.sao_chroma);
      if (on) sv = sk < 3 ? gld(j.sao + sci * 3 + sk) : (j.sao_rect ? gld(j.sao_rect + sci) : 0);
    }
  }

  // No next tile: every element defined all the same, so that the compiler
  // sees the registers dead between commit() and the next issue().
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int k = 0; k < PERC; k++) vc[c][k] = make_uint2(0, 0);
    slv.issue(nullptr, 0, mrect{0, 0, 0, 0});
    slh.issue(nullptr, 0, mrect{0, 0, 0, 0});
    scv0.issue(nullptr, 0, mrect{0, 0, 0, 0});
    scv1.issue(nullptr, 0, mrect{0, 0, 0, 0});
    sch0.issue(nullptr, 0, mrect{0, 0, 0, 0});
    sch1.issue(nullptr, 0, mrect{0, 0, 0, 0});
#pragma unroll
    for (int k = 0; k < PBK; k++) bk[k] = 0u;
    small = sv = 0;
  }

  // Writes the loaded data to LDS (luma through the LMCS inverse table) and
  // lists the nonzero edge descriptors.  The caller synchronises afterwards.
  __device__ __forceinline__ void commit(const dev_job& j, const tile_geom& g, const tile_ref& tr,
                                         fused_lds<TW, TH>& s) {
    const int tid = ltid();
    const int W = j.w, H = j.h;
    const lwin& L = g.L;
    const int nch = (L.w + 7) >> 3, total = L.h * NCH;
    // LMCS: samples below 1024 through the table staged in LDS, with no branch
    // per sample; chunks holding a larger sample (bit depth above 10, or
    // samples beyond the bit depth) are rewritten from the global table.
    const bool lut_lds = j.bd <= 10;
    bool far = false;
    uint32_t any = 0;  // OR of the samples written (the ALF's 16-bit lanes need them below 2^15)
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int i = tid + k * NT, yy = i / NCH, ch = i - __mul24(yy, NCH);
      if (i < total && ch < nch) {
        uint4 w = v[k];
        if (j.do_lmcs) {
          far |= !lut_lds || ((w.x | w.y | w.z | w.w) & 0xfc00fc00u) != 0;
          if (lut_lds)
            w = make_uint4(lut_pair(s.lut, w.x), lut_pair(s.lut, w.y), lut_pair(s.lut, w.z), lut_pair(s.lut, w.w));
        }
        any |= w.x | w.y | w.z | w.w;
        *reinterpret_cast<uint4*>(&s.L[yy * LD::LW + 8 * ch]) = w;
      }
    }
    if (far) {
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int i = tid + k * NT, yy = i / NCH, ch = i - __mul24(yy, NCH);
        const uint4 w = v[k];
        if (i < total && ch < nch && (!lut_lds || ((w.x | w.y | w.z | w.w) & 0xfc00fc00u) != 0)) {
          uint32_t w4[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
          for (int q = 0; q < 4; q++)
            w4[q] = (uint32_t)gld(j.lmcs_lut + (w4[q] & 0xffffu)) | ((uint32_t)gld(j.lmcs_lut + (w4[q] >> 16)) << 16);
          *reinterpret_cast<uint4*>(&s.L[yy * LD::LW + 8 * ch]) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        }
      }
    }
    if (g.chroma) {
      const int nchc = (g.Cw.w + 3) >> 2, per_plane = g.Cw.h * NC4;
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int k = 0; k < PERC; k++) {
          const int i = tid + k * NT, yy = i / NC4, ch = i - __mul24(yy, NC4);
          if (i < per_plane && ch < nchc) {
            *reinterpret_cast<uint2*>(&s.Cc[c][yy * LD::CW + 4 * ch]) = vc[c][k];
            any |= vc[c][k].x | vc[c][k].y;
          }
        }
    }
    // LMCS outputs and the deblocking / SAO outputs stay within the bit
    // depth, so a sample of 2^15 or more reaching the ALF is one written here
    if (any & 0x80008000u) s.wide = 1u;
    if (g.dbk) {
      slv.commit(s.m.lv, g.mlv, s.m.llv, &s.nlist[0], 0);
      slh.commit(s.m.lh, g.mlh, s.m.llh, &s.nlist[1], 0);
    }
    if (g.dbk_c) {
      scv0.commit(s.m.cv[0], g.mcv, s.m.lcv, &s.nlist[2], 0);
      scv1.commit(s.m.cv[1], g.mcv, s.m.lcv, &s.nlist[2], 1u << 15);
      sch0.commit(s.m.ch[0], g.mch, s.m.lch, &s.nlist[3], 0);
      sch1.commit(s.m.ch[1], g.mch, s.m.lch, &s.nlist[3], 1u << 15);
    }
    if ((tr.alf_cw & 0xff) != ALF_OFF8) {
#pragma unroll
      for (int k = 0; k < PBK; k++)
        if (tid + k * NT < ALF_BANK_WORDS) s.bank[tid + k * NT] = bk[k];
    }
    if (g.in_grid && tid < ALF_CBANK_WORDS) s.cbank[tid] = (uint32_t)small;
    else if (g.in_grid && tid < ALF_CBANK_WORDS + 2 * ALF_CCBANK_WORDS) s.ccbank[tid - ALF_CBANK_WORDS] = (int32_t)small;
    const int scell = tid - 128, sk = scell & 3;
    if (j.do_sao && scell >= 0 && scell < 36) {
      if (sk < 3) {
        s.saop[scell >> 2][sk] = sv;
        const sao_run rn = sao_run_of(sv);
        s.saorun[scell >> 2][sk] = make_uint4((uint32_t)rn.type | ((uint32_t)rn.aux << 8) | ((uint32_t)rn.fits << 16),
                                              (uint32_t)rn.btab, (uint32_t)(rn.btab >> 32), 0u);
      } else {
        s.saor[scell >> 2] = sv;
        // rectangles inside which a sample's whole 8-neighbourhood lies in
        // the picture and in this CTB's rectangle
        const uint64_t rr = j.sao_rect ? sv : ((uint64_t)W << 32) | ((uint64_t)H << 48);
        const int qx0 = (int)(rr & 0xffff), qy0 = (int)((rr >> 16) & 0xffff), qx1 = (int)((rr >> 32) & 0xffff),
                  qy1 = (int)(rr >> 48);
        for (int c = 0; c < 3; c++) {
          int ax0 = qx0, ay0 = qy0, ax1 = qx1, ay1 = qy1, pw = W, ph = H;
          if (c) { ax0 >>= 1; ay0 >>= 1; ax1 = (ax1 + 1) >> 1; ay1 = (ay1 + 1) >> 1; pw = j.wc; ph = j.hc; }
          const uint32_t lx = (uint32_t)max(ax0, 0) + 1, hx = (uint32_t)max(min(ax1, pw) - 1, 0);
          const uint32_t ly = (uint32_t)max(ay0, 0) + 1, hy = (uint32_t)max(min(ay1, ph) - 1, 0);
          s.safe[scell >> 2][c] = (uint64_t)lx | ((uint64_t)hx << 16) | ((uint64_t)ly << 32) | ((uint64_t)hy << 48);
        }
      }
    }
  }
};

// Filters tile `tr` whose loads are in `ld`: commit, deblock, SAO, ALF.  When
// `more` is set, the loads of the next tile `ntr` are issued into `ld` once
// the luma ALF is done, so that they are in flight during the chroma ALF and
// reach the registers by the next call.  (The next job is read from `jobs`
// directly, which keeps its fields on scalar loads.)
template <int TW, int TH, bool HBD>
__device__ __forceinline__ void fused_tile(const dev_job& j, fused_lds<TW, TH>& s, const tile_ref& tr,
                                           tile_loads<TW, TH>& ld, const dev_job* __restrict__ jobs, bool more,
                                           const tile_ref& ntr) {
  typedef fused_lds<TW, TH> LD;
  constexpr int NT = LD::NT;  // threads per workgroup
  const tile_geom g = tile_geometry<TW, TH>(j, s, tr);
  const int x0 = g.x0, y0 = g.y0, x1 = g.x1, y1 = g.y1;
  const int cx0 = g.cx0, cy0 = g.cy0, cx1 = g.cx1, cy1 = g.cy1;
  const int W = j.w, H = j.h;
  const int maxv = (1 << j.bd) - 1;
  const bool chroma = g.chroma, dbk = g.dbk, dbk_c = g.dbk_c, in_grid = g.in_grid;
  lwin L = g.L;
  const lwin Cw = g.Cw;
  const mrect mlv = g.mlv, mlh = g.mlh, mcv = g.mcv, mch = g.mch;
  const sao_tab tb = g.tb;
  // ALF parameters of the tile's CTB (from the tile record)
  const int ctb = 1 << j.ctb_log2;
  const int ry = y0 >> j.ctb_log2;
  const uint32_t cw = tr.alf_cw;
  const int bank_idx = (int)(cw & 0xff);
  const uint64_t r = tr.alf_rect;
  ld.commit(j, g, tr, s);
  __syncthreads();
  const bool wide = !HBD && s.wide != 0u;  // read before the store phase clears it
  // ---- deblocking: all vertical edges, then all horizontal edges
  if (dbk) {
    lds_deblock_luma<true, LD::LW>(j, L, s.m.lv, LD::MLV_W, mlv, s.m.llv, (int)s.nlist[0], maxv, s.tct);
    if (dbk_c)
      lds_deblock_chroma<true, LD::CW>(Cw, s.Cc[1], s.m.cv[0], s.m.cv[1], LD::MCV_W, mcv, s.m.lcv, (int)s.nlist[2], j.bd_dbk,
                               maxv, s.tct);
    __syncthreads();
    lds_deblock_luma<false, LD::LW>(j, L, s.m.lh, LD::MLH_W, mlh, s.m.llh, (int)s.nlist[1], maxv, s.tct);
    if (dbk_c)
      lds_deblock_chroma<false, LD::CW>(Cw, s.Cc[1], s.m.ch[0], s.m.ch[1], LD::MCH_W, mch, s.m.lch, (int)s.nlist[3], j.bd_dbk,
                                maxv, s.tct);
  }
  __syncthreads();  // deblocking done; the edge descriptors' LDS is reused
  // ---- SAO into the second buffers.  S stores luma [x0-4, x0+TW+4) x
  // [y0-3, y0+TH+3) and Sc chroma [cx0-2, cx0+TW/2+2) x [cy0-2, ...) at fixed
  // origins (so a 4x4 block's ALF neighbourhood is three aligned 8-byte reads
  // per row); the windows describe the part SAO produces (inside the picture).
  const int sox = x0 - 4, soy = y0 - 3, scox = cx0 - 2, scoy = cy0 - 2;
  lwin S{nullptr, LD::SW, max(0, x0 - 3), max(0, y0 - 3), 0, 0};
  S.w = min(W, x1 + 3) - S.ox;
  S.h = min(H, y1 + 3) - S.oy;
  S.p = s.o.S + (S.oy - soy) * LD::SW + (S.ox - sox);
  if (j.do_sao) lds_sao<TW + 6, TH + 6, true, true, 1, NT>(j, s.saop, s.saorun, s.saor, s.safe, tb, L, S, 0, 0, 0, W, H);
  else lds_copy(L, S);
  if (chroma) {
    lwin Sc{nullptr, LD::SCW, max(0, cx0 - 2), max(0, cy0 - 2), 0, 0};
    Sc.w = min(j.wc, cx1 + 2) - Sc.ox;
    Sc.h = min(j.hc, cy1 + 2) - Sc.oy;
    Sc.p = s.o.Sc[0] + (Sc.oy - scoy) * LD::SCW + (Sc.ox - scox);
    lwin Cs = Cw;
    Cs.p = s.Cc[0];
    constexpr int SD1 = LD::CH * LD::CW, DD1 = LD::SCH * LD::SCW;  // Cr after Cb
    if (j.do_sao) {
      lds_sao<TW / 2 + 4, TH / 2 + 4, false, false, 2, NT>(j, s.saop, s.saorun, s.saor, s.safe, tb, Cs, Sc, 1, SD1, DD1, j.wc, j.hc);
    } else {
      for (int c = 0; c < 2; c++) {
        lwin a = Cs, b = Sc;
        a.p += c * SD1;
        b.p += c * DD1;
        lds_copy(a, b);
      }
    }
  }
  // ---- ALF: this tile lies inside one CTB
  const int rx0 = (int)(r & 0xffff), ry0 = (int)((r >> 16) & 0xffff), rx1 = (int)((r >> 32) & 0xffff) - 1,
            ry1 = (int)(r >> 48) - 1;
  const int bw = (x1 - x0 + 3) / 4, bh = (y1 - y0 + 3) / 4;
  const int xe = x0 + 4 * bw, ye = y0 + 4 * bh;  // block-aligned end of the tile
  const int ccx0 = rx0 >> 1, ccy0 = ry0 >> 1, ccx1 = ((rx1 + 2) >> 1) - 1, ccy1 = ((ry1 + 2) >> 1) - 1;
  const bool luma_interior = x0 - 3 >= rx0 && xe + 2 <= rx1 && y0 - 3 >= ry0 && ye + 2 <= ry1;
  const bool chroma_interior = cx0 - 2 >= ccx0 && cx1 + 1 <= ccx1 && cy0 - 2 >= ccy0 && cy1 + 1 <= ccy1;
  if (!luma_interior || (chroma && !chroma_interior)) {
    __syncthreads();  // SAO output complete
    if (!luma_interior) pad_to_rect(s.o.S, LD::SW, sox, soy, x0 - 3, xe + 3, y0 - 3, ye + 3, rx0, rx1, ry0, ry1);
    if (chroma && !chroma_interior)
      for (int c = 0; c < 2; c++)
        pad_to_rect(s.o.Sc[c], LD::SCW, scox, scoy, cx0 - 2, cx1 + 2, cy0 - 2, cy1 + 2, ccx0, ccx1, ccy0, ccy1);
  }
  __syncthreads();
  // padded luma sample (x, y), |x - [x0, xe)| <= 3, |y - [y0, ye)| <= 3
  auto Sl = [&](int x, int y) -> int { return s.o.S[__mul24(y - soy, LD::SW) + (x - sox)]; };
  // luma: one 4x4 block per thread
  {
    const int amax = (1 << j.alf_bd) - 1;
    const int y_ctb = ry * ctb;
    const bool carve = !(y_ctb + ctb >= H) || (H - y_ctb) > ctb - 4;
    // results go to the tile's output buffer in LDS (pitch TW, in the input
    // window's space, dead since SAO) and leave in one pass at the end
    uint16_t* const ob = s.L;
    // blocks on a grid of BW per row (compile-time divisor); blocks past the
    // tile's width are skipped
    constexpr int BW = TW / 4;
#pragma unroll 1
    for (int b = ltid(); b < BW * bh; b += NT) {
      const int bx = b % BW, x4 = x0 + 4 * bx, y4 = y0 + 4 * (b / BW);
      if (bx >= bw) continue;
      uint32_t outp[4][2];
      if (!HBD && wide)  // 16-bit lanes could wrap: the 32-bit form
        alf_luma_block_lds<LD::SW, true>(s.o.S, sox, soy, x4, y4, s.bank, bank_idx != (int)ALF_OFF8, y_ctb, ctb,
                                         carve, j.alf_bd, outp);
      else
        alf_luma_block_lds<LD::SW, HBD>(s.o.S, sox, soy, x4, y4, s.bank, bank_idx != (int)ALF_OFF8, y_ctb, ctb,
                                        carve, j.alf_bd, outp);
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
        *reinterpret_cast<uint2*>(&ob[(y4 - y0 + rr) * TW + (x4 - x0)]) = make_uint2(outp[rr][0], outp[rr][1]);
    }
  }
  // the next tile's loads, in flight during the chroma ALF
  if (more) {
    const dev_job nj = jobs[ntr.job];  // one burst of scalar loads
    ld.issue(nj, tile_geometry<TW, TH>(nj, s, ntr), ntr);
  } else {
    ld.clear();
  }
  // chroma ALF + CC-ALF: two horizontally adjacent samples per thread.
  // Sample x of Sc sits at an even storage column when x - cx0 is even, so
  // pairs (x, x+1) are aligned 32-bit reads and (x-1, x) / (x+1, x+2) one
  // alignbit each.
  if (chroma) {
    // pairs on a grid of PW per row (compile-time divisor); pairs past the
    // tile's chroma width are skipped
    constexpr int PW = TW / 4;
    const int cwid = cx1 - cx0, chei = cy1 - cy0, npair = PW * chei;
    const int ctbc = ctb >> 1;
    const int amax = (1 << j.alf_bd) - 1, half = 1 << (j.alf_bd - 1);
    const s16x2 zero2 = {0, 0}, vmax2 = {(short)amax, (short)amax};
    // each component's CC-ALF tap sum for this tile (one filter per CTB)
    int ccsum0 = 0, ccsum1 = 0;
    if (in_grid) {
      const int f0 = (int)((cw >> 24) & 0xf), f1 = (int)((cw >> 28) & 0xf);
#pragma unroll
      for (int t = 0; t < 7; t++) {
        if (f0 != (int)ALF_OFF4) ccsum0 += s.ccbank[f0 * 7 + t];
        if (f1 != (int)ALF_OFF4) ccsum1 += s.ccbank[ALF_CCBANK_WORDS + f1 * 7 + t];
      }
    }
    for (int i = ltid(); i < 2 * npair; i += NT) {
      const int c = i >= npair ? 1 : 0, k = i - c * npair;
      const int yy = k / PW, x = cx0 + 2 * (k - __mul24(yy, PW)), y = cy0 + yy;
      if (x >= cx0 + cwid) continue;
      const uint16_t* sc = s.o.Sc[c];
      // pair of padded chroma samples (xa, xa+1), xa - cx0 even
      auto P2 = [&](int xa, int ya) -> uint32_t {
        return *reinterpret_cast<const uint32_t*>(&sc[__mul24(ya - scoy, LD::SCW) + (xa - scox)]);
      };
      uint32_t outp = P2(x, y);
      if (in_grid) {
        const int alt = (int)((cw >> (c == 0 ? 8 : 16)) & 0xff);
        if (alt != (int)ALF_OFF8 && !HBD && !wide) {
          const uint32_t* f = s.cbank + alt * 6;
          int sft, d1, d2;
          alf_t46(y - ry * ctbc, ctbc, sft, d1, d2);
          const uint32_t m2 = P2(x - 2, y), m0 = outp, p2 = P2(x + 2, y);
          const uint32_t am2 = P2(x - 2, y + d1), a0 = P2(x, y + d1), ap2 = P2(x + 2, y + d1);
          const uint32_t bm2 = P2(x - 2, y - d1), b0 = P2(x, y - d1), bp2 = P2(x + 2, y - d1);
          const uint32_t u2 = P2(x, y + d2), l2 = P2(x, y - d2);
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
          outp = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(__builtin_elementwise_add_sat(cur, d), zero2),
                                                   vmax2));
        } else if (alt != (int)ALF_OFF8) {
          // above 12 bits: 32-bit arithmetic per sample
          const uint32_t* f = s.cbank + alt * 6;
          int sft, d1, d2;
          alf_t46(y - ry * ctbc, ctbc, sft, d1, d2);
          auto B = [&](int xx, int yy2) -> int { return sc[(yy2 - scoy) * LD::SCW + (xx - scox)]; };
          uint32_t o2[2];
          for (int h = 0; h < 2; h++) {
            const int xs = x + h, c0 = B(xs, y);
            int sum = alf_pair(f[0], B(xs, y + d2), B(xs, y - d2), c0);
            sum += alf_pair(f[1], B(xs + 1, y + d1), B(xs - 1, y - d1), c0);
            sum += alf_pair(f[2], B(xs, y + d1), B(xs, y - d1), c0);
            sum += alf_pair(f[3], B(xs - 1, y + d1), B(xs + 1, y - d1), c0);
            sum += alf_pair(f[4], B(xs + 2, y), B(xs - 2, y), c0);
            sum += alf_pair(f[5], B(xs + 1, y), B(xs - 1, y), c0);
            o2[h] = (uint32_t)clampi(c0 + ((sum + (1 << (sft - 1))) >> sft), 0, amax);
          }
          outp = o2[0] | (o2[1] << 16);
        }
        const int ccf = (int)((cw >> (c == 0 ? 24 : 28)) & 0xf);
        if (ccf != (int)ALF_OFF4) {
          const int32_t* cf = s.ccbank + c * ALF_CCBANK_WORDS + ccf * 7;
          const int yl = 2 * y;
          const int y_ctb_c = ry * ctbc;
          const bool lbb = !(yl + ctb >= H && (H - y_ctb_c * 2) <= ctb - 4);
          int yp1, yp2;
          alf_t47((y - y_ctb_c) * 2, ctb, lbb, yp1, yp2);
          const cc_taps k(cf, c == 0 ? ccsum0 : ccsum1);
          uint32_t o2[2];
          for (int h = 0; h < 2; h++) {
            const int xl = 2 * (x + h);
            const int sum = k.sum(Sl(xl, yl - yp1), Sl(xl - 1, yl), Sl(xl + 1, yl), Sl(xl - 1, yl + yp1),
                                  Sl(xl, yl + yp1), Sl(xl + 1, yl + yp1), Sl(xl, yl + yp2), Sl(xl, yl));
            const int base = (int)((outp >> (16 * h)) & 0xffffu);
            o2[h] = (uint32_t)clampi(base + clampi((sum + 64) >> 7, -half, half - 1), 0, amax);
          }
          outp = o2[0] | (o2[1] << 16);
        }
      }
      // output buffer [2][TH/2][TW/2] in the chroma input windows' space
      uint16_t* const oc = &s.Cc[0][0];
      *reinterpret_cast<uint32_t*>(&oc[(c * (TH / 2) + (y - cy0)) * (TW / 2) + (x - cx0)]) = outp;
    }
  }
  // ---- store the tile: 16-byte chunks of whole rows; issued after the next
  // tile's loads, so that waiting for those loads does not wait for these
  // stores (one in-order counter covers both)
  __syncthreads();
  if (ltid() == 0) s.wide = 0u;  // for the next tile's commit (a barrier follows)
  {
    const int tid = ltid();
    auto put = [&](const uint16_t* src, uint16_t* dst, int xa, int xb) {
      const uint4 v = *reinterpret_cast<const uint4*>(src);
      if (xa + 8 <= xb) {
        gst(reinterpret_cast<uint4*>(dst), v);
      } else {
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (xa + k < xb) gst(dst + k, (uint16_t)(w[k >> 1] >> (16 * (k & 1))));
      }
    };
    constexpr int LCH = TW / 8, NLC = (TH * LCH + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < NLC; q++) {
      const int i = tid + q * NT, yy = i / LCH, xx = 8 * (i - __mul24(yy, LCH));
      if (i < TH * LCH && y0 + yy < y1 && x0 + xx < x1)
        put(&s.L[yy * TW + xx], j.out.y + (size_t)(y0 + yy) * j.py + x0 + xx, x0 + xx, x1);
    }
    if (chroma) {
      // both planes: 2 * (TH/2) rows; Cr lies at a fixed distance from Cb
      // (one allocation), which keeps the plane choice arithmetic
      constexpr int CCH = TW / 16, NCC = (TH * CCH + NT - 1) / NT;
      const uint16_t* const oc = &s.Cc[0][0];
      const ptrdiff_t dcr = j.out.cr - j.out.cb;
#pragma unroll
      for (int q = 0; q < NCC; q++) {
        const int i = tid + q * NT, row = i / CCH, xx = 8 * (i - __mul24(row, CCH));
        const int c = row >= TH / 2 ? 1 : 0, yy = row - c * (TH / 2);
        if (i < TH * CCH && cy0 + yy < cy1 && cx0 + xx < cx1)
          put(&oc[row * (TW / 2) + xx], j.out.cb + c * dcr + (size_t)(cy0 + yy) * j.pc + cx0 + xx, cx0 + xx, cx1);
      }
    }
  }
  if (ltid() < 4) s.nlist[ltid()] = 0;  // descriptor lists of the next tile
  __syncthreads();  // LDS is reused by the next tile
}

// Persistent launch: the tile list is cut into 8 contiguous chunks, one per
// XCD (workgroups are dispatched round-robin over the 8 XCDs), so spatially
// adjacent tiles - which share halo samples - hit the same L2.
template <int TW, int TH, bool HBD>
__global__ void __launch_bounds__((fused_lds<TW, TH>::NT), 3) k_fused(const dev_job* __restrict__ jobs, const tile_ref* __restrict__ tiles,
                                               uint32_t n_tiles) {
  __shared__ fused_lds<TW, TH> s;
  const uint32_t xcd = blockIdx.x & 7, k = blockIdx.x >> 3, per_xcd = gridDim.x >> 3;
  const uint32_t chunk = (n_tiles + 7) >> 3;
  const uint32_t begin = xcd * chunk, end = min(n_tiles, begin + chunk);
  int cur_job = -1;
  if (ltid() < 4) s.nlist[ltid()] = 0;
  if (ltid() == 0) s.wide = 0u;
  if (ltid() < TC_TAB_LEN) s.tct[ltid()] = kTc[ltid()];
  __syncthreads();
  // Software pipeline: the loads of tile t+1 are issued while tile t is
  // filtered (inside fused_tile), those of the first tile here.
  tile_loads<TW, TH> ld;
  tile_ref tr{};
  if (begin + k < end) {
    tr = tiles[begin + k];
    const dev_job& j0 = jobs[tr.job];
    ld.issue(j0, tile_geometry<TW, TH>(j0, s, tr), tr);
  }
  for (uint32_t t = begin + k; t < end; t += per_xcd) {
    const bool more = t + per_xcd < end;
    tile_ref ntr{};
    if (more) ntr = tiles[t + per_xcd];
    const dev_job& j = jobs[tr.job];
    if ((int)tr.job != cur_job) {
      // LMCS inverse table in LDS for bit depths <= 10
      if (j.do_lmcs && j.bd <= 10)
        for (int i = ltid(); i < 1024; i += blockDim.x) s.lut[i] = gld(j.lmcs_lut + i);
      cur_job = (int)tr.job;
      __syncthreads();
    }
    fused_tile<TW, TH, HBD>(j, s, tr, ld, jobs, more, ntr);
    tr = ntr;
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

namespace {
template <int TW, int TH>
int occupancy_of(bool hbd) {
  int n = 0;
  constexpr int NT = fused_lds<TW, TH>::NT;
  hipError_t e = hbd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<TW, TH, true>, NT, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fused<TW, TH, false>, NT, 0);
  return e == hipSuccess ? n : 0;
}
template <int TW, int TH>
void launch_shape(bool hbd, const dev_job* jobs, const tile_ref* tiles, uint32_t n_tiles, uint32_t grid,
                  hipStream_t s) {
  constexpr int NT = fused_lds<TW, TH>::NT;
  if (hbd) hipLaunchKernelGGL((k_fused<TW, TH, true>), dim3(grid), dim3(NT), 0, s, jobs, tiles, n_tiles);
  else hipLaunchKernelGGL((k_fused<TW, TH, false>), dim3(grid), dim3(NT), 0, s, jobs, tiles, n_tiles);
}
}  // namespace

int fused_threads(int tw, int th) { return tw * th > 64 * 64 ? fused_lds<128, 64>::NT : fused_lds<64, 64>::NT; }

int fused_max_blocks_per_cu(int tw, int th, bool hbd) {
  if (tw == 64 && th == 64) return occupancy_of<64, 64>(hbd);
  if (tw == 128 && th == 64) return occupancy_of<128, 64>(hbd);
  if (tw == 64 && th == 32) return occupancy_of<64, 32>(hbd);
  if (tw == 32 && th == 32) return occupancy_of<32, 32>(hbd);
  return 0;
}

hipError_t launch_fused(int tw, int th, bool hbd, const dev_job* jobs, const tile_ref* tiles, uint32_t n_tiles,
                        uint32_t grid, hipStream_t s) {
  grid = (grid + 7) & ~7u;
  if (tw == 64 && th == 64) launch_shape<64, 64>(hbd, jobs, tiles, n_tiles, grid, s);
  else if (tw == 128 && th == 64) launch_shape<128, 64>(hbd, jobs, tiles, n_tiles, grid, s);
  else if (tw == 64 && th == 32) launch_shape<64, 32>(hbd, jobs, tiles, n_tiles, grid, s);
  else if (tw == 32 && th == 32) launch_shape<32, 32>(hbd, jobs, tiles, n_tiles, grid, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace vvclf
