// vvclf_enc.hip — the encoder's per-CTB ALF luma decision as batched GPU
// filtering: alf_decide_and_apply_clipped (src/alf_enc.rs:185-316) and
// alf_decide_and_apply_with_aps_clipped (:350-470).
//
// The reference filters the whole picture once per candidate (16 fixed
// sets, then each APS set), measures each CTB's luma SSE against the source
// after every trial, keeps per CTB the candidate with the lowest SSE (off
// first, then the candidates in order; ties keep the earlier one) and copies
// that trial's samples back.  A CTB's ALF output depends only on the pre-ALF
// picture, so here:
//   k_alf_enc<TRIAL>   one thread per 4x4 block classifies once and filters
//                      the block with every candidate's bank; per candidate
//                      the block's squared error against the source is
//                      reduced over the wave and added to sse[ctb][cand]
//                      (a wave lies in one CTB); candidate 0 is "off"
//   k_alf_pick         per CTB the first minimum
//   k_alf_enc<COMMIT>  the chosen candidate's output (or the pre-ALF
//                      samples) written once
// and the clip-index search of an APS (design_clip_rdo_for_luma_aps,
// src/alf_aps_design.rs:697-768) on the same kernel:
//   k_alf_enc<CLASS>   the block filtered with each candidate bank, the
//                      squared error summed per class (the SSE splits over
//                      classes, since a block's output depends only on its
//                      class's row)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/vvclf.h"
#include "vvclf_filters.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

namespace {

struct enc_args {
  const uint16_t* rec;  // pre-ALF luma (pitch)
  const uint16_t* src;  // source luma (pitch)
  uint16_t* out;        // commit output (pitch)
  int pitch, w, h;
  int ctb_log2, wctb, hctb, bd;
  const uint64_t* rect;        // per CTB fetch rectangle (pack_rect)
  const uint32_t* banks;       // [n_cand][ALF_BANK_WORDS]
  int n_cand;                  // 16 + number of APS sets
  const uint8_t* ok;           // [n_cand]: 0 for an absent or unsignalled APS
  unsigned long long* sse;     // [n_ctbs][1 + n_cand]
  const int32_t* choice;       // [n_ctbs]: chosen trial, -1 off
};

enum { ENC_TRIAL = 0, ENC_COMMIT = 1, ENC_CLASS = 2 };
constexpr int CLASS_MAX_CAND = 4;  // candidate banks of one ENC_CLASS launch

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int vb_kind(int yb, int ctb) { return yb == ctb - 8 ? 1 : (yb == ctb - 4 ? 2 : 0); }

// the class of a block from its 32-bit window (virtual-boundary variant by vb)
__device__ __forceinline__ int class_w(const int (&W)[10][10], int vb, bool carve, int bd) {
  if (vb == 0) return alf_block_class<0, false>(W, bd);
  if (vb == 1) return carve ? alf_block_class<1, true>(W, bd) : alf_block_class<1, false>(W, bd);
  return carve ? alf_block_class<2, true>(W, bd) : alf_block_class<2, false>(W, bd);
}

template <int MODE, bool HBD>
__global__ void __launch_bounds__(256) k_alf_enc(enc_args a, int tw_log2, int th_log2) {
  const int tb = 1 << (tw_log2 - 2), nby = 1 << (th_log2 - 2);
  const int bx = threadIdx.x % tb, by = threadIdx.x / tb;
  const int tx0 = blockIdx.x << tw_log2, ty0 = blockIdx.y << th_log2;
  const int ctb = 1 << a.ctb_log2;
  const int rx = tx0 >> a.ctb_log2, ry = ty0 >> a.ctb_log2;
  const size_t ci = (size_t)ry * a.wctb + rx;
  const int x4 = tx0 + bx * 4, y4 = ty0 + by * 4;
  // lanes past the tile or the picture take part in the wave reductions with
  // zero error and touch no memory
  const bool live = by < nby && x4 < a.w && y4 < a.h;
  const uint64_t r = a.rect[ci];
  const int cx0 = (int)(r & 0xffff), cy0 = (int)((r >> 16) & 0xffff), cx1 = (int)((r >> 32) & 0xffff) - 1,
            cy1 = (int)(r >> 48) - 1;
  const int y_ctb = ry * ctb;
  const bool carve = !(y_ctb + ctb >= a.h) || (a.h - y_ctb) > ctb - 4;
  const int vb = vb_kind(y4 - y_ctb, ctb);
  const int amax = (1 << a.bd) - 1;
  // the block's valid samples (inside the picture)
  const int nr = live ? min(4, a.h - y4) : 0, nc = live ? min(4, a.w - x4) : 0;
  // 10x10 window of the block (rect-clamped), packed or 32-bit
  uint32_t P[10][6];
  int W[10][10];
  bool wide = HBD;
  if (live) {
    if (!HBD && x4 - 3 >= cx0 && x4 + 6 <= cx1 && y4 - 3 >= cy0 && y4 + 6 <= cy1) {
      const uint16_t* w0 = a.rec + (size_t)(y4 - 3) * a.pitch + x4 - 4;
#pragma unroll
      for (int rr = 0; rr < 10; rr++) {
        const uint2* q = reinterpret_cast<const uint2*>(w0 + (size_t)rr * a.pitch);
        const uint2 u = q[0], v = q[1], t = q[2];
        P[rr][0] = u.x; P[rr][1] = u.y; P[rr][2] = v.x; P[rr][3] = v.y; P[rr][4] = t.x; P[rr][5] = t.y;
      }
    } else {
#pragma unroll
      for (int rr = 0; rr < 10; rr++) {
        const uint16_t* row = a.rec + (size_t)clampi(y4 - 3 + rr, cy0, cy1) * a.pitch;
        uint32_t v[12];
        v[0] = 0;
        v[11] = 0;
#pragma unroll
        for (int q = 1; q <= 10; q++) v[q] = row[clampi(x4 - 4 + q, cx0, cx1)];
#pragma unroll
        for (int k = 0; k < 6; k++) P[rr][k] = v[2 * k] | (v[2 * k + 1] << 16);
      }
    }
    uint32_t any = 0;
#pragma unroll
    for (int rr = 0; rr < 10; rr++)
#pragma unroll
      for (int k = 0; k < 6; k++) any |= P[rr][k];
    wide = HBD || (any & 0x80008000u);
  } else {
#pragma unroll
    for (int rr = 0; rr < 10; rr++)
#pragma unroll
      for (int k = 0; k < 6; k++) P[rr][k] = 0;
  }
  if (wide) {
#pragma unroll
    for (int rr = 0; rr < 10; rr++)
#pragma unroll
      for (int cc = 0; cc < 10; cc++) W[rr][cc] = (int)((P[rr][(cc + 1) >> 1] >> (16 * ((cc + 1) & 1))) & 0xffffu);
  }
  const int cls = wide ? class_w(W, vb, carve, a.bd) : alf_pk_block_class(P, vb, carve, a.bd);
  const int row_of_bank = ((cls & 31) * 4 + (cls >> 5)) * 12;

  // the block's 16 filtered samples with candidate s (out[r][c])
  auto filter = [&](int s, int (&o)[4][4]) {
    const uint32_t* f = a.banks + (size_t)s * ALF_BANK_WORDS + row_of_bank;
    if (wide) {
      uint32_t fw[12];
#pragma unroll
      for (int k = 0; k < 12; k++) fw[k] = f[k];
      if (vb == 0) alf_block_filter<0>(W, fw, amax, o);
      else if (vb == 1) alf_block_filter<1>(W, fw, amax, o);
      else alf_block_filter<2>(W, fw, amax, o);
    } else {
      uint32_t op[4][2];
      if (vb == 0) alf_pk_filter<0>(P, f, amax, op);
      else if (vb == 1) alf_pk_filter<1>(P, f, amax, op);
      else alf_pk_filter<2>(P, f, amax, op);
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++) o[rr][c] = (int)((op[rr][c >> 1] >> (16 * (c & 1))) & 0xffffu);
    }
  };
  auto centre = [&](int rr, int c) {
    return wide ? W[rr + 3][c + 3] : (int)((P[rr + 3][(c + 4) >> 1] >> (16 * ((c + 4) & 1))) & 0xffffu);
  };

  if constexpr (MODE == ENC_CLASS) {
    // sse: [n_cand][25], per class; ok and the off candidate unused
    __shared__ unsigned long long acc[CLASS_MAX_CAND * 25];
    for (int i = threadIdx.x; i < CLASS_MAX_CAND * 25; i += blockDim.x) acc[i] = 0;
    __syncthreads();
    if (live) {
      const int k = cls & 31;
      for (int s = 0; s < a.n_cand; s++) {
        int o[4][4];
        filter(s, o);
        unsigned long long es = 0;
#pragma unroll
        for (int rr = 0; rr < 4; rr++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            if (rr < nr && c < nc) {
              const long long d = (long long)a.src[(size_t)(y4 + rr) * a.pitch + x4 + c] - o[rr][c];
              es += (unsigned long long)(d * d);
            }
        if (es) atomicAdd(&acc[s * 25 + k], es);
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.n_cand * 25; i += blockDim.x)
      if (acc[i]) atomicAdd(a.sse + i, acc[i]);
  } else if constexpr (MODE == ENC_TRIAL) {
    int S[4][4];
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
#pragma unroll
      for (int c = 0; c < 4; c++) S[rr][c] = (rr < nr && c < nc) ? a.src[(size_t)(y4 + rr) * a.pitch + x4 + c] : 0;
    unsigned long long* row = a.sse + ci * (size_t)(1 + a.n_cand);
    // candidate 0: off (the pre-ALF samples)
    unsigned long long e = 0;
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (rr < nr && c < nc) {
          const long long d = (long long)S[rr][c] - centre(rr, c);
          e += (unsigned long long)(d * d);
        }
    e = wave_sum64(e);
    if ((threadIdx.x & 63) == 0 && e) atomicAdd(row, e);
    for (int s = 0; s < a.n_cand; s++) {
      if (!a.ok[s]) continue;
      int o[4][4];
      filter(s, o);
      unsigned long long es = 0;
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (rr < nr && c < nc) {
            const long long d = (long long)S[rr][c] - o[rr][c];
            es += (unsigned long long)(d * d);
          }
      es = wave_sum64(es);
      if ((threadIdx.x & 63) == 0 && es) atomicAdd(row + 1 + s, es);
    }
  } else {
    if (!live) return;
    const int s = a.choice[ci];
    int o[4][4];
    if (s >= 0) {
      filter(s, o);
    } else {
#pragma unroll
      for (int rr = 0; rr < 4; rr++)
#pragma unroll
        for (int c = 0; c < 4; c++) o[rr][c] = centre(rr, c);
    }
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (rr < nr && c < nc) a.out[(size_t)(y4 + rr) * a.pitch + x4 + c] = (uint16_t)o[rr][c];
  }
}

// per CTB the first candidate with the lowest SSE (off first); a candidate
// naming an unusable APS leaves the CTB unfiltered, so its error is the off
// error
__global__ void __launch_bounds__(256) k_alf_pick(unsigned long long* sse, const uint8_t* ok, int n_ctbs, int n_cand,
                                                  int32_t* choice) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_ctbs) return;
  unsigned long long* row = sse + (size_t)i * (1 + n_cand);
  unsigned long long best = row[0];
  int t = -1;
  for (int s = 0; s < n_cand; s++) {
    if (!ok[s]) row[1 + s] = row[0];
    if (row[1 + s] < best) {
      best = row[1 + s];
      t = s;
    }
  }
  choice[i] = t;
}

// Tiles of k_alf_enc, inside one CTB: 128 x 32 for CTBs of 128 and more (a
// wave holds two block rows, so only one wave per CTB column runs both
// virtual-boundary variants), else 64 x 64 or the CTB
struct alf_tiling {
  int twl, thl;
  dim3 grid;
  unsigned threads;
  alf_tiling(uint32_t W, uint32_t H, uint32_t ctb_log2) {
    twl = ctb_log2 < 6 ? (int)ctb_log2 : (ctb_log2 >= 7 ? 7 : 6);
    thl = ctb_log2 < 6 ? (int)ctb_log2 : (ctb_log2 >= 7 ? 5 : 6);
    grid = dim3((W + (1u << twl) - 1) >> twl, (H + (1u << thl) - 1) >> thl);
    threads = std::max(64u, (1u << (twl - 2)) << (thl - 2));
  }
};

struct dbuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  hipError_t alloc(size_t n, hipStream_t st) {
    s = st;
    return hipMallocAsync(&p, std::max<size_t>(n, 4), st);
  }
  ~dbuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
};

// ---------------------------------------------------------------------------
// Chroma decisions (4:2:0): chroma ALF alternatives
// (chroma_alf_decide_and_apply_clipped, src/alf_enc.rs:628-806) and CC-ALF
// filters (cc_alf_decide_and_apply_clipped, :935-1140).  One thread per chroma
// sample over T x T tiles, T = min(chroma CTB, 16): a tile, and so a wave,
// lies in one CTB.  Per candidate the sample's squared error against the
// source is reduced over the wave into sse[ctb][1 + cand]; candidate 0 is off
// (the input sample).
// ---------------------------------------------------------------------------
struct cenc_args {
  const uint16_t* rec;   // the component before the decision (pitch cp)
  const uint16_t* src;   // source component (pitch cp)
  uint16_t* out;         // commit output (pitch cp)
  const uint16_t* luma;  // CC-ALF: the luma plane before luma ALF (pitch lp)
  int cp, lp, wc, hc, w, h;
  int ctb_log2, wctb, bd;
  const uint64_t* rect;   // per CTB luma fetch rectangle (pack_rect)
  const uint32_t* cbank;  // chroma alternatives, 6 tap words each
  const int32_t* ccbank;  // CC filters of the component, 7 taps each
  int n_cand;
  unsigned long long* sse;  // [n_ctbs][1 + n_cand]
  const int32_t* choice;    // [n_ctbs]
};

// Chroma ALF of sample (xs, y) with alternative f (the per-sample form of
// k_alf_chroma, src/alf.rs:1014-1065): taps of the 5x5 diamond read clamped
// to the CTB's chroma rectangle, Table 46 rows near the virtual boundary.
__device__ __forceinline__ int chroma_alf_at(const cenc_args& a, int xs, int y, int cx0, int cy0, int cx1, int cy1,
                                             const uint32_t* f, int sft, int y1, int y2, int maxv) {
  auto S = [&](int dx, int dy) -> int {
    return a.rec[(size_t)clampi(y + dy, cy0, cy1) * a.cp + clampi(xs + dx, cx0, cx1)];
  };
  const int c0 = S(0, 0);
  int sum = alf_pair(f[0], S(0, y2), S(0, -y2), c0);
  sum += alf_pair(f[1], S(1, y1), S(-1, -y1), c0);
  sum += alf_pair(f[2], S(0, y1), S(0, -y1), c0);
  sum += alf_pair(f[3], S(-1, y1), S(1, -y1), c0);
  sum += alf_pair(f[4], S(2, 0), S(-2, 0), c0);
  sum += alf_pair(f[5], S(1, 0), S(-1, 0), c0);
  return clampi(c0 + ((sum + (1 << (sft - 1))) >> sft), 0, maxv);
}

// CC-ALF correction of chroma sample (xs, y) (src/alf.rs:447-523): the seven
// luma taps around (2 xs, 2 y), rows by Table 47, clamped to the CTB's luma
// rectangle [lx0, lx1) x [ly0, ly1).
__device__ __forceinline__ int cc_delta_at(const cenc_args& a, int xs, int y, int lx0, int ly0, int lx1, int ly1,
                                           const cc_taps& k, int yp1, int yp2, int half) {
  const int yl = 2 * y, xl = 2 * xs;
  auto row = [&](int dy) { return a.luma + (size_t)clampi(yl + dy, ly0, ly1 - 1) * a.lp; };
  const uint16_t *la = row(-yp1), *lc = row(0), *lb = row(yp1), *ld = row(yp2);
  auto cx = [&](int dx) { return clampi(xl + dx, lx0, lx1 - 1); };
  const int sum = k.sum(la[cx(0)], lc[cx(-1)], lc[cx(1)], lb[cx(-1)], lb[cx(0)], lb[cx(1)], ld[cx(0)], lc[cx(0)]);
  return clampi((sum + 64) >> 7, -half, half - 1);
}

template <bool TRIAL, bool CC>
__global__ void __launch_bounds__(256) k_chroma_enc(cenc_args a, int tile_log2) {
  const int T = 1 << tile_log2;
  const int tx = threadIdx.x % T, ty = threadIdx.x / T;
  const int x = blockIdx.x * T + tx, y = blockIdx.y * T + ty;
  const bool live = ty < T && x < a.wc && y < a.hc;
  const int ctb = 1 << a.ctb_log2, ctbc = ctb >> 1;
  // the tile's CTB (the tile lies in one)
  const int rx = (2 * (int)(blockIdx.x * T)) >> a.ctb_log2, ry = (2 * (int)(blockIdx.y * T)) >> a.ctb_log2;
  const size_t ci = (size_t)ry * a.wctb + rx;
  const uint64_t r = a.rect[ci];
  const int lx0 = (int)(r & 0xffff), ly0 = (int)((r >> 16) & 0xffff), lx1 = (int)((r >> 32) & 0xffff),
            ly1 = (int)(r >> 48);
  const int maxv = (1 << a.bd) - 1;
  const int cur = live ? (int)a.rec[(size_t)y * a.cp + x] : 0;
  // per-candidate geometry
  int sft = 7, y1 = 1, y2 = 2, yp1 = 1, yp2 = 2;
  const int cx0 = lx0 >> 1, cy0 = ly0 >> 1, cx1 = ((lx1 + 1) >> 1) - 1, cy1 = ((ly1 + 1) >> 1) - 1;
  if (CC) {
    const int yl = 2 * y, y_ctb_c = ry * ctbc;
    const bool lbb = !(yl + ctb >= a.h && (a.h - y_ctb_c * 2) <= ctb - 4);
    alf_t47((y - y_ctb_c) * 2, ctb, lbb, yp1, yp2);
  } else {
    alf_t46(y - ry * ctbc, ctbc, sft, y1, y2);
  }
  const int half = 1 << (a.bd - 1);
  auto candidate = [&](int c) -> int {
    if (CC) {
      const cc_taps k(a.ccbank + c * 7);
      return clampi(cur + cc_delta_at(a, x, y, lx0, ly0, lx1, ly1, k, yp1, yp2, half), 0, maxv);
    }
    return chroma_alf_at(a, x, y, cx0, cy0, cx1, cy1, a.cbank + c * 6, sft, y1, y2, maxv);
  };
  if (TRIAL) {
    const int s = live ? (int)a.src[(size_t)y * a.cp + x] : 0;
    unsigned long long* row = a.sse + ci * (size_t)(1 + a.n_cand);
    long long d = s - cur;
    unsigned long long e = wave_sum64((unsigned long long)(d * d));
    if ((threadIdx.x & 63) == 0 && e) atomicAdd(row, e);
    for (int c = 0; c < a.n_cand; c++) {
      d = live ? (long long)s - candidate(c) : 0;
      e = wave_sum64((unsigned long long)(d * d));
      if ((threadIdx.x & 63) == 0 && e) atomicAdd(row + 1 + c, e);
    }
  } else {
    if (!live) return;
    const int c = a.choice[ci];
    a.out[(size_t)y * a.cp + x] = (uint16_t)(c >= 0 ? candidate(c) : cur);
  }
}

// ---------------------------------------------------------------------------
// SAO decision (sao_decide_picture, src/sao_enc.rs:385-462): per CTB and
// component, band statistics (count and sum of src - rec per band,
// band = rec >> (BitDepth - 5)) and edge statistics (per edge class and
// category 1..4, neighbours taken from rec inside the plane) over the CTB's
// region, then the reference's per-CTB choice from them.
// k_sao_stats: a workgroup per T x T tile of one component (inside one CTB),
// counters in LDS, one 64-bit atomic per counter and tile into the CTB's row.
// ---------------------------------------------------------------------------
constexpr int SAO_STAT_WORDS = 100;  // 32 band counts, 32 band sums, 16 edge counts, 16 edge sums, overflow
enum { SB_CNT = 0, SB_DIF = 32, SE_CNT = 64, SE_DIF = 80, S_OVER = 96 };

struct sao_args {
  const uint16_t* rec[3];
  const uint16_t* src[3];
  int pitch[3], pw[3], ph[3];  // plane pitch and size
  int w, h, ctb_log2, wctb, bd, luma, chroma;
  long long* stats;  // [n_ctbs][3][SAO_STAT_WORDS]
};

__device__ __forceinline__ int eo_cat(int r, int n0, int n1) {
  const int e = (r > n0) - (r < n0) + (r > n1) - (r < n1);
  return e == 2 ? 4 : (e == 1 ? 3 : (e == -1 ? 2 : (e == -2 ? 1 : 0)));
}

__global__ void __launch_bounds__(256) k_sao_stats(sao_args a, int tile_log2_luma) {
  __shared__ int cnt[SAO_STAT_WORDS];
  const int c = blockIdx.z;
  if (c == 0 ? !a.luma : !a.chroma) return;
  for (int i = threadIdx.x; i < SAO_STAT_WORDS; i += 256) cnt[i] = 0;
  __syncthreads();
  const int tl = c ? tile_log2_luma - 1 : tile_log2_luma, T = 1 << tl;
  const int x0 = blockIdx.x * T, y0 = blockIdx.y * T;
  const int pw = a.pw[c], ph = a.ph[c];
  if (x0 < pw && y0 < ph) {  // (workgroup-uniform)
    const int ctb = 1 << a.ctb_log2, sh = c ? 1 : 0;
    const int rx = (x0 << sh) >> a.ctb_log2, ry = (y0 << sh) >> a.ctb_log2;
    // the CTB's region: luma CTB clipped to the picture; chroma half of it,
    // rounded down (src/sao_enc.rs:404-432)
    const int lx = rx * ctb, ly = ry * ctb;
    const int rw = min(ctb, a.w - lx) >> sh, rh = min(ctb, a.h - ly) >> sh;
    const int ex = min((lx >> sh) + rw, pw), ey = min((ly >> sh) + rh, ph);
    const uint16_t* R = a.rec[c];
    const uint16_t* S = a.src[c];
    const int P = a.pitch[c], shift = a.bd - 5;
    for (int i = threadIdx.x; i < T * T; i += 256) {
      const int x = x0 + (i & (T - 1)), y = y0 + (i >> tl);
      if (x >= ex || y >= ey) continue;
      const int r = R[(size_t)y * P + x], d = (int)S[(size_t)y * P + x] - r;
      const int b = r >> shift;
      if (b >= 32) {
        atomicOr(&cnt[S_OVER], 1);
      } else {
        atomicAdd(&cnt[SB_CNT + b], 1);
        atomicAdd(&cnt[SB_DIF + b], d);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int dx = k == 1 ? 0 : (k == 3 ? 1 : -1), dy = k == 0 ? 0 : -1;  // n0 = (x+dx, y+dy), n1 = mirror
        const int x_0 = x + dx, y_0 = y + dy, x_1 = x - dx, y_1 = y - dy;
        if (x_0 < 0 || y_0 < 0 || x_0 >= pw || y_0 >= ph || x_1 < 0 || y_1 < 0 || x_1 >= pw || y_1 >= ph) continue;
        const int cat = eo_cat(r, R[(size_t)y_0 * P + x_0], R[(size_t)y_1 * P + x_1]);
        if (cat) {
          atomicAdd(&cnt[SE_CNT + k * 4 + cat - 1], 1);
          atomicAdd(&cnt[SE_DIF + k * 4 + cat - 1], d);
        }
      }
    }
    __syncthreads();
    long long* row = a.stats + (((size_t)ry * a.wctb + rx) * 3 + c) * SAO_STAT_WORDS;
    for (int i = threadIdx.x; i < SAO_STAT_WORDS; i += 256)
      if (cnt[i]) atomicAdd(reinterpret_cast<unsigned long long*>(row + i), (unsigned long long)(long long)cnt[i]);
  }
}

__device__ void sao_offsets(const int32_t (&absv)[4], const int (&sign)[4], int bd, int32_t (&ov)[5]) {
  const int sh = bd - min(bd, 10);
  ov[0] = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) ov[i + 1] = (1 - 2 * (sign[i] & 1)) * (absv[i] << sh);
}

// try_band_offset (src/sao_enc.rs:77-170) from the statistics; false when no
// band window improves
__device__ bool sao_band_pick(const long long* st, int bd, vvclf_sao_ctb& o) {
  long long best_delta = 0;
  int best_start = 0;
  int32_t best_off[4] = {0, 0, 0, 0};
  const long long cmax = (1ll << (min(bd, 10) - 5)) - 1;
  for (int start = 0; start < 28; start++) {
    long long delta = 0;
    int32_t off[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
      const long long n = st[SB_CNT + start + i], dsum = st[SB_DIF + start + i];
      if (n > 0) {
        long long opt = (dsum + n / 2) / n;
        opt = opt < -cmax ? -cmax : (opt > cmax ? cmax : opt);
        off[i] = (int32_t)opt;
        delta -= dsum * opt;
      }
    }
    if (delta < best_delta) {
      best_delta = delta;
      best_start = start;
      for (int i = 0; i < 4; i++) best_off[i] = off[i];
    }
  }
  if (best_delta >= 0) return false;
  int32_t ab[4];
  int sg[4];
  for (int i = 0; i < 4; i++) {
    ab[i] = best_off[i] < 0 ? -best_off[i] : best_off[i];
    sg[i] = best_off[i] < 0;
  }
  o.sao_type_idx = VVCLF_SAO_BAND;
  o.eo_class = 0;
  o.band_position = (uint8_t)(best_start & 31);
  o.pad = 0;
  sao_offsets(ab, sg, bd, o.offset_val);
  return true;
}

// try_edge_offset_class (src/sao_enc.rs:211-284): the class's candidate and
// its delta
__device__ long long sao_eo_pick(const long long* st, int cls, int bd, vvclf_sao_ctb& o) {
  int32_t off[4] = {0, 0, 0, 0};
  long long delta = 0;
  for (int i = 0; i < 4; i++) {
    const long long n = st[SE_CNT + cls * 4 + i], dsum = st[SE_DIF + cls * 4 + i];
    if (n > 0) {
      const long long av = dsum < 0 ? -dsum : dsum;
      long long opt = (av + n / 2) / n;
      opt = opt < 0 ? 0 : (opt > 7 ? 7 : opt);
      off[i] = (int32_t)opt;
      delta -= av * opt;
    }
  }
  const int sg[4] = {0, 0, 1, 1};
  o.sao_type_idx = VVCLF_SAO_EDGE;
  o.eo_class = (uint8_t)cls;
  o.band_position = 0;
  o.pad = 0;
  sao_offsets(off, sg, bd, o.offset_val);
  return delta;
}

__device__ void sao_clear(vvclf_sao_ctb& o) {
  o.sao_type_idx = 0;
  o.eo_class = 0;
  o.band_position = 0;
  o.pad = 0;
  for (int i = 0; i < 5; i++) o.offset_val[i] = 0;
}

// sao_decide_ctb (src/sao_enc.rs:44-71): edge offset when a class improves,
// else band offset when a window improves, else not applied
__device__ void sao_decide_one(const long long* st, int bd, vvclf_sao_ctb& o) {
  vvclf_sao_ctb bo, t;
  const bool has_bo = sao_band_pick(st, bd, bo);
  long long best = 0;
  int best_cls = -1;
  for (int k = 0; k < 4; k++) {
    const long long d = sao_eo_pick(st, k, bd, t);
    if (d < best) {
      best = d;
      best_cls = k;
    }
  }
  if (best_cls >= 0) sao_eo_pick(st, best_cls, bd, o);
  else if (has_bo) o = bo;
  else sao_clear(o);
}

// one thread per CTB; *bad is set when a sample of 2^BitDepth or more would
// reach a band statistic the decision reads (the reference panics there)
__global__ void __launch_bounds__(256) k_sao_decide(sao_args a, int n_ctbs, vvclf_sao_ctb_params* out,
                                                    unsigned* bad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_ctbs) return;
  const long long* st = a.stats + (size_t)i * 3 * SAO_STAT_WORDS;
  vvclf_sao_ctb_params p;
  sao_clear(p.luma);
  sao_clear(p.cb);
  sao_clear(p.cr);
  if (a.luma) {
    if (st[S_OVER]) atomicOr(bad, 1u);
    sao_decide_one(st, a.bd, p.luma);
  }
  const int ctb = 1 << a.ctb_log2, rx = i % a.wctb, ry = i / a.wctb;
  const int cw = min(ctb, a.w - rx * ctb) / 2, ch = min(ctb, a.h - ry * ctb) / 2;
  if (a.chroma && cw > 0 && ch > 0) {
    const long long *sb = st + SAO_STAT_WORDS, *sr = st + 2 * SAO_STAT_WORDS;
    if (sb[S_OVER]) atomicOr(bad, 1u);
    sao_decide_one(sb, a.bd, p.cb);
    // sao_decide_ctb_constrained (src/sao_enc.rs:286-347)
    if (p.cb.sao_type_idx == VVCLF_SAO_BAND) {
      if (sr[S_OVER]) atomicOr(bad, 1u);
      if (!sao_band_pick(sr, a.bd, p.cr)) {
        sao_clear(p.cr);
        p.cr.sao_type_idx = VVCLF_SAO_BAND;
      }
    } else if (p.cb.sao_type_idx == VVCLF_SAO_EDGE) {
      vvclf_sao_ctb t;
      if (sao_eo_pick(sr, p.cb.eo_class, a.bd, t) < 0) {
        p.cr = t;
      } else {
        p.cr.sao_type_idx = VVCLF_SAO_EDGE;
        p.cr.eo_class = p.cb.eo_class;
      }
    }
  }
  out[i] = p;
}

// ---------------------------------------------------------------------------
// Chroma SAO merge decision (apply_chroma_sao_merge, src/sao_enc.rs:525-660).
// A CTB's choice reads its left and above neighbours' final parameters, so
// the CTBs are decided in anti-diagonal waves (one launch per diagonal, every
// CTB of a diagonal in parallel); a workgroup per CTB evaluates the three
// candidates (independent, merge left, merge above) over the chroma CTB in
// one pass: SAO of each candidate on the pre-SAO samples (as apply_sao_ctb,
// src/sao.rs:504-626, without tile clips), Cb + Cr SSE against the source.
// ---------------------------------------------------------------------------
struct merge_args {
  const uint16_t* rec[2];  // pre-SAO Cb, Cr (pitch)
  const uint16_t* src[2];
  int pitch, cw, ch;        // chroma planes (both the same size)
  int wctb, hctb, cctb, bd;
  vvclf_sao_ctb_params* merged;  // in: independent parameters; out: merged
  uint8_t* map;                  // 0 independent, 1 merge left, 2 merge above
};

// SAO of chroma sample (x, y) with parameters p, reading the pre-SAO plane
__device__ __forceinline__ int sao_sample(const uint16_t* pre, int pitch, int pw, int ph, int x, int y,
                                          const vvclf_sao_ctb& p, int bd) {
  const int maxv = (1 << bd) - 1;
  const int v = pre[(size_t)y * pitch + x];
  if (p.sao_type_idx == VVCLF_SAO_EDGE) {
    const int c = p.eo_class & 3;
    const int dx = c == 1 ? 0 : (c == 3 ? 1 : -1), dy = c == 0 ? 0 : -1;  // n0; n1 mirrored
    const int x0 = x + dx, y0 = y + dy, x1 = x - dx, y1 = y - dy;
    if (x0 < 0 || y0 < 0 || x1 < 0 || y1 < 0 || x0 >= pw || y0 >= ph || x1 >= pw || y1 >= ph) return min(v, maxv);
    const int n0 = pre[(size_t)y0 * pitch + x0], n1 = pre[(size_t)y1 * pitch + x1];
    int e = 2 + (v > n0) - (v < n0) + (v > n1) - (v < n1);
    if (e <= 2) e = e == 2 ? 0 : e + 1;
    return clampi(v + (e ? p.offset_val[e] : 0), 0, maxv);
  }
  if (p.sao_type_idx == VVCLF_SAO_BAND) {
    const int k = (((v >> (bd - 5)) & 31) - p.band_position) & 31;
    return clampi(v + (k < 4 ? p.offset_val[k + 1] : 0), 0, maxv);
  }
  return v;
}

__global__ void __launch_bounds__(256) k_sao_merge_step(merge_args a, int d, int rx_lo) {
  __shared__ unsigned long long part[4][3];
  const int rx = rx_lo + blockIdx.x, ry = d - rx;
  const size_t i = (size_t)ry * a.wctb + rx;
  const vvclf_sao_ctb_params in = a.merged[i];
  const vvclf_sao_ctb_params lf = rx > 0 ? a.merged[i - 1] : in;
  const vvclf_sao_ctb_params ab = ry > 0 ? a.merged[i - a.wctb] : in;
  const int x0 = rx * a.cctb, y0 = ry * a.cctb;
  const int cw = max(0, min(a.cctb, a.cw - x0)), ch = max(0, min(a.cctb, a.ch - y0));
  unsigned long long s0 = 0, s1 = 0, s2 = 0;
  for (int t = threadIdx.x; t < cw * ch; t += 256) {
    const int x = x0 + t % cw, y = y0 + t / cw;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int sv = a.src[c][(size_t)y * a.pitch + x];
      const vvclf_sao_ctb &pi = c ? in.cr : in.cb, &pl = c ? lf.cr : lf.cb, &pa = c ? ab.cr : ab.cb;
      long long dd = sv - sao_sample(a.rec[c], a.pitch, a.cw, a.ch, x, y, pi, a.bd);
      s0 += (unsigned long long)(dd * dd);
      dd = sv - sao_sample(a.rec[c], a.pitch, a.cw, a.ch, x, y, pl, a.bd);
      s1 += (unsigned long long)(dd * dd);
      dd = sv - sao_sample(a.rec[c], a.pitch, a.cw, a.ch, x, y, pa, a.bd);
      s2 += (unsigned long long)(dd * dd);
    }
  }
  s0 = wave_sum64(s0);
  s1 = wave_sum64(s1);
  s2 = wave_sum64(s2);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[w][0] = s0;
    part[w][1] = s1;
    part[w][2] = s2;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  unsigned long long t0 = 0, t1 = 0, t2 = 0;
  for (int k = 0; k < 4; k++) {
    t0 += part[k][0];
    t1 += part[k][1];
    t2 += part[k][2];
  }
  auto bits = [](const vvclf_sao_ctb& c) -> long long {
    return c.sao_type_idx == VVCLF_SAO_BAND ? 16 : (c.sao_type_idx == VVCLF_SAO_EDGE ? 10 : 1);
  };
  long long best = (long long)t0 + 8 * (1 + bits(in.cb) + bits(in.cr));
  int choice = 0;
  vvclf_sao_ctb bcb = in.cb, bcr = in.cr;
  if (rx > 0 && (long long)t1 + 8 < best) {
    best = (long long)t1 + 8;
    choice = 1;
    bcb = lf.cb;
    bcr = lf.cr;
  }
  if (ry > 0 && (long long)t2 + 16 < best) {
    choice = 2;
    bcb = ab.cb;
    bcr = ab.cr;
  }
  a.map[i] = (uint8_t)choice;
  a.merged[i].cb = bcb;
  a.merged[i].cr = bcr;
}

}  // namespace

extern "C" int vvclf_alf_decide_luma(const vvclf_picture* src, vvclf_picture* rec, uint32_t ctb_log2, uint32_t bd,
                                     uint32_t cfi, const vvclf_alf_aps* const* apses, uint32_t n_aps,
                                     const vvclf_bounds* tb, vvclf_alf_ctb* out, uint64_t* sse_out) {
  clear_error();
  if (!src || !rec || !out || (n_aps && !apses)) return fail(VVCLF_ERR_INVALID, "alf_decide_luma: NULL argument");
  const vvclf_plane &S = src->luma, &R = rec->luma;
  const uint32_t W = R.width, H = R.height;
  if (!S.samples || !R.samples || S.width != W || S.height != H || W == 0 || H == 0)
    return fail(VVCLF_ERR_INVALID, "alf_decide_luma: source and reconstruction luma differ");
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "alf_decide_luma: bit depth %u", bd);
  if (ctb_log2 < 3 || ctb_log2 > 8) return fail(VVCLF_ERR_UNSUPPORTED, "alf_decide_luma: ctb_log2_size_y %u", ctb_log2);
  if (W > 65535 || H > 65535) return fail(VVCLF_ERR_UNSUPPORTED, "alf_decide_luma: picture above 65535 samples");
  if (n_aps > 239) return fail(VVCLF_ERR_INVALID, "alf_decide_luma: %u APS sets (at most 239)", n_aps);
  for (uint32_t i = 0; i < n_aps; i++)
    if (apses[i] && apses[i]->alf_luma_filter_signal_flag && apses[i]->n_luma_coeff < 25)
      return fail(VVCLF_ERR_INVALID, "alf_decide_luma: APS %u carries %u luma filter rows", i, apses[i]->n_luma_coeff);
  // the filters hold luma coefficients in 16-bit lanes (conformant streams
  // keep |c| <= 128); larger values are refused rather than filtered
  // differently
  for (uint32_t i = 0; i < n_aps; i++) {
    const vvclf_alf_aps* ap = apses[i];
    if (!ap || !ap->alf_luma_filter_signal_flag) continue;
    for (int c = 0; c < 25; c++)
      for (int k = 0; k < 12; k++)
        if (ap->luma_coeff[c][k] < -32768 || ap->luma_coeff[c][k] > 32767)
          return fail(VVCLF_ERR_UNSUPPORTED, "alf_decide_luma: luma coefficient %d outside 16 bits", ap->luma_coeff[c][k]);
  }
  const uint32_t ctb = 1u << ctb_log2, wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  const uint32_t n_ctbs = wctb * hctb, n_cand = 16 + n_aps;
  // banks: 16 fixed sets, then the APS sets (an absent or unsignalled APS
  // leaves its bank zero, i.e. the identity, as the reference skips the CTB)
  vvclf_alf_config ac;
  memset(&ac, 0, sizeof(ac));
  ac.alf_enabled = 1;
  ac.bit_depth = bd;
  ac.ctb_log2_size_y = ctb_log2;
  ac.chroma_format_idc = cfi;
  vvclf_alf_aps_binding b;
  memset(&b, 0, sizeof(b));
  b.luma_apses = apses;
  b.n_luma_apses = n_aps;
  host_tables t;
  build_alf_banks(ac, b, t);
  std::vector<uint8_t> ok(n_cand, 1);
  for (uint32_t i = 0; i < n_aps; i++) ok[16 + i] = apses[i] && apses[i]->alf_luma_filter_signal_flag;
  std::vector<uint64_t> rects(n_ctbs);
  for (uint32_t ry = 0; ry < hctb; ry++)
    for (uint32_t rx = 0; rx < wctb; rx++) {
      rects[(size_t)ry * wctb + rx] = ctb_fetch_rect(tb, rx, ry, ctb, W, H);
      if (!rects[(size_t)ry * wctb + rx]) return fail(VVCLF_ERR_INVALID, "alf_decide_luma: empty clip rectangle");
    }
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const uint32_t pitch = (W + 63) / 64 * 64;
  const size_t plane = (size_t)pitch * H * 2;
  dbuf drec, dsrc, dout, drect, dbank, dsse, dchoice, dok;
  hipError_t e = drec.alloc(plane, st);
  if (e == hipSuccess) e = dsrc.alloc(plane, st);
  if (e == hipSuccess) e = dout.alloc(plane, st);
  if (e == hipSuccess) e = drect.alloc(rects.size() * 8, st);
  if (e == hipSuccess) e = dbank.alloc(t.alf_bank.size() * 4, st);
  if (e == hipSuccess) e = dsse.alloc((size_t)n_ctbs * (1 + n_cand) * 8, st);
  if (e == hipSuccess) e = dchoice.alloc((size_t)n_ctbs * 4, st);
  if (e == hipSuccess) e = dok.alloc(n_cand, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "alf_decide_luma: device buffers: %s", hipGetErrorString(e));
  e = copy_rows(drec.p, (size_t)pitch * 2, R.samples, (size_t)R.stride * 2, (size_t)W * 2, H,
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = copy_rows(dsrc.p, (size_t)pitch * 2, S.samples, (size_t)S.stride * 2, (size_t)W * 2, H,
                         hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(drect.p, rects.data(), rects.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dbank.p, t.alf_bank.data(), t.alf_bank.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dok.p, ok.data(), n_cand, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(dsse.p, 0, (size_t)n_ctbs * (1 + n_cand) * 8, st);
  enc_args a;
  a.rec = (const uint16_t*)drec.p;
  a.src = (const uint16_t*)dsrc.p;
  a.out = (uint16_t*)dout.p;
  a.pitch = (int)pitch;
  a.w = (int)W;
  a.h = (int)H;
  a.ctb_log2 = (int)ctb_log2;
  a.wctb = (int)wctb;
  a.hctb = (int)hctb;
  a.bd = (int)bd;
  a.rect = (const uint64_t*)drect.p;
  a.banks = (const uint32_t*)dbank.p;
  a.n_cand = (int)n_cand;
  a.ok = (const uint8_t*)dok.p;
  a.sse = (unsigned long long*)dsse.p;
  a.choice = (const int32_t*)dchoice.p;
  const alf_tiling tg(W, H, ctb_log2);
  const bool hbd = bd > 12;
  if (e == hipSuccess) {
    if (hbd) hipLaunchKernelGGL((k_alf_enc<ENC_TRIAL, true>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    else hipLaunchKernelGGL((k_alf_enc<ENC_TRIAL, false>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    hipLaunchKernelGGL(k_alf_pick, dim3((n_ctbs + 255) / 256), dim3(256), 0, st, (unsigned long long*)dsse.p,
                       (const uint8_t*)dok.p, (int)n_ctbs, (int)n_cand, (int32_t*)dchoice.p);
    if (hbd) hipLaunchKernelGGL((k_alf_enc<ENC_COMMIT, true>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    else hipLaunchKernelGGL((k_alf_enc<ENC_COMMIT, false>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    e = hipGetLastError();
  }
  std::vector<int32_t> choice(n_ctbs);
  std::vector<uint64_t> sse;
  if (sse_out) sse.resize((size_t)n_ctbs * (1 + n_cand));
  if (e == hipSuccess) e = hipMemcpyAsync(choice.data(), dchoice.p, (size_t)n_ctbs * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && sse_out)
    e = hipMemcpyAsync(sse.data(), dsse.p, sse.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = copy_rows(R.samples, (size_t)R.stride * 2, dout.p, (size_t)pitch * 2, (size_t)W * 2, H,
                         hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "alf_decide_luma");
  for (uint32_t i = 0; i < n_ctbs; i++) {
    memset(&out[i], 0, sizeof(out[i]));
    if (choice[i] >= 0) {
      out[i].luma_on = 1;
      out[i].luma_filt_set_idx = (uint8_t)choice[i];
    }
  }
  if (sse_out) memcpy(sse_out, sse.data(), sse.size() * 8);
  return VVCLF_OK;
}

namespace {

// ---------------------------------------------------------------------------
// Clip-index search of a luma APS (design_clip_rdo_for_luma_aps,
// src/alf_aps_design.rs:697-768).  The reference walks filtIdx 0..24 and tap
// j 0..11, replaying the whole picture for clip indices 1, 2, 3 of the tap
// and keeping the first strict improvement of the picture's SSE.  The SSE is
// a sum over classes and a trial of row k changes only class k's part, so
// "trial < best" compares class k's SSE alone and the rows' searches are
// independent: per tap j, candidate bank c (c = 1, 2, 3) holds every row's
// current clip indices with tap j set to c, one k_alf_enc<ENC_CLASS> pass
// gives every (c, class) SSE, and k_clip_pick makes the 25 choices.
// ---------------------------------------------------------------------------
struct clip_tables {
  uint8_t perm[4][12];  // APS tap read at bank position k, per transposeIdx
  uint32_t clipw[4];    // clip word per clip index
};

// banks[b] (b < n_banks): rows with the current clip indices; with j >= 0,
// tap j of every row takes clip index b + 1
__global__ void __launch_bounds__(256) k_clip_banks(const int32_t* coeff, const uint8_t* clip, int j, clip_tables t,
                                                    uint32_t* banks, int n_banks) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_banks * ALF_BANK_WORDS) return;
  const int b = i / ALF_BANK_WORDS, w = i % ALF_BANK_WORDS;
  const int k = w / 48, tr = (w / 12) & 3, q = w % 12;
  const int src = t.perm[tr][q];
  const int idx = (j >= 0 && src == j) ? b + 1 : clip[k * 12 + src];
  banks[i] = ((uint32_t)(uint16_t)(int16_t)coeff[k * 12 + src]) | (t.clipw[idx] << 16);
}

// per class k: with j < 0 the baseline (trial row 0 becomes the current
// SSE); else the first of indices 1..3 whose SSE is strictly below the
// current one, or 0.  Consumed trial rows are cleared for the next pass.
__global__ void k_clip_pick(unsigned long long* trial, unsigned long long* cur, unsigned long long* base,
                            uint8_t* clip, int j) {
  const int k = threadIdx.x;
  if (k >= 25) return;
  if (j < 0) {
    cur[k] = base[k] = trial[k];
    trial[k] = 0;
    return;
  }
  unsigned long long best = cur[k];
  int bc = 0;
  for (int c = 1; c <= 3; c++) {
    const unsigned long long v = trial[(c - 1) * 25 + k];
    if (v < best) {
      best = v;
      bc = c;
    }
    trial[(c - 1) * 25 + k] = 0;
  }
  clip[k * 12 + j] = (uint8_t)bc;
  cur[k] = best;
}

}  // namespace

extern "C" int vvclf_alf_clip_rdo_luma(const vvclf_picture* src, const vvclf_picture* rec,
                                       const int32_t coeff[VVCLF_ALF_NUM_FILTERS][VVCLF_ALF_LUMA_NUM_COEFF],
                                       uint32_t ctb_log2, uint32_t bd, uint32_t cfi, vvclf_alf_clip_rdo* out) {
  clear_error();
  (void)cfi;  // the search filters luma only
  if (!src || !rec || !coeff || !out) return fail(VVCLF_ERR_INVALID, "alf_clip_rdo_luma: NULL argument");
  const vvclf_plane &S = src->luma, &R = rec->luma;
  const uint32_t W = R.width, H = R.height;
  if (!S.samples || !R.samples || S.width != W || S.height != H || W == 0 || H == 0)
    return fail(VVCLF_ERR_INVALID, "alf_clip_rdo_luma: source and reconstruction luma differ");
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "alf_clip_rdo_luma: bit depth %u", bd);
  if (ctb_log2 < 3 || ctb_log2 > 8)
    return fail(VVCLF_ERR_UNSUPPORTED, "alf_clip_rdo_luma: ctb_log2_size_y %u", ctb_log2);
  if (W > 65535 || H > 65535) return fail(VVCLF_ERR_UNSUPPORTED, "alf_clip_rdo_luma: picture above 65535 samples");
  for (int c = 0; c < 25; c++)
    for (int k = 0; k < 12; k++)
      if (coeff[c][k] < -32768 || coeff[c][k] > 32767)
        return fail(VVCLF_ERR_UNSUPPORTED, "alf_clip_rdo_luma: luma coefficient %d outside 16 bits", coeff[c][k]);
  const uint32_t ctb = 1u << ctb_log2, wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  const uint32_t n_ctbs = wctb * hctb;
  std::vector<uint64_t> rects(n_ctbs);
  for (uint32_t ry = 0; ry < hctb; ry++)
    for (uint32_t rx = 0; rx < wctb; rx++) rects[(size_t)ry * wctb + rx] = ctb_fetch_rect(nullptr, rx, ry, ctb, W, H);
  clip_tables tabs;
  alf_luma_bank_tables(bd, tabs.perm, tabs.clipw);
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const uint32_t pitch = (W + 63) / 64 * 64;
  const size_t plane = (size_t)pitch * H * 2;
  const size_t trial_bytes = CLASS_MAX_CAND * 25 * 8;
  dbuf drec, dsrc, drect, dcoeff, dclip, dbank, dtrial, dcur, dbase;
  hipError_t e = drec.alloc(plane, st);
  if (e == hipSuccess) e = dsrc.alloc(plane, st);
  if (e == hipSuccess) e = drect.alloc(rects.size() * 8, st);
  if (e == hipSuccess) e = dcoeff.alloc(25 * 12 * 4, st);
  if (e == hipSuccess) e = dclip.alloc(25 * 12, st);
  if (e == hipSuccess) e = dbank.alloc((size_t)3 * ALF_BANK_WORDS * 4, st);
  if (e == hipSuccess) e = dtrial.alloc(trial_bytes, st);
  if (e == hipSuccess) e = dcur.alloc(25 * 8, st);
  if (e == hipSuccess) e = dbase.alloc(25 * 8, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "alf_clip_rdo_luma: device buffers: %s", hipGetErrorString(e));
  e = copy_rows(drec.p, (size_t)pitch * 2, R.samples, (size_t)R.stride * 2, (size_t)W * 2, H,
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = copy_rows(dsrc.p, (size_t)pitch * 2, S.samples, (size_t)S.stride * 2, (size_t)W * 2, H,
                         hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(drect.p, rects.data(), rects.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dcoeff.p, coeff, 25 * 12 * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(dclip.p, 0, 25 * 12, st);
  if (e == hipSuccess) e = hipMemsetAsync(dtrial.p, 0, trial_bytes, st);
  enc_args a;
  memset(&a, 0, sizeof(a));
  a.rec = (const uint16_t*)drec.p;
  a.src = (const uint16_t*)dsrc.p;
  a.pitch = (int)pitch;
  a.w = (int)W;
  a.h = (int)H;
  a.ctb_log2 = (int)ctb_log2;
  a.wctb = (int)wctb;
  a.hctb = (int)hctb;
  a.bd = (int)bd;
  a.rect = (const uint64_t*)drect.p;
  a.banks = (const uint32_t*)dbank.p;
  a.sse = (unsigned long long*)dtrial.p;
  const alf_tiling tg(W, H, ctb_log2);
  const bool hbd = bd > 12;
  for (int j = -1; j < 12 && e == hipSuccess; j++) {
    const int nb = j < 0 ? 1 : 3;
    a.n_cand = nb;
    hipLaunchKernelGGL(k_clip_banks, dim3((nb * ALF_BANK_WORDS + 255) / 256), dim3(256), 0, st,
                       (const int32_t*)dcoeff.p, (const uint8_t*)dclip.p, j, tabs, (uint32_t*)dbank.p, nb);
    if (hbd) hipLaunchKernelGGL((k_alf_enc<ENC_CLASS, true>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    else hipLaunchKernelGGL((k_alf_enc<ENC_CLASS, false>), tg.grid, dim3(tg.threads), 0, st, a, tg.twl, tg.thl);
    hipLaunchKernelGGL(k_clip_pick, dim3(1), dim3(64), 0, st, (unsigned long long*)dtrial.p,
                       (unsigned long long*)dcur.p, (unsigned long long*)dbase.p, (uint8_t*)dclip.p, j);
    e = hipGetLastError();
  }
  uint8_t clip[25 * 12];
  uint64_t cur[25], base[25];
  if (e == hipSuccess) e = hipMemcpyAsync(clip, dclip.p, sizeof(clip), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(cur, dcur.p, sizeof(cur), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(base, dbase.p, sizeof(base), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "alf_clip_rdo_luma");
  memset(out, 0, sizeof(*out));
  memcpy(out->coeff, coeff, sizeof(out->coeff));
  memcpy(out->clip_idx, clip, sizeof(clip));
  for (int k = 0; k < 25; k++) {
    out->baseline_sse_y += base[k];
    out->post_clip_sse_y += cur[k];
  }
  for (int i = 0; i < 25 * 12; i++)
    if (clip[i]) out->alf_luma_clip_flag = 1;
  return VVCLF_OK;
}

namespace {

// Shared host part of the two chroma decisions.
int chroma_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma, vvclf_alf_ctb* apply,
                  const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2, uint32_t bd, uint32_t cfi,
                  const vvclf_bounds* tb, uint64_t* sse_out, bool cc) {
  const char* name = cc ? "cc_alf_decide" : "alf_decide_chroma";
  if (!src || !rec || !apply || !aps || (cc && !pre_luma)) return fail(VVCLF_ERR_INVALID, "%s: NULL argument", name);
  if (c_idx != 1 && c_idx != 2) return fail(VVCLF_ERR_INVALID, "%s: component %u (1 Cb, 2 Cr)", name, c_idx);
  if (cfi == 0) return VVCLF_OK;  // monochrome: nothing to decide (as the reference)
  if (cfi != 1) return fail(VVCLF_ERR_UNSUPPORTED, "%s: chroma_format_idc %u (4:2:0 only)", name, cfi);
  uint32_t n;
  if (cc) {
    n = c_idx == 1 ? aps->n_cc_cb_coeff : aps->n_cc_cr_coeff;
    if (n > VVCLF_ALF_MAX_CC_FILTERS) return fail(VVCLF_ERR_INVALID, "%s: %u CC-ALF filters", name, n);
  } else {
    if (!aps->alf_chroma_filter_signal_flag) return VVCLF_OK;
    n = aps->n_chroma_coeff;
    if (n > VVCLF_ALF_MAX_CHROMA_ALT) return fail(VVCLF_ERR_INVALID, "%s: %u chroma alternatives", name, n);
  }
  if (n == 0) return VVCLF_OK;
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "%s: bit depth %u", name, bd);
  if (ctb_log2 < 3 || ctb_log2 > 8) return fail(VVCLF_ERR_UNSUPPORTED, "%s: ctb_log2_size_y %u", name, ctb_log2);
  const vvclf_plane &SP = c_idx == 1 ? src->cb : src->cr, &RP = c_idx == 1 ? rec->cb : rec->cr;
  const uint32_t W = rec->luma.width, H = rec->luma.height, Wc = RP.width, Hc = RP.height;
  // 4:2:0 chroma planes of half the luma size, rounded either way
  if (!SP.samples || !RP.samples || SP.width != Wc || SP.height != Hc || (Wc != W / 2 && Wc != (W + 1) / 2) ||
      (Hc != H / 2 && Hc != (H + 1) / 2) || Wc == 0 || Hc == 0)
    return fail(VVCLF_ERR_INVALID, "%s: chroma planes do not match a 4:2:0 %ux%u picture", name, W, H);
  if (W > 65535 || H > 65535) return fail(VVCLF_ERR_UNSUPPORTED, "%s: picture above 65535 samples", name);
  if (cc && (!pre_luma->samples || pre_luma->width != W || pre_luma->height != H || pre_luma->stride < W))
    return fail(VVCLF_ERR_INVALID, "%s: the pre-ALF luma plane does not match the picture", name);
  // coefficient ranges of the filter lanes (as the ALF stage)
  for (uint32_t k = 0; k < n; k++)
    for (int t = 0; t < (cc ? VVCLF_ALF_CC_NUM_COEFF : 6); t++) {
      const int32_t v = cc ? (c_idx == 1 ? aps->cc_cb_coeff[k][t] : aps->cc_cr_coeff[k][t]) : aps->chroma_coeff[k][t];
      if (cc ? (v < -(1 << 20) || v >= (1 << 20)) : (v < -32768 || v > 32767))
        return fail(VVCLF_ERR_UNSUPPORTED, "%s: coefficient %d outside the filter lanes", name, v);
    }
  const uint32_t ctb = 1u << ctb_log2, wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  const uint32_t n_ctbs = wctb * hctb;
  vvclf_alf_config ac;
  memset(&ac, 0, sizeof(ac));
  ac.alf_enabled = 1;
  ac.bit_depth = bd;
  ac.ctb_log2_size_y = ctb_log2;
  ac.chroma_format_idc = cfi;
  vvclf_alf_aps_binding b;
  memset(&b, 0, sizeof(b));
  if (!cc) b.chroma_aps = aps;
  else if (c_idx == 1) b.cc_cb_aps = aps;
  else b.cc_cr_aps = aps;
  host_tables t;
  build_alf_banks(ac, b, t);
  std::vector<uint64_t> rects(n_ctbs);
  for (uint32_t ry = 0; ry < hctb; ry++)
    for (uint32_t rx = 0; rx < wctb; rx++) {
      rects[(size_t)ry * wctb + rx] = ctb_fetch_rect(tb, rx, ry, ctb, W, H);
      if (!rects[(size_t)ry * wctb + rx]) return fail(VVCLF_ERR_INVALID, "%s: empty clip rectangle", name);
    }
  std::vector<uint8_t> ok(n, 1);
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const uint32_t cpitch = (Wc + 63) / 64 * 64, lpitch = (W + 63) / 64 * 64;
  const size_t cplane = (size_t)cpitch * Hc * 2, lplane = cc ? (size_t)lpitch * H * 2 : 0;
  dbuf drec, dsrc, dout, dl, drect, dcb, dccb, dsse, dchoice, dok;
  hipError_t e = drec.alloc(cplane, st);
  if (e == hipSuccess) e = dsrc.alloc(cplane, st);
  if (e == hipSuccess) e = dout.alloc(cplane, st);
  if (e == hipSuccess) e = dl.alloc(lplane, st);
  if (e == hipSuccess) e = drect.alloc(rects.size() * 8, st);
  if (e == hipSuccess) e = dcb.alloc(t.alf_cbank.size() * 4, st);
  if (e == hipSuccess) e = dccb.alloc(t.alf_ccbank.size() * 4, st);
  if (e == hipSuccess) e = dsse.alloc((size_t)n_ctbs * (1 + n) * 8, st);
  if (e == hipSuccess) e = dchoice.alloc((size_t)n_ctbs * 4, st);
  if (e == hipSuccess) e = dok.alloc(n, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "%s: device buffers: %s", name, hipGetErrorString(e));
  e = copy_rows(drec.p, (size_t)cpitch * 2, RP.samples, (size_t)RP.stride * 2, (size_t)Wc * 2, Hc,
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = copy_rows(dsrc.p, (size_t)cpitch * 2, SP.samples, (size_t)SP.stride * 2, (size_t)Wc * 2, Hc,
                         hipMemcpyHostToDevice, st);
  if (e == hipSuccess && cc)
    e = copy_rows(dl.p, (size_t)lpitch * 2, pre_luma->samples, (size_t)pre_luma->stride * 2, (size_t)W * 2, H,
                         hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(drect.p, rects.data(), rects.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dcb.p, t.alf_cbank.data(), t.alf_cbank.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dccb.p, t.alf_ccbank.data(), t.alf_ccbank.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dok.p, ok.data(), n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(dsse.p, 0, (size_t)n_ctbs * (1 + n) * 8, st);
  cenc_args a;
  a.rec = (const uint16_t*)drec.p;
  a.src = (const uint16_t*)dsrc.p;
  a.out = (uint16_t*)dout.p;
  a.luma = (const uint16_t*)dl.p;
  a.cp = (int)cpitch;
  a.lp = (int)lpitch;
  a.wc = (int)Wc;
  a.hc = (int)Hc;
  a.w = (int)W;
  a.h = (int)H;
  a.ctb_log2 = (int)ctb_log2;
  a.wctb = (int)wctb;
  a.bd = (int)bd;
  a.rect = (const uint64_t*)drect.p;
  a.cbank = (const uint32_t*)dcb.p;
  a.ccbank = (const int32_t*)dccb.p + (c_idx - 1) * ALF_CCBANK_WORDS;
  a.n_cand = (int)n;
  a.sse = (unsigned long long*)dsse.p;
  a.choice = (const int32_t*)dchoice.p;
  const int tl = std::min<int>((int)ctb_log2 - 1, 4), T = 1 << tl;
  const dim3 grid((Wc + T - 1) / T, (Hc + T - 1) / T);
  const unsigned threads = std::max(64u, (unsigned)(T * T));
  if (e == hipSuccess) {
    if (cc) hipLaunchKernelGGL((k_chroma_enc<true, true>), grid, dim3(threads), 0, st, a, tl);
    else hipLaunchKernelGGL((k_chroma_enc<true, false>), grid, dim3(threads), 0, st, a, tl);
    hipLaunchKernelGGL(k_alf_pick, dim3((n_ctbs + 255) / 256), dim3(256), 0, st, (unsigned long long*)dsse.p,
                       (const uint8_t*)dok.p, (int)n_ctbs, (int)n, (int32_t*)dchoice.p);
    if (cc) hipLaunchKernelGGL((k_chroma_enc<false, true>), grid, dim3(threads), 0, st, a, tl);
    else hipLaunchKernelGGL((k_chroma_enc<false, false>), grid, dim3(threads), 0, st, a, tl);
    e = hipGetLastError();
  }
  std::vector<int32_t> choice(n_ctbs);
  std::vector<uint64_t> sse;
  if (sse_out) sse.resize((size_t)n_ctbs * (1 + n));
  if (e == hipSuccess) e = hipMemcpyAsync(choice.data(), dchoice.p, (size_t)n_ctbs * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && sse_out) e = hipMemcpyAsync(sse.data(), dsse.p, sse.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = copy_rows(RP.samples, (size_t)RP.stride * 2, dout.p, (size_t)cpitch * 2, (size_t)Wc * 2, Hc,
                         hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, name);
  for (uint32_t i = 0; i < n_ctbs; i++) {
    const int k = choice[i];
    if (k < 0) continue;  // the CTB's record keeps what it held
    if (cc) {
      if (c_idx == 1) apply[i].cc_cb_idc = (uint8_t)(k + 1);
      else apply[i].cc_cr_idc = (uint8_t)(k + 1);
    } else if (c_idx == 1) {
      apply[i].cb_on = 1;
      apply[i].cb_alt_idx = (uint8_t)k;
    } else {
      apply[i].cr_on = 1;
      apply[i].cr_alt_idx = (uint8_t)k;
    }
  }
  if (sse_out) memcpy(sse_out, sse.data(), sse.size() * 8);
  return VVCLF_OK;
}

}  // namespace

extern "C" int vvclf_alf_decide_chroma(const vvclf_picture* src, vvclf_picture* rec, vvclf_alf_ctb* apply,
                                       const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2, uint32_t bd,
                                       uint32_t cfi, const vvclf_bounds* tb, uint64_t* sse) {
  clear_error();
  return chroma_decide(src, rec, nullptr, apply, aps, c_idx, ctb_log2, bd, cfi, tb, sse, false);
}

extern "C" int vvclf_cc_alf_decide(const vvclf_picture* src, vvclf_picture* rec, const vvclf_plane* pre_luma,
                                   vvclf_alf_ctb* apply, const vvclf_alf_aps* aps, uint32_t c_idx, uint32_t ctb_log2,
                                   uint32_t bd, uint32_t cfi, const vvclf_bounds* tb, uint64_t* sse) {
  clear_error();
  return chroma_decide(src, rec, pre_luma, apply, aps, c_idx, ctb_log2, bd, cfi, tb, sse, true);
}

extern "C" int vvclf_sao_decide_picture(const vvclf_picture* src, const vvclf_picture* rec, uint32_t ctb_log2,
                                        uint32_t bd, uint32_t sao_luma, uint32_t sao_chroma,
                                        vvclf_sao_ctb_params* out) {
  clear_error();
  if (!src || !rec || !out) return fail(VVCLF_ERR_INVALID, "sao_decide_picture: NULL argument");
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "sao_decide_picture: bit depth %u", bd);
  if (ctb_log2 < 3 || ctb_log2 > 8)
    return fail(VVCLF_ERR_UNSUPPORTED, "sao_decide_picture: ctb_log2_size_y %u", ctb_log2);
  const uint32_t W = rec->luma.width, H = rec->luma.height;
  if (W == 0 || H == 0 || W > 65535 || H > 65535)
    return fail(VVCLF_ERR_UNSUPPORTED, "sao_decide_picture: picture %ux%u", W, H);
  const vvclf_plane* rp[3] = {&rec->luma, &rec->cb, &rec->cr};
  const vvclf_plane* sp[3] = {&src->luma, &src->cb, &src->cr};
  const bool use[3] = {sao_luma != 0, sao_chroma != 0, sao_chroma != 0};
  for (int c = 0; c < 3; c++) {
    if (!use[c]) continue;
    if (sp[c]->width != rp[c]->width || sp[c]->height != rp[c]->height ||
        (rp[c]->width && (!rp[c]->samples || !sp[c]->samples)))
      return fail(VVCLF_ERR_INVALID, "sao_decide_picture: plane %d of source and reconstruction differ", c);
    if (c && rp[c]->width && ((rp[c]->width != W / 2 && rp[c]->width != (W + 1) / 2) ||
                              (rp[c]->height != H / 2 && rp[c]->height != (H + 1) / 2)))
      return fail(VVCLF_ERR_UNSUPPORTED, "sao_decide_picture: chroma planes other than 4:2:0");
  }
  const uint32_t ctb = 1u << ctb_log2, wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  const uint32_t n_ctbs = wctb * hctb;
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  sao_args a;
  memset(&a, 0, sizeof(a));
  dbuf dr[3], ds[3], dst, dout, dbad;
  hipError_t e = hipSuccess;
  for (int c = 0; c < 3 && e == hipSuccess; c++) {
    const uint32_t pw = use[c] ? rp[c]->width : 0, ph = use[c] ? rp[c]->height : 0;
    const uint32_t pitch = (pw + 63) / 64 * 64;
    a.pitch[c] = (int)pitch;
    a.pw[c] = (int)pw;
    a.ph[c] = (int)ph;
    e = dr[c].alloc((size_t)pitch * ph * 2, st);
    if (e == hipSuccess) e = ds[c].alloc((size_t)pitch * ph * 2, st);
    if (e == hipSuccess && pw && ph)
      e = copy_rows(dr[c].p, (size_t)pitch * 2, rp[c]->samples, (size_t)rp[c]->stride * 2, (size_t)pw * 2, ph,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess && pw && ph)
      e = copy_rows(ds[c].p, (size_t)pitch * 2, sp[c]->samples, (size_t)sp[c]->stride * 2, (size_t)pw * 2, ph,
                           hipMemcpyHostToDevice, st);
    a.rec[c] = (const uint16_t*)dr[c].p;
    a.src[c] = (const uint16_t*)ds[c].p;
  }
  const size_t stat_bytes = (size_t)n_ctbs * 3 * SAO_STAT_WORDS * 8;
  if (e == hipSuccess) e = dst.alloc(stat_bytes, st);
  if (e == hipSuccess) e = dout.alloc((size_t)n_ctbs * sizeof(vvclf_sao_ctb_params), st);
  if (e == hipSuccess) e = dbad.alloc(4, st);
  if (e == hipSuccess) e = hipMemsetAsync(dst.p, 0, stat_bytes, st);
  if (e == hipSuccess) e = hipMemsetAsync(dbad.p, 0, 4, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "sao_decide_picture: device buffers: %s", hipGetErrorString(e));
  a.w = (int)W;
  a.h = (int)H;
  a.ctb_log2 = (int)ctb_log2;
  a.wctb = (int)wctb;
  a.bd = (int)bd;
  a.luma = use[0];
  a.chroma = use[1] && rp[1]->width > 0;
  a.stats = (long long*)dst.p;
  // luma tiles of min(CTB, 32), chroma tiles of half that (inside one CTB)
  const int tl = std::min<int>((int)ctb_log2, 5), T = 1 << tl;
  const dim3 grid((W + T - 1) / T, (H + T - 1) / T, 3);
  hipLaunchKernelGGL(k_sao_stats, grid, dim3(256), 0, st, a, tl);
  hipLaunchKernelGGL(k_sao_decide, dim3((n_ctbs + 255) / 256), dim3(256), 0, st, a, (int)n_ctbs,
                     (vvclf_sao_ctb_params*)dout.p, (unsigned*)dbad.p);
  e = hipGetLastError();
  unsigned bad = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout.p, (size_t)n_ctbs * sizeof(vvclf_sao_ctb_params), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad.p, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "sao_decide_picture");
  if (bad)
    return fail(VVCLF_ERR_INVALID, "sao_decide_picture: a reconstruction sample of 2^%u or more in a band statistic",
                bd);
  return VVCLF_OK;
}

extern "C" int vvclf_sao_chroma_merge(const vvclf_picture* src, const vvclf_picture* rec,
                                      const vvclf_sao_ctb_params* independent, uint32_t bd, uint32_t ctb_log2,
                                      vvclf_sao_ctb_params* merged, uint8_t* merge_map) {
  clear_error();
  if (!src || !rec || !independent || !merged || !merge_map)
    return fail(VVCLF_ERR_INVALID, "sao_chroma_merge: NULL argument");
  if (bd < 8 || bd > 16) return fail(VVCLF_ERR_INVALID, "sao_chroma_merge: bit depth %u", bd);
  if (ctb_log2 < 3 || ctb_log2 > 8) return fail(VVCLF_ERR_UNSUPPORTED, "sao_chroma_merge: ctb_log2_size_y %u", ctb_log2);
  const uint32_t W = rec->luma.width, H = rec->luma.height;
  if (W == 0 || H == 0 || W > 65535 || H > 65535)
    return fail(VVCLF_ERR_UNSUPPORTED, "sao_chroma_merge: picture %ux%u", W, H);
  const vvclf_plane* rp[2] = {&rec->cb, &rec->cr};
  const vvclf_plane* sp[2] = {&src->cb, &src->cr};
  const uint32_t cw = rp[0]->width, ch = rp[0]->height;
  for (int c = 0; c < 2; c++)
    if (rp[c]->width != cw || rp[c]->height != ch || sp[c]->width != cw || sp[c]->height != ch ||
        (cw && (!rp[c]->samples || !sp[c]->samples)))
      return fail(VVCLF_ERR_INVALID, "sao_chroma_merge: chroma planes of source and reconstruction differ");
  if (cw && ((cw != W / 2 && cw != (W + 1) / 2) || (ch != H / 2 && ch != (H + 1) / 2)))
    return fail(VVCLF_ERR_UNSUPPORTED, "sao_chroma_merge: chroma planes other than 4:2:0");
  const uint32_t ctb = 1u << ctb_log2, wctb = (W + ctb - 1) / ctb, hctb = (H + ctb - 1) / ctb;
  const size_t n = (size_t)wctb * hctb;
  vvclf_ctx* ctx = default_ctx();
  if (!ctx) return VVCLF_ERR_HIP;
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const uint32_t pitch = (cw + 63) / 64 * 64;
  dbuf dr[2], ds[2], dp, dm;
  hipError_t e = hipSuccess;
  merge_args a;
  memset(&a, 0, sizeof(a));
  for (int c = 0; c < 2 && e == hipSuccess; c++) {
    e = dr[c].alloc((size_t)pitch * ch * 2, st);
    if (e == hipSuccess) e = ds[c].alloc((size_t)pitch * ch * 2, st);
    if (e == hipSuccess && cw && ch)
      e = copy_rows(dr[c].p, (size_t)pitch * 2, rp[c]->samples, (size_t)rp[c]->stride * 2, (size_t)cw * 2, ch,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess && cw && ch)
      e = copy_rows(ds[c].p, (size_t)pitch * 2, sp[c]->samples, (size_t)sp[c]->stride * 2, (size_t)cw * 2, ch,
                           hipMemcpyHostToDevice, st);
    a.rec[c] = (const uint16_t*)dr[c].p;
    a.src[c] = (const uint16_t*)ds[c].p;
  }
  if (e == hipSuccess) e = dp.alloc(n * sizeof(vvclf_sao_ctb_params), st);
  if (e == hipSuccess) e = dm.alloc(n, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dp.p, independent, n * sizeof(vvclf_sao_ctb_params), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail(VVCLF_ERR_NOMEM, "sao_chroma_merge: device buffers: %s", hipGetErrorString(e));
  a.pitch = (int)pitch;
  a.cw = (int)cw;
  a.ch = (int)ch;
  a.wctb = (int)wctb;
  a.hctb = (int)hctb;
  a.cctb = (int)(ctb / 2);
  a.bd = (int)bd;
  a.merged = (vvclf_sao_ctb_params*)dp.p;
  a.map = (uint8_t*)dm.p;
  // anti-diagonals d = rx + ry in order: each CTB's neighbours are decided
  for (int d = 0; d <= (int)(wctb + hctb) - 2; d++) {
    const int lo = std::max(0, d - ((int)hctb - 1)), hi = std::min(d, (int)wctb - 1);
    hipLaunchKernelGGL(k_sao_merge_step, dim3(hi - lo + 1), dim3(256), 0, st, a, d, lo);
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(merged, dp.p, n * sizeof(vvclf_sao_ctb_params), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(merge_map, dm.p, n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "sao_chroma_merge");
  return VVCLF_OK;
}

