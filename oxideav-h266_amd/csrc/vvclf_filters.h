// vvclf_filters.h — per-line / per-block filter arithmetic shared by the
// plain stage kernels and the fused tile kernel.  All integer, bit-exact
// restatements of the reference sample filters; the callers decide where the
// samples come from (HBM or LDS) and where they go.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vvclf_internal.h"

namespace vvclf {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Loads and stores through a pointer known to address global memory (HBM).
// A pointer read from memory (the job records of the batched kernel) is
// generic to the compiler, which then emits flat accesses; those also count
// against the LDS / scalar-load counter, so that every wait for an LDS read
// or a scalar load would also wait for all HBM traffic in flight.
#define VVCLF_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const VVCLF_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(VVCLF_GLOBAL T*)p = v;
}
// uint2 / uint4 are classes whose copy constructor takes a generic reference,
// so they move as native vectors here
typedef uint32_t vu32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t vu32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 gld(const uint2* p) {
  const vu32x2 v = *(const VVCLF_GLOBAL vu32x2*)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 gld(const uint4* p) {
  const vu32x4 v = *(const VVCLF_GLOBAL vu32x4*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(uint2* p, uint2 v) {
  *(VVCLF_GLOBAL vu32x2*)p = vu32x2{v.x, v.y};
}
__device__ __forceinline__ void gst(uint4* p, uint4 v) {
  *(VVCLF_GLOBAL vu32x4*)p = vu32x4{v.x, v.y, v.z, v.w};
}

// Table 43 (spec data): beta' and tC' indexed by the clamped Q.  beta' is
// evaluated in closed form on the device (0 below 16, Q - 10 up to 28, then
// 2Q - 38), checked against the table at compile time; tC' is read from a
// table (kTc in constant memory, or a copy in LDS).
constexpr int16_t kBetaTab[64] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,
                                  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18, 20, 22, 24,
                                  26, 28, 30, 32, 34, 36, 38, 40, 42, 44, 46, 48, 50, 52, 54, 56,
                                  58, 60, 62, 64, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84, 86, 88};
__host__ __device__ constexpr int beta_prime(int q) { return q < 16 ? 0 : (q <= 28 ? q - 10 : 2 * q - 38); }
constexpr bool beta_closed_form() {
  for (int q = 0; q < 64; q++)
    if (beta_prime(q) != kBetaTab[q]) return false;
  return true;
}
static_assert(beta_closed_form(), "beta' closed form");
__constant__ static const int16_t kTc[66] = {0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   0,   0,   0,
                                            0,  0,  0,  0,  3,  4,   4,   4,   4,   5,   5,   5,   5,   7,
                                            7,  8,  9,  10, 10, 11,  13,  14,  15,  17,  19,  21,  24,  25,
                                            29, 33, 36, 41, 45, 51,  57,  64,  71,  80,  89,  100, 112, 125,
                                            141, 157, 177, 198, 222, 250, 280, 314, 352, 395};
constexpr int TC_TAB_LEN = 66;

// beta and tC of a packed edge descriptor (eqs. 1276-1279, src/deblock.rs:439-452);
// tct: the tC' table (kTc, or its copy in LDS).
__device__ __forceinline__ void edge_thresholds(uint32_t e, int bd, const int16_t* tct, int& beta, int& tc) {
  int qb = (e >> E_QB_SHIFT) & 63, qt = (e >> E_QT_SHIFT) & 127;
  beta = beta_prime(qb) * (1 << (bd - 8));
  int tcp = tct[qt];
  if (bd < 10) {
    tc = (tcp + (1 << (9 - bd))) >> (10 - bd);
  } else {
    tc = tcp * (1 << (bd - 10));
  }
}

// ---------------------------------------------------------------------------
// Luma deblocking, one sample line across the edge: p[0..7] (p0 nearest the
// edge) and q[0..7].  The decision (src/deblock.rs:1224-1409) combines the
// per-line terms of lines 0 and 3 of the 4-line segment.
// ---------------------------------------------------------------------------
struct luma_terms {
  int dp, dq, sp, sq, spq, dpl, dql, sp7, sq7, spm, sqm;
};

__device__ __forceinline__ luma_terms luma_line_terms(const int* p, const int* q, int mflp, int mflq) {
  luma_terms t;
  t.dp = abs(p[2] - 2 * p[1] + p[0]);
  t.dq = abs(q[2] - 2 * q[1] + q[0]);
  t.sp = abs(p[3] - p[0]);
  t.sq = abs(q[0] - q[3]);
  t.spq = abs(p[0] - q[0]);
  t.dpl = abs(p[5] - 2 * p[4] + p[3]);
  t.dql = abs(q[5] - 2 * q[4] + q[3]);
  t.sp7 = abs(p[7] - p[6] - p[5] + p[4]);
  t.sq7 = abs(q[4] - q[5] - q[6] + q[7]);
  // |p3 - p_mflP| of eq. 1369; only consumed when the side is large (mfl 5 or 7).
  t.spm = abs(p[3] - (mflp >= 7 ? p[7] : p[5]));
  t.sqm = abs(q[3] - (mflq >= 7 ? q[7] : q[5]));
  return t;
}

// The terms of line L (0 or 3) of the lane's segment, for segments held by
// aligned quads of lanes (DPP quad_perm [L,L,L,L]; all four lanes of a quad
// are active together).
template <int L>
__device__ __forceinline__ luma_terms quad_terms(const luma_terms& t) {
  constexpr int ctrl = L * 0x55;
  luma_terms r;
  r.dp = __builtin_amdgcn_update_dpp(0, t.dp, ctrl, 0xf, 0xf, false);
  r.dq = __builtin_amdgcn_update_dpp(0, t.dq, ctrl, 0xf, 0xf, false);
  r.sp = __builtin_amdgcn_update_dpp(0, t.sp, ctrl, 0xf, 0xf, false);
  r.sq = __builtin_amdgcn_update_dpp(0, t.sq, ctrl, 0xf, 0xf, false);
  r.spq = __builtin_amdgcn_update_dpp(0, t.spq, ctrl, 0xf, 0xf, false);
  r.dpl = __builtin_amdgcn_update_dpp(0, t.dpl, ctrl, 0xf, 0xf, false);
  r.dql = __builtin_amdgcn_update_dpp(0, t.dql, ctrl, 0xf, 0xf, false);
  r.sp7 = __builtin_amdgcn_update_dpp(0, t.sp7, ctrl, 0xf, 0xf, false);
  r.sq7 = __builtin_amdgcn_update_dpp(0, t.sq7, ctrl, 0xf, 0xf, false);
  r.spm = __builtin_amdgcn_update_dpp(0, t.spm, ctrl, 0xf, 0xf, false);
  r.sqm = __builtin_amdgcn_update_dpp(0, t.sqm, ctrl, 0xf, 0xf, false);
  return r;
}


struct luma_decision {
  int dE, dEp, dEq, mflp, mflq;
};

// luma_edge_decision, src/deblock.rs:1224-1409, over the terms of lines 0 and 3.
__device__ __forceinline__ luma_decision luma_decide(const luma_terms& a, const luma_terms& b, int beta, int tc,
                                                     int mflp, int mflq, bool ctb_row) {
  bool sp_large = mflp > 3, sq_large = mflq > 3;
  if (ctb_row) sp_large = false;
  bool dsam0 = false, dsam3 = false;
  if (sp_large || sq_large) {
    int dp0l, dp3l, dq0l, dq3l;
    if (sp_large) {
      dp0l = (a.dp + a.dpl + 1) >> 1;
      dp3l = (b.dp + b.dpl + 1) >> 1;
    } else {
      mflp = 3;
      dp0l = a.dp;
      dp3l = b.dp;
    }
    if (sq_large) {
      dq0l = (a.dq + a.dql + 1) >> 1;
      dq3l = (b.dq + b.dql + 1) >> 1;
    } else {
      dq0l = a.dq;
      dq3l = b.dq;
    }
    int sp0l = a.sp, sp3l = b.sp, sq0l = a.sq, sq3l = b.sq;
    if (mflp == 7) {
      sp0l += a.sp7;
      sp3l += b.sp7;
    }
    if (mflq == 7) {
      sq0l += a.sq7;
      sq3l += b.sq7;
    }
    int dl = (dp0l + dq0l) + (dp3l + dq3l);
    if (dl < beta) {
      int thr1 = (3 * beta) >> 5, thr2 = beta >> 4, tcl = (5 * tc + 1) >> 1;
      int spin0 = sp_large ? (sp0l + a.spm + 1) >> 1 : sp0l;
      int sqin0 = sq_large ? (sq0l + a.sqm + 1) >> 1 : sq0l;
      int spin3 = sp_large ? (sp3l + b.spm + 1) >> 1 : sp3l;
      int sqin3 = sq_large ? (sq3l + b.sqm + 1) >> 1 : sq3l;
      dsam0 = 2 * (dp0l + dq0l) < thr2 && spin0 + sqin0 < thr1 && a.spq < tcl;
      dsam3 = 2 * (dp3l + dq3l) < thr2 && spin3 + sqin3 < thr1 && b.spq < tcl;
    }
  }
  luma_decision d;
  d.dE = 0;
  d.dEp = 0;
  d.dEq = 0;
  if (dsam0 && dsam3) {
    d.dE = 3;
    d.dEp = 1;
    d.dEq = 1;
    d.mflp = mflp;
    d.mflq = mflq;
    return d;
  }
  int dpq0 = a.dp + a.dq, dpq3 = b.dp + b.dq, dp = a.dp + b.dp, dq = a.dq + b.dq, dd = dpq0 + dpq3;
  bool ssam0 = false, ssam3 = false;
  if (dd < beta && mflp > 2 && mflq > 2) {
    int t1 = beta >> 3, t2 = beta >> 2, tcl = (5 * tc + 1) >> 1;
    ssam0 = 2 * dpq0 < t2 && a.sp + a.sq < t1 && a.spq < tcl;
    ssam3 = 2 * dpq3 < t2 && b.sp + b.sq < t1 && b.spq < tcl;
  }
  if (dd < beta) {
    d.dE = 1;
    if (ssam0 && ssam3) {
      d.dE = 2;
      mflp = 3;
      mflq = 3;
    }
    if (mflp > 1 && mflq > 1) {
      int thr = (beta + (beta >> 1)) >> 3;
      if (dp < thr) d.dEp = 1;
      if (dq < thr) d.dEq = 1;
    }
    if (d.dE == 1) {
      mflp = 1 + d.dEp;
      mflq = 1 + d.dEq;
    }
  }
  d.mflp = mflp;
  d.mflq = mflq;
  return d;
}

// Filters one line in registers: np[]/nq[] receive the new values, `nwp` /
// `nwq` how many samples per side changed (short: src/deblock.rs:1680-1776,
// long: :1519-1673).  Palette sides are handled by the caller (write skip).
__device__ __forceinline__ void luma_filter_line(const int* p, const int* q, const luma_decision& d, int tc,
                                                 int* np, int* nq, int& nwp, int& nwq) {
  nwp = 0;
  nwq = 0;
  if (d.dE == 3) {
    int mp = d.mflp, mq = d.mflq;
    int rm;
    if (mp == 5 && mq == 5) {
      rm = (p[4] + p[3] + 2 * (p[2] + p[1] + p[0] + q[0] + q[1] + q[2]) + q[3] + q[4] + 8) >> 4;
    } else if (mp == mq) {
      // mp == mq here is 7 or 3; eq. 1390 reads up to p6/q6, which the
      // reference leaves at 0 beyond the fetched depth (mfl + 1 samples).
      int p4 = mp >= 4 ? p[4] : 0, p5 = mp >= 5 ? p[5] : 0, p6 = mp >= 6 ? p[6] : 0;
      int q4 = mq >= 4 ? q[4] : 0, q5 = mq >= 5 ? q[5] : 0, q6 = mq >= 6 ? q[6] : 0;
      rm = (p6 + p5 + p4 + p[3] + p[2] + p[1] + 2 * (p[0] + q[0]) + q[1] + q[2] + q[3] + q4 + q5 + q6 + 8) >> 4;
    } else if ((mp == 7 && mq == 5) || (mp == 5 && mq == 7)) {
      rm = (p[5] + p[4] + p[3] + p[2] + 2 * (p[1] + p[0] + q[0] + q[1]) + q[2] + q[3] + q[4] + q[5] + 8) >> 4;
    } else if ((mp == 5 && mq == 3) || (mp == 3 && mq == 5)) {
      rm = (p[3] + p[2] + p[1] + p[0] + q[0] + q[1] + q[2] + q[3] + 4) >> 3;
    } else if (mp == 3 && mq == 7) {
      rm = (2 * (p[2] + p[1] + p[0] + q[0]) + p[0] + p[1] + q[1] + q[2] + q[3] + q[4] + q[5] + q[6] + 8) >> 4;
    } else {
      // eq. 1394 arm; p6..p1 beyond the fetched depth read as 0.
      int pp[8];
#pragma unroll
      for (int i = 0; i < 8; i++) pp[i] = i <= mp ? p[i] : 0;
      int qq[3];
#pragma unroll
      for (int j = 0; j < 3; j++) qq[j] = j <= mq ? q[j] : 0;
      rm = (pp[6] + pp[5] + pp[4] + pp[3] + pp[2] + pp[1] + 2 * (qq[2] + qq[1] + qq[0] + pp[0]) + qq[0] + qq[1] + 8) >>
           4;
    }
    // refP / refQ (eqs. 1395/1396) without a runtime array index
    int refp = (p[1] + p[0] + 1) >> 1, refq = (q[1] + q[0] + 1) >> 1;
#pragma unroll
    for (int k = 2; k < 8; k++) {
      if (mp == k) refp = (p[k] + p[k - 1] + 1) >> 1;
      if (mq == k) refq = (q[k] + q[k - 1] + 1) >> 1;
    }
    const int F7[7] = {59, 50, 41, 32, 23, 14, 5}, F5[5] = {58, 45, 32, 19, 6}, F3[3] = {53, 32, 11};
    const int T7[7] = {6, 5, 4, 3, 2, 1, 1}, T5[5] = {6, 5, 4, 3, 2}, T3[3] = {6, 4, 2};
#pragma unroll
    for (int i = 0; i < 7; i++) {
      if (i < mp) {
        int f = mp == 7 ? F7[i] : (mp == 5 ? F5[i < 5 ? i : 0] : F3[i < 3 ? i : 0]);
        int t = mp == 7 ? T7[i] : (mp == 5 ? T5[i < 5 ? i : 0] : T3[i < 3 ? i : 0]);
        int lo = p[i] - ((tc * t) >> 1), hi = p[i] + ((tc * t) >> 1);
        np[i] = clampi((rm * f + refp * (64 - f) + 32) >> 6, lo, hi);
      }
      if (i < mq) {
        int g = mq == 7 ? F7[i] : (mq == 5 ? F5[i < 5 ? i : 0] : F3[i < 3 ? i : 0]);
        int t = mq == 7 ? T7[i] : (mq == 5 ? T5[i < 5 ? i : 0] : T3[i < 3 ? i : 0]);
        int lo = q[i] - ((tc * t) >> 1), hi = q[i] + ((tc * t) >> 1);
        nq[i] = clampi((rm * g + refq * (64 - g) + 32) >> 6, lo, hi);
      }
    }
    nwp = mp;
    nwq = mq;
  } else if (d.dE == 2) {
    int p3 = p[3], p2 = p[2], p1 = p[1], p0 = p[0], q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    np[0] = clampi((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3, p0 - 3 * tc, p0 + 3 * tc);
    np[1] = clampi((p2 + p1 + p0 + q0 + 2) >> 2, p1 - 2 * tc, p1 + 2 * tc);
    np[2] = clampi((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - tc, p2 + tc);
    nq[0] = clampi((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3, q0 - 3 * tc, q0 + 3 * tc);
    nq[1] = clampi((p0 + q0 + q1 + q2 + 2) >> 2, q1 - 2 * tc, q1 + 2 * tc);
    nq[2] = clampi((p0 + q0 + q1 + 3 * q2 + 2 * q3 + 4) >> 3, q2 - tc, q2 + tc);
    nwp = 3;
    nwq = 3;
  } else if (d.dE == 1) {
    int p2 = p[2], p1 = p[1], p0 = p[0], q0 = q[0], q1 = q[1], q2 = q[2];
    int delta_raw = (9 * (q0 - p0) - 3 * (q1 - p1) + 8) >> 4;
    if (abs(delta_raw) < tc * 10) {
      int delta = clampi(delta_raw, -tc, tc);
      np[0] = p0 + delta;
      nq[0] = q0 - delta;
      nwp = 1;
      nwq = 1;
      if (d.dEp == 1) {
        np[1] = p1 + clampi((((p2 + p0 + 1) >> 1) - p1 + delta) >> 1, -(tc >> 1), tc >> 1);
        nwp = 2;
      }
      if (d.dEq == 1) {
        nq[1] = q1 + clampi((((q2 + q0 + 1) >> 1) - q1 - delta) >> 1, -(tc >> 1), tc >> 1);
        nwq = 2;
      }
    }
  }
}

// Two samples (low and high half of v, each below 1024) through a 1024-entry
// table.
__device__ __forceinline__ uint32_t lut_pair(const uint16_t* lut, uint32_t v) {
  return (uint32_t)lut[v & 0x3ffu] | ((uint32_t)lut[(v >> 16) & 0x3ffu] << 16);
}

// ---------------------------------------------------------------------------
// Chroma deblocking, one line: c[0..7] = p3 p2 p1 p0 q0 q1 q2 q3.
// ---------------------------------------------------------------------------
// chroma_strong_decision_v/h, src/deblock.rs:1932-1982
__device__ __forceinline__ bool chroma_strong_row(const int* c, int beta, int tc, bool short_p) {
  int p1 = c[2];
  int p3 = short_p ? p1 : c[0];
  int p2 = short_p ? p1 : c[1];
  int p0 = c[3], q0 = c[4], q1 = c[5], q2 = c[6], q3 = c[7];
  int dpq = 2 * (abs(p2 - 2 * p1 + p0) + abs(q2 - 2 * q1 + q0));
  int edge = abs(p3 - p0) + abs(q0 - q3);
  return dpq < (beta >> 2) && edge < (beta >> 3) && abs(p0 - q0) < ((5 * tc + 1) >> 1);
}

// Strong / weak chroma line filter (src/deblock.rs:1803-1926, :1986-2074).
// n[] receives new values at the same indices as c[]; the returned mask has
// bit i set when sample i is written.
__device__ __forceinline__ uint32_t chroma_filter_line(const int* c, bool strong, bool short_p, int tc, int* n) {
  int p3 = c[0], p2 = c[1], p1 = c[2], p0 = c[3], q0 = c[4], q1 = c[5], q2 = c[6], q3 = c[7];
#pragma unroll
  for (int i = 0; i < 8; i++) n[i] = c[i];
  if (strong) {
    if (short_p) {
      n[3] = clampi((3 * p1 + 2 * p0 + q0 + q1 + q2 + 4) >> 3, p0 - tc, p0 + tc);
      n[4] = clampi((2 * p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4) >> 3, q0 - tc, q0 + tc);
      n[5] = clampi((p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4) >> 3, q1 - tc, q1 + tc);
      n[6] = clampi((p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4) >> 3, q2 - tc, q2 + tc);
      return 0x78u;
    }
    n[3] = clampi((p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2 + 4) >> 3, p0 - tc, p0 + tc);
    n[2] = clampi((2 * p3 + p2 + 2 * p1 + p0 + q0 + q1 + 4) >> 3, p1 - tc, p1 + tc);
    n[1] = clampi((3 * p3 + 2 * p2 + p1 + p0 + q0 + 4) >> 3, p2 - tc, p2 + tc);
    n[4] = clampi((p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4) >> 3, q0 - tc, q0 + tc);
    n[5] = clampi((p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4) >> 3, q1 - tc, q1 + tc);
    n[6] = clampi((p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4) >> 3, q2 - tc, q2 + tc);
    return 0x7eu;
  }
  int delta = clampi((((q0 - p0) * 4) + p1 - q1 + 4) >> 3, -tc, tc);
  n[3] = p0 + delta;
  n[4] = q0 - delta;
  return 0x18u;
}

// ---------------------------------------------------------------------------
// SAO (src/sao.rs:504-626) for one sample, given the pre-SAO sample and its
// two EO neighbours (ignored for band offset).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sao_offset(int off_sel, uint64_t prm) {
  // offsets are 14-bit two's complement fields
  int raw = (int)((prm >> (8 + 14 * (off_sel - 1))) & 0x3fff);
  return (raw ^ 0x2000) - 0x2000;
}

__device__ __forceinline__ int sao_edge_sample(int p, int n0, int n1, uint64_t prm, int maxv) {
  int e = 2 + ((p > n0) - (p < n0)) + ((p > n1) - (p < n1));
  if (e <= 2) e = e == 2 ? 0 : e + 1;
  int off = e >= 1 ? sao_offset(e, prm) : 0;
  return clampi(p + off, 0, maxv);
}

__device__ __forceinline__ int sao_band_sample(int p, uint64_t prm, int bd, int maxv) {
  int pos = (int)((prm >> 2) & 31);
  int band = bd >= 5 ? ((p >> (bd - 5)) & 31) : (p & 31);
  int k = (band - pos) & 31;
  int off = k < 4 ? sao_offset(k + 1, prm) : 0;
  return clampi(p + off, 0, maxv);
}

// ---------------------------------------------------------------------------
// SAO over runs of 8 samples sharing one CTB (src/sao.rs:504-626), two
// samples per 32-bit word (low half first).  The four offsets of a CTB
// component are folded into a table of 8 bytes, each offset T biased to
// T + 128 and unused entries 128 (offset 0): edge offset byte e for the raw
// index e = 2 + sign(p-n0) + sign(p-n1) (remapped 0,1,2,3,4 -> offsets
// 1,2,-,3,4, index 2 adds nothing), band offset byte k for k = band -
// band_position in 0..3 (k >= 4: byte 4).  One v_perm_b32 looks the byte up
// for both samples of a word.  Valid when every offset fits in i8 (10-bit:
// |offset| <= 31; 12-bit: <= 124) and the SAO bit depth is at most 15;
// otherwise the per-sample path above is used.
//
// The packed steps saturate at 0 and 65535 where the integer ones do not:
// sign(p - n) + 1 is min(sat(sat(p + 1) - n), 2) and the output
// min(sat(sat(p + T') - 128), maxv).  They differ from the integer result
// only for p = 65535 or p + T' > 65535, i.e. p > 65280, whose output is maxv
// either way when maxv <= 32767.
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

struct sao_run {
  int type;       // 0 off, 1 band, 2 edge
  int aux;        // edge class, or band position
  uint64_t btab;  // biased offset bytes (see above)
  bool fits;      // all offsets fit in i8
};

__device__ __forceinline__ sao_run sao_run_of(uint64_t prm) {
  sao_run r;
  r.type = (int)(prm & 3);
  r.aux = (int)((prm >> 2) & 31);
  const int o1 = sao_offset(1, prm), o2 = sao_offset(2, prm), o3 = sao_offset(3, prm), o4 = sao_offset(4, prm);
  r.fits = max(max(abs(o1), abs(o2)), max(abs(o3), abs(o4))) <= 127;
  auto b = [](int v, int k) -> uint64_t { return (uint64_t)(uint8_t)(v + 128) << (8 * k); };
  const uint64_t unused = 0x8080808080808080ull;
  r.btab = r.type == 2 ? ((unused & ~0xffff00ffffull) | b(o1, 0) | b(o2, 1) | b(o3, 3) | b(o4, 4))
                       : ((unused & ~0xffffffffull) | b(o1, 0) | b(o2, 1) | b(o3, 2) | b(o4, 3));
  return r;
}

// Eight samples x8..x8+7 of a row, packed, with the samples x8-1 (l) and
// x8+8 (r) in the low 16 bits of their words.
struct sao_row8 {
  uint32_t w[4];
  uint32_t l, r;
};

// Samples 2k-1, 2k and 2k+1, 2k+2 of a run as packed words.
__device__ __forceinline__ uint32_t pk_prev(const sao_row8& R, int k) {
  return k == 0 ? ((R.w[0] << 16) | R.l) : __builtin_amdgcn_alignbit(R.w[k], R.w[k - 1], 16);
}
__device__ __forceinline__ uint32_t pk_next(const sao_row8& R, int k) {
  return __builtin_amdgcn_alignbit(k == 3 ? R.r : R.w[k + 1], R.w[k], 16);
}

// Both samples of p plus their biased table bytes idx (per half, <= 7),
// clipped to [0, maxv]; maxv2 holds maxv in both halves.
__device__ __forceinline__ uint32_t sao_pk_add(uint32_t p, uint32_t idx, uint64_t btab, uint32_t maxv2) {
  const uint32_t t = __builtin_amdgcn_perm((uint32_t)(btab >> 32), (uint32_t)btab, idx | 0x0c000c00u);
  const u16x2 x = __builtin_elementwise_sub_sat(__builtin_elementwise_add_sat(as_u16x2(p), as_u16x2(t)), (u16x2)(128));
  return as_u32(__builtin_elementwise_min(x, as_u16x2(maxv2)));
}

// Edge offset of a run: A, P, B are rows y-1, y, y+1.  Neighbours by class
// (Table 44).
template <int CLS>
__device__ __forceinline__ void sao_eo8(const sao_row8& A, const sao_row8& P, const sao_row8& B, uint64_t btab,
                                        uint32_t maxv2, uint32_t (&out)[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t p = P.w[k];
    const uint32_t n0 = CLS == 0 ? pk_prev(P, k) : (CLS == 1 ? A.w[k] : (CLS == 2 ? pk_prev(A, k) : pk_next(A, k)));
    const uint32_t n1 = CLS == 0 ? pk_next(P, k) : (CLS == 1 ? B.w[k] : (CLS == 2 ? pk_next(B, k) : pk_prev(B, k)));
    const u16x2 pp1 = __builtin_elementwise_add_sat(as_u16x2(p), (u16x2)(1));
    const u16x2 c0 = __builtin_elementwise_min(__builtin_elementwise_sub_sat(pp1, as_u16x2(n0)), (u16x2)(2));
    const u16x2 c1 = __builtin_elementwise_min(__builtin_elementwise_sub_sat(pp1, as_u16x2(n1)), (u16x2)(2));
    out[k] = sao_pk_add(p, as_u32(c0) + as_u32(c1), btab, maxv2);
  }
}

__device__ __forceinline__ void sao_eo8_any(int cls, const sao_row8& A, const sao_row8& P, const sao_row8& B,
                                            uint64_t btab, uint32_t maxv2, uint32_t (&out)[4]) {
  switch (cls & 3) {
    case 0: sao_eo8<0>(A, P, B, btab, maxv2, out); break;
    case 1: sao_eo8<1>(A, P, B, btab, maxv2, out); break;
    case 2: sao_eo8<2>(A, P, B, btab, maxv2, out); break;
    default: sao_eo8<3>(A, P, B, btab, maxv2, out); break;
  }
}

// Band offset of a run.
__device__ __forceinline__ void sao_bo8(const sao_row8& P, int pos, uint64_t btab, int bd, uint32_t maxv2,
                                        uint32_t (&out)[4]) {
  const unsigned short sh = (unsigned short)(bd >= 5 ? bd - 5 : 0);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const u16x2 band = as_u16x2(P.w[k]) >> (u16x2)(sh);
    const u16x2 kk = __builtin_elementwise_min((band - (u16x2)((unsigned short)pos)) & (u16x2)(31), (u16x2)(4));
    out[k] = sao_pk_add(P.w[k], as_u32(kk), btab, maxv2);
  }
}

// Sample i (-1..8) of a run as an integer.
__device__ __forceinline__ int sao_at(const sao_row8& R, int i) {
  return i < 0 ? (int)R.l : (i > 7 ? (int)R.r : (int)((R.w[i >> 1] >> (16 * (i & 1))) & 0xffffu));
}

// Sample i of o replaced by v.
__device__ __forceinline__ void sao_put(uint32_t (&o)[4], int i, int v) {
  o[i >> 1] = (i & 1) ? ((o[i >> 1] & 0xffffu) | ((uint32_t)v << 16)) : ((o[i >> 1] & 0xffff0000u) | (uint32_t)v);
}

// Edge offset of a run sample by sample, for any class (dx0, dy0, dx1, dy1:
// Table 44 neighbour offsets) and any offsets (prm), each neighbour checked
// against [cx0, cx1) x [cy0, cy1): a sample with a neighbour outside is only
// clipped.  x8, y: position of the run's first sample.
__device__ __forceinline__ void sao_eo8_checked(int dx0, int dy0, int dx1, int dy1, const sao_row8& A,
                                                const sao_row8& P, const sao_row8& B, int x8, int y, int cx0,
                                                int cy0, int cx1, int cy1, uint64_t prm, int maxv, uint32_t (&o)[4]) {
  uint32_t n0w[4], n1w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const sao_row8& R0 = dy0 == 0 ? P : A;
    const sao_row8& R1 = dy1 == 0 ? P : B;
    n0w[k] = dx0 < 0 ? pk_prev(R0, k) : (dx0 > 0 ? pk_next(R0, k) : R0.w[k]);
    n1w[k] = dx1 < 0 ? pk_prev(R1, k) : (dx1 > 0 ? pk_next(R1, k) : R1.w[k]);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int sh = 16 * (i & 1);
    const int pv = (int)((P.w[i >> 1] >> sh) & 0xffffu), n0 = (int)((n0w[i >> 1] >> sh) & 0xffffu),
              n1 = (int)((n1w[i >> 1] >> sh) & 0xffffu);
    const int x = x8 + i, nx0 = x + dx0, ny0 = y + dy0, nx1 = x + dx1, ny1 = y + dy1;
    const bool outside = nx0 < cx0 || ny0 < cy0 || nx1 < cx0 || ny1 < cy0 || nx0 >= cx1 || ny0 >= cy1 ||
                         nx1 >= cx1 || ny1 >= cy1;
    sao_put(o, i, outside ? clampi(pv, 0, maxv) : sao_edge_sample(pv, n0, n1, prm, maxv));
  }
}

// Table 44 neighbour offsets per EO class.
__device__ __forceinline__ void eo_offsets(int cls, int& h0, int& v0, int& h1, int& v1) {
  switch (cls & 3) {
    case 0: h0 = -1; v0 = 0; h1 = 1; v1 = 0; break;
    case 1: h0 = 0; v0 = -1; h1 = 0; v1 = 1; break;
    case 2: h0 = -1; v0 = -1; h1 = 1; v1 = 1; break;
    default: h0 = 1; v0 = -1; h1 = -1; v1 = 1; break;
  }
}

// ---------------------------------------------------------------------------
// ALF helpers
// ---------------------------------------------------------------------------
// table_45, src/alf.rs:1117-1130 (applyAlfLineBufBoundary = 1): returns
// shift and the three vertical offsets.
__device__ __forceinline__ void alf_t45(int j, int ctb, int& sh, int& y1, int& y2, int& y3) {
  sh = 7; y1 = 1; y2 = 2; y3 = 3;
  if (j == ctb - 5 || j == ctb - 4) { sh = 10; y1 = 0; y2 = 0; y3 = 0; }
  else if (j == ctb - 6 || j == ctb - 3) { y2 = 1; y3 = 1; }
  else if (j == ctb - 7 || j == ctb - 2) { y3 = 2; }
}

// table_46, src/alf.rs:1133-1143
__device__ __forceinline__ void alf_t46(int j, int h, int& sh, int& y1, int& y2) {
  sh = 7; y1 = 1; y2 = 2;
  if (j == h - 2 || j == h - 3) { sh = 10; y1 = 0; y2 = 0; }
  else if (j == h - 1 || j == h - 4) { y2 = 1; }
}

// cc_alf_table_47, src/alf.rs:528-539
__device__ __forceinline__ void alf_t47(int yl, int ctb, bool lbb, int& y1, int& y2) {
  y1 = 1; y2 = 2;
  if (lbb) {
    if (yl == ctb - 5 || yl == ctb - 4) { y1 = 0; y2 = 0; }
    else if (yl == ctb - 6 || yl == ctb - 3) { y1 = 1; y2 = 1; }
  }
}

__device__ __forceinline__ int lo16(uint32_t w) { return (int)(w & 0xffffu); }
__device__ __forceinline__ int hi16(uint32_t w) { return (int)(w >> 16); }

// The seven CC-ALF taps of one filter (src/alf.rs:447-523, eqs. 1513-1518).
// The tap sum  sum_k f[k] * (L_k - cen)  is evaluated as
// sum_k f[k] * L_k - cen * sum_k f[k]: the two agree modulo 2^32, the width
// the sum is kept in, and the second needs no per-tap difference.  With
// |f[k]| < 2^20 (checked on the host) every factor fits the signed 24-bit
// multiplier: samples are below 2^16 and |sum_k f[k]| < 7 * 2^20.
struct cc_taps {
  int f[7], fsum;
  __device__ __forceinline__ explicit cc_taps(const int32_t* cf) {
    fsum = 0;
#pragma unroll
    for (int t = 0; t < 7; t++) {
      f[t] = cf[t];
      fsum += f[t];
    }
  }
  // with the tap sum already known
  __device__ __forceinline__ cc_taps(const int32_t* cf, int sum_of_taps) : fsum(sum_of_taps) {
#pragma unroll
    for (int t = 0; t < 7; t++) f[t] = cf[t];
  }
  // taps in eq. 1513 order: (0,-yP1), (-1,0), (1,0), (-1,yP1), (0,yP1), (1,yP1), (0,yP2)
  __device__ __forceinline__ int sum(int a, int l, int r, int bm, int b0, int bp, int d, int cen) const {
    const uint32_t s = (uint32_t)__mul24(f[0], a) + (uint32_t)__mul24(f[1], l) + (uint32_t)__mul24(f[2], r) +
                       (uint32_t)__mul24(f[3], bm) + (uint32_t)__mul24(f[4], b0) + (uint32_t)__mul24(f[5], bp) +
                       (uint32_t)__mul24(f[6], d);
    return (int)(s - (uint32_t)__mul24(fsum, cen));
  }
};

// One clipped tap pair: f * (clamp(a - c, -k, k) + clamp(b - c, -k, k)).
__device__ __forceinline__ int alf_pair(uint32_t w, int a, int b, int c) {
  int f = (int)(int16_t)(w & 0xffff);
  int k = (int)(w >> 16);
  return f * (clampi(a - c, -k, k) + clampi(b - c, -k, k));
}

// The activity and transpose tables of the classification (src/alf.rs:822-1010)
// as register constants: entry i of VAR_TAB in bits 4i..4i+3 of ALF_VAR_NIB,
// entry i of TRANSPOSE in bits 2i..2i+1 of ALF_TRANSPOSE_2B (constant arrays
// would be global loads, and their waits would also wait for stores in
// flight).
constexpr uint8_t ALF_VAR_TAB[16] = {0, 1, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 4};
constexpr uint8_t ALF_TRANSPOSE[8] = {0, 1, 0, 2, 2, 3, 1, 3};
constexpr uint64_t ALF_VAR_NIB = 0x4333333332222210ull;
constexpr uint32_t ALF_TRANSPOSE_2B = 0xde84u;
constexpr bool alf_tables_packed() {
  for (int i = 0; i < 16; i++)
    if (((ALF_VAR_NIB >> (4 * i)) & 0xf) != ALF_VAR_TAB[i]) return false;
  for (int i = 0; i < 8; i++)
    if (((ALF_TRANSPOSE_2B >> (2 * i)) & 3) != ALF_TRANSPOSE[i]) return false;
  return true;
}
static_assert(alf_tables_packed(), "packed ALF classification tables");

// 4x4-block classification (src/alf.rs:822-1010) from the Laplacian sums.
__device__ __forceinline__ int alf_class_from_sums(int sh, int sv, int sd0, int sd1, int ac, int bd) {
  int hv1, hv0, dir_hv, d1, d0, dir_d;
  if (sv > sh) { hv1 = sv; hv0 = sh; dir_hv = 1; } else { hv1 = sh; hv0 = sv; dir_hv = 3; }
  if (sd0 > sd1) { d1 = sd0; d0 = sd1; dir_d = 0; } else { d1 = sd1; d0 = sd0; dir_d = 2; }
  bool prefer_d = (int64_t)d1 * hv0 > (int64_t)hv1 * d0;
  int hvd1 = prefer_d ? d1 : hv1, hvd0 = prefer_d ? d0 : hv0;
  int dir1 = prefer_d ? dir_d : dir_hv, dir2 = prefer_d ? dir_hv : dir_d;
  int dir_s = hvd1 * 2 > 9 * hvd0 ? 2 : (hvd1 > 2 * hvd0 ? 1 : 0);
  int64_t prod = (int64_t)(sh + sv) * ac;
  int64_t shifted = prod >> (bd - 1);
  shifted = shifted < 0 ? 0 : (shifted > 15 ? 15 : shifted);
  int filt = (int)((ALF_VAR_NIB >> (4 * (int)shifted)) & 0xf);
  int tr = (int)((ALF_TRANSPOSE_2B >> (2 * (dir1 * 2 + (dir2 >> 1)))) & 3);
  if (dir_s != 0) filt += (((dir1 & 1) << 1) + dir_s) * 5;
  return filt | (tr << 5);
}

}  // namespace vvclf

namespace vvclf {

// ---------------------------------------------------------------------------
// ALF luma for one 4x4 block from a 10x10 register window
// W[r][c] = sample(x4 - 3 + c, y4 - 3 + r), already clamped to the fetch
// rectangle.  VB selects the block's position against the ALF virtual
// boundary (4 rows above the CTB bottom):
//   0: away from it; 1: block at CTB row CtbSizeY - 8; 2: at CtbSizeY - 4.
// Classification (src/alf.rs:822-1010): the VB padding of row reads becomes a
// clamp of the window row index (VB 1: rows > y4+3 read y4+3; VB 2: rows <
// y4 read y4); CARVE is carve_out_active, which shortens the sum window.
// Filtering (src/alf.rs:652-724): the Table 45 vertical offsets are
// (min(1,m), min(2,m), min(3,m)) with m per block row; shift 10 where m = 0.
// All row / column indices are compile-time constants after unrolling.
// ---------------------------------------------------------------------------
template <int VB, bool CARVE>
__device__ __forceinline__ int alf_block_class(const int (&W)[10][10], int bd) {
  constexpr int jmin = (VB == 2 && CARVE) ? 0 : -2;
  constexpr int jmax = (VB == 1 && CARVE) ? 3 : 5;
  constexpr int ac = ((VB == 1 || VB == 2) && CARVE) ? 3 : 2;
  int sh = 0, sv = 0, sd0 = 0, sd1 = 0;
#pragma unroll
  for (int jj = jmin; jj <= jmax; jj++) {
#pragma unroll
    for (int ii = -2; ii <= 5; ii++) {
      if (((ii + 4) & 1) != ((jj + 4) & 1)) continue;
      // window row indices of rows y4+jj, y4+jj-1, y4+jj+1 after VB padding
      int rc = jj + 3, ru = jj + 2, rd = jj + 4;
      if (VB == 1) { rc = rc > 6 ? 6 : rc; ru = ru > 6 ? 6 : ru; rd = rd > 6 ? 6 : rd; }
      if (VB == 2) { rc = rc < 3 ? 3 : rc; ru = ru < 3 ? 3 : ru; rd = rd < 3 ? 3 : rd; }
      const int cc = ii + 3;
      const int c2 = W[rc][cc] * 2;
      sh += abs(c2 - W[rc][cc - 1] - W[rc][cc + 1]);
      sv += abs(c2 - W[ru][cc] - W[rd][cc]);
      sd0 += abs(c2 - W[ru][cc - 1] - W[rd][cc + 1]);
      sd1 += abs(c2 - W[ru][cc + 1] - W[rd][cc - 1]);
    }
  }
  return alf_class_from_sums(sh, sv, sd0, sd1, ac, bd);
}

template <int VB>
__device__ __forceinline__ void alf_block_filter(const int (&W)[10][10], const uint32_t (&fw)[12], int maxv,
                                                 int (&out)[4][4]) {
  constexpr int MT[3][4] = {{3, 3, 3, 3}, {3, 2, 1, 0}, {0, 1, 2, 3}};
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int m = MT[VB][r];
    const int y1 = m < 1 ? m : 1, y2 = m < 2 ? m : 2, y3 = m < 3 ? m : 3;
    const int sft = m == 0 ? 10 : 7;
    const int R = r + 3;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int C = c + 3;
      const int cur = W[R][C];
      int sum = alf_pair(fw[0], W[R + y3][C], W[R - y3][C], cur);
      sum += alf_pair(fw[1], W[R + y2][C + 1], W[R - y2][C - 1], cur);
      sum += alf_pair(fw[2], W[R + y2][C], W[R - y2][C], cur);
      sum += alf_pair(fw[3], W[R + y2][C - 1], W[R - y2][C + 1], cur);
      sum += alf_pair(fw[4], W[R + y1][C + 2], W[R - y1][C - 2], cur);
      sum += alf_pair(fw[5], W[R + y1][C + 1], W[R - y1][C - 1], cur);
      sum += alf_pair(fw[6], W[R + y1][C], W[R - y1][C], cur);
      sum += alf_pair(fw[7], W[R + y1][C - 1], W[R - y1][C + 1], cur);
      sum += alf_pair(fw[8], W[R + y1][C - 2], W[R - y1][C + 2], cur);
      sum += alf_pair(fw[9], W[R][C + 3], W[R][C - 3], cur);
      sum += alf_pair(fw[10], W[R][C + 2], W[R][C - 2], cur);
      sum += alf_pair(fw[11], W[R][C + 1], W[R][C - 1], cur);
      out[r][c] = clampi(cur + ((sum + (1 << (sft - 1))) >> sft), 0, maxv);
    }
  }
}

// Classification + filter of one block: picks the VB / carve variant.
// `bank` holds the CTB's 25 x 4 x 12 tap words; yb = y4 - yCtb.
__device__ __forceinline__ void alf_block(const int (&W)[10][10], const uint32_t* bank, int yb, int ctb, bool carve,
                                          int bd, int maxv, int (&out)[4][4]) {
  int cls;
  int vb = yb == ctb - 8 ? 1 : (yb == ctb - 4 ? 2 : 0);
  if (vb == 0) cls = alf_block_class<0, false>(W, bd);
  else if (vb == 1) cls = carve ? alf_block_class<1, true>(W, bd) : alf_block_class<1, false>(W, bd);
  else cls = carve ? alf_block_class<2, true>(W, bd) : alf_block_class<2, false>(W, bd);
  const uint32_t* f = bank + ((cls & 31) * 4 + (cls >> 5)) * 12;
  uint32_t fw[12];
#pragma unroll
  for (int k = 0; k < 12; k++) fw[k] = f[k];
  if (vb == 0) alf_block_filter<0>(W, fw, maxv, out);
  else if (vb == 1) alf_block_filter<1>(W, fw, maxv, out);
  else alf_block_filter<2>(W, fw, maxv, out);
}

// ---------------------------------------------------------------------------
// Packed ALF luma block, for sample bit depths <= 12.
// The block's neighbourhood is held as u16 pairs: P[r][k] holds the samples at
// x4-4+2k (low half) and x4-3+2k (high half) of row y4-3+r, so window column c
// (x = x4-3+c, c = 0..9) is half (c+1)&1 of word (c+1)>>1.  Two horizontally
// adjacent output samples are filtered together: sample differences (|d| <
// 2^12), their clips and the sum of a symmetric tap pair (|s| < 2^13) fit in
// i16 lanes; the products accumulate in 32 bits through v_dot2.
// Classification and filter follow the same equations as alf_block_class /
// alf_block_filter above (src/alf.rs:822-1010, :1052-1180).
// ---------------------------------------------------------------------------
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// sample at window row r, column c
template <int N>
__device__ __forceinline__ uint32_t pw_at(const uint32_t (&P)[N][6], int r, int c) {
  const int q = c + 1;
  return (q & 1) ? (P[r][q >> 1] >> 16) : (P[r][q >> 1] & 0xffffu);
}
// samples at window columns c, c+1 of row r as one pair
template <int N>
__device__ __forceinline__ uint32_t pw_pair(const uint32_t (&P)[N][6], int r, int c) {
  const int q = c + 1;
  return (q & 1) ? __builtin_amdgcn_alignbit(P[r][(q >> 1) + 1], P[r][q >> 1], 16) : P[r][q >> 1];
}
__device__ __forceinline__ uint32_t absdiff_acc(uint32_t a, uint32_t b, uint32_t acc) {
  return acc + (a > b ? a - b : b - a);
}

template <int VB, bool CARVE>
__device__ __forceinline__ int alf_pk_class(const uint32_t (&P)[10][6], int bd) {
  constexpr int jmin = (VB == 2 && CARVE) ? 0 : -2;
  constexpr int jmax = (VB == 1 && CARVE) ? 3 : 5;
  constexpr int ac = ((VB == 1 || VB == 2) && CARVE) ? 3 : 2;
  uint32_t sh = 0, sv = 0, sd0 = 0, sd1 = 0;
#pragma unroll
  for (int jj = jmin; jj <= jmax; jj++) {
#pragma unroll
    for (int ii = -2; ii <= 5; ii++) {
      if (((ii + 4) & 1) != ((jj + 4) & 1)) continue;
      int rc = jj + 3, ru = jj + 2, rd = jj + 4;
      if (VB == 1) { rc = rc > 6 ? 6 : rc; ru = ru > 6 ? 6 : ru; rd = rd > 6 ? 6 : rd; }
      if (VB == 2) { rc = rc < 3 ? 3 : rc; ru = ru < 3 ? 3 : ru; rd = rd < 3 ? 3 : rd; }
      const int cc = ii + 3;
      const uint32_t c2 = pw_at(P, rc, cc) * 2;
      sh = absdiff_acc(c2, pw_at(P, rc, cc - 1) + pw_at(P, rc, cc + 1), sh);
      sv = absdiff_acc(c2, pw_at(P, ru, cc) + pw_at(P, rd, cc), sv);
      sd0 = absdiff_acc(c2, pw_at(P, ru, cc - 1) + pw_at(P, rd, cc + 1), sd0);
      sd1 = absdiff_acc(c2, pw_at(P, ru, cc + 1) + pw_at(P, rd, cc - 1), sd1);
    }
  }
  return alf_class_from_sums((int)sh, (int)sv, (int)sd0, (int)sd1, ac, bd);
}

// One tap pair of the 7x7 diamond applied to two output samples: cur holds
// the pair of centre samples, a / b the pairs at the two symmetric positions;
// w is the tap word (coefficient i16 | clip << 16).
__device__ __forceinline__ void alf_pk_tap(uint32_t c0, uint32_t c1, uint32_t clip, s16x2 cur, uint32_t a, uint32_t b,
                                           int& acc0, int& acc1) {
  const s16x2 cl = as_s16x2(clip);
  const s16x2 hi = {cl.x, cl.x}, lo = {cl.y, cl.y};
  const s16x2 da = __builtin_elementwise_min(__builtin_elementwise_max(as_s16x2(a) - cur, lo), hi);
  const s16x2 db = __builtin_elementwise_min(__builtin_elementwise_max(as_s16x2(b) - cur, lo), hi);
  const s16x2 s = da + db;
  acc0 = __builtin_amdgcn_sdot2(s, as_s16x2(c0), acc0, false);
  acc1 = __builtin_amdgcn_sdot2(s, as_s16x2(c1), acc1, false);
}

// Diamond tap geometry: tap k reads (R + dy*Y, C + dx) and (R - dy*Y, C - dx)
// where Y is the row distance allowed by the virtual boundary for |dy|.
__device__ constexpr int kTapDy[12] = {3, 2, 2, 2, 1, 1, 1, 1, 1, 0, 0, 0};
__device__ constexpr int kTapDx[12] = {0, 1, 0, -1, 2, 1, 0, -1, -2, 3, 2, 1};

// Output rows R0 .. R0+NR-1 of the block: out[r][h] receives the filtered
// samples (x4+2h, x4+2h+1) of row y4+r as a u16 pair.  Q[i] holds window row
// R0 + i (N >= NR + 6 rows).  Taps are the outer loop so that only one tap's
// constants are live.
// NOCLIP: the caller has checked that the block's clip values all exceed
// maxv and that every window sample is at most maxv, so that no clip acts
// (|a - cur| <= maxv): the taps are summed unclipped, (a + b) - 2 cur.  This
// is the case of the fixed filter sets and of APS filters without clipping
// (AlfClip = 2^BitDepth, src/alf.rs:627-644) on decoded samples.
template <int VB, int R0, int NR, int N, bool NOCLIP = false>
__device__ __forceinline__ void alf_pk_filter_rows(const uint32_t (&Q)[N][6], const uint32_t* f, int maxv,
                                                   uint32_t (&out)[4][2]) {
  constexpr int MT[3][4] = {{3, 3, 3, 3}, {3, 2, 1, 0}, {0, 1, 2, 3}};
  int acc0[NR][2], acc1[NR][2];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int h = 0; h < 2; h++) acc0[r][h] = acc1[r][h] = MT[VB][R0 + r] == 0 ? 512 : 64;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const uint32_t w = f[k], kk = w >> 16;
    const uint32_t c0 = w & 0xffffu, c1 = w << 16, clip = kk | ((0u - kk) << 16);
    if (NOCLIP) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int m = MT[VB][R0 + r];
        const int dy = kTapDy[k] < m ? kTapDy[k] : m;
        const int R = r + 3;
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int C = 3 + 2 * h;
          const s16x2 cur = as_s16x2(pw_pair(Q, R, C));
          // samples <= 4095 here: the sums stay inside i16
          const s16x2 sum = as_s16x2(pw_pair(Q, R + dy, C + kTapDx[k])) +
                            as_s16x2(pw_pair(Q, R - dy, C - kTapDx[k])) - (cur + cur);
          acc0[r][h] = __builtin_amdgcn_sdot2(sum, as_s16x2(c0), acc0[r][h], false);
          acc1[r][h] = __builtin_amdgcn_sdot2(sum, as_s16x2(c1), acc1[r][h], false);
        }
      }
    } else {
      // Without a virtual boundary, the pair dy rows and dx / 2 pairs
      // further on (even dx) has this pair's centre as its b sample and this
      // pair's a samples as its centre: its clipped b difference is the
      // negation of this pair's clipped a difference (the clip is
      // symmetric), taken from da[] instead of recomputed; for dx = +-1 the
      // pair at column 4 + dx takes it from the halves of two pairs.
      const s16x2 cl = as_s16x2(clip);
      const s16x2 hi = {cl.x, cl.x}, lo = {cl.y, cl.y};
      s16x2 da[NR][2];
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int m = MT[VB][R0 + r];
        const int dy = kTapDy[k] < m ? kTapDy[k] : m;
        const int dx = kTapDx[k];
        const int R = r + 3;  // centre row in Q
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int C = 3 + 2 * h;
          const s16x2 cur = as_s16x2(pw_pair(Q, R, C));
          da[r][h] = __builtin_elementwise_min(
              __builtin_elementwise_max(as_s16x2(pw_pair(Q, R + dy, C + dx)) - cur, lo), hi);
          const int rs = r - dy, hs = h - dx / 2;
          s16x2 sum;
          if (VB == 0 && (dx & 1) == 0 && (dy | dx) != 0 && rs >= 0 && hs >= 0 && hs < 2) {
            sum = da[r][h] - da[rs][hs];
          } else if (VB == 0 && (dx & 1) != 0 && 2 * h == dx + 1 && rs >= 0) {
            // odd dx: the mirrored samples are columns 4 and 5 of row rs,
            // the high half of pair 0 and the low half of pair 1
            sum = da[r][h] - as_s16x2(__builtin_amdgcn_alignbit(as_u32(da[rs][1]), as_u32(da[rs][0]), 16));
          } else {
            const s16x2 db = __builtin_elementwise_min(
                __builtin_elementwise_max(as_s16x2(pw_pair(Q, R - dy, C - dx)) - cur, lo), hi);
            sum = da[r][h] + db;
          }
          acc0[r][h] = __builtin_amdgcn_sdot2(sum, as_s16x2(c0), acc0[r][h], false);
          acc1[r][h] = __builtin_amdgcn_sdot2(sum, as_s16x2(c1), acc1[r][h], false);
        }
      }
    }
  }
  const s16x2 zero = {0, 0}, vmax = {(short)maxv, (short)maxv};
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int sft = MT[VB][R0 + r] == 0 ? 10 : 7;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      // the offsets saturate to i16; the saturating add and the clip to
      // [0, maxv] then give the same result as the exact sum
      const s16x2 d = __builtin_amdgcn_cvt_pk_i16(acc0[r][h] >> sft, acc1[r][h] >> sft);
      const s16x2 o = __builtin_elementwise_add_sat(as_s16x2(pw_pair(Q, r + 3, 3 + 2 * h)), d);
      out[R0 + r][h] = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(o, zero), vmax));
    }
  }
}

template <int VB, bool NOCLIP = false>
__device__ __forceinline__ void alf_pk_filter(const uint32_t (&P)[10][6], const uint32_t* f, int maxv,
                                              uint32_t (&out)[4][2]) {
  alf_pk_filter_rows<VB, 0, 4, 10, NOCLIP>(P, f, maxv, out);
}

// Classification reading the window from LDS row by row (three window rows
// live at a time) instead of holding all ten; base points at sample
// (x4-4, y4-3), rows PITCH samples apart, 8-byte aligned.  Same sums as
// alf_pk_class.
template <int PITCH>
__device__ __forceinline__ void lds_row(const uint16_t* base, int r, uint32_t (&R)[6]) {
  const uint2* q = reinterpret_cast<const uint2*>(base + r * PITCH);
  const uint2 a = q[0], b = q[1], c = q[2];
  R[0] = a.x; R[1] = a.y; R[2] = b.x; R[3] = b.y; R[4] = c.x; R[5] = c.y;
}
__device__ __forceinline__ uint32_t row_at(const uint32_t (&R)[6], int c) {
  const int q = c + 1;
  return (q & 1) ? (R[q >> 1] >> 16) : (R[q >> 1] & 0xffffu);
}

template <int VB, bool CARVE, int PITCH>
__device__ __forceinline__ int alf_lds_class(const uint16_t* base, int bd) {
  constexpr int jmin = (VB == 2 && CARVE) ? 0 : -2;
  constexpr int jmax = (VB == 1 && CARVE) ? 3 : 5;
  constexpr int ac = ((VB == 1 || VB == 2) && CARVE) ? 3 : 2;
  uint32_t sh = 0, sv = 0, sd0 = 0, sd1 = 0;
#pragma unroll
  for (int jj = jmin; jj <= jmax; jj++) {
    int rc = jj + 3, ru = jj + 2, rd = jj + 4;
    if (VB == 1) { rc = rc > 6 ? 6 : rc; ru = ru > 6 ? 6 : ru; rd = rd > 6 ? 6 : rd; }
    if (VB == 2) { rc = rc < 3 ? 3 : rc; ru = ru < 3 ? 3 : ru; rd = rd < 3 ? 3 : rd; }
    uint32_t U[6], Cr[6], D[6];
    lds_row<PITCH>(base, ru, U);
    lds_row<PITCH>(base, rc, Cr);
    lds_row<PITCH>(base, rd, D);
#pragma unroll
    for (int ii = -2; ii <= 5; ii++) {
      if (((ii + 4) & 1) != ((jj + 4) & 1)) continue;
      const int cc = ii + 3;
      const uint32_t c2 = row_at(Cr, cc) * 2;
      sh = absdiff_acc(c2, row_at(Cr, cc - 1) + row_at(Cr, cc + 1), sh);
      sv = absdiff_acc(c2, row_at(U, cc) + row_at(D, cc), sv);
      sd0 = absdiff_acc(c2, row_at(U, cc - 1) + row_at(D, cc + 1), sd0);
      sd1 = absdiff_acc(c2, row_at(U, cc + 1) + row_at(D, cc - 1), sd1);
    }
    __asm__ volatile("" ::: "memory");  // keep each row's loads next to their use
  }
  return alf_class_from_sums((int)sh, (int)sv, (int)sd0, (int)sd1, ac, bd);
}

template <int PITCH>
__device__ __forceinline__ int alf_lds_block_class(const uint16_t* base, int vb, bool carve, int bd) {
  if (vb == 0) return alf_lds_class<0, false, PITCH>(base, bd);
  if (vb == 1) return carve ? alf_lds_class<1, true, PITCH>(base, bd) : alf_lds_class<1, false, PITCH>(base, bd);
  return carve ? alf_lds_class<2, true, PITCH>(base, bd) : alf_lds_class<2, false, PITCH>(base, bd);
}

// Virtual-boundary variant of the block at yb = y4 - yCtb: 0 none, 1 the block
// above the boundary row, 2 the block below it.
__device__ __forceinline__ int alf_vb_kind(int yb, int ctb) { return yb == ctb - 8 ? 1 : (yb == ctb - 4 ? 2 : 0); }

// N rows of 12 samples (6 aligned pairs) starting at `base` (8-byte aligned).
template <int PITCH, int N>
__device__ __forceinline__ void load_pairs(const uint16_t* base, uint32_t (&P)[N][6]) {
#pragma unroll
  for (int rr = 0; rr < N; rr++) {
    const uint2* q = reinterpret_cast<const uint2*>(base + rr * PITCH);
    const uint2 a = q[0], b = q[1], c = q[2];
    P[rr][0] = a.x; P[rr][1] = a.y; P[rr][2] = b.x; P[rr][3] = b.y; P[rr][4] = c.x; P[rr][5] = c.y;
  }
}

// ALF luma of the 4x4 block at (x4, y4) from an LDS buffer S (pitch SW,
// origin (sox, soy)) holding the SAO output of [x4-4, x4+8) x [y4-3, y4+7)
// with every position outside the CTB's fetch rectangle already replaced by
// the clamped sample.  outp[r] receives row y4+r as two u16 pairs.  The
// classification streams three window rows at a time and the filter runs in
// two halves of two output rows, each reloading its 8 window rows: the
// compiler barriers keep the loads apart so that fewer registers are live.
// alf_on = false copies the block.  (src/alf.rs:822-1010, :1052-1180.)
template <int SW, bool HBD>
__device__ __forceinline__ void alf_luma_block_lds(const uint16_t* S, int sox, int soy, int x4, int y4,
                                                   const uint32_t* bank, bool alf_on, int y_ctb, int ctb, bool carve,
                                                   int alf_bd, uint32_t (&outp)[4][2]) {
  const int amax = (1 << alf_bd) - 1;
  if (!alf_on) {
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const uint2 q = *reinterpret_cast<const uint2*>(&S[__mul24(y4 + rr - soy, SW) + (x4 - sox)]);
      outp[rr][0] = q.x;
      outp[rr][1] = q.y;
    }
  } else if (HBD) {
    int Wn[10][10];
#pragma unroll
    for (int rr = 0; rr < 10; rr++)
#pragma unroll
      for (int cc = 0; cc < 10; cc++) Wn[rr][cc] = S[__mul24(y4 - 3 + rr - soy, SW) + (x4 - 3 + cc - sox)];
    int outv[4][4];
    alf_block(Wn, bank, y4 - y_ctb, ctb, carve, alf_bd, amax, outv);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      outp[rr][0] = (uint32_t)outv[rr][0] | ((uint32_t)outv[rr][1] << 16);
      outp[rr][1] = (uint32_t)outv[rr][2] | ((uint32_t)outv[rr][3] << 16);
    }
  } else {
    const uint16_t* base = &S[__mul24(y4 - 3 - soy, SW) + (x4 - 4 - sox)];
    const int vb = alf_vb_kind(y4 - y_ctb, ctb);
    const int cls = alf_lds_block_class<SW>(base, vb, carve, alf_bd);
    const uint32_t* f = bank + ((cls & 31) * 4 + (cls >> 5)) * 12;
    __asm__ volatile("" ::: "memory");
    {
      uint32_t Q[8][6];
      load_pairs<SW>(base, Q);
      if (vb == 0) alf_pk_filter_rows<0, 0, 2, 8>(Q, f, amax, outp);
      else if (vb == 1) alf_pk_filter_rows<1, 0, 2, 8>(Q, f, amax, outp);
      else alf_pk_filter_rows<2, 0, 2, 8>(Q, f, amax, outp);
    }
    __asm__ volatile("" ::: "memory");
    {
      uint32_t Q[8][6];
      load_pairs<SW>(base + 2 * SW, Q);
      if (vb == 0) alf_pk_filter_rows<0, 2, 2, 8>(Q, f, amax, outp);
      else if (vb == 1) alf_pk_filter_rows<1, 2, 2, 8>(Q, f, amax, outp);
      else alf_pk_filter_rows<2, 2, 2, 8>(Q, f, amax, outp);
    }
  }
}

// Class of a block at yb = y4 - yCtb (virtual-boundary variant picked here).
__device__ __forceinline__ int alf_pk_block_class(const uint32_t (&P)[10][6], int vb, bool carve, int bd) {
  if (vb == 0) return alf_pk_class<0, false>(P, bd);
  if (vb == 1) return carve ? alf_pk_class<1, true>(P, bd) : alf_pk_class<1, false>(P, bd);
  return carve ? alf_pk_class<2, true>(P, bd) : alf_pk_class<2, false>(P, bd);
}

// Classification + filter of one block from the packed window; `bank` holds
// the CTB's 25 x 4 x 12 tap words, yb = y4 - yCtb.
__device__ __forceinline__ void alf_block_pk(const uint32_t (&P)[10][6], const uint32_t* bank, int yb, int ctb,
                                             bool carve, int bd, int maxv, uint32_t (&out)[4][2],
                                             bool in_range = false) {
  int cls;
  const int vb = yb == ctb - 8 ? 1 : (yb == ctb - 4 ? 2 : 0);
  if (vb == 0) cls = alf_pk_class<0, false>(P, bd);
  else if (vb == 1) cls = carve ? alf_pk_class<1, true>(P, bd) : alf_pk_class<1, false>(P, bd);
  else cls = carve ? alf_pk_class<2, true>(P, bd) : alf_pk_class<2, false>(P, bd);
  const uint32_t* f = bank + ((cls & 31) * 4 + (cls >> 5)) * 12;
  // the unclipped form where it holds for every lane of the wave (a
  // wave-uniform branch: fixed-set CTBs and APS sets without clipping)
  bool noclip = in_range;
#pragma unroll
  for (int k = 0; k < 12; k++) noclip = noclip && (f[k] >> 16) > (uint32_t)maxv;
  if (__all(noclip)) {
    if (vb == 0) alf_pk_filter<0, true>(P, f, maxv, out);
    else if (vb == 1) alf_pk_filter<1, true>(P, f, maxv, out);
    else alf_pk_filter<2, true>(P, f, maxv, out);
  } else {
    if (vb == 0) alf_pk_filter<0>(P, f, maxv, out);
    else if (vb == 1) alf_pk_filter<1>(P, f, maxv, out);
    else alf_pk_filter<2>(P, f, maxv, out);
  }
}

// ---------------------------------------------------------------------------
// ALF luma of one block through a sample accessor at(dx, dy) -> the sample at
// (x4 + dx, y4 + dy) clamped to the fetch rectangle, with run-time loops: the
// rare 32-bit path (windows holding samples of 2^15 or more), kept light on
// registers so that the packed path keeps its occupancy.  Same sums, virtual
// boundary rules and taps as alf_block_class / alf_block_filter
// (src/alf.rs:822-1010, :652-724); put(c, r, v) receives output sample
// (x4 + c, y4 + r).
// ---------------------------------------------------------------------------
template <class AT>
__device__ __forceinline__ int alf_class_at(AT at, int vb, bool carve, int bd) {
  const int jmin = (vb == 2 && carve) ? 0 : -2, jmax = (vb == 1 && carve) ? 3 : 5;
  const int ac = (vb != 0 && carve) ? 3 : 2;
  // row offset after the virtual-boundary padding
  auto rowv = [&](int r) { return vb == 1 ? (r > 3 ? 3 : r) : (vb == 2 ? (r < 0 ? 0 : r) : r); };
  int sh = 0, sv = 0, sd0 = 0, sd1 = 0;
#pragma unroll 1
  for (int jj = jmin; jj <= jmax; jj++) {
    const int rc = rowv(jj), ru = rowv(jj - 1), rd = rowv(jj + 1);
#pragma unroll 1
    for (int ii = -2 + (jj & 1); ii <= 5; ii += 2) {  // positions with ii, jj of one parity
      const int c2 = 2 * at(ii, rc);
      sh += abs(c2 - at(ii - 1, rc) - at(ii + 1, rc));
      sv += abs(c2 - at(ii, ru) - at(ii, rd));
      sd0 += abs(c2 - at(ii - 1, ru) - at(ii + 1, rd));
      sd1 += abs(c2 - at(ii + 1, ru) - at(ii - 1, rd));
    }
  }
  return alf_class_from_sums(sh, sv, sd0, sd1, ac, bd);
}

template <class AT, class PUT>
__device__ __forceinline__ void alf_filter_at(AT at, const uint32_t* f, int vb, int amax, PUT put) {
#pragma unroll 1
  for (int r = 0; r < 4; r++) {
    const int m = vb == 0 ? 3 : (vb == 1 ? 3 - r : r);
    const int sft = m == 0 ? 10 : 7;
#pragma unroll 1
    for (int c = 0; c < 4; c++) {
      const int cur = at(c, r);
      int sum = 0;
#pragma unroll
      for (int k = 0; k < 12; k++) {
        const int dy = kTapDy[k] < m ? kTapDy[k] : m, dx = kTapDx[k];
        sum += alf_pair(f[k], at(c + dx, r + dy), at(c - dx, r - dy), cur);
      }
      put(c, r, clampi(cur + ((sum + (1 << (sft - 1))) >> sft), 0, amax));
    }
  }
}

}  // namespace vvclf
