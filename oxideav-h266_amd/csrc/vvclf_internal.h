// vvclf_internal.h — device-side layout shared by the host library and the
// HIP kernels.  Everything a kernel reads about one picture sits in one POD
// (`dev_job`) passed by value; all pointers are device pointers into HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vvclf {

// ---------------------------------------------------------------------------
// Deblocking edge descriptors: one u32 per 4-sample luma segment or 2-sample
// chroma segment, derived on the host from the CU records (the
// deblock_cu_dir walk, src/deblock.rs:707-983).  0 means "not filtered"
// (bS == 0, tC == 0, or the chroma (1,1)/bS<2 no-op of src/deblock.rs:1823).
//   bits 0-1   bS (1..2)
//   bits 2-4   maxFilterLengthP  (0..7)
//   bits 5-7   maxFilterLengthQ
//   bit  8     P side palette (write skip)
//   bit  9     Q side palette
//   bit  10    luma EDGE_HOR on a CTB row (sidePisLargeBlk forced 0)
//   bits 11-16 beta table index Q (0..63)
//   bits 17-23 tC table index Q (0..65)
// ---------------------------------------------------------------------------
enum : uint32_t {
  E_BS_SHIFT = 0,
  E_MFLP_SHIFT = 2,
  E_MFLQ_SHIFT = 5,
  E_PLTP_BIT = 1u << 8,
  E_PLTQ_BIT = 1u << 9,
  E_CTBROW_BIT = 1u << 10,
  E_QB_SHIFT = 11,
  E_QT_SHIFT = 17,
};

__host__ __device__ inline uint32_t pack_edge(uint32_t bs, uint32_t mflp, uint32_t mflq, bool pltp, bool pltq,
                                              bool ctbrow, uint32_t qb, uint32_t qt) {
  return (bs << E_BS_SHIFT) | (mflp << E_MFLP_SHIFT) | (mflq << E_MFLQ_SHIFT) | (pltp ? E_PLTP_BIT : 0u) |
         (pltq ? E_PLTQ_BIT : 0u) | (ctbrow ? E_CTBROW_BIT : 0u) | (qb << E_QB_SHIFT) | (qt << E_QT_SHIFT);
}

// SAO per CTB and component, packed in a u64:
//   bits 0-1 type (0 off, 1 band, 2 edge), bits 2-6 eo class / band position,
//   bits 8-63: four offsets, 14-bit two's complement each (bits 8+14*i).
static inline uint64_t pack_sao(uint32_t type, uint32_t param, const int32_t off[4]) {
  uint64_t v = (uint64_t)(type & 3) | ((uint64_t)(param & 31) << 2);
  for (int i = 0; i < 4; i++) v |= ((uint64_t)((uint32_t)off[i] & 0x3fffu)) << (8 + 14 * i);
  return v;
}

// Per-CTB fetch rectangle (luma samples, half-open), u16 each:
// x0 | y0 << 16 | x1 << 32 | y1 << 48.
static inline uint64_t pack_rect(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  return (uint64_t)x0 | ((uint64_t)y0 << 16) | ((uint64_t)x1 << 32) | ((uint64_t)y1 << 48);
}

// ALF per CTB, packed in a u32 (each field 0xff = off):
//   bits 0-7 luma bank index, 8-15 Cb alt filter, 16-23 Cr alt filter,
//   bits 24-27 CC Cb filter (0xf off), 28-31 CC Cr filter (0xf off).
enum : uint32_t { ALF_OFF8 = 0xffu, ALF_OFF4 = 0xfu };

// Luma filter bank entry: low 16 bits coefficient (i16), high 16 bits the
// clip value min(AlfClip, 2^BitDepth - 1) (u16; identical clamping since a
// sample difference never exceeds 2^BitDepth - 1).  Layout per bank:
// [25 classes][4 transposeIdx][12 taps], taps already permuted by the
// transposeIdx table (eqs. 1442-1445), so a kernel reads tap k directly.
static constexpr int ALF_BANK_WORDS = 25 * 4 * 12;
// Chroma bank: [8 alt filters][6 taps] same packing.
static constexpr int ALF_CBANK_WORDS = 8 * 6;
// CC bank: [4 filters][7 taps] as i32 per component.
static constexpr int ALF_CCBANK_WORDS = 4 * 7;

// One output tile of the fused batched form: job index, tile origin, and the
// ALF parameters of the CTB holding the tile (its alf_ctb word, 0xffffffff
// where ALF is off or outside the CTB grid, and its alf_rect), filled on the
// host so that the kernel starts the tile without a dependent load.
struct tile_ref {
  uint32_t job;
  uint16_t x0, y0;
  uint32_t alf_cw;
  uint32_t pad;
  uint64_t alf_rect;
};

struct dev_planes {
  uint16_t* y;
  uint16_t* cb;
  uint16_t* cr;
};

// Everything the kernels need about one picture.  Sample planes: luma pitch
// `py`, chroma pitch `pc` (in samples, multiples of 64).
struct dev_job {
  // geometry
  int32_t w, h;          // luma
  int32_t wc, hc;        // chroma (0 for 4:0:0)
  int32_t py, pc;        // pitches
  int32_t bd;            // BitDepth of the planes (write clip)
  int32_t bd_dbk;        // DeblockParams::bit_depth (thresholds)
  int32_t ctb_log2;      // CtbLog2SizeY for ALF
  int32_t wctb, hctb;    // ALF CTB grid (AlfPicture dims)
  int32_t sao_ctb_log2;  // CtbLog2SizeY for SAO
  int32_t swctb, shctb;  // SAO CTB grid (SaoPicture dims)
  int32_t sao_bd, alf_bd;
  // flags
  int32_t do_lmcs, do_dbk, do_dbk_chroma, do_sao, do_alf;
  int32_t sao_luma, sao_chroma;
  // buffers
  dev_planes in;   // pristine input (never written by runs)
  dev_planes a;    // work buffer A (LMCS + deblock in place)
  dev_planes b;    // work buffer B (SAO output)
  dev_planes out;  // final output
  const uint16_t* lmcs_lut;  // 65536 entries: every u16 sample value
  // deblock edge maps
  int32_t gw, gh;            // luma 4x4 grid dims (ceil(w/4), ceil(h/4))
  const uint32_t* e_lv;      // [gh][gw] vertical edge at x=4*gx, rows 4*gy..+3
  const uint32_t* e_lh;      // [gh][gw] horizontal edge at y=4*gy, cols 4*gx..+3
  int32_t cvw, cvh;          // chroma vertical map dims (ceil(wc/8), ceil(hc/2))
  int32_t chw, chh;          // chroma horizontal map dims (ceil(wc/2), ceil(hc/8))
  const uint32_t* e_cv[2];   // Cb, Cr vertical edges
  const uint32_t* e_ch[2];   // Cb, Cr horizontal edges
  // the nonzero descriptors of each direction as lists (luma gy << 16 | gx,
  // chroma (comp - 1) << 31 | sy << 16 | sx), for the per-picture kernels
  const uint32_t* l_lv;
  const uint32_t* l_lh;
  const uint32_t* l_cv;
  const uint32_t* l_ch;
  int32_t n_lv, n_lh, n_cv, n_ch;
  // SAO
  const uint64_t* sao;       // [hctb*wctb][3]
  const uint64_t* sao_rect;  // [hctb*wctb] (luma coords), NULL = picture
  // ALF
  const uint32_t* alf_ctb;   // [hctb*wctb]
  const uint64_t* alf_rect;  // [hctb*wctb] luma fetch rectangle
  const uint32_t* alf_bank;  // [n_banks][ALF_BANK_WORDS]
  const uint32_t* alf_cbank; // ALF_CBANK_WORDS
  const int32_t* alf_ccbank; // [2][ALF_CCBANK_WORDS]
  int32_t alf_chroma_on, cc_on;
};

}  // namespace vvclf
