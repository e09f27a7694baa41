/* synth.h — seeded synthetic pictures with reference-shaped side information
 * (CU records, motion grid, SAO / ALF pictures, APS sets, LMCS tables) for
 * the benchmark workloads and parity tests.  Pure host C++; deterministic for
 * a given spec on every machine (splitmix64). */
#ifndef VVCLF_SYNTH_H
#define VVCLF_SYNTH_H

#include "../../include/vvclf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vvclf_synth_spec {
  uint32_t width, height, bit_depth, ctb_log2;
  uint64_t seed;
  uint32_t min_cu_w, min_cu_h; /* smallest CU (8x8, or 4x8 for dense grids) */
  float split_scale;           /* multiplies the per-size split probabilities */
  float p_intra;               /* CU coded intra (bS = 2 against anything) */
  float p_mv_change;           /* inter CU carries a distinct motion vector */
  float p_coded;               /* inter CU has coded luma residual */
  uint32_t qp_min, qp_max;
  uint32_t rich;       /* also draw sub-block, TB-split, multi-TB, palette,
                          BDPCM, IBC, CIIP, explicit chroma QP records,
                          slice offsets and a dual chroma tree */
  uint32_t lmcs;       /* random monotone 16-piece LMCS */
  uint32_t sao;        /* per-CTB SAO, type uniform over off / band / edge */
  uint32_t alf;        /* ALF luma */
  float p_alf_luma;    /* per-CTB luma ALF enable */
  uint32_t n_luma_aps; /* APS luma filter sets (slots 16..) */
  float p_fixed_set;   /* share of luma-on CTBs using a fixed filter set */
  uint32_t chroma_alf; /* chroma ALF with alternative filters */
  uint32_t cc_alf;     /* CC-ALF for Cb and Cr */
  float p_cc;          /* per-CTB CC-ALF enable */
  uint32_t tiles;      /* 2x2 tile grid with loop filtering across tiles off */
  uint64_t aps_seed;   /* nonzero: APS / LMCS tables drawn from this seed, so
                          pictures of one stream share them */
} vvclf_synth_spec;

typedef struct vvclf_synth_frame vvclf_synth_frame;

/* Fill `spec` with the defaults of workload `config` (0..4) for a picture of
 * width x height. */
void vvclf_synth_default_spec(uint32_t config, uint32_t width, uint32_t height, uint64_t seed,
                              vvclf_synth_spec* spec);
int vvclf_synth_make(const vvclf_synth_spec* spec, vvclf_synth_frame** out);
void vvclf_synth_free(vvclf_synth_frame* f);
/* The descriptor; its picture planes are owned by the synth frame. */
vvclf_frame* vvclf_synth_desc(vvclf_synth_frame* f);
/* Restore the picture planes to the generated input samples. */
void vvclf_synth_reset(vvclf_synth_frame* f);
/* Pristine generated input planes (read-only copy). */
const vvclf_picture* vvclf_synth_input(const vvclf_synth_frame* f);
/* Counts of CU records, intra CUs, and deblock segments by bS per stage. */
void vvclf_synth_stats(const vvclf_synth_frame* f, uint64_t out[8]);

#ifdef __cplusplus
}
#endif
#endif
