// The host's 65536-entry LMCS inverse table (build_lmcs_lut, walked once over
// the samples) against the oracle's per-sample idx_y_inv + inverse map
// (src/lmcs.rs:209-238) for every u16 input: payloads from the derivation at
// bit depths 8..16 with every min / max bin, and arbitrary non-decreasing
// pivots with bins the derivation would not produce (empty bins, a max bin
// below the min bin).  Built and run by tests/test_host_walk.py; exits
// non-zero on a difference.
#include <cstdio>
#include <random>

#include "vvclf_host_internal.h"
extern "C" {
#include "vvclf_oracle.h"
}

int main() {
  std::mt19937_64 rng(2024);
  int checked = 0, derived = 0;
  for (int it = 0; it < 600; it++) {
    vvclf_lmcs_derived d{};
    vvclf_lmcs_binding b{};
    const uint32_t bd = 8 + (uint32_t)(rng() % 9);
    if (it % 2 == 0) {
      vvclf_lmcs_data a{};
      a.lmcs_min_bin_idx = (uint8_t)(rng() % 8);
      a.lmcs_delta_max_bin_idx = (uint8_t)(rng() % 8);
      a.lmcs_delta_cw_prec_minus1 = (uint8_t)(rng() % 14);
      for (int i = 0; i < VVCLF_LMCS_NUM_BINS; i++) {
        a.lmcs_delta_abs_cw[i] = (uint32_t)(rng() % 48);
        a.lmcs_delta_sign_cw_flag[i] = (uint8_t)(rng() % 2);
      }
      if (oracle_derive_lmcs(&a, bd, &d) != 0) continue;
      derived++;
      b.min_bin_idx = a.lmcs_min_bin_idx;
      b.max_bin_idx = (uint8_t)(15 - a.lmcs_delta_max_bin_idx);
    } else {
      d.bit_depth = bd;
      uint32_t p = (uint32_t)(rng() % 3000);
      for (int i = 0; i <= VVCLF_LMCS_NUM_BINS; i++) {
        d.lmcs_pivot[i] = p;
        p += (rng() % 4 == 0) ? 0u : (uint32_t)(rng() % 9000);
      }
      for (int i = 0; i < VVCLF_LMCS_NUM_BINS; i++) {
        d.inv_scale_coeff[i] = (uint32_t)(rng() % 70000);
        d.input_pivot[i] = (uint32_t)(rng() % 70000);
      }
      b.min_bin_idx = (uint8_t)(rng() % 16);
      b.max_bin_idx = (uint8_t)(rng() % 16);
    }
    b.derived = &d;
    std::vector<uint16_t> lut;
    vvclf::build_lmcs_lut(b, lut);
    for (uint32_t s = 0; s < 65536; s++) {
      const uint32_t idx = oracle_lmcs_idx_y_inv(&d, s, b.min_bin_idx, b.max_bin_idx);
      const uint32_t want = oracle_lmcs_inverse_map(&d, s, idx);
      if (lut[s] != want) {
        printf("payload %d (bit depth %u, bins %u..%u): sample %u maps to %u, the oracle to %u\n", it, bd,
               b.min_bin_idx, b.max_bin_idx, s, lut[s], want);
        return 1;
      }
    }
    checked++;
  }
  printf("lmcs tables identical: %d payloads (%d derived)\n", checked, derived);
  return checked > 300 && derived > 100 ? 0 : 3;
}
