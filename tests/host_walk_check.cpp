// Host CU walk on several threads against the same walk on one thread: the
// deblocking descriptor maps, work lists and bS histogram must be identical
// (the multi-threaded walk claims segments atomically, builds the CU grid
// in row bands and finishes the maps in row bands).  Built and run by
// tests/test_host_walk.py; exits non-zero on a difference.
#include <cstdio>

#include "synth.h"
#include "vvclf_host_internal.h"

using namespace vvclf;

int main() {
  int bad = 0;
  struct pic {
    int cfg;
    uint32_t w, h;
    int rich;
  } pics[] = {{1, 3840, 2160, 0}, {0, 1920, 1080, 1}, {3, 1920, 1088, 1}};
  for (const pic& p : pics) {
    vvclf_synth_spec sp;
    vvclf_synth_default_spec(p.cfg, p.w, p.h, 77, &sp);
    sp.rich = p.rich;
    vvclf_synth_frame* sf = nullptr;
    if (vvclf_synth_make(&sp, &sf) != VVCLF_OK) return 2;
    const vvclf_frame* f = vvclf_synth_desc(sf);
    host_tables a;
    const int ra = derive_deblock_maps(*f, p.w, p.h, p.w / 2, p.h / 2, a, 1);
    for (int nt : {3, 8, 16}) {
      host_tables b;
      const int rb = derive_deblock_maps(*f, p.w, p.h, p.w / 2, p.h / 2, b, nt);
      bool same = ra == rb && a.e_lv == b.e_lv && a.e_lh == b.e_lh && a.l_lv == b.l_lv && a.l_lh == b.l_lh &&
                  a.l_cv == b.l_cv && a.l_ch == b.l_ch;
      for (int c = 0; c < 2; c++) same = same && a.e_cv[c] == b.e_cv[c] && a.e_ch[c] == b.e_ch[c];
      for (int i = 0; i < 3; i++) same = same && a.bs_hist[i] == b.bs_hist[i];
      printf("cfg %d %ux%u, %d threads: rc %d/%d, lists %zu %zu %zu %zu: %s\n", p.cfg, p.w, p.h, nt, ra, rb,
             b.l_lv.size(), b.l_lh.size(), b.l_cv.size(), b.l_ch.size(), same ? "identical" : "DIFFERENT");
      bad += !same;
    }
    vvclf_synth_free(sf);
  }
  return bad ? 1 : 0;
}
