// parallel_for called from several threads at once: every index of every
// call runs exactly once and every call returns.  Built and run by
// tests/test_host_walk.py; exits non-zero on a miss.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "vvclf_host_internal.h"

int main() {
  std::atomic<long> bad{0};
  std::vector<std::thread> callers;
  for (int c = 0; c < 6; c++)
    callers.emplace_back([&, c] {
      for (int rep = 0; rep < 2000; rep++) {
        const int n = 1 + (rep + c) % 17;
        std::vector<std::atomic<int>> hit(n);
        for (auto& h : hit) h = 0;
        vvclf::parallel_for(n, [&](int i) { hit[i]++; });
        for (int i = 0; i < n; i++)
          if (hit[i] != 1) bad++;
      }
    });
  for (auto& t : callers) t.join();
  printf("missed or repeated indices: %ld\n", bad.load());
  return bad != 0;
}
