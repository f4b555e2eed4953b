// Checks the engine's host random streams against the standard library:
// MT19937 against std::mt19937 word for word, the uniform sampler's
// unique-index draws against the same draws made with std::mt19937, the
// glibc random() restatement against the C library's rand() after srand(),
// and the importance sampler's draws against std::discrete_distribution over
// std::mt19937.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../graph-cut-ransac_amd/csrc/host_samplers.hpp"

int main() {
  const uint32_t seeds[] = {0u, 1u, 5489u, 0xdeadbeefu, 0x5bd1e995u};
  for (uint32_t sd : seeds) {
    gcr::MT19937 a(sd);
    std::mt19937 b(sd);
    for (long i = 0; i < 3000000; ++i)
      if (a() != b()) {
        std::printf("MT19937 differs: seed %u at %ld\n", sd, i);
        return 1;
      }
    // sampler stream: 7 unique indices out of n, 50000 samples
    const size_t n = 1000 + sd % 4000;
    gcr::UniformSampler s(sd, n);
    std::mt19937 g(sd);
    std::vector<int32_t> out(7);
    for (int c = 0; c < 50000; ++c) {
      s.sample(nullptr, n, out.data(), 7);
      std::uniform_int_distribution<size_t> d(0, n - 1);
      size_t r[7];
      for (size_t i = 0; i < 7; i++) {
        r[i] = d(g);
        for (int j = (int)i - 1; j >= 0; j--)
          if (r[i] == r[j]) {
            i--;
            break;
          }
      }
      for (int k = 0; k < 7; ++k)
        if ((size_t)out[k] != r[k]) {
          std::printf("sampler differs: seed %u sample %d\n", sd, c);
          return 1;
        }
    }
  }
  // glibc random(): rand() after srand(seed), including seed 0 and seeds >= 2^31
  const uint32_t gseeds[] = {0u, 1u, 2u, 12345u, 0x7fffffffu, 0x80000000u, 0xdeadbeefu, 0xffffffffu};
  for (uint32_t sd : gseeds) {
    gcr::GlibcRandom g(sd);
    std::srand(sd);
    for (long i = 0; i < 1000000; ++i) {
      const int a = g.next(), b = std::rand();
      if (a != b) {
        std::printf("glibc random differs: seed %u at %ld (%d vs %d)\n", sd, i, a, b);
        return 1;
      }
    }
  }
  // importance sampler: independent std::discrete_distribution draws
  {
    std::vector<double> pr(3000);
    std::mt19937 pg(7);
    for (auto& v : pr) v = std::uniform_real_distribution<double>(0.0, 1.0)(pg);
    gcr::ImportanceSampler s(99u, pr.data(), pr.size());
    std::mt19937 g(99u);
    std::discrete_distribution<int> d(pr.begin(), pr.end());
    std::vector<int32_t> out(7);
    for (int c = 0; c < 20000; ++c) {
      s.sample(nullptr, pr.size(), out.data(), 7);
      for (int k = 0; k < 7; ++k)
        if (out[k] != d(g)) {
          std::printf("importance sampler differs at sample %d\n", c);
          return 1;
        }
    }
  }
  std::printf("ok\n");
  return 0;
}
