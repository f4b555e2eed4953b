// Host thread pool of the engine: every item of every job runs exactly once,
// jobs of different sizes follow each other with and without idle gaps, two
// pools run jobs concurrently, and the first exception of a job reaches its
// caller after the job has drained.  Prints "ok" on success.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../graph-cut-ransac_amd/csrc/host_pool.hpp"

static int check_pool(gcr::HostPool& pool, int jobs, int gap_us) {
  std::vector<std::atomic<int>> hits(257);
  for (int j = 0; j < jobs; ++j) {
    const size_t n = (size_t)(j % 257);
    for (size_t i = 0; i < n; ++i) hits[i].store(0);
    pool.run(n, [&](size_t i) { hits[i].fetch_add(1); });
    for (size_t i = 0; i < n; ++i)
      if (hits[i].load() != 1) return 1;
    if (gap_us && j % 7 == 0) std::this_thread::sleep_for(std::chrono::microseconds(gap_us));
  }
  std::atomic<int> ran{0};
  bool caught = false;
  try {
    pool.run(64, [&](size_t i) {
      ran.fetch_add(1);
      if (i == 5) throw std::runtime_error("item 5");
    });
  } catch (const std::runtime_error&) {
    caught = true;
  }
  if (!caught || ran.load() != 64) return 2;
  return 0;
}

int main() {
  int rc_a = 0, rc_b = 0;
  std::thread a([&] { rc_a = check_pool(gcr::HostPool::group(0, 2), 3000, 1500); });
  std::thread b([&] { rc_b = check_pool(gcr::HostPool::group(1, 2), 3000, 0); });
  a.join();
  b.join();
  const int rc_c = check_pool(gcr::HostPool::get(), 500, 0);
  if (rc_a || rc_b || rc_c) {
    std::printf("failed %d %d %d\n", rc_a, rc_b, rc_c);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
