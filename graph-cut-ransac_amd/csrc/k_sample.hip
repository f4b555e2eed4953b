// The uniform sampler's draws on the device (uniform_sampler.h:97-115 through
// uniform_random_generator.h:43-125): std::mt19937 seeded with the pair's
// seed, each index drawn by libstdc++'s uniform_int_distribution<size_t> on
// [0, range - 1] (Lemire's multiply-shift with rejection: the 64-bit product
// of a 32-bit word and the range; the draw takes the next word while the low
// half of the product is below 2^32 mod range), and a sample's index that
// repeats an earlier index of the same sample is drawn again.
//
// One wavefront per pair.  The generator's 624-word state lives in LDS during
// the launch (two buffers: the current block and the one it was generated
// from) and in HBM between launches, with the position of the next word.  A
// block is regenerated into the other buffer in four lane-parallel phases
// (words 0..226 read only old words; 227..453 also new words 0..226;
// 454..622 new words 227..395; 623 new words 0 and 396) and tempered as it
// is read.  Draws go through a queue: up to 64 words at a time are
// multiplied out, and the accepted values are appended in word order (ballot
// prefix) with the index of the word they came from.  Samples are then formed
// 64 at a time, one per lane, each from the next m queued values; the first
// lane whose m values hold a repeat and every lane after it are discarded,
// that sample is formed value by value with the repeat rule, and the next
// batch starts after it.  So every sample consumes exactly the values the
// sequential sampler consumes.  Words read ahead are given back at the end:
// the stored position is the word after the last value consumed, in the
// current block or the one before it, or (after long runs of rejected words)
// in a block regenerated from the launch's first one.
#include "kernels.hpp"

namespace gcr {

__global__ void __launch_bounds__(64) k_mt_seed(const uint32_t* __restrict__ seeds, const int32_t* __restrict__ pairs,
                                                int num, DevMT* __restrict__ st) {
  const int i = blockIdx.x;
  if (i >= num || threadIdx.x != 0) return;
  DevMT* s = st + pairs[i];
  uint32_t v = seeds[i];
  s->x[0] = v;
  for (uint32_t k = 1; k < 624; ++k) {
    v = 1812433253u * (v ^ (v >> 30)) + k;
    s->x[k] = v;
  }
  s->pos = 624;
}

namespace {

constexpr uint32_t kUpper = 0x80000000u, kLower = 0x7fffffffu, kMatA = 0x9908b0dfu;
constexpr int kQueue = 2048;  // ring of accepted values (a power of two)

__device__ __forceinline__ uint32_t mt_next(uint32_t xi, uint32_t xi1, uint32_t xm) {
  const uint32_t y = (xi & kUpper) | (xi1 & kLower);
  return xm ^ (y >> 1) ^ ((0u - (y & 1u)) & kMatA);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t z) {
  z ^= z >> 11;
  z ^= (z << 7) & 0x9d2c5680u;
  z ^= (z << 15) & 0xefc60000u;
  z ^= z >> 18;
  return z;
}

// nx = the block after ox (std::mt19937's twist)
__device__ void mt_twist(const uint32_t* ox, uint32_t* nx) {
  const int lane = threadIdx.x;
  for (int i = lane; i < 227; i += 64) nx[i] = mt_next(ox[i], ox[i + 1], ox[i + 397]);
  __syncthreads();
  for (int i = 227 + lane; i < 454; i += 64) nx[i] = mt_next(ox[i], ox[i + 1], nx[i - 227]);
  __syncthreads();
  for (int i = 454 + lane; i < 623; i += 64) nx[i] = mt_next(ox[i], ox[i + 1], nx[i - 227]);
  __syncthreads();
  if (lane == 0) nx[623] = mt_next(ox[623], nx[0], nx[396]);
  __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(64) k_sample_uniform(const DevSampleItem* __restrict__ items, DevMT* __restrict__ st,
                                                       int32_t* __restrict__ samples) {
  __shared__ uint32_t xs[2][624];
  __shared__ uint32_t q[kQueue];
  __shared__ int32_t qw[kQueue];
  const DevSampleItem it = items[blockIdx.x];
  const int lane = threadIdx.x;
  const int m = it.m;
  const int64_t count = it.count;
  int32_t* const out = samples + it.base * m;
  const uint32_t r = it.range;
  if ((uint32_t)m > r || m > 8) {
    // the sampler refuses before drawing: every sample is -1
    for (int64_t k = lane; k < count * m; k += 64) out[k] = -1;
    return;
  }
  if (count <= 0) return;
  DevMT* const s = st + it.pair;
  for (int i = lane; i < 624; i += 64) xs[0][i] = s->x[i];
  int pos = s->pos, cur = 0;
  int32_t blk_base = -pos;  // word index (from this launch's first word) of word 0 of the current block
  __syncthreads();
  const uint32_t thr = (uint32_t)((0x100000000ull - (uint64_t)r) % (uint64_t)r);
  uint32_t qh = 0, qt = 0;  // queue head and tail (values consumed / appended)
  auto produce = [&]() {
    if (pos >= 624) {
      mt_twist(xs[cur], xs[cur ^ 1]);
      cur ^= 1;
      pos = 0;
      blk_base += 624;
    }
    const int k = 624 - pos < 64 ? 624 - pos : 64;
    bool acc = false;
    uint32_t val = 0;
    if (lane < k) {
      const uint64_t prod = (uint64_t)mt_temper(xs[cur][pos + lane]) * (uint64_t)r;
      acc = (uint32_t)prod >= thr;
      val = (uint32_t)(prod >> 32);
    }
    const uint64_t b = __ballot(acc);
    if (acc) {
      const uint32_t at = (qt + (uint32_t)__popcll(b & ((1ull << lane) - 1ull))) & (kQueue - 1);
      q[at] = val;
      qw[at] = blk_base + pos + lane;
    }
    qt += (uint32_t)__popcll(b);
    pos += k;
    __syncthreads();
  };
  // ranges of 2^28 and more reject up to half the words: one sample per batch
  const int batch = r >= (1u << 28) ? 1 : 64;
  int64_t done = 0;
  while (done < count) {
    const int nb = count - done < batch ? (int)(count - done) : batch;
    while (qt - qh < (uint32_t)(m * nb)) produce();
    uint32_t v[8];
    bool clean = true;
    if (lane < nb) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < m) v[k] = q[(qh + (uint32_t)(m * lane + k)) & (kQueue - 1)];
#pragma unroll
      for (int i = 1; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j)
          if (i < m && v[i] == v[j]) clean = false;
    }
    const uint64_t dirty = __ballot(lane < nb && !clean);
    const int d = dirty ? __ffsll((unsigned long long)dirty) - 1 : nb;
    if (lane < d) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < m) out[(done + lane) * m + k] = (int32_t)v[k];
    }
    qh += (uint32_t)(m * d);
    done += d;
    if (d < nb) {
      // sample `done` value by value: a repeat is drawn again
      uint32_t w[8];
      int have = 0;
      while (have < m) {
        while (qt == qh) produce();
        const uint32_t val = q[qh & (kQueue - 1)];
        ++qh;
        bool rep = false;
#pragma unroll
        for (int j = 0; j < 8; ++j) rep |= j < have && w[j] == val;
        if (!rep) {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (j == have) w[j] = val;
          ++have;
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < m && lane == k) out[done * m + k] = (int32_t)w[k];
      ++done;
    }
  }
  // the generator continues after the word of the last value consumed
  const int32_t target = qw[(qh - 1) & (kQueue - 1)] + 1;
  int from = cur, npos = target - blk_base;
  if (target < blk_base && target >= blk_base - 624) {
    from = cur ^ 1;
    npos = target - (blk_base - 624);
  } else if (target < blk_base) {
    __syncthreads();
    for (int i = lane; i < 624; i += 64) xs[0][i] = s->x[i];
    __syncthreads();
    int32_t b = -s->pos;
    from = 0;
    while (target - b > 624) {
      mt_twist(xs[from], xs[from ^ 1]);
      from ^= 1;
      b += 624;
    }
    npos = target - b;
  }
  __syncthreads();
  for (int i = lane; i < 624; i += 64) s->x[i] = xs[from][i];
  if (lane == 0) s->pos = npos;
}

void launch_mt_seed(hipStream_t st, const uint32_t* seeds, const int32_t* pairs, int num, DevMT* states) {
  if (num <= 0) return;
  hipLaunchKernelGGL(k_mt_seed, dim3(num), dim3(64), 0, st, seeds, pairs, num, states);
}

void launch_sample_uniform(hipStream_t st, const DevSampleItem* items, int num_items, DevMT* states,
                           int32_t* samples) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_sample_uniform, dim3(num_items), dim3(64), 0, st, items, states, samples);
}

}  // namespace gcr
