// The uniform and PROSAC samplers' draws on the device (uniform_sampler.h:
// 97-115, prosac_sampler.h:176-237, both through
// uniform_random_generator.h:43-125): std::mt19937 seeded with the pair's
// seed, each index drawn by libstdc++'s uniform_int_distribution<size_t> on
// [0, range - 1] (Lemire's multiply-shift with rejection: the 64-bit product
// of a 32-bit word and the range; the draw takes the next word while the low
// half of the product is below 2^32 mod range), and an index that repeats an
// earlier index of the same sample is drawn again.
//
// A pair's chunk is a list of segments of samples that share the range, the
// number of drawn indices and an optional fixed last index (PROSAC's
// newest point of the subset); the host advances PROSAC's deterministic
// schedule (subset growth, convergence) and sends the segments, the device
// draws.  A batch of 64 samples may span segments: each lane takes its
// sample's segment, and the words of lane l start after the draws of the
// lanes before it.
//
// One wavefront per pair.  The generator's 624-word state lives in LDS during
// the launch (two buffers: the current block and the one it was generated
// from) and in HBM between launches, with the position of the next word.  A
// block is regenerated into the other buffer in four lane-parallel phases
// (words 0..226 read only old words; 227..453 also new words 0..226;
// 454..622 new words 227..395; 623 new words 0 and 396) and tempered as it
// is read.  Words go through a queue, 64 at a time.  Samples are formed 64
// at a time, one per lane, each from the next `draws` queued words; the
// first lane whose words hold a rejected word or a repeated index and every
// lane after it are discarded, that sample is formed word by word with the
// rejection and repeat rules, and the next batch starts after it.  So every
// sample consumes exactly the words the sequential sampler consumes.  Words
// read ahead are given back at the end: the stored position is the word
// after the last one consumed.
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
constexpr int kQueue = 1024;  // ring of tempered words (a power of two)

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

__global__ void __launch_bounds__(64) k_sample_streams(const DevSampleItem* __restrict__ items,
                                                       const DevSampleSeg* __restrict__ segs, DevMT* __restrict__ st,
                                                       int32_t* __restrict__ samples) {
  __shared__ uint32_t xs[2][624];
  __shared__ uint32_t q[kQueue];
  const DevSampleItem it = items[blockIdx.x];
  // segments: consecutive, starts relative to the item's first sample
  const int lane = threadIdx.x;
  const int m = it.m;
  int32_t* const out = samples + it.base * m;
  if (it.count <= 0) return;
  DevMT* const s = st + it.pair;
  for (int i = lane; i < 624; i += 64) xs[0][i] = s->x[i];
  int pos = s->pos, cur = 0;
  int32_t blk_base = -pos;  // index (from this launch's first word) of word 0 of the current block
  __syncthreads();
  uint32_t qh = 0, qt = 0;  // words consumed / read (queue entry j is the launch's word j)
  // append words until the queue holds `want` unconsumed ones
  auto produce = [&](uint32_t want) {
    while (qt - qh < want) {
      if (pos >= 624) {
        mt_twist(xs[cur], xs[cur ^ 1]);
        cur ^= 1;
        pos = 0;
        blk_base += 624;
      }
      const uint32_t short_by = want - (qt - qh);
      const int k = (uint32_t)(624 - pos) < short_by ? 624 - pos : (int)short_by;
      for (int j = lane; j < k; j += 64) q[(qt + (uint32_t)j) & (kQueue - 1)] = mt_temper(xs[cur][pos + j]);
      qt += (uint32_t)k;
      pos += k;
      __syncthreads();
    }
  };
  // a window of the item's segments in LDS for batches that span segments
  __shared__ DevSampleSeg wseg[128];
  int wbase = -1;
  const DevSampleSeg* const sg = segs + it.seg_off;
  int g = 0;  // segment of sample `done`
  // the current segment's end, range, rejection threshold (2^32 mod range),
  // draws and fixed index, as wave-uniform values read once per segment
  int64_t cs_end = 0;
  uint32_t cs_range = 0, cs_thr = 0;
  int cs_draws = 0, cs_fixed = -1;
  auto take_seg = [&](int idx) {
    const DevSampleSeg* const t = sg + __builtin_amdgcn_readfirstlane(idx);
    const int64_t e = t->start + t->count;
    cs_end = (int64_t)(((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)e >> 32)) << 32) |
                       (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)e));
    cs_range = __builtin_amdgcn_readfirstlane(t->range);
    cs_draws = __builtin_amdgcn_readfirstlane(t->draws);
    cs_fixed = __builtin_amdgcn_readfirstlane(t->fixed);
    cs_thr = cs_range ? (0u - cs_range) % cs_range : 0u;
  };
  take_seg(0);
  int64_t done = 0;
  while (done < it.count) {
    while (cs_end <= done) take_seg(++g);
    const int nb = it.count - done < 64 ? (int)(it.count - done) : 64;
    // lane l forms sample done + l with its segment's parameters
    uint32_t r, thr = 0;
    int dr, fixed;
    if (cs_end >= done + nb) {
      // the whole batch in segment g
      r = cs_range;
      thr = cs_thr;
      dr = r == 0 ? 0 : cs_draws;
      fixed = cs_fixed;
    } else {
      if (wbase < 0 || g + 64 > wbase + 128) {
        __syncthreads();
        wbase = g;
        for (int j = lane; j < 128 && wbase + j < it.nseg; j += 64) wseg[j] = sg[wbase + j];
        __syncthreads();
      }
      const int64_t smp = done + (lane < nb ? lane : nb - 1);
      int gl = g - wbase;
      while (wseg[gl].start + wseg[gl].count <= smp) ++gl;
      r = wseg[gl].range;
      dr = r == 0 ? 0 : wseg[gl].draws;
      fixed = wseg[gl].fixed;
      if (r != 0) thr = (0u - r) % r;
    }
    const bool refused = r == 0;
    // ranges of 2^28 and more reject up to half the words: such a sample
    // ends its batch (no lane after it reads ahead)
    const uint64_t wide = __ballot(lane < nb && r >= (1u << 28));
    const int nbw = wide ? (__ffsll((unsigned long long)wide)) : nb;  // lanes up to the first wide one
    // words of lane l start after the draws of lanes 0 .. l - 1
    const int dr0 = __shfl(dr, 0, 64);
    int excl, need;
    if (__ballot(lane < nbw && dr != dr0) == 0) {
      excl = dr0 * (lane < nbw ? lane : nbw);
      need = dr0 * nbw;
    } else {
      int incl = lane < nbw ? dr : 0;
#pragma unroll
      for (int sft = 1; sft < 64; sft <<= 1) {
        const int o = __shfl_up(incl, sft, 64);
        if (lane >= sft) incl += o;
      }
      excl = incl - (lane < nbw ? dr : 0);
      need = __shfl(incl, 63, 64);
    }
    produce((uint32_t)need);
    uint32_t v[8];
    bool clean = true;
    if (lane < nbw) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < dr) {
          const uint64_t prod = (uint64_t)q[(qh + (uint32_t)(excl + k)) & (kQueue - 1)] * (uint64_t)r;
          clean = clean && (uint32_t)prod >= thr;
          v[k] = (uint32_t)(prod >> 32);
        }
#pragma unroll
      for (int i = 1; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j)
          if (i < dr && v[i] == v[j]) clean = false;
    }
    const uint64_t dirty = __ballot(lane < nbw && !clean);
    const int d = dirty ? __ffsll((unsigned long long)dirty) - 1 : nbw;
    if (lane < d) {
      int32_t* o = out + (done + lane) * m;
      if (refused) {
        for (int k = 0; k < m; ++k) o[k] = -1;
      } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < dr) o[k] = (int32_t)v[k];
        if (fixed >= 0) o[m - 1] = fixed;
      }
    }
    // words of lanes 0 .. d - 1 (lanes from nbw on add no draws)
    qh += (uint32_t)(d < 64 ? __shfl(excl, d, 64) : need);
    done += d;
    if (d < nbw) {
      // sample `done` word by word with lane d's parameters: a rejected word
      // or a repeat is drawn again
      const uint32_t rd = __shfl(r, d, 64), thd = __shfl(thr, d, 64);
      const int drd = __shfl(dr, d, 64), fxd = __shfl(fixed, d, 64);
      uint32_t w[8];
      int have = 0;
      while (have < drd) {
        if (qt == qh) produce(64);
        const uint64_t prod = (uint64_t)q[qh & (kQueue - 1)] * (uint64_t)rd;
        ++qh;
        if ((uint32_t)prod < thd) continue;
        const uint32_t val = (uint32_t)(prod >> 32);
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
      int32_t* o = out + done * m;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < drd && lane == k) o[k] = (int32_t)w[k];
      if (lane == 0 && fxd >= 0) o[m - 1] = fxd;
      ++done;
    }
  }
  // the generator continues after the last word consumed.  At most one
  // batch's draws (<= 8 * 64 < 624 words) are read ahead, so that word lies
  // in the current block or the one before it.
  const int32_t target = (int32_t)qh;
  int from = cur, npos = target - blk_base;
  if (target < blk_base) {
    from = cur ^ 1;
    npos = target - (blk_base - 624);
  }
  __syncthreads();
  for (int i = lane; i < 624; i += 64) s->x[i] = xs[from][i];
  if (lane == 0) s->pos = npos;
}

void launch_mt_seed(hipStream_t st, const uint32_t* seeds, const int32_t* pairs, int num, DevMT* states) {
  if (num <= 0) return;
  hipLaunchKernelGGL(k_mt_seed, dim3(num), dim3(64), 0, st, seeds, pairs, num, states);
}

void launch_sample_streams(hipStream_t st, const DevSampleItem* items, const DevSampleSeg* segs, int num_items,
                           DevMT* states, int32_t* samples) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_sample_streams, dim3(num_items), dim3(64), 0, st, items, segs, states, samples);
}

}  // namespace gcr
