// Host-side samplers of the engine.
//
// The reference's samplers (inc/uniform_random_generator.h:43-125,
// inc/samplers/uniform_sampler.h:97-115, inc/samplers/prosac_sampler.h:51-237)
// draw with std::mt19937 through libstdc++'s uniform_int_distribution<size_t>.
// The engine draws the same numbers (MT19937 below equals std::mt19937), so a
// given seed yields the reference's sample stream; only the seeding differs (the reference seeds from
// std::random_device, the engine from gcr_params_t::seed).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <vector>

namespace gcr {

// The 32-bit Mersenne Twister MT19937 (Matsumoto & Nishimura 1998) with the
// standard seeding, producing the same numbers as std::mt19937 (a URBG, so
// std::uniform_int_distribution draws from it exactly as from std::mt19937).
// The 624-word state is regenerated and tempered a whole block at a time, in
// loops the compiler vectorises, instead of one tempering per call.
class MT19937 {
 public:
  using result_type = uint32_t;
  static constexpr result_type min() { return 0u; }
  static constexpr result_type max() { return 0xffffffffu; }
  explicit MT19937(uint32_t seed) {
    x_[0] = seed;
    for (uint32_t i = 1; i < kN; ++i) x_[i] = 1812433253u * (x_[i - 1] ^ (x_[i - 1] >> 30)) + i;
    pos_ = kN;
  }
  result_type operator()() {
    if (pos_ >= kN) refill();
    return out_[pos_++];
  }

 private:
  static constexpr uint32_t kN = 624, kM = 397, kA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;
  void refill() {
    uint32_t i = 0;
    for (; i < kN - kM; ++i) {
      const uint32_t y = (x_[i] & kUpper) | (x_[i + 1] & kLower);
      x_[i] = x_[i + kM] ^ (y >> 1) ^ ((0u - (y & 1u)) & kA);
    }
    for (; i < kN - 1; ++i) {
      const uint32_t y = (x_[i] & kUpper) | (x_[i + 1] & kLower);
      x_[i] = x_[i + kM - kN] ^ (y >> 1) ^ ((0u - (y & 1u)) & kA);
    }
    const uint32_t y = (x_[kN - 1] & kUpper) | (x_[0] & kLower);
    x_[kN - 1] = x_[kM - 1] ^ (y >> 1) ^ ((0u - (y & 1u)) & kA);
    for (uint32_t k = 0; k < kN; ++k) {
      uint32_t z = x_[k];
      z ^= z >> 11;
      z ^= (z << 7) & 0x9d2c5680u;
      z ^= (z << 15) & 0xefc60000u;
      z ^= z >> 18;
      out_[k] = z;
    }
    pos_ = 0;
  }
  uint32_t x_[kN];
  uint32_t out_[kN];
  uint32_t pos_;
};

class RandomGenerator {
 public:
  explicit RandomGenerator(uint32_t seed) : gen_(seed) {}
  void reset(size_t lo, size_t hi) {
    dist_ = std::uniform_int_distribution<size_t>(lo, hi);
    lo_ = lo;
    span_ = (uint64_t)(hi - lo) + 1u;
    lemire_ = hi >= lo && span_ < 0xffffffffull;
    if (lemire_) thr_ = (uint32_t)(0u - (uint32_t)span_) % (uint32_t)span_;
  }
  // one draw of uniform_int_distribution<size_t>(lo, hi) over this 32-bit
  // engine: libstdc++ 11 takes Lemire's multiply-shift for spans below
  // 2^32 - 1 (the product's high word; the next word while the low word is
  // below 2^32 mod span), written out here; other spans go through the
  // library's distribution itself
  size_t draw() {
    if (!lemire_) return dist_(gen_);
    uint64_t prod = (uint64_t)gen_() * span_;
    if ((uint32_t)prod < (uint32_t)span_)
      while ((uint32_t)prod < thr_) prod = (uint64_t)gen_() * span_;
    return lo_ + (size_t)(prod >> 32);
  }
  // m unique values, rejection of repeats as the reference does (a value
  // equal to an earlier one of the sample is drawn again)
  void unique_set(size_t* out, size_t m) {
    size_t k = 0;
    while (k < m) {
      const size_t v = draw();
      bool rep = false;
      for (size_t j = 0; j < k; ++j) rep |= out[j] == v;
      if (!rep) out[k++] = v;
    }
  }

 private:
  MT19937 gen_;
  std::uniform_int_distribution<size_t> dist_;
  size_t lo_ = 0;
  uint64_t span_ = 1;
  uint32_t thr_ = 0;
  bool lemire_ = false;
};

inline uint32_t main_seed(uint64_t seed) { return (uint32_t)(seed & 0xffffffffu); }
inline uint32_t lo_seed(uint64_t seed) { return (uint32_t)((seed & 0xffffffffu) ^ 0x5bd1e995u); }

// A run of consecutive samples whose first `draws` indices are unique draws
// in [0, range) and whose last index is `fixed` when fixed >= 0.  range 0:
// the sampler refuses (no draws, every index -1).
struct SampleSeg {
  int64_t count;
  uint32_t range;
  int32_t draws;
  int32_t fixed;
};

class Sampler {
 public:
  virtual ~Sampler() {}
  // Samplers whose random draws the device can make: advance everything but
  // the random stream over `want` samples of m indices from a pool of pool_n
  // and describe them as segments; false when the sampler cannot.
  virtual bool plan(size_t want, size_t pool_n, size_t m, std::vector<SampleSeg>& segs) {
    (void)want;
    (void)pool_n;
    (void)m;
    (void)segs;
    return false;
  }
  // Draw m indices from pool[0..pool_n) (pool == nullptr: identity pool).
  // Samplers whose state the main loop's update() calls change apply that
  // update to each sample as they draw it (the loop updates every drawn
  // sample exactly once, in draw order, and the update does not depend on the
  // sample's score).
  virtual bool sample(const int32_t* pool, size_t pool_n, int32_t* out, size_t m) = 0;
  // count samples of m indices from the identity pool of pool_n into
  // out[c * m ..], a refused sample as m times -1 (the same draws as calling
  // sample() count times)
  virtual void sample_chunk(size_t pool_n, int32_t* out, size_t count, size_t m) {
    for (size_t c = 0; c < count; ++c)
      if (!sample(nullptr, pool_n, out + c * m, m))
        for (size_t k = 0; k < m; ++k) out[c * m + k] = -1;
  }
  // true when update() changes the state (the loop must then replay a chunk
  // when it updates a sample a second time, see update_again)
  virtual bool has_updates() const { return false; }
  // one more update of a drawn sample: the reference's loop updates the last
  // sample of an iteration twice when all 100 draws of the iteration failed
  // (inc/GCRANSAC.h:331-335 then :474-478)
  virtual void update_again(const int32_t* s, size_t m) {
    (void)s;
    (void)m;
  }
  virtual std::unique_ptr<Sampler> clone() const { return nullptr; }
  // true when the sampler's next samples do not depend on anything the loop
  // does meanwhile, so they may be drawn ahead of the request (on another
  // host thread, while the device works)
  virtual bool draws_ahead() const { return false; }
};

// glibc's random() (the generator behind rand()): the TYPE_3 additive
// feedback generator r[i] = r[i-3] + r[i-31] (mod 2^32) with output r[i] >> 1,
// its 31-word state seeded by the Park-Miller recurrence
// r[i] = 16807 r[i-1] mod (2^31 - 1) (Schrage's method on the signed 32-bit
// word, so seeds >= 2^31 start from a negative word) from the seed (0 taken
// as 1) and 310 outputs discarded, as srandom() does.  rand() after srand(seed) gives the
// same numbers (checked word for word against the C library).
class GlibcRandom {
 public:
  static constexpr int32_t kRandMax = 2147483647;
  explicit GlibcRandom(uint32_t seed) {
    if (seed == 0) seed = 1;
    r_[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;  // a signed 32-bit word, as srandom keeps it
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = word / 127773, lo = word % 127773;
      int64_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      word = (int32_t)w;
      r_[i] = word;
    }
    f_ = 3;
    b_ = 0;
    for (int i = 0; i < 310; ++i) next();
  }
  int32_t next() {
    const uint32_t v = (uint32_t)r_[f_] + (uint32_t)r_[b_];
    r_[f_] = (int32_t)v;
    f_ = f_ + 1 == 31 ? 0 : f_ + 1;
    b_ = b_ + 1 == 31 ? 0 : b_ + 1;
    return (int32_t)(v >> 1);
  }

 private:
  int32_t r_[31];
  int f_, b_;
};

class UniformSampler : public Sampler {
 public:
  bool draws_ahead() const override { return true; }
  UniformSampler(uint32_t seed, size_t rows) : rng_(seed) { rng_.reset(0, rows); }
  bool plan(size_t want, size_t pool_n, size_t m, std::vector<SampleSeg>& segs) override {
    segs.clear();
    if (m > pool_n) segs.push_back(SampleSeg{(int64_t)want, 0u, 0, -1});
    else segs.push_back(SampleSeg{(int64_t)want, (uint32_t)pool_n, (int32_t)m, -1});
    return true;
  }
  bool sample(const int32_t* pool, size_t pool_n, int32_t* out, size_t m) override {
    return draw_one(pool, pool_n, out, m);
  }
  void sample_chunk(size_t pool_n, int32_t* out, size_t count, size_t m) override {
    for (size_t c = 0; c < count; ++c)
      if (!draw_one(nullptr, pool_n, out + c * m, m))
        for (size_t k = 0; k < m; ++k) out[c * m + k] = -1;
  }

 private:
  bool draw_one(const int32_t* pool, size_t pool_n, int32_t* out, size_t m) {
    if (m > pool_n) return false;
    if (pool_n != range_n_) {
      rng_.reset(0, pool_n - 1);
      range_n_ = pool_n;
    }
    size_t small[16];
    if (m > 16) tmp_.resize(m);
    size_t* t = m > 16 ? tmp_.data() : small;
    rng_.unique_set(t, m);
    for (size_t i = 0; i < m; ++i) out[i] = pool ? pool[t[i]] : (int32_t)t[i];
    return true;
  }
  RandomGenerator rng_;
  std::vector<size_t> tmp_;
  size_t range_n_ = 0;
};

// PROSAC: points sorted by decreasing quality; returns raw indices.
class ProsacSampler : public Sampler {
 public:
  bool draws_ahead() const override { return true; }
  ProsacSampler(uint32_t seed, size_t n, size_t m, size_t convergence = 100000)
      : m_(m), n_(n), conv_(convergence), k_(1), growth_(growth_table(n, m, convergence)), rng_(seed) {
    largest_ = m_;
    subset_ = m_;
    rng_.reset(0, subset_ - 1);
    range_ = subset_;
  }
  bool plan(size_t want, size_t, size_t m, std::vector<SampleSeg>& segs) override {
    segs.clear();
    for (size_t c = 0; c < want; ++c) {
      SampleSeg g{1, 0u, 0, -1};
      if (m != m_) {
        increment();
        g.range = 0;  // refused: no draws, indices -1
        g.draws = 0;
      } else if (k_ > conv_) {
        g.range = (uint32_t)range_;
        g.draws = (int32_t)m_;
      } else {
        g.range = (uint32_t)range_;
        g.draws = (int32_t)m_ - 1;
        g.fixed = (int32_t)(subset_ - 1);
        increment();
      }
      if (!segs.empty() && segs.back().range == g.range && segs.back().draws == g.draws &&
          segs.back().fixed == g.fixed)
        ++segs.back().count;
      else
        segs.push_back(g);
    }
    return true;
  }
  bool sample(const int32_t*, size_t, int32_t* out, size_t m) override { return draw_one(out, m); }
  void sample_chunk(size_t, int32_t* out, size_t count, size_t m) override {
    for (size_t c = 0; c < count; ++c)
      if (!draw_one(out + c * m, m))
        for (size_t k = 0; k < m; ++k) out[c * m + k] = -1;
  }

 private:
  bool draw_one(int32_t* out, size_t m) {
    if (m != m_) {
      increment();
      return false;
    }
    size_t small[16];
    if (m > 16) tmp_.resize(m);
    size_t* t = m > 16 ? tmp_.data() : small;
    if (k_ > conv_) {
      rng_.unique_set(t, m);
      for (size_t i = 0; i < m; ++i) out[i] = (int32_t)t[i];
      return true;
    }
    rng_.unique_set(t, m_ - 1);
    for (size_t i = 0; i + 1 < m_; ++i) out[i] = (int32_t)t[i];
    out[m_ - 1] = (int32_t)(subset_ - 1);
    increment();
    return true;
  }
  void increment() {
    ++k_;
    if (k_ > conv_) {
      rng_.reset(0, n_ - 1);
      range_ = n_;
    } else if (k_ > (*growth_)[subset_ - 1]) {
      ++subset_;
      if (subset_ > n_) subset_ = n_;
      if (largest_ < subset_) largest_ = subset_;
      rng_.reset(0, subset_ - 2);
      range_ = subset_ - 1;
    }
  }
  size_t m_, n_, conv_, k_, largest_ = 0, subset_ = 0, range_ = 0;
  std::shared_ptr<const std::vector<size_t>> growth_;

  // The growth function T'_n for n = 1 .. N (prosac_sampler.h:142-164),
  // computed once per (N, m, convergence) and shared by the pairs of a batch.
  static std::shared_ptr<const std::vector<size_t>> growth_table(size_t n, size_t m, size_t conv) {
    static std::mutex mu;
    static std::map<std::array<size_t, 3>, std::shared_ptr<const std::vector<size_t>>> cache;
    std::lock_guard<std::mutex> lock(mu);
    std::shared_ptr<const std::vector<size_t>>& slot = cache[{n, m, conv}];
    if (!slot) {
      auto g = std::make_shared<std::vector<size_t>>(n, 0);
      double T_n = (double)conv;
      for (size_t i = 0; i < m; i++) T_n *= static_cast<double>(m - i) / (n - i);
      size_t T_n_prime = 1;
      for (size_t i = 0; i < n; ++i) {
        if (i + 1 <= m) {
          (*g)[i] = T_n_prime;
          continue;
        }
        const double Tn_plus1 = static_cast<double>(i + 1) * T_n / (i + 1 - m);
        (*g)[i] = T_n_prime + (unsigned int)std::ceil(Tn_plus1 - T_n);
        T_n = Tn_plus1;
        T_n_prime = (*g)[i];
      }
      slot = std::move(g);
    }
    return slot;
  }
  RandomGenerator rng_;
  std::vector<size_t> tmp_;
};

// NG-RANSAC importance sampling (inc/samplers/importance_sampler.h:51-135):
// each index of a sample is an independent draw of
// std::discrete_distribution over the point probabilities (repeats allowed),
// through the same MT19937 stream the other samplers use.
// A caller's host sampler (the template API's user-defined sampler classes,
// samplers/sampler.h:44-87) behind a C callback: fn(ctx, pool, pool_n, out,
// m) writes m indices and returns 1, or returns 0 for a refused sample.
class CallbackSampler : public Sampler {
 public:
  using Fn = int32_t (*)(void* ctx, const int32_t* pool, int64_t pool_n, int32_t* out, int32_t m);
  CallbackSampler(Fn fn, void* ctx) : fn_(fn), ctx_(ctx) {}
  bool sample(const int32_t* pool, size_t pool_n, int32_t* out, size_t m) override {
    return fn_(ctx_, pool, (int64_t)pool_n, out, (int32_t)m) != 0;
  }

 private:
  Fn fn_;
  void* ctx_;
};

class ImportanceSampler : public Sampler {
 public:
  ImportanceSampler(uint32_t seed, const double* probabilities, size_t n)
      : gen_(seed), dist_(probabilities, probabilities + n) {}
  bool sample(const int32_t*, size_t, int32_t* out, size_t m) override {
    for (size_t i = 0; i < m; ++i) out[i] = dist_(gen_);
    return true;
  }

 private:
  MT19937 gen_;
  std::discrete_distribution<int> dist_;
};

// Adaptive reordering (inc/samplers/adaptive_reordering_sampler.h:46-166): a
// max-heap of (probability, index); a sample is the heap's top m entries, and
// each sampled point's probability is then re-estimated from its Beta prior
// (a, b from the initial probability and the estimator variance) and its
// appearance count, jittered by glibc's rand() stream (seeded here from the
// run's seed; the reference's rand() is the process-global, unseeded one), and
// pushed back.  MIN / MAX follow OpenCV's macros (MIN(a, b) = a > b ? b : a,
// MAX(a, b) = a < b ? b : a).
class AdaptiveReorderingSampler : public Sampler {
 public:
  AdaptiveReorderingSampler(uint32_t seed, const double* probabilities, size_t n, double estimator_variance,
                            double randomness = 0.01)
      : rand_(seed),
        randomness_2_(randomness / 2.0),
        randomness_rand_max_(randomness / static_cast<double>(GlibcRandom::kRandMax)) {
    probs_.reserve(n);
    for (size_t i = 0; i < n; ++i) {
      double p = probabilities[i];
      if (p == 1.0) p -= 1e-6;
      const double a = p * p * (1.0 - p) / estimator_variance - p;
      const double b = a * (1.0 - p) / p;
      probs_.push_back(Entry{p, 0, a, b});
      queue_.emplace(p, i);
    }
  }
  bool sample(const int32_t*, size_t, int32_t* out, size_t m) override {
    for (size_t i = 0; i < m; ++i) {
      out[i] = (int32_t)queue_.top().second;
      queue_.pop();
    }
    update(out, m);
    return true;
  }
  bool has_updates() const override { return true; }
  void update_again(const int32_t* s, size_t m) override { update(s, m); }
  std::unique_ptr<Sampler> clone() const override { return std::make_unique<AdaptiveReorderingSampler>(*this); }
  // the rand() stream, which the SPRT test's pool shuffle draws from before
  // the first sample
  GlibcRandom& stream() { return rand_; }

 private:
  void update(const int32_t* s, size_t m) {
    for (size_t i = 0; i < m; ++i) {
      Entry& e = probs_[(size_t)s[i]];
      ++e.appearances;
      double v = std::abs(e.a / (e.a + e.b + (double)e.appearances)) + randomness_rand_max_ * (double)rand_.next() -
                 randomness_2_;
      const double mn = 0.999 > v ? v : 0.999;
      v = 0.0 < mn ? mn : 0.0;
      e.p = v;
      queue_.emplace(v, (size_t)s[i]);
    }
  }
  struct Entry {
    double p;
    size_t appearances;
    double a, b;
  };
  GlibcRandom rand_;
  double randomness_2_, randomness_rand_max_;
  std::vector<Entry> probs_;
  std::priority_queue<std::pair<double, size_t>, std::vector<std::pair<double, size_t>>> queue_;
};

}  // namespace gcr
