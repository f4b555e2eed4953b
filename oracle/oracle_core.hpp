// CPU oracle for the Graph-Cut RANSAC hot path -- TEST INFRASTRUCTURE ONLY.
//
// This file is a plain C++17 restatement of the reference algorithm
// (danini/graph-cut-ransac) used as the parity checker for the HIP engine and
// as the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it.  The shipped engine never links it.
//
// Citations are relative to the reference repository root, with
//   inc/ = src/pygcransac/include/
//   py/  = src/pygcransac/src/
//
// Shared pieces: point container, model type, seeded samplers, grid
// neighbourhood, MSAC scoring, small dense linear algebra and the exact
// integer min-cut used for the graph-cut labelling.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <random>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Data containers
// ---------------------------------------------------------------------------
// N x cols row-major FP64 matrix: the reference's cv::Mat(N, 4, CV_64F)
// (py/gcransac_python.cpp:4274).
struct Points {
  const double* data = nullptr;
  size_t rows = 0;
  size_t cols = 4;
  const double* row(size_t i) const { return data + i * cols; }
};

// 3x3 model descriptor stored row-major: m[3*i+j] == descriptor(i, j)
// (inc/model.h:43-58 holds an Eigen::MatrixXd; element access is logical).
struct Model {
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double& operator()(int i, int j) { return m[3 * i + j]; }
  double operator()(int i, int j) const { return m[3 * i + j]; }
};

// inc/scoring_function.h:48-72.  operator< compares value only.
struct Score {
  size_t inlier_number = 0;
  double value = 0.0;
  bool operator<(const Score& o) const { return value < o.value; }
};

// inc/settings.h:44-86 (defaults), trimmed to the fields the loop reads.
struct Settings {
  bool do_final_iterated_least_squares = true;
  bool do_local_optimization = true;
  double desired_fps = -1;
  size_t max_local_optimization_number = 10;
  size_t max_graph_cut_number = 10;
  size_t max_least_squares_iterations = 10;
  size_t min_iteration_number_before_lo = 20;
  size_t min_iteration_number = 20;
  size_t max_iteration_number = std::numeric_limits<size_t>::max();
  size_t max_unsuccessful_model_generations = 100;
  double spatial_coherence_weight = 0.14;
  double threshold = 2.0;
  double confidence = 0.95;
};

// inc/statistics.h:43-71 (fields used by the loop).
struct Statistics {
  size_t iteration_number = 0;
  size_t graph_cut_number = 0;
  size_t local_optimization_number = 0;
  size_t accepted_models = 0;
  size_t rejected_models = 0;
  double score = 0.0;
  std::vector<size_t> inliers;
};

// ---------------------------------------------------------------------------
// Random sampling
// ---------------------------------------------------------------------------
// inc/uniform_random_generator.h:43-125.  The reference seeds std::mt19937
// from std::random_device (:51-54); the restatement takes an explicit seed so
// runs are reproducible.  Draws go through libstdc++'s
// uniform_int_distribution<size_t>, exactly as the reference does.
class UniformRandomGenerator {
 public:
  explicit UniformRandomGenerator(uint32_t seed) : generator(seed) {}
  void resetGenerator(size_t lo, size_t hi) {
    generate = std::uniform_int_distribution<size_t>(lo, hi);
  }
  size_t draw() { return generate(generator); }
  // :90-103
  void generateUniqueRandomSet(size_t* sample, size_t sample_size, size_t max_) {
    resetGenerator(0, max_);
    for (size_t i = 0; i < sample_size; i++) {
      sample[i] = generate(generator);
      for (int j = (int)i - 1; j >= 0; j--) {
        if (sample[i] == sample[j]) {
          i--;
          break;
        }
      }
    }
  }
  // :73-88 (no explicit range; uses the current distribution)
  void generateUniqueRandomSet(size_t* sample, size_t sample_size) {
    for (size_t i = 0; i < sample_size; i++) {
      sample[i] = generate(generator);
      for (int j = (int)i - 1; j >= 0; j--) {
        if (sample[i] == sample[j]) {
          i--;
          break;
        }
      }
    }
  }

 private:
  std::mt19937 generator;
  std::uniform_int_distribution<size_t> generate;
};

struct Sampler {
  virtual ~Sampler() {}
  virtual bool sample(const std::vector<size_t>& pool, size_t* subset, size_t m) = 0;
  // inc/samplers/sampler.h: called by the main loop after every drawn sample
  virtual void update(const size_t* subset, size_t m) {
    (void)subset;
    (void)m;
  }
};

// inc/samplers/uniform_sampler.h:97-115
class UniformSampler : public Sampler {
 public:
  UniformSampler(uint32_t seed, size_t rows) : gen(seed) { gen.resetGenerator(0, rows); }
  bool sample(const std::vector<size_t>& pool, size_t* subset, size_t m) override {
    if (m > pool.size()) return false;
    gen.generateUniqueRandomSet(subset, m, pool.size() - 1);
    for (size_t i = 0; i < m; ++i) subset[i] = pool[subset[i]];
    return true;
  }

 private:
  UniformRandomGenerator gen;
};

// inc/samplers/prosac_sampler.h:51-237.  Points are assumed sorted by
// decreasing quality; the sampler returns raw indices (the pool is ignored,
// as in the reference).
class ProsacSampler : public Sampler {
 public:
  ProsacSampler(uint32_t seed, size_t point_number, size_t sample_size,
                size_t ransac_convergence_iterations = 100000)
      : sample_size_(sample_size),
        point_number_(point_number),
        ransac_convergence_iterations_(ransac_convergence_iterations),
        kth_sample_number_(1),
        gen_(seed) {
    // :129-174 growth function
    growth_function_.assign(point_number_, 0);
    double T_n = (double)ransac_convergence_iterations_;
    for (size_t i = 0; i < sample_size_; i++)
      T_n *= static_cast<double>(sample_size_ - i) / (point_number_ - i);
    size_t T_n_prime = 1;
    for (size_t i = 0; i < point_number_; ++i) {
      if (i + 1 <= sample_size_) {
        growth_function_[i] = T_n_prime;
        continue;
      }
      double Tn_plus1 = static_cast<double>(i + 1) * T_n / (i + 1 - sample_size_);
      growth_function_[i] = T_n_prime + (unsigned int)std::ceil(Tn_plus1 - T_n);
      T_n = Tn_plus1;
      T_n_prime = growth_function_[i];
    }
    largest_sample_size_ = sample_size_;
    subset_size_ = sample_size_;
    gen_.resetGenerator(0, subset_size_ - 1);
  }
  const std::vector<size_t>& growthFunction() const { return growth_function_; }

  bool sample(const std::vector<size_t>& /*pool*/, size_t* subset, size_t m) override {
    if (m != sample_size_) {
      incrementIterationNumber();
      return false;
    }
    if (kth_sample_number_ > ransac_convergence_iterations_) {
      gen_.generateUniqueRandomSet(subset, m);
      return true;
    }
    gen_.generateUniqueRandomSet(subset, sample_size_ - 1);
    subset[sample_size_ - 1] = subset_size_ - 1;
    incrementIterationNumber();
    return true;
  }

 private:
  // :66-88
  void incrementIterationNumber() {
    ++kth_sample_number_;
    if (kth_sample_number_ > ransac_convergence_iterations_)
      gen_.resetGenerator(0, point_number_ - 1);
    else if (kth_sample_number_ > growth_function_[subset_size_ - 1]) {
      ++subset_size_;
      if (subset_size_ > point_number_) subset_size_ = point_number_;
      if (largest_sample_size_ < subset_size_) largest_sample_size_ = subset_size_;
      gen_.resetGenerator(0, subset_size_ - 2);
    }
  }
  size_t sample_size_, point_number_, ransac_convergence_iterations_, kth_sample_number_;
  size_t largest_sample_size_ = 0, subset_size_ = 0;
  std::vector<size_t> growth_function_;
  UniformRandomGenerator gen_;
};

// Seeds of the two samplers of one run.  The reference draws both from
// std::random_device; the restatement fixes them as a function of one seed.
// inc/samplers/importance_sampler.h:51-135 (NG-RANSAC importance sampling):
// every index of a sample is an independent draw of std::discrete_distribution
// over the given point probabilities (repeats allowed), from std::mt19937.
class ImportanceSampler : public Sampler {
 public:
  ImportanceSampler(uint32_t seed, const double* probabilities, size_t n)
      : generator(seed), dist(probabilities, probabilities + n) {}
  bool sample(const std::vector<size_t>&, size_t* subset, size_t m) override {
    for (size_t i = 0; i < m; ++i) subset[i] = (size_t)dist(generator);
    return true;
  }

 private:
  std::mt19937 generator;
  std::discrete_distribution<int> dist;
};

// inc/samplers/adaptive_reordering_sampler.h:46-166: points ordered by a
// running inlier-probability estimate (a max-heap of (probability, index));
// a sample is the heap's top m entries; update() re-estimates each sampled
// point's probability from its Beta prior (a, b from the initial probability
// and the estimator variance) and its appearance count, jittered by the C
// library's rand(), and pushes it back.  The reference calls rand() without
// seeding it; here srand(seed) is called when the sampler is built, so a run
// is reproducible (rand() is the C library's, global state).  MIN / MAX are
// OpenCV's macros: MIN(a, b) = a > b ? b : a, MAX(a, b) = a < b ? b : a.
class AdaptiveReorderingSampler : public Sampler {
 public:
  AdaptiveReorderingSampler(uint32_t seed, const double* probabilities, size_t n, double estimator_variance,
                            double randomness = 0.01)
      : randomness_2(randomness / 2.0), randomness_rand_max(randomness / static_cast<double>(RAND_MAX)) {
    srand(seed);
    probs.reserve(n);
    for (size_t i = 0; i < n; ++i) {
      double p = probabilities[i];
      if (p == 1.0) p -= 1e-6;
      const double a = p * p * (1.0 - p) / estimator_variance - p;
      const double b = a * (1.0 - p) / p;
      probs.push_back(Entry{p, 0, a, b});
      queue.emplace(p, i);
    }
  }
  bool sample(const std::vector<size_t>&, size_t* subset, size_t m) override {
    for (size_t i = 0; i < m; ++i) {
      subset[i] = queue.top().second;
      queue.pop();
    }
    return true;
  }
  void update(const size_t* subset, size_t m) override {
    for (size_t i = 0; i < m; ++i) {
      Entry& e = probs[subset[i]];
      ++e.appearances;
      double v = std::abs(e.a / (e.a + e.b + (double)e.appearances)) + randomness_rand_max * (double)rand() -
                 randomness_2;
      const double mn = 0.999 > v ? v : 0.999;
      v = 0.0 < mn ? mn : 0.0;
      e.p = v;
      queue.emplace(v, subset[i]);
    }
  }

 private:
  struct Entry {
    double p;
    size_t appearances;
    double a, b;
  };
  std::vector<Entry> probs;
  std::priority_queue<std::pair<double, size_t>, std::vector<std::pair<double, size_t>>> queue;
  double randomness_2, randomness_rand_max;
};

inline uint32_t main_sampler_seed(uint64_t seed) { return (uint32_t)(seed & 0xffffffffu); }
inline uint32_t lo_sampler_seed(uint64_t seed) {
  return (uint32_t)((seed & 0xffffffffu) ^ 0x5bd1e995u);
}

// ---------------------------------------------------------------------------
// Grid neighbourhood: inc/neighborhood/grid_neighborhood_graph.h:229-301
// ---------------------------------------------------------------------------
// The neighbour lists the graph-cut labelling walks
// (inc/neighborhood/neighborhood_graph.h:45-99).
class Neighborhood {
 public:
  virtual ~Neighborhood() {}
  virtual const std::vector<size_t>& getNeighbors(size_t i) const = 0;
  virtual size_t neighborNumber() const = 0;
};

// Cell index along axis d = floor(x_d / cell_size_d) converted to size_t;
// key = sum_d idx_d * n^d.  Points sharing a key form a clique.  The cell
// member lists keep insertion (= ascending point) order.  The cells live in
// the same container type as the reference's (an unordered_map from key to
// (points, axis indices of the cell's first point), filled in point order),
// so iterating it meets the cells in the reference's order, which the space
// partitioning selector depends on.
class GridNeighborhood : public Neighborhood {
 public:
  using CellMap = std::unordered_map<size_t, std::pair<std::vector<size_t>, std::vector<size_t>>>;
  GridNeighborhood() {}
  GridNeighborhood(const Points& pts, const double cell_sizes[4], size_t divisions) {
    initialize(pts, cell_sizes, divisions);
  }
  void initialize(const Points& pts, const double cell_sizes[4], size_t divisions) {
    cells_.clear();
    cell_of_.assign(pts.rows, 0);
    neighbor_number_ = 0;
    divisions_ = divisions;
    cell_sizes_.assign(cell_sizes, cell_sizes + 4);
    std::vector<size_t> idx(4);
    for (size_t r = 0; r < pts.rows; ++r) {
      const double* p = pts.row(r);
      size_t key = 0, offset = 1;
      for (int d = 0; d < 4; ++d) {
        idx[(size_t)d] = (size_t)(long long)std::floor(p[d] / cell_sizes[d]);
        key += offset * idx[(size_t)d];
        offset *= divisions;
      }
      auto& cell = cells_[key];
      if (cell.second.empty()) cell.second = idx;
      cell.first.push_back(r);
      cell_of_[r] = key;
    }
    for (auto& kv : cells_) {
      const size_t n = kv.second.first.size();
      neighbor_number_ += n * (n - 1) / 2;
    }
  }
  const std::vector<size_t>& getNeighbors(size_t i) const override { return cells_.at(cell_of_[i]).first; }
  size_t cellKey(size_t i) const { return cell_of_[i]; }
  size_t filledCellNumber() const { return cells_.size(); }
  size_t neighborNumber() const override { return neighbor_number_; }
  bool empty() const { return cells_.empty(); }
  const CellMap& getCells() const { return cells_; }
  const std::vector<double>& getCellSizes() const { return cell_sizes_; }
  size_t getDivisionNumber() const { return divisions_; }

 private:
  CellMap cells_;
  std::vector<size_t> cell_of_;
  size_t neighbor_number_ = 0;
  size_t divisions_ = 0;
  std::vector<double> cell_sizes_;
};

// ---------------------------------------------------------------------------
// Radius neighbourhood: inc/neighborhood/flann_neighborhood_graph.h:100-139
// ---------------------------------------------------------------------------
// The reference converts the (x1, y1, x2, y2) rows to float and calls
// OpenCV's FlannBasedMatcher::radiusMatch of the set against itself with
// radius r.  OpenCV passes r*r (in float) to FLANN's radius search over
// squared L2 distances in float, d2 = ((dx0^2 + dx1^2) + dx2^2) + dx3^2, which
// keeps the points with d2 <= r*r, ordered by (d2, index); the reference
// drops each list's first entry (the point itself, or an equal point of lower
// index) and keeps the rest.  FLANN answers the search approximately (four
// randomised kd-trees whose splits come from OpenCV's process-wide RNG, six
// leaf checks per query); this is the search it approximates: every point
// within the radius, found exhaustively.  The graph is initialised when at
// least one list is non-empty (:131-138).
class RadiusNeighborhood : public Neighborhood {
 public:
  RadiusNeighborhood() {}
  RadiusNeighborhood(const Points& pts, double radius) { initialize(pts, radius); }

  // FLANN's L2<float> distance: the difference of the float coordinates,
  // squared and summed left to right in float.
  static float squared_distance(const float* a, const float* b) {
    const float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2], d3 = a[3] - b[3];
    float s = d0 * d0 + d1 * d1;
    s = s + d2 * d2;
    s = s + d3 * d3;
    return 0.0f + s;
  }
  static float squared_radius(double radius) {
    const float r = (float)radius;
    return r * r;
  }

  void initialize(const Points& pts, double radius) {
    const size_t n = pts.rows;
    std::vector<float> f(4 * n);
    for (size_t i = 0; i < n; ++i)
      for (int d = 0; d < 4; ++d) f[4 * i + d] = (float)pts.row(i)[d];
    const float r2 = squared_radius(radius);
    lists_.assign(n, {});
    neighbor_number_ = 0;
    std::vector<std::pair<float, size_t>> found;
    for (size_t i = 0; i < n; ++i) {
      found.clear();
      for (size_t j = 0; j < n; ++j) {
        const float d2 = squared_distance(&f[4 * i], &f[4 * j]);
        if (d2 <= r2) found.emplace_back(d2, j);
      }
      std::sort(found.begin(), found.end());
      if (found.empty()) continue;
      lists_[i].reserve(found.size() - 1);
      for (size_t k = 1; k < found.size(); ++k) lists_[i].push_back(found[k].second);
      neighbor_number_ += found.size() - 1;
    }
  }
  const std::vector<size_t>& getNeighbors(size_t i) const override { return lists_[i]; }
  size_t neighborNumber() const override { return neighbor_number_; }
  bool initialized() const { return neighbor_number_ > 0; }

 private:
  std::vector<std::vector<size_t>> lists_;
  size_t neighbor_number_ = 0;
};

// ---------------------------------------------------------------------------
// Small dense linear algebra
// ---------------------------------------------------------------------------
// 3x3 product, element (i,j) = (a(i,0)b(0,j) + a(i,1)b(1,j)) + a(i,2)b(2,j).
inline void mat3_mul(const double* a, const double* b, double* c) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      t[3 * i + j] = a[3 * i + 0] * b[0 + j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  std::memcpy(c, t, sizeof(t));
}

inline void mat3_transpose(const double* a, double* t) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[3 * j + i] = a[3 * i + j];
}

// 3x3 inverse by the adjugate (cofactor) formula.
inline void mat3_inverse(const double* a, double* inv) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / det;
  double t[9];
  t[0] = c00 * id;
  t[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  t[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  t[3] = c01 * id;
  t[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  t[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  t[6] = c02 * id;
  t[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  t[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  std::memcpy(inv, t, sizeof(t));
}

// Determinant by partial-pivot LU, as Eigen does for dynamic-size matrices
// (inc/estimators/homography_estimator.h:334-335 calls MatrixXd::determinant).
inline double det_partial_pivot(const double* a_in, int n) {
  double a[16 * 16];
  std::memcpy(a, a_in, sizeof(double) * n * n);
  double det = 1.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > best) {
        best = std::fabs(a[i * n + k]);
        p = i;
      }
    if (best == 0.0) return 0.0;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      det = -det;
    }
    const double piv = a[k * n + k];
    det *= piv;
    for (int i = k + 1; i < n; ++i) {
      const double f = a[i * n + k] / piv;
      for (int j = k + 1; j < n; ++j) a[i * n + j] = a[i * n + j] - f * a[k * n + j];
    }
  }
  return det;
}

// Pivoting Gauss elimination on [A | b], _Size x (_Size+1), row-major.
// Follows inc/math_utils.h:45-87 operation by operation: an up-front row
// pre-sort by |a_ii|, elimination without further pivoting, back-substitution.
template <int S>
inline void gaussElimination(double (&m)[S][S + 1], double (&x)[S]) {
  int i, j, k;
  for (i = 0; i < S; i++)
    for (k = i + 1; k < S; k++)
      if (std::fabs(m[i][i]) < std::fabs(m[k][i]))
        for (j = 0; j <= S; j++) {
          const double t = m[i][j];
          m[i][j] = m[k][j];
          m[k][j] = t;
        }
  for (i = 0; i < S - 1; i++)
    for (k = i + 1; k < S; k++) {
      const double t = m[k][i] / m[i][i];
      for (j = 0; j <= S; j++) m[k][j] = m[k][j] - t * m[i][j];
    }
  for (i = S - 1; i >= 0; i--) {
    x[i] = m[i][S];
    for (j = i + 1; j < S; j++) x[i] = x[i] - m[i][j] * x[j];
    x[i] = x[i] / m[i][i];
  }
}

// Summation order of every reduction the restatement performs inside a
// library call of the reference (Eigen's QR, products, norms), where the
// reference's own order is internal to Eigen: term i goes to partial i mod 64
// (ascending i), then partials combine as p[l] += p[l + s] for s = 32..1.
// Reductions the reference writes out as loops keep the reference's
// sequential order instead.
template <class F>
inline double sum64(size_t begin, size_t end, F term) {
  double p[64];
  for (int l = 0; l < 64; ++l) p[l] = 0.0;
  for (size_t i = begin; i < end; ++i) p[i & 63] += term(i);
  for (int s = 32; s >= 1; s >>= 1)
    for (int l = 0; l < s; ++l) p[l] += p[l + s];
  return p[0];
}

// sum64 of K terms at once: term(i, out) writes the K terms of index i.
// Equivalent to K separate sum64 calls.
template <int K, class F>
inline void sum64_multi(size_t begin, size_t end, F term, double* out) {
  double p[K][64];
  for (int k = 0; k < K; ++k)
    for (int l = 0; l < 64; ++l) p[k][l] = 0.0;
  double t[K];
  for (size_t i = begin; i < end; ++i) {
    term(i, t);
    for (int k = 0; k < K; ++k) p[k][i & 63] += t[k];
  }
  for (int k = 0; k < K; ++k) {
    for (int s = 32; s >= 1; s >>= 1)
      for (int l = 0; l < s; ++l) p[k][l] += p[k][l + s];
    out[k] = p[k][0];
  }
}

// Summation order of the fundamental-matrix fits' sums (eight-point A^T A,
// LM cost and Jacobian sums): term i goes to partial i mod 256 (ascending
// i), then p[l] += p[l + s] for s = 128..1.  Partials are never -0.0 (they
// start at +0.0), so for up to 128 terms this is bit-identical to sum64.
template <class F>
inline double sum256(size_t begin, size_t end, F term) {
  double p[256];
  for (int l = 0; l < 256; ++l) p[l] = 0.0;
  for (size_t i = begin; i < end; ++i) p[i & 255] += term(i);
  for (int s = 128; s >= 1; s >>= 1)
    for (int l = 0; l < s; ++l) p[l] += p[l + s];
  return p[0];
}

template <int K, class F>
inline void sum256_multi(size_t begin, size_t end, F term, double* out) {
  thread_local std::vector<double> p;
  p.assign((size_t)K * 256, 0.0);
  double t[K];
  for (size_t i = begin; i < end; ++i) {
    term(i, t);
    for (int k = 0; k < K; ++k) p[(size_t)k * 256 + (i & 255)] += t[k];
  }
  for (int k = 0; k < K; ++k) {
    double* q = p.data() + (size_t)k * 256;
    for (int s = 128; s >= 1; s >>= 1)
      for (int l = 0; l < s; ++l) q[l] += q[l + s];
    out[k] = q[0];
  }
}

// Least squares min ||A x - b|| by Householder QR with column pivoting
// (the role Eigen::ColPivHouseholderQR::solve plays at
// inc/estimators/solver_homography_four_point.h:255-256).  A is m x n,
// column-major (column j at A + j*m).  Pivot: the column of largest remaining
// squared norm (first on ties), norms recomputed after each step.  Rank
// cut-off as Eigen's default: |R_kk| <= eps * min(m,n) * |R_00| ends the rank;
// the solution entries beyond the rank are zero.  A and b are overwritten.
// Column sums use the sum256 order (the sum64 bits for up to 128 rows).
inline void colpiv_qr_solve(double* A, double* b, int m, int n, double* x) {
  int perm[16];
  double norms[16];
  for (int j = 0; j < n; ++j) {
    perm[j] = j;
    const double* c = A + (size_t)j * m;
    norms[j] = sum256(0, m, [&](size_t i) { return c[i] * c[i]; });
  }
  const int kmax = std::min(m, n);
  for (int k = 0; k < kmax; ++k) {
    int p = k;
    for (int j = k + 1; j < n; ++j)
      if (norms[j] > norms[p]) p = j;
    if (p != k) {
      for (int i = 0; i < m; ++i) std::swap(A[(size_t)k * m + i], A[(size_t)p * m + i]);
      std::swap(norms[k], norms[p]);
      std::swap(perm[k], perm[p]);
    }
    double* col = A + (size_t)k * m;
    const double alpha = col[k];
    const double sigma = sum256(k + 1, m, [&](size_t i) { return col[i] * col[i]; });
    double tau = 0.0, beta = alpha;
    if (sigma != 0.0) {
      const double nrm = std::sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
      for (int i = k + 1; i < m; ++i) col[i] = col[i] * scale;
      col[k] = beta;
    }
    if (tau != 0.0) {
      for (int j = k + 1; j < n; ++j) {
        double* cj = A + (size_t)j * m;
        const double w = cj[k] + sum256(k + 1, m, [&](size_t i) { return col[i] * cj[i]; });
        const double tw = tau * w;
        cj[k] = cj[k] - tw;
        for (int i = k + 1; i < m; ++i) cj[i] = cj[i] - tw * col[i];
      }
      const double w = b[k] + sum256(k + 1, m, [&](size_t i) { return col[i] * b[i]; });
      const double tw = tau * w;
      b[k] = b[k] - tw;
      for (int i = k + 1; i < m; ++i) b[i] = b[i] - tw * col[i];
    }
    for (int j = k + 1; j < n; ++j) {
      const double* cj = A + (size_t)j * m;
      norms[j] = sum256(k + 1, m, [&](size_t i) { return cj[i] * cj[i]; });
    }
  }
  const double thr = std::numeric_limits<double>::epsilon() * (double)kmax;
  const double r00 = std::fabs(A[0]);
  int rank = 0;
  for (int k = 0; k < kmax; ++k) {
    if (std::fabs(A[(size_t)k * m + k]) > thr * r00) ++rank;
    else break;
  }
  double z[16];
  for (int k = 0; k < n; ++k) z[k] = 0.0;
  for (int k = rank - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < rank; ++j) s = s - A[(size_t)j * m + k] * z[j];
    z[k] = s / A[(size_t)k * m + k];
  }
  for (int k = 0; k < n; ++k) x[perm[k]] = z[k];
}

// ---------------------------------------------------------------------------
// Exact min-cut for the graph-cut labelling
// ---------------------------------------------------------------------------
// The reference minimises the energy with Boykov-Kolmogorov max-flow in FP64
// (inc/maxflow.cpp:472-604) and labels a node SINK iff it is in the sink
// search tree at termination (inc/graph.h:478-488): exactly the nodes that
// still reach the sink through residual arcs.  That set is the same for every
// maximum flow, so the restatement computes it with a different max-flow
// algorithm (Dinic) on capacities quantised to 2^-S fixed point, S chosen from
// the graph size so no sum can overflow, then a reverse BFS from the sink.
inline int mincut_scale_bits(size_t nodes, size_t edges) {
  // Every capacity is <= 1 + (node degree); total <= nodes + 3*edges + 1.
  const double bound = (double)nodes + 3.0 * (double)edges + 1.0;
  int lg = 0;
  while (std::ldexp(1.0, lg) < bound) ++lg;
  return 61 - lg;
}
// Positive capacities never round to zero, so the set of arcs with positive
// residual capacity is the same as in FP64.
inline int64_t quantise_cap(double c, int S) {
  if (!(c > 0.0)) return 0;
  const int64_t q = (int64_t)std::llrint(std::ldexp(c, S));
  return q < 1 ? 1 : q;
}

class DinicMinCut {
 public:
  // Nodes 0..n-1 are the variables, n = source, n+1 = sink.
  explicit DinicMinCut(int n) : n_(n), head_(n + 2, -1) {}
  void add_arc_pair(int u, int v, int64_t cap_uv, int64_t cap_vu) {
    to_.push_back(v); cap_.push_back(cap_uv); nxt_.push_back(head_[u]); head_[u] = (int)to_.size() - 1;
    to_.push_back(u); cap_.push_back(cap_vu); nxt_.push_back(head_[v]); head_[v] = (int)to_.size() - 1;
  }
  int64_t maxflow() {
    const int s = n_, t = n_ + 1;
    int64_t flow = 0;
    std::vector<int> level(n_ + 2), it(n_ + 2), q(n_ + 2);
    while (true) {
      std::fill(level.begin(), level.end(), -1);
      int qh = 0, qt = 0;
      q[qt++] = s;
      level[s] = 0;
      while (qh < qt) {
        const int u = q[qh++];
        for (int a = head_[u]; a != -1; a = nxt_[a])
          if (cap_[a] > 0 && level[to_[a]] < 0) {
            level[to_[a]] = level[u] + 1;
            q[qt++] = to_[a];
          }
      }
      if (level[t] < 0) break;
      for (int i = 0; i < n_ + 2; ++i) it[i] = head_[i];
      while (int64_t f = dfs(s, t, std::numeric_limits<int64_t>::max(), level, it)) flow += f;
    }
    return flow;
  }
  // SINK labels: nodes that reach the sink in the residual graph.
  std::vector<uint8_t> sink_side() const {
    const int t = n_ + 1;
    std::vector<uint8_t> mark(n_ + 2, 0);
    std::vector<int> q;
    q.push_back(t);
    mark[t] = 1;
    for (size_t qh = 0; qh < q.size(); ++qh) {
      const int v = q[qh];
      // arcs (u -> v) with residual > 0 are the sisters of v's out-arcs
      for (int a = head_[v]; a != -1; a = nxt_[a]) {
        const int u = to_[a];
        if (!mark[u] && cap_[a ^ 1] > 0) {
          mark[u] = 1;
          q.push_back(u);
        }
      }
    }
    mark.resize(n_);
    return mark;
  }

 private:
  int64_t dfs(int u, int t, int64_t f, std::vector<int>& level, std::vector<int>& it) {
    if (u == t) return f;
    for (int& a = it[u]; a != -1; a = nxt_[a]) {
      const int v = to_[a];
      if (cap_[a] > 0 && level[v] == level[u] + 1) {
        const int64_t d = dfs(v, t, std::min(f, cap_[a]), level, it);
        if (d > 0) {
          cap_[a] -= d;
          cap_[a ^ 1] += d;
          return d;
        }
      }
    }
    return 0;
  }
  int n_;
  std::vector<int> head_, to_, nxt_;
  std::vector<int64_t> cap_;
};

// One graph-cut problem as the reference builds it in GCRANSAC::labeling
// (inc/GCRANSAC.h:913-1021) through Energy/Graph (inc/energy.h:204-254,
// inc/graph.h:388-426): per-node terminal residual tr_cap (source - sink)
// accumulated with add_tweights' exact operation sequence, and one arc pair
// per same-cell point pair p<q.
struct GraphCutProblem {
  std::vector<double> tr_cap;   // per node
  std::vector<int> edge_p, edge_q;
  std::vector<double> cap_pq, cap_qp;
  double flow_const = 0.0;      // Graph::flow accumulated by add_tweights
};

// inc/graph.h:388-397
inline void add_tweights(GraphCutProblem& g, int i, double cap_source, double cap_sink) {
  const double delta = g.tr_cap[i];
  if (delta > 0) cap_source += delta;
  else cap_sink -= delta;
  g.flow_const += (cap_source < cap_sink) ? cap_source : cap_sink;
  g.tr_cap[i] = cap_source - cap_sink;
}

// General add_term2 of inc/energy.h:211-254 (all three regularity cases).
inline void add_term2(GraphCutProblem& g, int x, int y, double A, double B, double C, double D) {
  add_tweights(g, x, D, A);
  B -= A;
  C -= D;
  if (B < 0) {
    add_tweights(g, x, 0, B);
    add_tweights(g, y, 0, -B);
    g.edge_p.push_back(x); g.edge_q.push_back(y); g.cap_pq.push_back(0); g.cap_qp.push_back(B + C);
  } else if (C < 0) {
    add_tweights(g, x, 0, -C);
    add_tweights(g, y, 0, C);
    g.edge_p.push_back(x); g.edge_q.push_back(y); g.cap_pq.push_back(B + C); g.cap_qp.push_back(0);
  } else {
    g.edge_p.push_back(x); g.edge_q.push_back(y); g.cap_pq.push_back(B); g.cap_qp.push_back(C);
  }
}

// Solve and return the SINK labels (1 = SINK = inlier).
inline std::vector<uint8_t> solve_graph_cut(const GraphCutProblem& g) {
  const int n = (int)g.tr_cap.size();
  const int S = mincut_scale_bits((size_t)n, g.edge_p.size());
  DinicMinCut mc(n);
  for (int i = 0; i < n; ++i) {
    const double c = g.tr_cap[i];
    if (c > 0) mc.add_arc_pair(n, i, quantise_cap(c, S), 0);
    else if (c < 0) mc.add_arc_pair(i, n + 1, quantise_cap(-c, S), 0);
  }
  for (size_t e = 0; e < g.edge_p.size(); ++e)
    mc.add_arc_pair(g.edge_p[e], g.edge_q[e], quantise_cap(g.cap_pq[e], S), quantise_cap(g.cap_qp[e], S));
  mc.maxflow();
  return mc.sink_side();
}

}  // namespace orc
