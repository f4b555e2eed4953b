// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// SPRT preemptive model verification: inc/preemption/preemption_sprt.h:
// 206-379, restated.
//
// Two inputs of the reference are not reproducible and are fixed here:
//  - t_M.  The reference times one minimal solve and m residuals with the
//    wall clock, but divides the residual time by m and then by the solve time
//    without re-reading the clock in between (:121-132), so t_M = (x / m) / x
//    for the solve time x: 1/m up to the last bit, which is what is used.
//    (When the first sample yields no model the reference keeps the raw solve
//    time in milliseconds; that case is not restated.)
//  - the random point pool.  The reference shuffles it with the C library's
//    rand() (:162-176), whose stream nothing seeds; here the stream is seeded
//    with the main sampler's seed when the main sampler is built (the adaptive
//    reordering sampler seeds it itself and later draws continue it).
#pragma once

#include <cmath>
#include <cstdlib>
#include <vector>

#include "oracle_core.hpp"

namespace orc {

class SprtPreemption {
 public:
  struct History {
    double epsilon, delta, A;
    size_t k;
  };

  // point_number >= sample_size (the drivers check the row count first).
  SprtPreemption(size_t point_number, size_t sample_size, size_t max_solutions, double minimum_inlier_ratio)
      : t_M_(1.0 / (double)sample_size), m_S_((double)max_solutions), n_(point_number) {
    pool_.resize(n_);
    for (size_t i = 0; i < n_; ++i) pool_[i] = i;
    int max = (int)n_;
    for (size_t i = 0; i < n_; ++i) {
      const size_t r = (size_t)(std::rand() % max);
      const size_t t = pool_[r];
      --max;
      pool_[r] = pool_[(size_t)max];
      pool_[(size_t)max] = t;
    }
    random_pool_idx_ = 0;
    History h;
    h.delta = minimum_inlier_ratio / 10.0;
    h.epsilon = minimum_inlier_ratio;
    h.A = estimateThresholdA(h.epsilon, h.delta);
    h.k = 0;
    histories_.push_back(h);
  }

  // :299-330 (the C_{n+1} recursion of Chum & Matas' adaptive SPRT)
  double estimateThresholdA(double epsilon, double delta) const {
    const double C = (1 - delta) * std::log((1 - delta) / (1 - epsilon)) + delta * (std::log(delta / epsilon));
    const double K = (t_M_ * C) / m_S_ + 1;
    double An_1 = K;
    double An = K;
    for (unsigned int i = 0; i < 10; ++i) {
      An = K + std::log(An_1);
      if (std::fabs(An - An_1) < 1.5e-8) break;
      An_1 = An;
    }
    return An;
  }

  // :218-296 without index sets: the points in pool order from the running
  // pool index; a model is rejected as soon as the likelihood ratio exceeds
  // A.  Writes the tested inliers (pool order) and their truncated-quadratic
  // score; updates the test's history.
  template <class E>
  bool verify(const Model& model, const E& est, double threshold, size_t iteration, const Score& best,
              const Points& pts, std::vector<size_t>& inliers, Score& score) {
    inliers.clear();
    const double squared_threshold = threshold * threshold;
    const History cur = histories_.back();
    double lambda_new, lambda = 1;
    score = Score();
    bool valid = true;
    for (size_t tested = 0; tested < n_; ++tested) {
      const size_t idx = pool_[random_pool_idx_];
      const double r2 = est.squaredResidual(pts.row(idx), model);
      if (r2 < squared_threshold) {
        lambda_new = lambda * (cur.delta / cur.epsilon);
        ++score.inlier_number;
        score.value += 1.0 - r2 / squared_threshold;
        inliers.push_back(idx);
      } else {
        lambda_new = lambda * ((1 - cur.delta) / (1 - cur.epsilon));
      }
      if (++random_pool_idx_ >= n_) random_pool_idx_ = 0;
      if (lambda_new > cur.A * additional_model_probability_) {
        valid = false;
        break;
      }
      lambda = lambda_new;
    }
    update_history(valid, score, iteration, best, cur);
    return valid;
  }

 private:
  // :259-296
  void update_history(bool valid, const Score& score, size_t iteration, const Score& best, const History& cur) {
    if (valid) {
      if (score.value > best.value) {
        History h;
        h.epsilon = (double)score.inlier_number / n_;
        h.delta = cur.delta;
        h.A = estimateThresholdA(h.epsilon, cur.delta);
        h.k = iteration - last_sprt_update_;
        last_sprt_update_ = iteration;
        histories_.push_back(h);
      }
    } else {
      const double delta_estimated = (double)score.inlier_number / n_;
      if (delta_estimated > 0 && std::fabs(cur.delta - delta_estimated) / cur.delta > 0.05) {
        History h;
        h.epsilon = cur.epsilon;
        h.delta = delta_estimated;
        h.A = estimateThresholdA(cur.epsilon, delta_estimated);
        h.k = iteration - last_sprt_update_;
        last_sprt_update_ = iteration;
        histories_.push_back(h);
      }
    }
  }

 public:

  // :241-258: the same test over index sets (the space partitioning
  // selector's cells, in order); the pool index does not move.  A rejection
  // leaves only the current set: the remaining sets are still walked, point
  // by point, from the last accepted ratio (the reference's break exits the
  // inner loop), and their tested inliers count toward the score.
  template <class E>
  bool verify_sets(const Model& model, const E& est, double threshold, size_t iteration, const Score& best,
                   const Points& pts, const std::vector<const std::vector<size_t>*>& sets,
                   std::vector<size_t>& inliers, Score& score) {
    inliers.clear();
    const double squared_threshold = threshold * threshold;
    const History cur = histories_.back();
    double lambda_new, lambda = 1;
    score = Score();
    bool valid = true;
    for (const std::vector<size_t>* set : sets)
      for (const size_t idx : *set) {
        const double r2 = est.squaredResidual(pts.row(idx), model);
        if (r2 < squared_threshold) {
          lambda_new = lambda * (cur.delta / cur.epsilon);
          ++score.inlier_number;
          score.value += 1.0 - r2 / squared_threshold;
          inliers.push_back(idx);
        } else {
          lambda_new = lambda * ((1 - cur.delta) / (1 - cur.epsilon));
        }
        if (lambda_new > cur.A * additional_model_probability_) {
          valid = false;
          break;
        }
        lambda = lambda_new;
      }
    update_history(valid, score, iteration, best, cur);
    return valid;
  }

  const std::vector<size_t>& pool() const { return pool_; }
  size_t poolIndex() const { return random_pool_idx_; }
  const History& current() const { return histories_.back(); }
  size_t historyLength() const { return histories_.size(); }

 private:
  double t_M_, m_S_;
  size_t n_;
  std::vector<size_t> pool_;
  size_t random_pool_idx_ = 0;
  std::vector<History> histories_;
  size_t last_sprt_update_ = 0;
  double additional_model_probability_ = 1.0;
};

}  // namespace orc

namespace orc {

// x86's truncating double -> int32 conversion (cvttsd2si), which the
// reference's implicit conversions compile to: out of range or NaN gives
// INT_MIN.
inline int x86_cvtt_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return std::numeric_limits<int>::min();
  return (int)v;
}

// inc/inlier_selectors/space_partitioning_ransac.h:339-504, homographies:
// the grid cells (in the grid's iteration order) that a model can map
// points into.  A cell of fewer than 4 points is always kept; otherwise its
// corners in the first image are projected by H and the cell is kept when a
// projected corner falls into the cell's second-image cell, or when the
// rectangle spanned by its projected (0,0) and (1,1) corners, grown by the
// threshold, passes the reference's overlap test.  The corner projections
// are pure functions of the corner, so the reference's per-model cache is
// not restated.  Cells with an axis index at or beyond the division count
// (points outside the image) make the reference read corner projections it
// did not compute for this model; such cells are kept here.
class SpacePartitioning {
 public:
  explicit SpacePartitioning(const GridNeighborhood* grid) : grid_(grid) {}

  struct Corner {
    int ix, iy;
    double x, y;
  };
  static Corner project(const double* H, size_t h, size_t v, const std::vector<double>& cs) {
    const double kX1 = (double)h * cs[0], kY1 = (double)v * cs[1];
    double x2p = kX1 * H[0] + kY1 * H[1] + H[2], y2p = kX1 * H[3] + kY1 * H[4] + H[5],
           h2p = kX1 * H[6] + kY1 * H[7] + H[8];
    x2p /= h2p;
    y2p /= h2p;
    return Corner{x86_cvtt_i32(x2p / cs[2]), x86_cvtt_i32(y2p / cs[3]), x2p, y2p};
  }

  // Eigen's fixed-size 3x3 determinant (cofactors along the first row)
  static double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
  }

  static bool keep_cell(const double* H, const std::vector<size_t>& ci, size_t npts, const std::vector<double>& cs,
                        double P, double threshold) {
    if (npts < 4) return true;
    if ((double)ci[0] >= P || (double)ci[1] >= P) return true;
    static const int steps[8] = {0, 0, 0, 1, 1, 0, 1, 1};
    for (int s = 0; s < 8; s += 2) {
      const size_t h = ci[0] + (size_t)steps[s];
      if ((double)h >= P) continue;
      const size_t v = ci[1] + (size_t)steps[s + 1];
      if ((double)v >= P) continue;
      const Corner c = project(H, h, v, cs);
      if ((size_t)(long)c.ix == ci[2] && (size_t)(long)c.iy == ci[3]) return true;
    }
    const size_t h11 = ci[0] + 1, v11 = ci[1] + 1;
    const Corner c00 = project(H, ci[0], ci[1], cs);
    const Corner c11 = project(H, h11, v11, cs);
    const double l1x = c00.x - threshold, l1y = c00.y - threshold;
    const double r1x = c11.x + threshold, r1y = c11.y + threshold;
    const double l2x = cs[2] * (double)ci[2], l2y = cs[3] * (double)ci[3];
    const double r2x = l2x + cs[2], r2y = l2y + cs[3];
    return (l1x <= r2x && l2x <= r1x) || (r1y <= l2y && r2y <= l1y);
  }

  void run(const Model& model, double threshold, std::vector<const std::vector<size_t>*>& sets,
           size_t& count) const {
    sets.clear();
    count = 0;
    if (std::fabs(det3(model.m)) < 1e-3) return;
    const std::vector<double>& cs = grid_->getCellSizes();
    const double P = (double)grid_->getDivisionNumber();
    for (const auto& kv : grid_->getCells()) {
      const std::vector<size_t>& pts = kv.second.first;
      if (keep_cell(model.m, kv.second.second, pts.size(), cs, P, threshold)) {
        sets.push_back(&pts);
        count += pts.size();
      }
    }
  }

 private:
  const GridNeighborhood* grid_;
};

}  // namespace orc
