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
