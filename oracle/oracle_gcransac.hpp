// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Restatement of the GC-RANSAC loop, inc/GCRANSAC.h:157-1021, with the MSAC
// scoring function of inc/scoring_function.h:137-233, optional SPRT
// preemption (inc/preemption/preemption_sprt.h) and the space partitioning
// inlier selector (inc/inlier_selectors/space_partitioning_ransac.h); without
// them, the reference's empty defaults.
#pragma once

#include "oracle_core.hpp"
#include "oracle_preemption.hpp"

namespace orc {

// inc/scoring_function.h:137-233 (MSAC), full pass over all points.  The
// early-exit test at :182-183 can never fire (its left side is >= 1 while
// -best.value <= 0), so it is not restated.
template <class E>
Score msac_score(const Points& pts, const Model& model, const E& est, double squared_truncated_threshold,
                 std::vector<size_t>* inliers) {
  Score score;
  if (inliers) inliers->clear();
  for (size_t i = 0; i < pts.rows; ++i) {
    const double r2 = est.squaredResidual(pts.row(i), model);
    if (r2 < squared_truncated_threshold) {
      if (inliers) inliers->push_back(i);
      ++score.inlier_number;
      score.value -= r2;
    }
  }
  if (score.inlier_number == 0) return Score();
  score.value = (score.value + score.inlier_number * squared_truncated_threshold) / squared_truncated_threshold;
  return score;
}

// The same score over index sets (the space partitioning selector's cells,
// in order): scoring_function.h:186-214.
template <class E>
Score msac_score_sets(const Points& pts, const Model& model, const E& est, double squared_truncated_threshold,
                      const std::vector<const std::vector<size_t>*>& sets, std::vector<size_t>* inliers) {
  Score score;
  if (inliers) inliers->clear();
  for (const std::vector<size_t>* set : sets)
    for (const size_t i : *set) {
      const double r2 = est.squaredResidual(pts.row(i), model);
      if (r2 < squared_truncated_threshold) {
        if (inliers) inliers->push_back(i);
        ++score.inlier_number;
        score.value -= r2;
      }
    }
  if (score.inlier_number == 0) return Score();
  score.value = (score.value + score.inlier_number * squared_truncated_threshold) / squared_truncated_threshold;
  return score;
}

template <class E>
class GCRANSAC {
 public:
  Settings settings;

  const Statistics& getRansacStatistics() const { return statistics; }

  // inc/GCRANSAC.h:157-172
  static size_t getIterationNumber(size_t inlier_number, size_t point_number, size_t sample_size,
                                   double log_probability) {
    const double q = std::pow(static_cast<double>(inlier_number) / point_number, sample_size);
    const double log2 = std::log(1 - q);
    if (std::abs(log2) < std::numeric_limits<double>::epsilon()) return std::numeric_limits<size_t>::max();
    const double iter = log_probability / log2;
    return static_cast<size_t>(iter) + 1;
  }

  // inc/GCRANSAC.h:202-627
  void run(const Points& pts, const E& est, Sampler* main_sampler, Sampler* lo_sampler,
           const Neighborhood* neighborhood, Model& obtained_model) {
    statistics = Statistics();
    main_sampler_ = main_sampler;
    lo_sampler_ = lo_sampler;
    neighborhood_ = neighborhood;
    const size_t m = E::sampleSize();
    log_probability_ = std::log(1.0 - settings.confidence);
    size_t max_iteration = getIterationNumber(1, pts.rows, m, log_probability_);
    std::vector<size_t> current_sample(m);
    bool do_local_optimization = false;
    point_number_ = pts.rows;
    size_t inlier_container_offset = 0, inlier_container_idx;
    Model so_far_the_best_model;
    Score current_score, so_far_the_best_score;
    std::vector<std::vector<size_t>> temp_inner_inliers(2);
    truncated_threshold_ = 3.0 / 2.0 * settings.threshold;
    squared_truncated_threshold_ = truncated_threshold_ * truncated_threshold_;
    std::vector<size_t> pool(point_number_);
    for (size_t i = 0; i < point_number_; ++i) pool[i] = i;
    std::vector<Model> models;
    std::vector<const std::vector<size_t>*> selected_sets;
    size_t selected_point_number = 0;

    while (settings.min_iteration_number > statistics.iteration_number ||
           statistics.iteration_number < std::min(max_iteration, settings.max_iteration_number)) {
      do_local_optimization = false;
      ++statistics.iteration_number;
      models.clear();
      int unsuccessful_model_generations = -1;
      // the sampler's update() after each drawn sample (:300-336) and once
      // after the iteration (:474-478), whether or not a model came out
      while (++unsuccessful_model_generations < (int)settings.max_unsuccessful_model_generations) {
        if (!main_sampler_->sample(pool, current_sample.data(), m)) {
          main_sampler_->update(current_sample.data(), m);
          continue;
        }
        if (!est.isValidSample(pts, current_sample.data())) {
          main_sampler_->update(current_sample.data(), m);
          continue;
        }
        if (est.estimateModel(pts, current_sample.data(), &models)) break;
        main_sampler_->update(current_sample.data(), m);
      }
      statistics.iteration_number += unsuccessful_model_generations;

      for (auto& model : models) {
        // inc/GCRANSAC.h:343-367: the space partitioning selector's cells; a
        // model whose cells hold fewer points than the best model's inliers
        // is skipped
        if (sp_) {
          sp_->run(model, truncated_threshold_, selected_sets, selected_point_number);
          if (selected_point_number < so_far_the_best_score.inlier_number) continue;
        }
        // :370-425: with SPRT the test provides the score and the inliers
        // (pool or cell order), and a rejected model is skipped
        if (sprt_) {
          const bool ok =
              sp_ ? sprt_->verify_sets(model, est, truncated_threshold_, statistics.iteration_number,
                                       so_far_the_best_score, pts, selected_sets,
                                       temp_inner_inliers[inlier_container_offset], current_score)
                  : sprt_->verify(model, est, truncated_threshold_, statistics.iteration_number, so_far_the_best_score,
                                  pts, temp_inner_inliers[inlier_container_offset], current_score);
          if (!ok) continue;
        } else if (sp_) {
          current_score = msac_score_sets(pts, model, est, squared_truncated_threshold_, selected_sets,
                                          &temp_inner_inliers[inlier_container_offset]);
        } else {
          current_score = msac_score(pts, model, est, squared_truncated_threshold_,
                                     &temp_inner_inliers[inlier_container_offset]);
        }
        bool is_model_updated = false;
        if (so_far_the_best_score < current_score &&
            est.isValidModel(model, pts, temp_inner_inliers[inlier_container_offset], current_sample.data(),
                             truncated_threshold_, is_model_updated)) {
          if (is_model_updated)
            current_score = msac_score(pts, model, est, squared_truncated_threshold_,
                                       &temp_inner_inliers[inlier_container_offset]);
          inlier_container_offset = (inlier_container_offset + 1) % 2;
          so_far_the_best_model = model;
          so_far_the_best_score = current_score;
          do_local_optimization = statistics.iteration_number > settings.min_iteration_number_before_lo &&
                                  so_far_the_best_score.inlier_number > m;
          max_iteration = getIterationNumber(so_far_the_best_score.inlier_number, point_number_, m, log_probability_);
        }
      }

      main_sampler_->update(current_sample.data(), m);

      if (settings.do_local_optimization && do_local_optimization) {
        ++statistics.local_optimization_number;
        graphCutLocalOptimization(pts, temp_inner_inliers[inlier_container_offset], so_far_the_best_model,
                                  so_far_the_best_score, est, settings.max_local_optimization_number);
        max_iteration = getIterationNumber(so_far_the_best_score.inlier_number, point_number_, m, log_probability_);
      }
    }

    if (so_far_the_best_score.inlier_number <= m) return;  // model not found (:521-527)

    if (settings.do_local_optimization && statistics.local_optimization_number == 0) {
      ++statistics.local_optimization_number;
      graphCutLocalOptimization(pts, temp_inner_inliers[inlier_container_offset], so_far_the_best_model,
                                so_far_the_best_score, est, settings.max_local_optimization_number);
    }

    if (temp_inner_inliers[inlier_container_offset].size() != so_far_the_best_score.inlier_number)
      inlier_container_offset = (inlier_container_offset + 1) % 2;
    if (temp_inner_inliers[inlier_container_offset].size() != so_far_the_best_score.inlier_number)
      so_far_the_best_score = msac_score(pts, so_far_the_best_model, est, squared_truncated_threshold_,
                                         &temp_inner_inliers[inlier_container_offset]);

    bool iterative_refitting_applied = false;
    if (settings.do_final_iterated_least_squares) {
      Model model = so_far_the_best_model;
      const bool success = iteratedLeastSquaresFitting(pts, est, temp_inner_inliers[inlier_container_offset], model);
      if (success) {
        inlier_container_idx = (inlier_container_offset + 1) % 2;
        current_score = msac_score(pts, model, est, squared_truncated_threshold_,
                                   &temp_inner_inliers[inlier_container_idx]);
        if (so_far_the_best_score < current_score) {
          iterative_refitting_applied = true;
          so_far_the_best_model = model;
          inlier_container_offset = inlier_container_idx;
        }
      }
    }
    if (!iterative_refitting_applied) {
      models.clear();
      est.estimateModelNonminimal(pts, temp_inner_inliers[inlier_container_offset].data(),
                                  so_far_the_best_score.inlier_number, &models, nullptr);
      for (auto& model : models) {
        inlier_container_idx = (inlier_container_offset + 1) % 2;
        current_score = msac_score(pts, model, est, squared_truncated_threshold_,
                                   &temp_inner_inliers[inlier_container_idx]);
        if (so_far_the_best_score < current_score) {
          so_far_the_best_model = model;
          inlier_container_offset = inlier_container_idx;
        }
      }
      // :615-616 reads models[0] when models is empty (undefined behaviour);
      // the restatement keeps the previous model in that case.
    }
    statistics.inliers.swap(temp_inner_inliers[inlier_container_offset]);
    statistics.score = so_far_the_best_score.value;
    obtained_model = so_far_the_best_model;
  }

  // Scoring with the run's squared truncated threshold (the score's
  // squared_truncated_threshold is set once in run(), :253-255).
  Score score(const Points& pts, const Model& model, const E& est, std::vector<size_t>* inl) const {
    return msac_score(pts, model, est, squared_truncated_threshold_, inl);
  }

 protected:
  // inc/GCRANSAC.h:630-758.  Tukey weights are stored by point index while
  // the non-minimal solvers of H and F read them by position in the compacted
  // sample (the reference's behaviour; the estimator decides what it reads).
  bool iteratedLeastSquaresFitting(const Points& pts, const E& est, std::vector<size_t>& inliers, Model& model) {
    const size_t sample_size = E::sampleSize();
    if (inliers.size() <= sample_size) return false;
    size_t iterations = 0;
    std::vector<size_t> tmp_inliers;
    std::vector<double> weights(pts.rows, 0.0);
    Score best_score;
    while (++iterations < settings.max_least_squares_iterations) {
      std::vector<Model> models;
      if (E::isWeightingApplicable()) {
        for (size_t k = 0; k < inliers.size(); ++k) {
          const size_t idx = inliers[k];
          const double r2 = est.squaredResidual(pts.row(idx), model);
          const double v = 1.0 - r2 / squared_truncated_threshold_;
          const double w = 0.0 > v ? 0.0 : v;  // MAX(0.0, v) as the reference's macro
          weights[idx] = w * w;
        }
        est.estimateModelNonminimal(pts, inliers.data(), inliers.size(), &models, weights.data());
        for (size_t idx : inliers) weights[idx] = 0;
      } else {
        est.estimateModelNonminimal(pts, inliers.data(), inliers.size(), &models, nullptr);
      }
      if (models.empty()) break;
      if (models.size() == 1) {
        const Score s = score(pts, models[0], est, &tmp_inliers);
        if (tmp_inliers.size() < sample_size) break;
        if (s.inlier_number <= inliers.size()) break;
        model = models[0];
        inliers.swap(tmp_inliers);
      } else {
        bool updated = false;
        for (auto& mdl : models) {
          const Score s = score(pts, mdl, est, &tmp_inliers);
          if (tmp_inliers.size() < sample_size) continue;
          if (s.inlier_number <= inliers.size()) continue;
          if (s.inlier_number >= best_score.inlier_number) {
            updated = true;
            best_score = s;
            model = mdl;
            inliers.swap(tmp_inliers);
          }
        }
        if (!updated) break;
      }
    }
    return iterations > 1;
  }

  // inc/GCRANSAC.h:780-910
  bool graphCutLocalOptimization(const Points& pts, std::vector<size_t>& so_far_the_best_inliers,
                                 Model& so_far_the_best_model, Score& so_far_the_best_score, const E& est,
                                 size_t trial_number) {
    const size_t inlier_limit = est.inlierLimit();
    Score max_score = so_far_the_best_score;
    Model best_model = so_far_the_best_model;
    std::vector<Model> models;
    std::vector<size_t> best_inliers, inliers, tmp_inliers;
    bool updated;
    ++statistics.local_optimization_number;  // counted twice per LO (:485,805)
    std::vector<size_t> current_sample(inlier_limit);
    while (++statistics.graph_cut_number < settings.max_graph_cut_number) {
      updated = false;
      inliers.clear();
      labeling(pts, best_model, est, settings.spatial_coherence_weight, settings.threshold, inliers);
      const size_t sample_size = std::min(inlier_limit, inliers.size());
      for (size_t trial = 0; trial < trial_number; ++trial) {
        models.clear();
        if (sample_size < inliers.size()) {
          lo_sampler_->sample(inliers, current_sample.data(), sample_size);
          if (!est.estimateModelNonminimal(pts, current_sample.data(), sample_size, &models, nullptr)) continue;
        } else if (E::sampleSize() < inliers.size()) {
          if (!est.estimateModelNonminimal(pts, inliers.data(), inliers.size(), &models, nullptr)) break;
        } else
          break;
        for (auto& model : models) {
          const Score s = score(pts, model, est, &tmp_inliers);
          if (max_score < s) {
            updated = true;
            max_score = s;
            best_model = model;
            best_inliers.swap(tmp_inliers);
            tmp_inliers.clear();
          }
        }
      }
      if (!updated) break;
    }
    if (so_far_the_best_score < max_score) {
      so_far_the_best_score = max_score;
      so_far_the_best_model = best_model;
      so_far_the_best_inliers.swap(best_inliers);
      return true;
    }
    return false;
  }

 public:
  // inc/GCRANSAC.h:913-1021.  Builds the same energy and returns the SINK set.
  // The reference's O(N^2) used_edges matrix only deduplicates pairs.  Both
  // neighbourhoods are symmetric (q lists p whenever p lists q, and a list
  // never drops a later point), so every neighbouring pair p<q is added once,
  // when p is visited: p ascending, then p's list order (ascending q for the
  // grid; (distance, q) for the radius graph).
  void labeling(const Points& pts, const Model& model, const E& est, double lambda, double threshold,
                std::vector<size_t>& inliers) {
    GraphCutProblem g;
    build_graph_cut(pts, model, est, lambda, threshold, neighborhood_, g);
    const std::vector<uint8_t> sink = solve_graph_cut(g);
    inliers.clear();
    for (size_t i = 0; i < pts.rows; ++i)
      if (sink[i]) inliers.push_back(i);
  }

  static void build_graph_cut(const Points& pts, const Model& model, const E& est, double lambda, double threshold,
                              const Neighborhood* nb, GraphCutProblem& g) {
    const size_t n = pts.rows;
    g.tr_cap.assign(n, 0.0);
    g.flow_const = 0.0;
    std::vector<double> dist(n);
    const double sq_thr = threshold * threshold * 9 / 4;
    const double one_minus_lambda = 1.0 - lambda;
    for (size_t i = 0; i < n; ++i) {
      const double r2 = est.squaredResidual(pts.row(i), model);
      double d = r2 / sq_thr;
      d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
      dist[i] = d;
      const double e = 1 - d;
      // add_term1(i, A, B) == add_tweights(i, B, A) (inc/energy.h:204-209)
      if (r2 <= sq_thr) add_tweights(g, (int)i, 0, one_minus_lambda * e);
      else add_tweights(g, (int)i, one_minus_lambda * (1 - e), 0);
    }
    if (lambda > 0 && nb != nullptr) {
      for (size_t p = 0; p < n; ++p) {
        const double e1 = dist[p];
        for (size_t q : nb->getNeighbors(p)) {
          if (q <= p) continue;  // q == p skipped; q < p already used by q
          const double e00 = 0.5 * (e1 + dist[q]);
          // add_term2(p, q, A=e00*l, B=l, C=l, D=0) (inc/energy.h:211-254)
          double A = e00 * lambda, B = lambda, C = lambda;
          const double D = 0.0;
          add_tweights(g, (int)p, D, A);
          B -= A;
          C -= D;
          g.edge_p.push_back((int)p);
          g.edge_q.push_back((int)q);
          // B >= 0 and C >= 0 for every term this loop builds (e00 <= 1).
          g.cap_pq.push_back(B);
          g.cap_qp.push_back(C);
        }
      }
    }
  }

 protected:
  Statistics statistics;
  Sampler* main_sampler_ = nullptr;
  Sampler* lo_sampler_ = nullptr;
  const Neighborhood* neighborhood_ = nullptr;

 public:
  // SPRT preemptive verification of the main loop's models (nullptr: every
  // model is scored on all points)
  SprtPreemption* sprt_ = nullptr;
  // space partitioning inlier selector (homographies; nullptr: all points)
  const SpacePartitioning* sp_ = nullptr;

 private:
  size_t point_number_ = 0;
  double truncated_threshold_ = 0, squared_truncated_threshold_ = 0, log_probability_ = 0;
};

}  // namespace orc
