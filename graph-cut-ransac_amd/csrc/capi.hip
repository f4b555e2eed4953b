// C ABI of libgcr_hip.so (declared in include/gcr.h).
#include <cmath>
#include <chrono>
#include <cstring>
#include <limits>
#include <string>

#include "engine.hpp"

namespace {

thread_local std::string g_last_error;

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace

namespace gcr {
void set_last_error(const std::string& msg) { g_last_error = msg; }
}  // namespace gcr

namespace {

template <class F>
int32_t guarded(F&& f) {
  try {
    g_last_error.clear();
    return f();
  } catch (const gcr::EngineError& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(GCR_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(GCR_ERR_INTERNAL, e.what());
  }
}

int32_t device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int64_t min_rows(int32_t kind) {
  return kind == GCR_PROBLEM_HOMOGRAPHY ? 4 : (kind == GCR_PROBLEM_FUNDAMENTAL ? 7 : 5);
}

bool invertible3(const double* K) {
  const double d = K[0] * (K[4] * K[8] - K[5] * K[7]) - K[1] * (K[3] * K[8] - K[5] * K[6]) +
                   K[2] * (K[3] * K[7] - K[4] * K[6]);
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(K[i])) return false;
  return std::isfinite(d) && d != 0.0;
}

// check_problem / check_probabilities: the call is to return "no model"
// (0 inliers), as the reference drivers do; the message is in g_last_error.
constexpr int32_t kNoModel = 1;

// Rules on one problem's parameters and size; pre is prefixed to messages
// ("" for single-pair calls, "pair i: " in batches).
int32_t check_problem(const std::string& pre, int32_t kind, int64_t n, const gcr_params_t& p, const double* K1,
                      const double* K2) {
  if (kind < 0 || kind > 2)
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "unknown problem kind " + std::to_string(kind) +
                                              " (0 homography, 1 fundamental, 2 essential)");
  if (n < min_rows(kind))
    return fail(GCR_ERR_INVALID_ARGUMENT,
                pre + "correspondences should be an array with dims [n,4], n>=" + std::to_string(min_rows(kind)));
  if (p.solver != 0) return fail(GCR_ERR_UNSUPPORTED, pre + "only point correspondences (solver=0) are built");
  if (p.sample_fn && !pre.empty())
    return fail(GCR_ERR_UNSUPPORTED, pre + "a host sampler callback (sample_fn) is for single-pair calls");
  if (!p.sample_fn && p.sampler != 0 && p.sampler != 1 && p.sampler != 3 && p.sampler != 4)
    return fail(GCR_ERR_UNSUPPORTED, pre + "sampler " + std::to_string(p.sampler) +
                                         " is not built; accepted samplers are 0 (uniform), 1 (PROSAC), "
                                         "3 (importance) and 4 (adaptive reordering)");
  if (!p.sample_fn && p.sampler == 4 && !(p.sampler_variance > 0.0))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "sampler_variance must be positive for sampler 4");
  // The drivers build a neighbourhood for homographies always and for F / E
  // when the spatial weight is above eps; findHomography_ replaces a
  // non-grid neighbourhood by the grid when space partitioning is on
  // (py/gcransac_python.cpp:4282-4286).  An unknown id makes them return 0
  // inliers (:1681-1686, 2604-2609, 4302-4307).
  const bool lambda_on = p.spatial_coherence_weight > std::numeric_limits<double>::epsilon();
  const bool graph = kind == GCR_PROBLEM_HOMOGRAPHY || lambda_on;
  const int32_t nb = (kind == GCR_PROBLEM_HOMOGRAPHY && p.use_space_partitioning) ? 0 : p.neighborhood;
  if (graph && nb != 0 && nb != 1) {
    g_last_error = pre + "unknown neighborhood-graph identifier " + std::to_string(p.neighborhood) +
                   "; the accepted values are 0 (grid) and 1 (radius)";
    return kNoModel;
  }
  if (graph && nb == 0 && lambda_on && !(p.neighborhood_size >= 1.0))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "neighborhood_size must be >= 1 grid division");
  if (graph && nb == 1 && !std::isfinite(p.neighborhood_size))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "neighborhood_size (the radius) must be finite");
  if (!(p.confidence > 0.0 && p.confidence < 1.0))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "conf must be in (0, 1)");
  if (!(p.threshold > 0.0) || !std::isfinite(p.threshold))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "threshold must be positive and finite");
  if (p.use_sprt && kind != GCR_PROBLEM_HOMOGRAPHY &&
      !(p.min_inlier_ratio_for_sprt > 0.0 && p.min_inlier_ratio_for_sprt < 1.0))
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "min_inlier_ratio_for_sprt must be in (0, 1)");
  if (p.max_iters < 0 || p.min_iters < 0 || p.lo_number < 0)
    return fail(GCR_ERR_INVALID_ARGUMENT, pre + "max_iters, min_iters and lo_number must be >= 0");
  if (!(p.time_limit >= 0.0)) return fail(GCR_ERR_INVALID_ARGUMENT, pre + "time_limit must be >= 0 (0: none)");
  if (kind == GCR_PROBLEM_ESSENTIAL) {
    if (!K1 || !K2) return fail(GCR_ERR_INVALID_ARGUMENT, pre + "K1 and K2 must be 3x3");
    if (!invertible3(K1) || !invertible3(K2))
      return fail(GCR_ERR_INVALID_ARGUMENT, pre + "K1 and K2 must be finite and invertible");
  }
  return GCR_OK;
}

// Samplers 3 and 4 read one probability per correspondence.  The reference
// drivers return 0 inliers when the sampler cannot be initialised (a size
// mismatch, py/gcransac_python.cpp:4369-4379); probabilities that are not
// finite and non-negative with a positive maximum are refused here.
// Returns 1 when the call must return "no model", 0 to go on, < 0 on error.
int32_t check_probabilities(const gcr_params_t& p, const double* probs, int64_t n_probs, int64_t n) {
  if (p.sample_fn || (p.sampler != 3 && p.sampler != 4)) return 0;
  if (!probs || n_probs != n) {
    g_last_error = "sampler " + std::to_string(p.sampler) + " needs one probability per correspondence (" +
                   std::to_string(n_probs) + " given for " + std::to_string(n) + ")";
    return kNoModel;
  }
  double mx = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!(probs[i] >= 0.0) || !std::isfinite(probs[i]))
      return fail(GCR_ERR_INVALID_ARGUMENT, "probabilities must be finite and non-negative");
    mx = probs[i] > mx ? probs[i] : mx;
  }
  if (!(mx > 0.0)) return fail(GCR_ERR_INVALID_ARGUMENT, "probabilities must not all be zero");
  return 0;
}

int32_t no_model(int64_t n, double* out, uint8_t* mask, gcr_stats_t* stats) {
  std::memset(out, 0, sizeof(double) * 9);
  std::memset(mask, 0, (size_t)n);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return 0;
}

// Validation shared by the single-pair entry points (py/bindings.cpp:881-889,
// 301-309, 733-741).
// Returns GCR_OK, 1 when the call is to return "no model" (see
// check_probabilities), or an error code.
int32_t check_common(const double* corr, int64_t n, int32_t kind, const gcr_params_t* p, double* out, uint8_t* mask,
                     const double* K1 = nullptr, const double* K2 = nullptr, const double* probs = nullptr,
                     int64_t n_probs = 0) {
  if (!corr || !p || !out || !mask) return fail(GCR_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (int32_t rc = check_problem("", kind, n, *p, K1, K2)) return rc;
  if (int32_t rc = check_probabilities(*p, probs, n_probs, n)) return rc;
  if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
  return GCR_OK;
}

}  // namespace

int32_t gcr::check_pair(int index, int32_t kind, int64_t n, const gcr_params_t& p, const double* K1,
                        const double* K2) {
  const std::string pre = "pair " + std::to_string(index) + ": ";
  if (int32_t rc = check_problem(pre, kind, n, p, K1, K2))
    return rc == kNoModel ? fail(GCR_ERR_INVALID_ARGUMENT, g_last_error) : rc;
  if (p.sampler == 3 || p.sampler == 4)
    return fail(GCR_ERR_UNSUPPORTED, pre + "samplers 3 and 4 read per-point probabilities, which the batch entry "
                                           "points do not take; use the single-pair entry points");
  return GCR_OK;
}

namespace {

int32_t run_single(const gcr::PairSpec& spec, const double* corr, double* out, uint8_t* mask, gcr_stats_t* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<gcr::PairSpec> specs{spec};
  std::vector<gcr::PairResult> res;
  gcr::run_batch(specs, corr, false, res);
  const gcr::PairResult& r = res[0];
  std::memcpy(out, r.model, sizeof(double) * 9);
  std::memset(mask, 0, (size_t)spec.n);
  for (int32_t i : r.inliers) mask[i] = 1;
  if (stats) {
    *stats = r.stats;
    stats->processing_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (r.status < 0) g_last_error = r.error;
  return r.status;
}

}  // namespace

extern "C" {

void gcr_default_params(gcr_params_t* p, int32_t problem) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->threshold = 1.0;
  p->confidence = 0.99;
  p->max_iters = 10000;
  p->min_iters = 50;
  p->use_sprt = 0;
  p->sampler = 0;
  p->neighborhood = 0;
  p->lo_number = 50;
  p->sampler_variance = 0.1;
  p->solver = 0;
  p->seed = 0;
  switch (problem) {
    case GCR_PROBLEM_HOMOGRAPHY:  // py/bindings.cpp:1158-1178
      p->spatial_coherence_weight = 0.975;
      p->min_inlier_ratio_for_sprt = 0.00001;
      p->neighborhood_size = 4.0;
      p->use_space_partitioning = 0;
      break;
    case GCR_PROBLEM_FUNDAMENTAL:  // py/bindings.cpp:1003-1022
      p->spatial_coherence_weight = 0.0;
      p->min_inlier_ratio_for_sprt = 0.00001;
      p->neighborhood_size = 20.0;
      break;
    default:  // essential, py/bindings.cpp:1091-1112
      p->spatial_coherence_weight = 0.1;
      p->min_inlier_ratio_for_sprt = 0.01;
      p->neighborhood_size = 20.0;
      break;
  }
}

int32_t gcr_find_homography(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                            const double* probabilities, int64_t n_probabilities, const gcr_params_t* params,
                            double* H_out, uint8_t* mask_out, gcr_stats_t* stats) {
  return guarded([&]() -> int32_t {
    const int32_t rc =
        check_common(corr, n, GCR_PROBLEM_HOMOGRAPHY, params, H_out, mask_out, nullptr, nullptr, probabilities, n_probabilities);
    if (rc == kNoModel) return no_model(n, H_out, mask_out, stats);
    if (rc != GCR_OK) return rc;
    gcr::PairSpec s;
    s.kind = GCR_PROBLEM_HOMOGRAPHY;
    s.n = n;
    s.probs = probabilities;
    s.h1 = h1;
    s.w1 = w1;
    s.h2 = h2;
    s.w2 = w2;
    s.params = *params;
    return run_single(s, corr, H_out, mask_out, stats);
  });
}

int32_t gcr_find_fundamental_matrix(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                                    const double* probabilities, int64_t n_probabilities, const gcr_params_t* params,
                                    double* F_out, uint8_t* mask_out, gcr_stats_t* stats) {
  return guarded([&]() -> int32_t {
    const int32_t rc =
        check_common(corr, n, GCR_PROBLEM_FUNDAMENTAL, params, F_out, mask_out, nullptr, nullptr, probabilities, n_probabilities);
    if (rc == kNoModel) return no_model(n, F_out, mask_out, stats);
    if (rc != GCR_OK) return rc;
    gcr::PairSpec s;
    s.kind = GCR_PROBLEM_FUNDAMENTAL;
    s.n = n;
    s.probs = probabilities;
    s.h1 = h1;
    s.w1 = w1;
    s.h2 = h2;
    s.w2 = w2;
    s.params = *params;
    return run_single(s, corr, F_out, mask_out, stats);
  });
}

int32_t gcr_find_essential_matrix(const double* corr, int64_t n, const double* K1, const double* K2, int32_t h1,
                                  int32_t w1, int32_t h2, int32_t w2, const double* probabilities,
                                  int64_t n_probabilities, const gcr_params_t* params, double* E_out,
                                  uint8_t* mask_out, gcr_stats_t* stats) {
  return guarded([&]() -> int32_t {
    const int32_t rc =
        check_common(corr, n, GCR_PROBLEM_ESSENTIAL, params, E_out, mask_out, K1, K2, probabilities, n_probabilities);
    if (rc == kNoModel) return no_model(n, E_out, mask_out, stats);
    if (rc != GCR_OK) return rc;
    gcr::PairSpec s;
    s.kind = GCR_PROBLEM_ESSENTIAL;
    s.n = n;
    s.probs = probabilities;
    s.h1 = h1;
    s.w1 = w1;
    s.h2 = h2;
    s.w2 = w2;
    std::memcpy(s.K1, K1, sizeof(s.K1));
    std::memcpy(s.K2, K2, sizeof(s.K2));
    s.params = *params;
    return run_single(s, corr, E_out, mask_out, stats);
  });
}

int32_t gcr_run(int32_t kind, const double* pts, int64_t n, const double* grid_pts, const double* cell_sizes,
                int32_t divisions, int32_t flags, const gcr_params_t* params, double* model_out, uint8_t* mask_out,
                gcr_stats_t* stats) {
  return guarded([&]() -> int32_t {
    if (kind < 0 || kind > 2) return fail(GCR_ERR_INVALID_ARGUMENT, "unknown problem kind");
    // points are already K-normalised: intrinsics play no part here
    static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (params && !params->sample_fn && (params->sampler == 3 || params->sampler == 4))
      return fail(GCR_ERR_UNSUPPORTED, "gcr_run takes no point probabilities: samplers 3 and 4 are not available");
    gcr_params_t pr{};
    if (params) {
      pr = *params;
      pr.neighborhood = 0;  // the template API's grid of the given cells
    }
    const int32_t rc = check_common(pts, n, kind, params ? &pr : nullptr, model_out, mask_out, I3, I3);
    if (rc != GCR_OK) return rc;
    // the preemption and inlier-selector slots of the template
    // (GCRANSAC.h:50-51): SPRT with the caller's minimum inlier ratio, and
    // space partitioning on the caller's grid for homographies
    if (pr.use_sprt && !(pr.min_inlier_ratio_for_sprt > 0.0 && pr.min_inlier_ratio_for_sprt < 1.0))
      return fail(GCR_ERR_INVALID_ARGUMENT, "SPRT needs a minimum inlier ratio in (0, 1)");
    if (pr.use_space_partitioning && kind != 0)
      return fail(GCR_ERR_UNSUPPORTED, "space partitioning is built for homographies only");
    if (!cell_sizes || divisions < 1) return fail(GCR_ERR_INVALID_ARGUMENT, "cell sizes and divisions >= 1 required");
    gcr::PairSpec s;
    s.kind = kind;
    s.n = n;
    s.params = pr;
    s.params.neighborhood_size = (double)divisions;
    s.cell_override = true;
    for (int d = 0; d < 4; ++d) s.cell[d] = cell_sizes[d];
    s.driver_stage = false;
    s.degensac = (flags & GCR_RUN_NO_DEGENSAC) == 0;
    s.grid_corr = grid_pts;
    return run_single(s, pts, model_out, mask_out, stats);
  });
}

int32_t gcr_find_batch(int32_t num_pairs, const int32_t* kinds, const int64_t* offsets, const double* corr,
                       int32_t corr_is_device, const int32_t* sizes, const double* K, const gcr_params_t* params,
                       double* models_out, uint8_t* masks_out, int32_t* num_inliers_out, gcr_stats_t* stats) {
  return guarded([&]() -> int32_t {
    if (num_pairs <= 0) return fail(GCR_ERR_INVALID_ARGUMENT, "num_pairs must be positive");
    if (!kinds || !offsets || !corr || !sizes || !params || !models_out || !masks_out || !num_inliers_out)
      return fail(GCR_ERR_INVALID_ARGUMENT, "null pointer argument");
    for (int i = 0; i < num_pairs; ++i) {
      const double* K1 = K ? K + 18 * i : nullptr;
      if (int32_t rc = gcr::check_pair(i, kinds[i], offsets[i + 1] - offsets[i], params[i], K1, K1 ? K1 + 9 : nullptr))
        return rc;
    }
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<gcr::PairSpec> specs((size_t)num_pairs);
    for (int i = 0; i < num_pairs; ++i) {
      gcr::PairSpec& s = specs[(size_t)i];
      s.kind = kinds[i];
      s.offset = offsets[i];
      s.n = offsets[i + 1] - offsets[i];
      if (corr_is_device && i > 0 && offsets[i] != offsets[i - 1] + specs[(size_t)i - 1].n)
        return fail(GCR_ERR_INVALID_ARGUMENT, "device input requires contiguous pairs");
      s.h1 = sizes[4 * i + 0];
      s.w1 = sizes[4 * i + 1];
      s.h2 = sizes[4 * i + 2];
      s.w2 = sizes[4 * i + 3];
      if (K) {
        std::memcpy(s.K1, K + 18 * i, sizeof(s.K1));
        std::memcpy(s.K2, K + 18 * i + 9, sizeof(s.K2));
      }
      s.params = params[i];
    }
    const double* base = corr;
    if (corr_is_device) base = corr + 4 * offsets[0];
    std::vector<gcr::PairResult> res;
    gcr::run_batch(specs, base, corr_is_device != 0, res);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < num_pairs; ++i) {
      const gcr::PairResult& r = res[(size_t)i];
      std::memcpy(models_out + 9 * i, r.model, sizeof(double) * 9);
      uint8_t* mk = masks_out + (offsets[i] - offsets[0]);
      std::memset(mk, 0, (size_t)specs[(size_t)i].n);
      for (int32_t k : r.inliers) mk[k] = 1;
      num_inliers_out[i] = r.status;
      if (r.status < 0) g_last_error = r.error;
      if (stats) {
        stats[i] = r.stats;
        stats[i].processing_time = secs;
      }
    }
    return GCR_OK;
  });
}

int32_t gcr_score_models(int32_t kind, const double* pts, int64_t n, const double* models, int64_t k, double sq_thr,
                         int64_t* counts, double* values, uint8_t* masks) {
  return guarded([&]() -> int32_t {
    if (!pts || !models || !counts || !values || n <= 0 || k <= 0)
      return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    gcr::score_models(kind == 0 ? 0 : 1, pts, n, models, k, sq_thr, counts, values, masks);
    return GCR_OK;
  });
}

int32_t gcr_solve_minimal(int32_t kind, const double* pts, int64_t n, const int32_t* samples, int64_t c,
                          double* models, int32_t* status) {
  return guarded([&]() -> int32_t {
    if (!pts || !samples || !models || !status || n <= 0 || c <= 0)
      return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
    if (kind < 0 || kind > 2) return fail(GCR_ERR_UNSUPPORTED, "minimal solver kind not built");
    const int m = kind == 0 ? 4 : (kind == 1 ? 7 : 5);
    for (int64_t j = 0; j < c * m; ++j)
      if (samples[j] < 0 || samples[j] >= n) return fail(GCR_ERR_INVALID_ARGUMENT, "sample index out of range");
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    gcr::solve_minimal(kind, pts, n, samples, c, models, status);
    return GCR_OK;
  });
}

int32_t gcr_last_batch_groups(void) { return gcr::last_batch_groups(); }

int32_t gcr_sampler_draws(int32_t sampler, uint32_t seed, int64_t n, int32_t m, const int64_t* chunks,
                          int32_t num_chunks, int32_t* out) {
  return guarded([&]() -> int32_t {
    if (!chunks || !out || num_chunks <= 0 || n <= 0 || n > (int64_t)0x7fffffff || m < 2 || m > 8)
      return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
    if (sampler != 0 && sampler != 1) return fail(GCR_ERR_UNSUPPORTED, "device draws: uniform (0) or PROSAC (1)");
    if (sampler == 1 && n < m) return fail(GCR_ERR_INVALID_ARGUMENT, "PROSAC needs at least m points");
    for (int32_t i = 0; i < num_chunks; ++i)
      if (chunks[i] < 0) return fail(GCR_ERR_INVALID_ARGUMENT, "negative chunk size");
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    gcr::sampler_draws(sampler, seed, n, m, chunks, num_chunks, out);
    return GCR_OK;
  });
}

int32_t gcr_labeling(int32_t kind, const double* pts, int64_t n, const double* model, double lambda,
                     double threshold, const double* cell_sizes, int64_t divisions, uint8_t* sink_out) {
  return guarded([&]() -> int32_t {
    if (!pts || !model || !cell_sizes || !sink_out || n <= 0 || divisions < 1)
      return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    gcr::labeling(kind, pts, n, model, lambda, threshold, cell_sizes, divisions, sink_out);
    return GCR_OK;
  });
}

int32_t gcr_fit_nonminimal(int32_t kind, const double* pts, int64_t n, const int32_t* sample, int64_t ns,
                           const double* weights, int32_t weight_mode, int32_t normalize, double threshold,
                           double* models_out) {
  return guarded([&]() -> int32_t {
    if (!pts || !sample || !models_out || n <= 0 || ns <= 0) return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
    if (kind < 0 || kind > 2) return fail(GCR_ERR_UNSUPPORTED, "fit kind not built");
    if (weight_mode != 0 && !weights) return fail(GCR_ERR_INVALID_ARGUMENT, "weights missing");
    if (weight_mode == 2 && (kind != 1 || normalize)) return fail(GCR_ERR_INVALID_ARGUMENT, "weight_mode 2 is raw F only");
    if (weight_mode == 3 && (kind != 2 || ns < 6))
      return fail(GCR_ERR_INVALID_ARGUMENT, "weight_mode 3 is the essential bundle adjustment (>= 6 points)");
    if (kind == 2 && ns < 5) return fail(GCR_ERR_INVALID_ARGUMENT, "an essential fit needs >= 5 points");
    if (weight_mode == 1 && ns > n) return fail(GCR_ERR_INVALID_ARGUMENT, "weights are read by sample position");
    for (int64_t j = 0; j < ns; ++j)
      if (sample[j] < 0 || sample[j] >= n) return fail(GCR_ERR_INVALID_ARGUMENT, "sample index out of range");
    if (device_count() <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    return gcr::fit_nonminimal(kind, pts, n, sample, ns, weights, weight_mode, normalize, threshold, models_out);
  });
}

void gcr_profile_enable(int32_t on) { gcr::profile_enable(on != 0); }
void gcr_profile_reset(void) { gcr::profile_reset(); }
int32_t gcr_profile_read(const char* family, double* total_ms, int64_t* launches, double* work) {
  if (!family || !total_ms || !launches || !work) return -1;
  return gcr::profile_read(family, total_ms, launches, work) ? 0 : -1;
}

const char* gcr_last_error(void) { return g_last_error.c_str(); }

int32_t gcr_device_count(void) { return device_count(); }

}  // extern "C"
