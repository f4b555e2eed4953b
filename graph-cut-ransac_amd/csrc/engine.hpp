// Batched Graph-Cut RANSAC engine (host side).
//
// Each image pair runs the reference's GCRANSAC::run control flow
// (inc/GCRANSAC.h:202-1021) as a C++20 coroutine.  Whenever the control flow
// needs data-parallel work it fills a device request and suspends; the driver
// gathers every pending request of one type across all pairs, runs them with
// one batch of kernel launches and one stream synchronisation, and resumes
// the coroutines.  One pair and a thousand pairs use the same code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gcr.h"

namespace gcr {

struct EngineError : std::runtime_error {
  int code;
  EngineError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// An error confined to one pair: that pair ends with status = code and the
// message, the batch's other pairs run on.
struct PairError : EngineError {
  using EngineError::EngineError;
};

struct PairSpec {
  int kind = GCR_PROBLEM_HOMOGRAPHY;
  int64_t offset = 0;  // first row in the batch correspondence array
  int64_t n = 0;
  int h1 = 0, w1 = 0, h2 = 0, w2 = 0;
  double K1[9] = {0}, K2[9] = {0};
  gcr_params_t params{};
  bool cell_override = false;  // use cell[] instead of the driver's cell sizes
  double cell[4] = {0, 0, 0, 0};
  // false: GCRANSAC::run alone (inc/GCRANSAC.h:202-627) on the points as
  // given -- no K normalisation, no post-hoc bundle adjustment, no identity
  // model -- as the C++ template API runs it
  bool driver_stage = true;
  bool degensac = true;  // fundamental: DEGENSAC (the reference estimator's default)
  const double* grid_corr = nullptr;  // host n x 4 points the grid is built on (nullptr: the correspondences)
  const double* probs = nullptr;      // host n point probabilities (samplers 3 and 4)
};

struct PairResult {
  double model[9] = {0};
  std::vector<int32_t> inliers;  // ascending point indices
  gcr_stats_t stats{};
  int status = 0;  // >= 0 number of inliers, < 0 error
  std::string error;  // message of a failed pair
};

// Runs all pairs.  corr: total x 4 (host or device pointer).
void run_batch(const std::vector<PairSpec>& pairs, const double* corr, bool corr_on_device,
               std::vector<PairResult>& results);

// the number of concurrent groups the last batch of this process ran as
int last_batch_groups();

// num_chunks consecutive chunks (chunks[i] samples of m indices) of the
// uniform (sampler 0) or PROSAC (1) sampler over n points seeded with seed,
// each drawn by its own device launch; out: sum(chunks) x m.
void sampler_draws(int32_t sampler, uint32_t seed, int64_t n, int32_t m, const int64_t* chunks, int32_t num_chunks,
                   int32_t* out);

void check_hip(hipError_t e, const char* what);

// message returned by gcr_last_error() on the calling thread
void set_last_error(const std::string& msg);

// Host-side checks of one pair of a batch, the rules the single-pair entry
// points apply (py/bindings.cpp:881-889, 301-309, 733-741): problem kind,
// row count, solver / sampler / neighbourhood ids, confidence, threshold,
// iteration limits and, for essential pairs, non-NULL invertible K1 / K2.
// Returns GCR_OK or an error code with the message set (no device work).
int32_t check_pair(int index, int32_t kind, int64_t n, const gcr_params_t& p, const double* K1, const double* K2);

// component entry points (host buffers in, host buffers out)
void score_models(int kind, const double* pts, int64_t n, const double* models, int64_t k, double sq_thr,
                  int64_t* counts, double* values, uint8_t* masks);
void solve_minimal(int kind, const double* pts, int64_t n, const int32_t* samples, int64_t c, double* models,
                   int32_t* status);
void labeling(int kind, const double* pts, int64_t n, const double* model, double lambda, double threshold,
              const double* cell, int64_t divisions, uint8_t* sink);
int fit_nonminimal(int kind, const double* pts, int64_t n, const int32_t* sample, int64_t ns, const double* weights,
                   int weight_mode, int normalize, double threshold, double* models_out);

// kernel profiler (HIP events on the engine stream)
void profile_enable(bool on);
void profile_reset();
bool profile_read(const std::string& name, double* ms, int64_t* launches, double* work);

}  // namespace gcr
