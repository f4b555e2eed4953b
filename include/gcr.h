/*
 * gcr.h -- C ABI of the MI355X Graph-Cut RANSAC engine (libgcr_hip.so).
 *
 * Plain pointers and sizes only.  Every entry point is synchronous and
 * re-entrant per call; the caller owns every buffer and the library keeps no
 * pointer after returning.  Return values: >= 0 success (for find_* the
 * number of inliers; 0 means "no model"), < 0 an error code below, with a
 * message in gcr_last_error().
 *
 * Reference citations are relative to the danini/graph-cut-ransac root, with
 *   py/ = src/pygcransac/src/   inc/ = src/pygcransac/include/
 */
#ifndef GCR_H_
#define GCR_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCR_OK 0
#define GCR_ERR_INVALID_ARGUMENT (-1) /* bad shape / value (reference: std::invalid_argument) */
#define GCR_ERR_UNSUPPORTED (-2)      /* sampler / neighbourhood / solver id not built */
#define GCR_ERR_DEVICE (-3)           /* HIP runtime error or no device */
#define GCR_ERR_OUT_OF_MEMORY (-4)
#define GCR_ERR_INTERNAL (-5)         /* device min-cut did not converge, cell too large */

#define GCR_PROBLEM_HOMOGRAPHY 0
#define GCR_PROBLEM_FUNDAMENTAL 1
#define GCR_PROBLEM_ESSENTIAL 2

/* Keyword arguments of pygcransac.findHomography / findFundamentalMatrix /
 * findEssentialMatrix (py/bindings.cpp:1003-1022, 1091-1112, 1158-1178). */
/* A host sampler (the template API's user-defined sampler classes,
 * samplers/sampler.h:44-87): writes m indices drawn from pool[0, pool_n)
 * (pool == NULL: the identity pool 0 .. pool_n - 1) into out and returns 1,
 * or returns 0 when it cannot draw a sample (Sampler::sample's false). */
typedef int32_t (*gcr_sample_fn)(void* ctx, const int32_t* pool, int64_t pool_n, int32_t* out, int32_t m);

typedef struct gcr_params {
  double threshold;                /* inlier-outlier threshold (px) */
  double confidence;               /* RANSAC confidence */
  double spatial_coherence_weight; /* lambda of the graph-cut energy */
  int64_t max_iters;
  int64_t min_iters;
  int32_t use_sprt;                /* SPRT preemptive verification of the main loop's models
                                      (preemption_sprt.h; the pool shuffle's rand() stream is seeded
                                      with seed) */
  double min_inlier_ratio_for_sprt; /* SPRT's initial epsilon (F, E; homographies use 0.01 as
                                      findHomography_ does), in (0, 1) */
  int32_t sampler;                 /* 0 uniform, 1 PROSAC (points sorted by quality), 3 importance
                                      (inc/samplers/importance_sampler.h), 4 adaptive reordering
                                      (inc/samplers/adaptive_reordering_sampler.h); 3 and 4 read the
                                      probabilities argument (one per correspondence; a size mismatch
                                      returns 0 = no model, as the reference drivers do) and are not
                                      available in the batch entry points */
  int32_t neighborhood;            /* 0 grid, 1 radius graph (flann_neighborhood_graph.h, searched
                                      exhaustively: every pair within the radius in float
                                      coordinates); another id returns 0 = no model, as the
                                      reference drivers do (GCR_ERR_INVALID_ARGUMENT in batches).
                                      A radius graph without an edge also returns no model. */
  double neighborhood_size;        /* grid divisions per axis, or the radius in pixels */
  int32_t use_space_partitioning;  /* homographies: the space partitioning inlier selector
                                      (space_partitioning_ransac.h) on the grid; a non-grid
                                      neighbourhood is replaced by the grid, as findHomography_ does */
  int32_t lo_number;               /* inner RANSAC trials per graph-cut */
  double sampler_variance;         /* estimator variance of sampler 4 */
  int32_t solver;                  /* 0 point correspondences */
  uint64_t seed;                   /* seeds the main and local-optimisation samplers (and sampler 4's
                                      rand() stream, which the reference leaves unseeded) */
  double time_limit;               /* seconds, 0 = none: GCRANSAC::setFPS's 1 / fps; the main loop
                                      stops after the iteration that exceeds it, timed from the
                                      call's start (GCRANSAC.h:503-516); runs then depend on the
                                      host clock, as the reference's do */
  gcr_sample_fn sample_fn;         /* NULL, or the main loop's sampler in place of `sampler`: every
                                      minimal sample of the main loop comes from
                                      sample_fn(sample_ctx, NULL, n, out, m), in the reference's order
                                      (samples are drawn ahead in chunks, so a sampler whose draws
                                      depend on the loop's update() calls is not supported); the local
                                      optimisation keeps its uniform sampler.  Single-pair calls only
                                      (gcr_run, gcr_find_*); called on one host thread at a time. */
  void* sample_ctx;
} gcr_params_t;

/* Subset of utils::RANSACStatistics (inc/statistics.h:43-71) plus device
 * counters. */
typedef struct gcr_stats {
  int64_t iteration_number;
  int64_t local_optimization_number;
  int64_t graph_cut_number;
  int64_t num_inliers;
  double score;
  double processing_time;     /* seconds, host wall clock around the call */
  int64_t candidates_solved;  /* minimal samples solved on the device */
  int64_t models_scored;      /* hypotheses scored on the device (each over all points) */
} gcr_stats_t;

/* Fills the reference's keyword defaults for a problem kind, except
 * use_sprt = 0, sampler = 0 (uniform) and neighborhood = 0 (grid). */
void gcr_default_params(gcr_params_t* params, int32_t problem);

/* Replaces findHomography_ (py/gcransac_python.cpp:4221, declared at
 * inc/gcransac_python.h:676).  corr: n x 4 row-major (x1 y1 x2 y2).
 * H_out: 3x3 row-major.  mask_out: n bytes (1 = inlier). */
int32_t gcr_find_homography(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                            const double* probabilities, int64_t n_probabilities, const gcr_params_t* params,
                            double* H_out, uint8_t* mask_out, gcr_stats_t* stats);

/* Replaces findFundamentalMatrix_ (py/gcransac_python.cpp:1600, declared at
 * inc/gcransac_python.h:264). */
int32_t gcr_find_fundamental_matrix(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                                    const double* probabilities, int64_t n_probabilities, const gcr_params_t* params,
                                    double* F_out, uint8_t* mask_out, gcr_stats_t* stats);

/* Replaces findEssentialMatrix_ (py/gcransac_python.cpp:2519, declared at
 * inc/gcransac_python.h:411).  K1, K2: 3x3 row-major intrinsics. */
int32_t gcr_find_essential_matrix(const double* corr, int64_t n, const double* K1, const double* K2, int32_t h1,
                                  int32_t w1, int32_t h2, int32_t w2, const double* probabilities,
                                  int64_t n_probabilities, const gcr_params_t* params, double* E_out,
                                  uint8_t* mask_out, gcr_stats_t* stats);

/* GCRANSAC<Estimator, GridNeighborhoodGraph<4>>::run (inc/GCRANSAC.h:202-627)
 * as the C++ template API calls it (inc/GCRANSAC.h:69-84): on the points as
 * given (essential problems: already K-normalised, threshold in those
 * units), with no Python-driver step before or after (no normalisation, no
 * post-hoc bundle adjustment, no identity model).  The grid neighbourhood is
 * GridNeighborhoodGraph<4>(grid_pts, cell_sizes, divisions)
 * (grid_neighborhood_graph.h:229-301); grid_pts NULL means pts.  flags: GCR_RUN_*.  mask_out:
 * n bytes (1 = inlier of the returned model).  No reference counterpart as a
 * C function: it backs include/gcransac.hpp. */
#define GCR_RUN_NO_DEGENSAC 1 /* fundamental: FundamentalMatrixEstimator(ratio, use_degensac = false) */
int32_t gcr_run(int32_t kind, const double* pts, int64_t n, const double* grid_pts, const double* cell_sizes,
                int32_t divisions, int32_t flags, const gcr_params_t* params, double* model_out, uint8_t* mask_out,
                gcr_stats_t* stats);

/* Batched solve of num_pairs independent problems of uneven sizes in one
 * engine run (no reference counterpart: the reference solves one pair per
 * call).  Pair i uses rows offsets[i] .. offsets[i+1]-1 of corr (total x 4),
 * problem kind kinds[i], image sizes sizes[4*i .. 4*i+3] = h1 w1 h2 w2,
 * intrinsics K[18*i .. 18*i+17] = K1, K2 row-major (essential pairs only; K
 * may be NULL otherwise) and params[i].  models_out: 9 per pair; masks_out:
 * total bytes; num_inliers_out: per pair (a negative value is that pair's
 * error code, e.g. a min-cut that gave up; the other pairs' results stand,
 * and gcr_last_error() names the last failed pair); stats: per pair
 * (nullable).  corr_is_device != 0 means corr is a device pointer already
 * resident in HBM (then no host-to-device copy of the correspondences is
 * made). */
int32_t gcr_find_batch(int32_t num_pairs, const int32_t* kinds, const int64_t* offsets, const double* corr,
                       int32_t corr_is_device, const int32_t* sizes, const double* K, const gcr_params_t* params,
                       double* models_out, uint8_t* masks_out, int32_t* num_inliers_out, gcr_stats_t* stats);

/* ---- learned guidance ------------------------------------------------------
 * No reference counterpart.  The reference's hook is its PROSAC sampler,
 * which expects the correspondences ordered by a quality score
 * (inc/samplers/prosac_sampler.h:51-237; py/gcransac_python.cpp:4344-4359
 * passes the caller's order through).  These entry points compute such a
 * score with a per-correspondence network 8-128-128-128-1 (GELU, BF16
 * weights, FP32 accumulation on the matrix cores), train it, and run guided
 * batches.  Parameters: GCR_MLP_NUM_PARAMS floats, row-major W[out][in]:
 *   W1[128][8] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] w4[128] b4
 * Features (GCR_MLP_NUM_FEATURES floats per correspondence): x1 y1 x2 y2
 * scaled to [-1, 1] by the image sizes, the descriptor-ratio score, the
 * mean displacement (x2 - x1) of the 8 nearest neighbours in image 1 minus
 * the point's own, times 2 / max(w1, h1), and a zero pad. */
#define GCR_MLP_NUM_PARAMS 34305
#define GCR_MLP_NUM_FEATURES 8

int32_t gcr_mlp_num_params(void);

/* Features of num_pairs pairs (rows offsets[i] .. offsets[i+1]-1 of corr,
 * sizes 4 per pair = h1 w1 h2 w2, ratios one per row).  feat_out:
 * total x GCR_MLP_NUM_FEATURES, rows in input order. */
int32_t gcr_mlp_features(int32_t num_pairs, const int64_t* offsets, const double* corr, const int32_t* sizes,
                         const double* ratios, float* feat_out);

/* Logits of n feature rows; prob_out (nullable) = sigmoid(logit). */
int32_t gcr_mlp_predict(const float* params, const float* feat, int64_t n, float* logit_out, float* prob_out);

/* Mean binary cross-entropy of n labelled rows and its gradient
 * (GCR_MLP_NUM_PARAMS floats), one device forward + backward pass. */
int32_t gcr_mlp_gradient(const float* params, const float* feat, const float* labels, int64_t n, float* grad_out,
                         float* loss_out);

/* steps Adam steps (lr, betas 0.9 / 0.999, eps 1e-8) on mini-batches of
 * `batch` rows drawn with replacement from the n rows by a hash of (seed,
 * step, slot), all on the device.  params, adam_m, adam_v are updated in
 * place; adam_step = steps already taken; loss_out (nullable): per step. */
int32_t gcr_mlp_train(float* params, float* adam_m, float* adam_v, int64_t adam_step, const float* feat,
                      const float* labels, int64_t n, int32_t batch, int32_t steps, float lr, uint64_t seed,
                      float* loss_out);

/* gcr_find_batch with learned guidance: features, logits, a stable
 * descending sort of each pair's rows by logit, then the batched engine with
 * the PROSAC sampler (params[i].sampler is overridden) on the sorted rows.
 * Masks, models and statistics refer to the caller's row order.  ratios:
 * descriptor-ratio score per row.  inputs_on_device != 0: corr and ratios
 * are device pointers.  logits_out (nullable): total floats. */
int32_t gcr_find_batch_guided(int32_t num_pairs, const int32_t* kinds, const int64_t* offsets, const double* corr,
                              int32_t inputs_on_device, const int32_t* sizes, const double* K, const double* ratios,
                              const float* mlp_params, const gcr_params_t* params, double* models_out,
                              uint8_t* masks_out, int32_t* num_inliers_out, gcr_stats_t* stats, float* logits_out);

/* ---- component entry points (parity tests and benchmarks) ---------------- */

/* MSAC scores of k models (row-major 3x3 each) over n points at squared
 * truncated threshold sq_thr (scoring_function.h:137-233).  kind 0:
 * homography transfer error, 1: Sampson distance.  masks (k x n) nullable. */
int32_t gcr_score_models(int32_t kind, const double* pts, int64_t n, const double* models, int64_t k, double sq_thr,
                         int64_t* counts, double* values, uint8_t* masks);

/* gcr_score_models for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  pts (n x 4), models
 * (k x 9), counts (k, int64), values (k) and masks (k x n, nullable) are
 * DEVICE pointers; the work is queued on `stream` (a hipStream_t; NULL = the
 * null stream) and the call returns without waiting.  The scoring kernel's
 * index tables live in a fixed device array of the library (about 1 MiB per
 * GPU, held until the process ends): k is limited to
 * GCR_SCORE_NOALLOC_MAX_MODELS and larger k is refused.  Calls that overlap
 * on one GPU share that array and race; order them on one stream.  No
 * hipMalloc, hipFree, memset or host synchronisation is made. */
#define GCR_SCORE_NOALLOC_MAX_MODELS 65536
int32_t gcr_score_models_noalloc(int32_t kind, const double* pts, int64_t n, const double* models, int64_t k,
                                 double sq_thr, int64_t* counts, double* values, uint8_t* masks, void* stream);

/* Minimal solves of c samples.  kind 0: 4-point homography, samples c x 4,
 * models c x 9, status c (0 sample rejected, 1 solver failure, 2 ok).
 * kind 1: seven-point fundamental matrix with the oriented-epipolar filter,
 * samples c x 7, models c x 3 x 9, status c x 3 (2 = model kept, in order).
 * kind 2: five-point essential matrix (Nister) with the oriented filter on
 * K-normalised points, samples c x 5, models c x 10 x 9, status c x 10. */
int32_t gcr_solve_minimal(int32_t kind, const double* pts, int64_t n, const int32_t* samples, int64_t c,
                          double* models, int32_t* status);

/* The number of concurrent groups (engines with their own stream and host
 * threads) the last gcr_find_batch / gcr_find_batch_guided call of this
 * process ran its pairs as: the batch-shape autotuner's choice, or
 * GCR_BATCH_GROUPS.  No counterpart in the reference (it runs one pair per
 * call); for callers that report or pin the batch shape. */
int32_t gcr_last_batch_groups(void);

/* Sampler draws on the device (UniformSampler::sample,
 * inc/samplers/uniform_sampler.h:97-115; ProsacSampler::sample,
 * inc/samplers/prosac_sampler.h:176-237; both through
 * inc/uniform_random_generator.h:43-125): num_chunks consecutive chunks of
 * the stream of sampler 0 (uniform over n points; n < m: every index -1) or
 * 1 (PROSAC over n points sorted by quality) seeded with seed, chunks[i]
 * samples of m indices each, each chunk drawn by its own launch.
 * out: sum(chunks) x m. */
int32_t gcr_sampler_draws(int32_t sampler, uint32_t seed, int64_t n, int32_t m, const int64_t* chunks,
                          int32_t num_chunks, int32_t* out);

/* Graph-cut labelling of one model (GCRANSAC::labeling, inc/GCRANSAC.h:913):
 * grid neighbourhood with the given cell sizes / divisions, spatial weight
 * lambda, threshold t.  sink_out: n bytes (1 = SINK = inlier). */
int32_t gcr_labeling(int32_t kind, const double* pts, int64_t n, const double* model, double lambda,
                     double threshold, const double* cell_sizes, int64_t divisions, uint8_t* sink_out);

/* Non-minimal fit of one sample of ns point indices.  kind 0: homography
 * least squares (homography_estimator.h:136-173); kind 1: fundamental-matrix
 * bundle adjustment, Hartley-normalised when normalize != 0
 * (fundamental_estimator.h:574-618), on raw coordinates otherwise; kind 2:
 * essential matrix on K-normalised points, Stewenius solver
 * (essential_estimator.h:308-335).
 * weight_mode 0: unweighted; 1: weights from an n-entry array, read by sample
 * position i for kinds 0 and 1 (how the reference's solvers read IRLS weights
 * on compacted points) and by point index for kind 2; 2 (kind 1, raw):
 * weights from the 3x3 model passed in weights, max(0, 1 - r2 / (1.5 t)^2)
 * per sample point, as the driver's post-hoc adjustment computes them; 3
 * (kind 2): the driver's pose bundle adjustment
 * (py/gcransac_python.cpp:2766-2823) with weights from the model passed in
 * weights.  threshold: t.  models_out: up to 3 x 9 (kinds 0, 1), 10 x 9
 * (kind 2) or 40 x 9 (mode 3); returns the number of models (< 0 on error). */
int32_t gcr_fit_nonminimal(int32_t kind, const double* pts, int64_t n, const int32_t* sample, int64_t ns,
                           const double* weights, int32_t weight_mode, int32_t normalize, double threshold,
                           double* models_out);

/* Kernel profiler: HIP events on the engine's stream around each kernel
 * family ("solve_h4", "solve_f7", "score_h", "score_f", "energy", "mincut",
 * "fit").  read returns 0 and fills total milliseconds, launches and the
 * family's work count (hypothesis-points for score_*), -1 when unknown. */
void gcr_profile_enable(int32_t on);
void gcr_profile_reset(void);
int32_t gcr_profile_read(const char* family, double* total_ms, int64_t* launches, double* work);

/* Last error message of the calling thread. */
const char* gcr_last_error(void);

/* Number of visible HIP devices (0 when none). */
int32_t gcr_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* GCR_H_ */
