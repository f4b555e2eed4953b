// CPU oracle C entry points -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Loaded with ctypes by tests/ and by bench.py's cpu_baseline leg.  Each
// find_* function restates the corresponding reference driver
// (py/gcransac_python.cpp) with SPRT and space partitioning off and the grid
// neighbourhood, which is the configuration the engine implements.
#include <chrono>
#include <cstdio>

#include "oracle_essential.hpp"
#include "oracle_fundamental.hpp"
#include "oracle_gcransac.hpp"
#include "oracle_homography.hpp"
#include "oracle_mlp.hpp"

extern "C" {

struct oracle_params {
  double threshold;
  double confidence;
  double spatial_coherence_weight;
  int64_t max_iters;
  int64_t min_iters;
  int32_t sampler;  // 0 uniform, 1 PROSAC
  int32_t lo_number;
  double neighborhood_size;  // grid divisions per axis
  uint64_t seed;
  const double* probabilities;  // per point (samplers 3 and 4), nullable
  double sampler_variance;      // adaptive reordering sampler (4)
  int32_t neighborhood;         // 0 grid (neighborhood_size = divisions), 1 radius (= radius)
  int32_t use_sprt;             // SPRT preemptive verification in the main loop
  double min_inlier_ratio_for_sprt;
  int32_t use_space_partitioning;  // homographies: space partitioning selector on the grid
};

struct oracle_stats {
  int64_t iteration_number;
  int64_t local_optimization_number;
  int64_t graph_cut_number;
  int64_t num_inliers;
  double score;
  double seconds;
};
}

namespace {

// grid_pts: the points the neighbourhood is built on (nullptr: pts).
// Returns 0, -2 for a sampler it cannot build, or kNoGraph when the radius
// neighbourhood has no edge: the drivers then return 0 inliers before any
// sampling (py/gcransac_python.cpp:1683-1691, 2608-2619, 4309-4317).
constexpr int kNoGraph = 1;
// sprt_ratio > 0: the driver's SPRT test with that minimum inlier ratio
// (findHomography_ builds it with the class default 0.01,
// py/gcransac_python.cpp:4390-4393; the F / E drivers pass
// min_inlier_ratio_for_sprt, :1781-1784, 2714-2717).
template <class E>
int run_gcransac(const orc::Points& pts, const E& est, const oracle_params* p, const double cell_sizes[4],
                 bool build_grid, orc::Model& model, std::vector<size_t>& inliers, orc::Statistics& st,
                 const orc::Points* grid_pts = nullptr, double threshold = -1.0, double sprt_ratio = 0.0,
                 bool space_partitioning = false) {
  const size_t m = E::sampleSize();
  const size_t divisions = (size_t)p->neighborhood_size;
  orc::GridNeighborhood grid;
  orc::RadiusNeighborhood radius;
  const orc::Neighborhood* graph = nullptr;
  if (build_grid && p->neighborhood == 1 && !space_partitioning) {
    radius.initialize(grid_pts ? *grid_pts : pts, p->neighborhood_size);
    if (!radius.initialized()) return kNoGraph;
    graph = &radius;
  } else if (build_grid) {
    grid.initialize(grid_pts ? *grid_pts : pts, cell_sizes, divisions);
    graph = &grid;
  }
  std::unique_ptr<orc::Sampler> main_sampler;
  if (p->sampler == 0)
    main_sampler.reset(new orc::UniformSampler(orc::main_sampler_seed(p->seed), pts.rows));
  else if (p->sampler == 1)
    main_sampler.reset(new orc::ProsacSampler(orc::main_sampler_seed(p->seed), pts.rows, m));
  else if (p->sampler == 3 && p->probabilities)
    main_sampler.reset(new orc::ImportanceSampler(orc::main_sampler_seed(p->seed), p->probabilities, pts.rows));
  else if (p->sampler == 4 && p->probabilities) {
    // py/gcransac_python.cpp:4349-4358: probabilities divided by their maximum
    std::vector<double> pr(p->probabilities, p->probabilities + pts.rows);
    double mx = 0;
    for (double v : pr) mx = mx < v ? v : mx;
    for (double& v : pr) v /= mx;
    main_sampler.reset(new orc::AdaptiveReorderingSampler(orc::main_sampler_seed(p->seed), pr.data(), pts.rows,
                                                          p->sampler_variance));
  } else
    return -2;
  orc::UniformSampler lo_sampler(orc::lo_sampler_seed(p->seed), pts.rows);
  std::unique_ptr<orc::SprtPreemption> sprt;
  if (sprt_ratio > 0.0) {
    // the rand() stream: seeded when the main sampler is built (sampler 4
    // seeds it itself)
    if (p->sampler != 4) std::srand(orc::main_sampler_seed(p->seed));
    sprt.reset(new orc::SprtPreemption(pts.rows, m, E::maximumMinimalSolutions(), sprt_ratio));
  }
  orc::GCRANSAC<E> gc;
  gc.sprt_ = sprt.get();
  // findHomography_ runs the selector on the grid (py/gcransac_python.cpp:
  // 4282-4286, 4396-4423)
  std::unique_ptr<orc::SpacePartitioning> sp;
  if (space_partitioning && build_grid) {
    sp.reset(new orc::SpacePartitioning(&grid));
    gc.sp_ = sp.get();
  }
  gc.settings.threshold = threshold > 0.0 ? threshold : p->threshold;
  gc.settings.spatial_coherence_weight = p->spatial_coherence_weight;
  gc.settings.confidence = p->confidence;
  gc.settings.max_local_optimization_number = (size_t)p->lo_number;
  gc.settings.max_iteration_number = (size_t)p->max_iters;
  gc.settings.min_iteration_number = (size_t)p->min_iters;
  gc.run(pts, est, main_sampler.get(), &lo_sampler, graph, model);
  st = gc.getRansacStatistics();
  inliers = st.inliers;
  return 0;
}

int no_model(int64_t n, double* model_out, uint8_t* mask_out, oracle_stats* stats) {
  std::memset(model_out, 0, sizeof(double) * 9);
  std::memset(mask_out, 0, (size_t)n);
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return 0;
}

void fill_stats(oracle_stats* s, const orc::Statistics& st, size_t num_inliers, double score, double secs) {
  if (!s) return;
  s->iteration_number = (int64_t)st.iteration_number;
  s->local_optimization_number = (int64_t)st.local_optimization_number;
  s->graph_cut_number = (int64_t)st.graph_cut_number;
  s->num_inliers = (int64_t)num_inliers;
  s->score = score;
  s->seconds = secs;
}

}  // namespace

extern "C" {

// py/gcransac_python.cpp:4221-4550 (findHomography_), grid neighbourhood with
// cell size (w + eps) / divisions (:4290-4297).
int oracle_find_homography(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                           const oracle_params* p, double* H_out, uint8_t* mask_out, oracle_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  orc::Points pts{corr, (size_t)n, 4};
  const double eps = std::numeric_limits<double>::epsilon();
  const double div = (double)(size_t)p->neighborhood_size;
  const double cells[4] = {(w1 + eps) / div, (h1 + eps) / div, (w2 + eps) / div, (h2 + eps) / div};
  orc::HomographyEstimator est;
  orc::Model model;
  std::vector<size_t> inliers;
  orc::Statistics st;
  const int rc = run_gcransac(pts, est, p, cells, true, model, inliers, st, nullptr, -1.0, p->use_sprt ? 0.01 : 0.0,
                              p->use_space_partitioning != 0);
  if (rc < 0) return rc;
  if (rc == kNoGraph) return no_model(n, H_out, mask_out, stats);
  for (int i = 0; i < 9; ++i) H_out[i] = model.m[i];
  std::memset(mask_out, 0, (size_t)n);
  for (size_t i : inliers) mask_out[i] = 1;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, st, inliers.size(), st.score, secs);
  return (int)inliers.size();
}

// py/gcransac_python.cpp:1600-1921 (findFundamentalMatrix_).  Grid cells are
// w / divisions without the epsilon used for H (:1670-1677); a spatial weight
// <= eps builds an empty grid (:1658-1666).  The post-hoc weighted bundle
// adjustment (:1832-1887) follows the run.
int oracle_find_fundamental_matrix(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                                   const oracle_params* p, double* F_out, uint8_t* mask_out, oracle_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  orc::Points pts{corr, (size_t)n, 4};
  const double div = (double)(size_t)p->neighborhood_size;
  const double cells[4] = {w1 / div, h1 / div, w2 / div, h2 / div};
  const bool grid = p->spatial_coherence_weight > std::numeric_limits<double>::epsilon();
  orc::FundamentalEstimator est(p->seed);
  orc::Model model;
  std::vector<size_t> inliers;
  orc::Statistics st;
  const int rc = run_gcransac(pts, est, p, cells, grid, model, inliers, st, nullptr, -1.0,
                              p->use_sprt ? p->min_inlier_ratio_for_sprt : 0.0);
  if (rc < 0) return rc;
  if (rc == kNoGraph) return no_model(n, F_out, mask_out, stats);
  double score = st.score;
  orc::fundamental_post_bundle_adjustment(pts, est, p->threshold, model, inliers, score);
  for (int i = 0; i < 9; ++i) F_out[i] = model.m[i];
  std::memset(mask_out, 0, (size_t)n);
  for (size_t i : inliers) mask_out[i] = 1;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, st, inliers.size(), score, secs);
  return (int)inliers.size();
}

// py/gcransac_python.cpp:2519-2855 (findEssentialMatrix_).  RANSAC runs on
// K-normalised coordinates with threshold / (mean focal length); the grid is
// built on the pixel coordinates (:2562-2600); the post-hoc pose bundle
// adjustment (:2766-2823) follows the run.
int oracle_find_essential_matrix(const double* corr, int64_t n, const double* K1, const double* K2, int32_t h1,
                                 int32_t w1, int32_t h2, int32_t w2, const oracle_params* p, double* E_out,
                                 uint8_t* mask_out, oracle_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  orc::Points pix{corr, (size_t)n, 4};
  std::vector<double> norm(4 * (size_t)n);
  orc::normalise_correspondences(corr, (size_t)n, K1, K2, norm.data());
  orc::Points pts{norm.data(), (size_t)n, 4};
  const double normaliser = 0.25 * (K1[0] + K1[4] + K2[0] + K2[4]);
  const double div = (double)(size_t)p->neighborhood_size;
  const double cells[4] = {w1 / div, h1 / div, w2 / div, h2 / div};
  const bool grid = p->spatial_coherence_weight > std::numeric_limits<double>::epsilon();
  orc::EssentialEstimator est;
  orc::Model model;
  std::vector<size_t> inliers;
  orc::Statistics st;
  const int rc = run_gcransac(pts, est, p, cells, grid, model, inliers, st, &pix, p->threshold / normaliser,
                              p->use_sprt ? p->min_inlier_ratio_for_sprt : 0.0);
  if (rc < 0) return rc;
  if (rc == kNoGraph) return no_model(n, E_out, mask_out, stats);
  double score = st.score;
  orc::essential_post_bundle_adjustment(pts, est, 3.0 / 2.0 * p->threshold / normaliser, model, inliers, score);
  for (int i = 0; i < 9; ++i) E_out[i] = model.m[i];
  std::memset(mask_out, 0, (size_t)n);
  for (size_t i : inliers) mask_out[i] = 1;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, st, inliers.size(), score, secs);
  return (int)inliers.size();
}

// GCRANSAC<Estimator, GridNeighborhoodGraph<4>>::run (inc/GCRANSAC.h:202-627)
// as the template API calls it: the points as given (essential: already
// K-normalised, threshold in those units), the grid built on grid_pts (nullptr:
// pts) with the given cells, no driver step before or after.  kind 0 H,
// 1 F, 2 E.  Returns the inlier count; the model is left untouched when
// none is found.
int oracle_run(int32_t kind, const double* pts_, int64_t n, const double* grid_pts_, const double* cell_sizes,
               int32_t flags, const oracle_params* p, double* model_out, uint8_t* mask_out, oracle_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  orc::Points pts{pts_, (size_t)n, 4};
  orc::Points gp{grid_pts_ ? grid_pts_ : pts_, (size_t)n, 4};
  orc::Model model;
  std::vector<size_t> inliers;
  orc::Statistics st;
  // the template's grid of the given cells; its preemption slot (SPRT with
  // the object's minimum inlier ratio) and selector slot (space
  // partitioning, homographies)
  oracle_params q = *p;
  q.neighborhood = 0;
  const double sprt = q.use_sprt ? q.min_inlier_ratio_for_sprt : 0.0;
  int rc;
  if (kind == 0) {
    orc::HomographyEstimator est;
    rc = run_gcransac(pts, est, &q, cell_sizes, true, model, inliers, st, &gp, -1.0, sprt,
                      q.use_space_partitioning != 0);
  } else if (kind == 1) {
    orc::FundamentalEstimator est(q.seed, 0.5, (flags & 1) == 0);  // bit 0: no DEGENSAC
    rc = run_gcransac(pts, est, &q, cell_sizes, true, model, inliers, st, &gp, -1.0, sprt);
  } else if (kind == 2) {
    orc::EssentialEstimator est;
    rc = run_gcransac(pts, est, &q, cell_sizes, true, model, inliers, st, &gp, -1.0, sprt);
  } else {
    return -1;
  }
  if (rc < 0) return rc;
  std::memset(mask_out, 0, (size_t)n);
  if (rc == kNoGraph) {
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return 0;
  }
  for (size_t i : inliers) mask_out[i] = 1;
  if (!inliers.empty())
    for (int i = 0; i < 9; ++i) model_out[i] = model.m[i];
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(stats, st, inliers.size(), st.score, secs);
  return (int)inliers.size();
}

// ---- component-level entry points used by the parity tests ----------------

// kind: 0 homography, 1 fundamental / essential (Sampson).
// Scores k models (row-major 3x3 each) with the MSAC rule at the squared
// truncated threshold sq_thr; writes count, value and (optionally) the k x n
// inlier mask.
void oracle_score(int32_t kind, const double* pts_, int64_t n, const double* models, int64_t k, double sq_thr,
                  int64_t* counts, double* values, uint8_t* masks) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::HomographyEstimator he;
  orc::FundamentalEstimator fe(0);
  std::vector<size_t> inl;
  for (int64_t j = 0; j < k; ++j) {
    orc::Model m;
    std::memcpy(m.m, models + 9 * j, sizeof(m.m));
    const orc::Score s = kind == 0 ? orc::msac_score(pts, m, he, sq_thr, &inl) : orc::msac_score(pts, m, fe, sq_thr, &inl);
    counts[j] = (int64_t)s.inlier_number;
    values[j] = s.value;
    if (masks) {
      std::memset(masks + j * n, 0, (size_t)n);
      for (size_t i : inl) masks[j * n + i] = 1;
    }
  }
}

// 4-point minimal solves of c samples (4 indices each).  status: 0 invalid
// sample (oriented test), 1 solver failure (NaN), 2 model written.
void oracle_h4_solve(const double* pts_, int64_t n, const int32_t* samples, int64_t c, double* models,
                     int32_t* status) {
  orc::Points pts{pts_, (size_t)n, 4};
  for (int64_t j = 0; j < c; ++j) {
    size_t s[4];
    for (int i = 0; i < 4; ++i) s[i] = (size_t)samples[4 * j + i];
    if (!orc::h_is_valid_sample(pts, s)) {
      status[j] = 0;
      continue;
    }
    orc::Model m;
    if (!orc::h4_minimal(pts, s, m)) {
      status[j] = 1;
      continue;
    }
    std::memcpy(models + 9 * j, m.m, sizeof(m.m));
    status[j] = 2;
  }
}

// Seven-point solves with the orientation filter (fundamental_estimator.h:161-184).
// models: c x 3 x 9; status: c x 3 (2 = model kept, in the solver's order).
void oracle_f7_solve(const double* pts_, int64_t n, const int32_t* samples, int64_t c, double* models,
                     int32_t* status) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::FundamentalEstimator fe(0);
  for (int64_t j = 0; j < c; ++j) {
    size_t s[7];
    for (int i = 0; i < 7; ++i) s[i] = (size_t)samples[7 * j + i];
    std::vector<orc::Model> ms;
    fe.estimateModel(pts, s, &ms);
    for (int k = 0; k < 3; ++k) {
      status[3 * j + k] = k < (int)ms.size() ? 2 : 0;
      if (k < (int)ms.size()) std::memcpy(models + (3 * j + k) * 9, ms[k].m, sizeof(double) * 9);
    }
  }
}

// Non-minimal fit of a sample.  kind 0: homography LS; kind 1: the
// fundamental estimator's normalised bundle adjustment.  weight_mode 0: none;
// 1: weights[i] read by sample position i (the reference's indexing on its
// compacted points); 2 (kind 1, normalize == 0): the post-hoc bundle
// adjustment on raw coordinates, weights[k] by sample position.
int oracle_fit_nonminimal(int32_t kind, const double* pts_, int64_t n, const int64_t* sample, int64_t ns,
                          const double* weights, int32_t weight_mode, int32_t normalize, double threshold,
                          double* models_out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  std::vector<orc::Model> models;
  if (kind == 0) {
    orc::HomographyEstimator he;
    if (!he.estimateModelNonminimal(pts, s.data(), (size_t)ns, &models, weight_mode ? weights : nullptr)) return 0;
  } else if (normalize) {
    orc::FundamentalEstimator fe(0);
    if (!fe.estimateModelNonminimal(pts, s.data(), (size_t)ns, &models, weight_mode ? weights : nullptr)) return 0;
  } else {
    // weight_mode 2: weights from the model passed in `weights`, as the
    // driver's post-hoc adjustment computes them (py/gcransac_python.cpp:1838-1846)
    std::vector<double> w((size_t)ns, 1.0);
    if (weight_mode == 2) {
      orc::Model wm;
      std::memcpy(wm.m, weights, sizeof(wm.m));
      const double tt = 3.0 / 2.0 * threshold, sq = tt * tt;
      for (int64_t k = 0; k < ns; ++k) {
        const double v = 1.0 - orc::f_squared_sampson(pts.row(s[(size_t)k]), wm.m) / sq;
        w[(size_t)k] = 0 > v ? 0 : v;
      }
    } else if (weight_mode == 1) {
      for (int64_t k = 0; k < ns; ++k) w[(size_t)k] = weights[k];
    }
    auto row = [&](size_t k) { return pts.row(s[k]); };
    auto wt = [&](size_t k) { return w[k]; };
    orc::f_bundle_adjustment((size_t)ns, row, wt, weight_mode != 0, models);
  }
  for (size_t k = 0; k < models.size(); ++k) std::memcpy(models_out + 9 * k, models[k].m, sizeof(double) * 9);
  return (int)models.size();
}

// Weighted / unweighted homography LS fit of a sample (normalise, QR, denormalise).
int oracle_h_nonminimal(const double* pts_, int64_t n, const int64_t* sample, int64_t ns, const double* weights,
                        double* model_out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  orc::HomographyEstimator he;
  std::vector<orc::Model> models;
  if (!he.estimateModelNonminimal(pts, s.data(), (size_t)ns, &models, weights)) return 0;
  std::memcpy(model_out, models[0].m, sizeof(double) * 9);
  return 1;
}

// Draws of the reference's UniformSampler for a pool of size pool_n:
// `count` samples of m unique indices.
void oracle_uniform_draws(uint32_t seed, int64_t pool_n, int32_t m, int64_t count, int64_t* out) {
  orc::UniformSampler s(seed, (size_t)pool_n);
  std::vector<size_t> pool((size_t)pool_n);
  for (size_t i = 0; i < pool.size(); ++i) pool[i] = i;
  std::vector<size_t> sub((size_t)m);
  for (int64_t j = 0; j < count; ++j) {
    s.sample(pool, sub.data(), (size_t)m);
    for (int i = 0; i < m; ++i) out[j * m + i] = (int64_t)sub[i];
  }
}

void oracle_prosac_draws(uint32_t seed, int64_t n, int32_t m, int64_t count, int64_t* out) {
  orc::ProsacSampler s(seed, (size_t)n, (size_t)m);
  std::vector<size_t> pool;
  std::vector<size_t> sub((size_t)m);
  for (int64_t j = 0; j < count; ++j) {
    s.sample(pool, sub.data(), (size_t)m);
    for (int i = 0; i < m; ++i) out[j * m + i] = (int64_t)sub[i];
  }
}

void oracle_prosac_growth(int64_t n, int32_t m, int64_t* out) {
  orc::ProsacSampler s(0, (size_t)n, (size_t)m);
  for (int64_t i = 0; i < n; ++i) out[i] = (int64_t)s.growthFunction()[i];
}

// Grid cell keys (inc/neighborhood/grid_neighborhood_graph.h:229-292).
void oracle_grid_keys(const double* pts_, int64_t n, const double* cell_sizes, int64_t divisions, uint64_t* keys,
                      int64_t* edge_count) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::GridNeighborhood g(pts, cell_sizes, (size_t)divisions);
  for (int64_t i = 0; i < n; ++i) keys[i] = (uint64_t)g.cellKey((size_t)i);
  *edge_count = (int64_t)g.neighborNumber();
}

// Graph-cut labelling of one model (inc/GCRANSAC.h:913-1021).  kind as in
// oracle_score.  Writes the SINK mask and the problem's terminal capacities
// and arc capacities (for fixtures); returns the edge count.
}  // extern "C"

namespace {
int64_t labeling_on(const orc::Neighborhood& g, int32_t kind, const orc::Points& pts, const double* model,
                    double lambda, double threshold, uint8_t* sink_out, double* tr_cap_out, int64_t max_edges,
                    int32_t* ep, int32_t* eq, double* cpq, double* cqp) {
  const int64_t n = (int64_t)pts.rows;
  orc::Model m;
  std::memcpy(m.m, model, sizeof(m.m));
  orc::GraphCutProblem prob;
  if (kind == 0) {
    orc::HomographyEstimator he;
    orc::GCRANSAC<orc::HomographyEstimator>::build_graph_cut(pts, m, he, lambda, threshold, &g, prob);
  } else {
    orc::FundamentalEstimator fe(0);
    orc::GCRANSAC<orc::FundamentalEstimator>::build_graph_cut(pts, m, fe, lambda, threshold, &g, prob);
  }
  const std::vector<uint8_t> sink = orc::solve_graph_cut(prob);
  std::memcpy(sink_out, sink.data(), (size_t)n);
  if (tr_cap_out) std::memcpy(tr_cap_out, prob.tr_cap.data(), sizeof(double) * (size_t)n);
  const int64_t ne = (int64_t)prob.edge_p.size();
  if (ep && ne <= max_edges) {
    for (int64_t e = 0; e < ne; ++e) {
      ep[e] = prob.edge_p[e];
      eq[e] = prob.edge_q[e];
      cpq[e] = prob.cap_pq[e];
      cqp[e] = prob.cap_qp[e];
    }
  }
  return ne;
}
}  // namespace

extern "C" {

int64_t oracle_labeling(int32_t kind, const double* pts_, int64_t n, const double* model, double lambda,
                        double threshold, const double* cell_sizes, int64_t divisions, uint8_t* sink_out,
                        double* tr_cap_out, int64_t max_edges, int32_t* ep, int32_t* eq, double* cpq, double* cqp) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::GridNeighborhood g(pts, cell_sizes, (size_t)divisions);
  return labeling_on(g, kind, pts, model, lambda, threshold, sink_out, tr_cap_out, max_edges, ep, eq, cpq, cqp);
}

// The same labelling on the radius neighbourhood
// (inc/neighborhood/flann_neighborhood_graph.h:100-139).
int64_t oracle_labeling_radius(int32_t kind, const double* pts_, int64_t n, const double* model, double lambda,
                               double threshold, double radius, uint8_t* sink_out, double* tr_cap_out,
                               int64_t max_edges, int32_t* ep, int32_t* eq, double* cpq, double* cqp) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::RadiusNeighborhood g(pts, radius);
  return labeling_on(g, kind, pts, model, lambda, threshold, sink_out, tr_cap_out, max_edges, ep, eq, cpq, cqp);
}

// Radius neighbour lists (inc/neighborhood/flann_neighborhood_graph.h:
// 100-139, exhaustive): list i is nbr[off[i] .. off[i+1]).  Returns the total
// length (the reference's neighbor_number); the lists are written when it is
// at most max_total.
int64_t oracle_radius_neighbors(const double* pts_, int64_t n, double radius, int64_t* off, int32_t* nbr,
                                int64_t max_total) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::RadiusNeighborhood g(pts, radius);
  const int64_t total = (int64_t)g.neighborNumber();
  if (off) {
    int64_t o = 0;
    for (int64_t i = 0; i < n; ++i) {
      off[i] = o;
      for (size_t q : g.getNeighbors((size_t)i)) {
        if (nbr && total <= max_total) nbr[o] = (int32_t)q;
        ++o;
      }
    }
    off[n] = o;
  }
  return total;
}

float oracle_radius_sq_distance(const float* a, const float* b) {
  return orc::RadiusNeighborhood::squared_distance(a, b);
}

// Energy terms of the labelling of one model, in the reference's call order:
// unary (E(x=0), E(x=1)) per point and pairwise (A, B, C, D) per same-cell
// pair p<q (inc/GCRANSAC.h:945-1008).  Returns the edge count; arrays are
// filled when max_edges is large enough.
int64_t oracle_energy_terms(int32_t kind, const double* pts_, int64_t n, const double* model, double lambda,
                            double threshold, const double* cell_sizes, int64_t divisions, double* ua, double* ub,
                            int64_t max_edges, int32_t* ep, int32_t* eq, double* ta, double* tb, double* tc,
                            double* td) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::GridNeighborhood g(pts, cell_sizes, (size_t)divisions);
  orc::Model m;
  std::memcpy(m.m, model, sizeof(m.m));
  orc::HomographyEstimator he;
  orc::FundamentalEstimator fe(0);
  std::vector<double> dist((size_t)n);
  const double sq_thr = threshold * threshold * 9 / 4;
  const double olm = 1.0 - lambda;
  for (int64_t i = 0; i < n; ++i) {
    const double r2 = kind == 0 ? he.squaredResidual(pts.row((size_t)i), m) : fe.squaredResidual(pts.row((size_t)i), m);
    double d = r2 / sq_thr;
    d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
    dist[(size_t)i] = d;
    const double e = 1 - d;
    if (r2 <= sq_thr) { ua[i] = olm * e; ub[i] = 0; }
    else { ua[i] = 0; ub[i] = olm * (1 - e); }
  }
  int64_t ne = 0;
  if (lambda > 0)
    for (int64_t p = 0; p < n; ++p)
      for (size_t q : g.getNeighbors((size_t)p)) {
        if ((int64_t)q <= p) continue;
        if (ne < max_edges && ep) {
          ep[ne] = (int32_t)p; eq[ne] = (int32_t)q;
          ta[ne] = 0.5 * (dist[(size_t)p] + dist[q]) * lambda; tb[ne] = lambda; tc[ne] = lambda; td[ne] = 0.0;
        }
        ++ne;
      }
  return ne;
}

// SINK labels of an energy given as terms (same form as oracle_energy_terms).
void oracle_energy_mincut(int64_t n, const double* ua, const double* ub, int64_t ne, const int32_t* ep,
                          const int32_t* eq, const double* ta, const double* tb, const double* tc, const double* td,
                          uint8_t* sink_out) {
  orc::GraphCutProblem g;
  g.tr_cap.assign((size_t)n, 0.0);
  for (int64_t i = 0; i < n; ++i) orc::add_tweights(g, (int)i, ub[i], ua[i]);
  for (int64_t e = 0; e < ne; ++e) orc::add_term2(g, ep[e], eq[e], ta[e], tb[e], tc[e], td[e]);
  const std::vector<uint8_t> s = orc::solve_graph_cut(g);
  std::memcpy(sink_out, s.data(), (size_t)n);
}

// Min-cut of an explicit graph given as terminal residuals and arc pairs.
void oracle_mincut(int64_t n, const double* tr_cap, int64_t ne, const int32_t* ep, const int32_t* eq,
                   const double* cpq, const double* cqp, uint8_t* sink_out) {
  orc::GraphCutProblem g;
  g.tr_cap.assign(tr_cap, tr_cap + n);
  g.edge_p.assign(ep, ep + ne);
  g.edge_q.assign(eq, eq + ne);
  g.cap_pq.assign(cpq, cpq + ne);
  g.cap_qp.assign(cqp, cqp + ne);
  const std::vector<uint8_t> s = orc::solve_graph_cut(g);
  std::memcpy(sink_out, s.data(), (size_t)n);
}

// ---- essential-matrix components ---------------------------------------------

// Sturm<10>::solve on 11 ascending coefficients; returns the root count.
int oracle_sturm10(const double* coeffs, double* roots) {
  orc::Sturm<10> st;
  return st.solve(coeffs, roots);
}

// Five-point minimal solves of c samples (c x 5 point indices, int32) on
// normalised points.  orient != 0 applies the estimator's oriented filter.
// models: c x 10 x 9 row-major; nmodels: c.
void oracle_e5_solve(const double* pts_, int64_t n, const int32_t* samples, int64_t c, int32_t orient,
                     double* models, int32_t* nmodels) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::EssentialEstimator est;
  for (int64_t k = 0; k < c; ++k) {
    size_t s[5];
    for (int i = 0; i < 5; ++i) s[i] = (size_t)samples[5 * k + i];
    std::vector<orc::Model> ms;
    if (orient) est.estimateModel(pts, s, &ms);
    else orc::nister_solve(pts, s, ms);
    nmodels[k] = (int32_t)ms.size();
    for (size_t j = 0; j < ms.size() && j < 10; ++j) std::memcpy(models + (10 * k + (int64_t)j) * 9, ms[j].m, 72);
  }
}

// Stewenius non-minimal fit through the estimator (Frobenius-normalised,
// E33 >= 0).  weights: per point index (nullptr: unweighted).  models_out:
// up to 10 x 9.  Returns the count.
int oracle_e_nonminimal(const double* pts_, int64_t n, const int64_t* sample, int64_t ns, const double* weights,
                        double* models_out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  std::vector<orc::Model> ms;
  orc::EssentialEstimator est;
  if (!est.estimateModelNonminimal(pts, s.data(), (size_t)ns, &ms, weights)) return 0;
  for (size_t j = 0; j < ms.size() && j < 10; ++j) std::memcpy(models_out + 9 * j, ms[j].m, 72);
  return (int)ms.size();
}

// The driver's pose bundle adjustment on a sample with per-position weights:
// Stewenius initial models (init_out, up to 10 x 9, count in *n_init), the
// cheirality-checked poses of each, refined (out, up to 40 x 9).  Returns the
// number of refined models.
int oracle_e_bundle(const double* pts_, int64_t n, const int64_t* sample, int64_t ns, const double* w_pos,
                    double* init_out, int32_t* n_init, double* out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  std::vector<orc::Model> init;
  orc::stewenius_solve(pts, s.data(), (size_t)ns, [&](size_t i) { return w_pos ? w_pos[i] : 1.0; }, init);
  *n_init = (int32_t)init.size();
  for (size_t j = 0; j < init.size() && j < 10; ++j) std::memcpy(init_out + 9 * j, init[j].m, 72);
  const double* p0 = pts.row(s[0]);
  double x1[3] = {p0[0], p0[1], 1.0}, x2[3] = {p0[2], p0[3], 1.0};
  orc::normalise3(x1);
  orc::normalise3(x2);
  int cnt = 0;
  for (auto& m0 : init) {
    std::vector<orc::RelPose> poses;
    orc::motion_from_essential(m0.m, x1, x2, poses);
    for (auto& pose : poses) {
      orc::refine_relpose((size_t)ns, [&](size_t k) { return pts.row(s[k]); },
                          [&](size_t k) { return w_pos ? w_pos[k] : 1.0; }, pose);
      if (cnt < 40) orc::essential_from_motion(pose, out + 9 * cnt);
      ++cnt;
    }
  }
  return cnt;
}

// motion_from_essential: poses (q w x y z, t) of E passing the cheirality
// test of (x1, x2) (homogeneous, unit norm).  poses_out: up to 4 x 7.
int oracle_motion_from_essential(const double* E, const double* x1, const double* x2, double* poses_out) {
  std::vector<orc::RelPose> poses;
  orc::motion_from_essential(E, x1, x2, poses);
  for (size_t j = 0; j < poses.size() && j < 4; ++j) {
    std::memcpy(poses_out + 7 * j, poses[j].q, 32);
    std::memcpy(poses_out + 7 * j + 4, poses[j].t, 24);
  }
  return (int)poses.size();
}

// K^-1 normalisation of correspondences as the essential driver does it.
void oracle_normalise_correspondences(const double* corr, int64_t n, const double* K1, const double* K2,
                                      double* out) {
  orc::normalise_correspondences(corr, (size_t)n, K1, K2, out);
}

// The driver's post-hoc pose bundle adjustment of a sample (weights from
// `model` by position at the truncated threshold trunc_thr): the refined
// models in order (up to 40 x 9).  Returns their number.
int oracle_e_post_bundle(const double* pts_, int64_t n, const int64_t* sample, int64_t ns, const double* model,
                         double trunc_thr, double* out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  const double sq = trunc_thr * trunc_thr;
  std::vector<double> w((size_t)ns);
  for (int64_t k = 0; k < ns; ++k) {
    const double v = 1.0 - orc::f_squared_sampson(pts.row(s[(size_t)k]), model) / sq;
    w[(size_t)k] = 0 > v ? 0 : v;
  }
  std::vector<orc::Model> init;
  orc::stewenius_solve(pts, s.data(), (size_t)ns, [&](size_t i) { return w[i]; }, init);
  const double* p0 = pts.row(s[0]);
  double x1[3] = {p0[0], p0[1], 1.0}, x2[3] = {p0[2], p0[3], 1.0};
  orc::normalise3(x1);
  orc::normalise3(x2);
  int cnt = 0;
  for (auto& m0 : init) {
    std::vector<orc::RelPose> poses;
    orc::motion_from_essential(m0.m, x1, x2, poses);
    for (auto& pose : poses) {
      orc::refine_relpose((size_t)ns, [&](size_t k) { return pts.row(s[k]); }, [&](size_t k) { return w[k]; }, pose);
      if (cnt < 40) orc::essential_from_motion(pose, out + 9 * cnt);
      ++cnt;
    }
  }
  return cnt;
}

// ---- learned inlier-probability network (oracle_mlp.hpp) ----------------

void oracle_mlp_features(const double* corr, int64_t n, int32_t h1, int32_t w1, int32_t h2, int32_t w2,
                         const double* ratio, float* feat) {
  orc::mlp::knn_features(corr, n, w1, h1, w2, h2, ratio, feat);
}

void oracle_mlp_knn(const double* corr, int64_t n, int32_t* out) { orc::mlp::knn_indices(corr, n, out); }

int32_t oracle_mlp_num_params(void) { return orc::mlp::NPARAM; }

// exact != 0: double arithmetic without BF16 rounding (params read as float).
void oracle_mlp_logits(const float* params, const float* feat, int64_t n, int32_t exact, double* out) {
  if (exact) {
    orc::mlp::Net<orc::mlp::Exact> net(params);
    for (int64_t i = 0; i < n; ++i) out[i] = net.logit(feat + orc::mlp::IN * i);
  } else {
    orc::mlp::Net<orc::mlp::RoundBF16> net(params);
    for (int64_t i = 0; i < n; ++i) out[i] = net.logit(feat + orc::mlp::IN * i);
  }
}

// Mean BCE and its gradient; params_d (double, exact mode only) overrides
// params when non-null so finite differences can perturb below float
// resolution.
double oracle_mlp_loss_grad(const float* params, const double* params_d, const float* feat, const float* labels,
                            int64_t n, int32_t exact, double* grad) {
  if (exact) {
    if (params_d) {
      orc::mlp::Net<orc::mlp::Exact> net(params_d);
      return net.loss_grad(feat, labels, n, grad);
    }
    orc::mlp::Net<orc::mlp::Exact> net(params);
    return net.loss_grad(feat, labels, n, grad);
  }
  orc::mlp::Net<orc::mlp::RoundBF16> net(params);
  return net.loss_grad(feat, labels, n, grad);
}

void oracle_mlp_adam(float* params, float* m, float* v, const float* grad, int64_t t, float lr) {
  orc::mlp::adam_step(params, m, v, grad, t, lr);
}

void oracle_mlp_gelu(const double* z, int64_t n, double* g, double* dg) {
  for (int64_t i = 0; i < n; ++i) {
    g[i] = orc::mlp::gelu<orc::mlp::Exact>(z[i]);
    dg[i] = orc::mlp::gelu_grad<orc::mlp::Exact>(z[i]);
  }
}

void oracle_mlp_round_bf16(const float* x, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc::mlp::round_bf16(x[i]);
}


// ---- single-function entry points, used to pin the restatement ------------

// Seven-point solver without the orientation filter (seven_point.h:91-291):
// the models of the determinant cubic's real roots (models_out 3 x 9) and the
// cubic's coefficients c0..c3 of det(lambda f1 + f2) with f1 -= f2 applied
// (coeffs_out).  Returns the model count (0: sample rejected).
int oracle_f7_raw(const double* pts_, int64_t n, const int32_t* sample, double* models_out) {
  orc::Points pts{pts_, (size_t)n, 4};
  size_t s[7];
  for (int i = 0; i < 7; ++i) s[i] = (size_t)sample[i];
  std::vector<orc::Model> ms;
  if (!orc::f7_solve(pts, s, ms)) return 0;
  for (size_t k = 0; k < ms.size() && k < 3; ++k) std::memcpy(models_out + 9 * k, ms[k].m, 72);
  return (int)ms.size();
}

// Oriented epipolar constraint of F on ns sample points (fundamental_estimator.h:737-800).
int oracle_f_orientation_valid(const double* pts_, int64_t n, const double* F, const int32_t* sample, int64_t ns) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> s(sample, sample + ns);
  return orc::f_orientation_valid(F, pts, s.data(), (size_t)ns) ? 1 : 0;
}

// Squared Sampson and squared symmetric epipolar distances of every row
// (fundamental_estimator.h:195-246).
void oracle_f_residuals(const double* pts_, int64_t n, const double* F, double* sampson, double* symmetric) {
  for (int64_t i = 0; i < n; ++i) {
    sampson[i] = orc::f_squared_sampson(pts_ + 4 * i, F);
    symmetric[i] = orc::f_squared_symmetric(pts_ + 4 * i, F);
  }
}

// Eight-point null vector of the rows as given (eight_point.h:93-182), no
// normalisation; weights (nullable) per row.
void oracle_f8_raw(const double* rows, int64_t n, const double* weights, double* F_out) {
  auto row = [&](size_t i) { return rows + 4 * i; };
  auto wt = [&](size_t i) { return weights[i]; };
  orc::f8_solve((size_t)n, row, wt, weights != nullptr, F_out);
}

// The rank-2 factorisation the LM starts from and its first normal equations
// (jacobian_impl.h:468-614): U, V row-major, sigma, Jtr (7), JtJ (7 x 7,
// lower triangle filled, upper mirrored).
void oracle_lm_f_normal_equations(const double* rows, int64_t n, const double* weights, const double* F,
                                  double* U_out, double* V_out, double* sigma_out, double* Jtr_out,
                                  double* JtJ_out) {
  auto row = [&](size_t i) { return rows + 4 * i; };
  auto wt = [&](size_t i) { return weights[i]; };
  orc::FactorizedF FF;
  FF.from(F);
  double Jtr[7], JtJ[7][7];
  orc::lm_f_normal_equations((size_t)n, row, wt, weights != nullptr, FF, Jtr, JtJ);
  std::memcpy(U_out, FF.U, 72);
  std::memcpy(V_out, FF.V, 72);
  *sigma_out = FF.sigma;
  for (int i = 0; i < 7; ++i) {
    Jtr_out[i] = Jtr[i];
    for (int j = 0; j <= i; ++j) JtJ_out[7 * i + j] = JtJ_out[7 * j + i] = JtJ[i][j];
  }
}

// LM refinement of F on the rows as given (bundle.cpp:252-339).  F_io in and
// out; trace_out (up to max_trace) receives the cost before the first step,
// then the cost of each trial step, accepted_out (up to max_trace - 1) 1 for
// each trial step that was kept.  Returns the trace length.
int oracle_lm_f_refine(const double* rows, int64_t n, const double* weights, double* F_io, int32_t max_iterations,
                       double* trace_out, uint8_t* accepted_out, int32_t max_trace) {
  auto row = [&](size_t i) { return rows + 4 * i; };
  auto wt = [&](size_t i) { return weights[i]; };
  std::vector<double> tr;
  std::vector<char> acc;
  orc::lm_refine_f((size_t)n, row, wt, weights != nullptr, F_io, max_iterations, &tr, &acc);
  const int k = (int)std::min<size_t>(tr.size(), (size_t)max_trace);
  for (int i = 0; i < k; ++i) trace_out[i] = tr[(size_t)i];
  for (int i = 0; i + 1 < k; ++i) accepted_out[i] = (uint8_t)acc[(size_t)i];
  return (int)tr.size();
}

// Truncated Sampson cost of F on the rows (robust_loss.h:19-31).
double oracle_lm_f_cost(const double* rows, int64_t n, const double* weights, const double* F) {
  auto row = [&](size_t i) { return rows + 4 * i; };
  auto wt = [&](size_t i) { return weights[i]; };
  return orc::lm_f_cost((size_t)n, row, wt, weights != nullptr, F);
}

// isValidModel of the fundamental (kind 1, min_ratio 0.5, DEGENSAC off) or
// essential (kind 2, min_ratio 0.1) estimator on the given inlier list at
// threshold t (fundamental_estimator.h:287-323, essential_estimator.h:279-305).
int oracle_epipolar_is_valid_model(int32_t kind, const double* pts_, int64_t n, const double* model,
                                   const int64_t* inliers, int64_t ni, double threshold) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::vector<size_t> in(inliers, inliers + ni);
  orc::Model m;
  std::memcpy(m.m, model, 72);
  bool updated = false;
  const size_t s[7] = {0, 0, 0, 0, 0, 0, 0};
  if (kind == 1) {
    orc::FundamentalEstimator fe(0, 0.5, false);
    return fe.isValidModel(m, pts, in, s, threshold, updated) ? 1 : 0;
  }
  orc::EssentialEstimator ee;
  return ee.isValidModel(m, pts, in, s, threshold, updated) ? 1 : 0;
}

// Homography sample test (homography_estimator.h:346-381) and model test
// (homography_estimator.h:326-342).
int oracle_h_is_valid_sample(const double* pts_, int64_t n, const int32_t* sample) {
  orc::Points pts{pts_, (size_t)n, 4};
  const size_t s[4] = {(size_t)sample[0], (size_t)sample[1], (size_t)sample[2], (size_t)sample[3]};
  return orc::h_is_valid_sample(pts, s) ? 1 : 0;
}

int oracle_h_is_valid_model(const double* H) {
  orc::HomographyEstimator he;
  orc::Model m;
  std::memcpy(m.m, H, 72);
  bool updated = false;
  std::vector<size_t> none;
  orc::Points pts{nullptr, 0, 4};
  return he.isValidModel(m, pts, none, nullptr, 0.0, updated) ? 1 : 0;
}

double oracle_det(const double* a, int32_t n) { return orc::det_partial_pivot(a, n); }

// The driver's post-hoc weighted bundle adjustment of a fundamental matrix
// (py/gcransac_python.cpp:1832-1887): model_io, inliers_io (capacity n),
// score_io are updated in place.  Returns the inlier count afterwards.
int64_t oracle_f_post_ba(const double* pts_, int64_t n, double threshold, double* model_io, int64_t* inliers_io,
                         int64_t ni, double* score_io) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::FundamentalEstimator fe(0);
  orc::Model m;
  std::memcpy(m.m, model_io, 72);
  std::vector<size_t> in(inliers_io, inliers_io + ni);
  double score = *score_io;
  orc::fundamental_post_bundle_adjustment(pts, fe, threshold, m, in, score);
  std::memcpy(model_io, m.m, 72);
  for (size_t k = 0; k < in.size(); ++k) inliers_io[k] = (int64_t)in[k];
  *score_io = score;
  return (int64_t)in.size();
}

}  // extern "C"

extern "C" {

// A sequence of SPRT verifications (inc/preemption/preemption_sprt.h:
// 206-379) of k models against the best-score values given per call: per
// call the pool index it started from, the points it tested, its inlier
// count and score value, whether it accepted, and the test's (epsilon,
// delta, A) after the call.  kind 0 homography (transfer error), 1
// fundamental (Sampson).  The rand() stream is seeded with seed first.
void oracle_sprt_sequence(int32_t kind, const double* pts_, int64_t n, const double* models, int64_t k,
                          double threshold, double min_ratio, int32_t sample_size, int32_t max_solutions,
                          uint32_t seed, const double* best_values, int64_t* pool_out, int64_t* start_out,
                          int64_t* tested_out, int64_t* count_out, double* value_out, uint8_t* accepted_out,
                          double* eps_out, double* delta_out, double* A_out) {
  orc::Points pts{pts_, (size_t)n, 4};
  std::srand(seed);
  orc::SprtPreemption sprt((size_t)n, (size_t)sample_size, (size_t)max_solutions, min_ratio);
  for (int64_t i = 0; i < n; ++i) pool_out[i] = (int64_t)sprt.pool()[(size_t)i];
  orc::HomographyEstimator he;
  orc::FundamentalEstimator fe(0);
  std::vector<size_t> inl;
  for (int64_t j = 0; j < k; ++j) {
    orc::Model m;
    std::memcpy(m.m, models + 9 * j, sizeof(m.m));
    orc::Score best;
    best.value = best_values[j];
    orc::Score sc;
    const size_t start = sprt.poolIndex();
    const bool ok = kind == 0 ? sprt.verify(m, he, threshold, (size_t)j + 1, best, pts, inl, sc)
                              : sprt.verify(m, fe, threshold, (size_t)j + 1, best, pts, inl, sc);
    start_out[j] = (int64_t)start;
    const size_t end = sprt.poolIndex();
    tested_out[j] = ok ? n : (int64_t)((end + (size_t)n - start) % (size_t)n);
    if (!ok && tested_out[j] == 0) tested_out[j] = n;  // rejected at the last point of a full cycle
    count_out[j] = (int64_t)sc.inlier_number;
    value_out[j] = sc.value;
    accepted_out[j] = ok ? 1 : 0;
    eps_out[j] = sprt.current().epsilon;
    delta_out[j] = sprt.current().delta;
    A_out[j] = sprt.current().A;
  }
}

double oracle_sprt_threshold_A(double epsilon, double delta, int32_t sample_size, int32_t max_solutions) {
  std::srand(1);
  orc::SprtPreemption sprt((size_t)sample_size, (size_t)sample_size, (size_t)max_solutions, 0.5);
  return sprt.estimateThresholdA(epsilon, delta);
}
}

extern "C" {

// The space partitioning selector (inc/inlier_selectors/
// space_partitioning_ransac.h:339-504) for one homography on the grid of the
// given cells: the cells in the grid's iteration order (key, point count,
// axis indices, kept flag); returns the number of filled cells and writes
// the kept point count.
int64_t oracle_sp_select(const double* pts_, int64_t n, const double* cell_sizes, int64_t divisions,
                         const double* model, double threshold, int64_t max_cells, uint64_t* keys_out,
                         int64_t* npts_out, int64_t* idx_out, uint8_t* kept_out, int64_t* kept_points) {
  orc::Points pts{pts_, (size_t)n, 4};
  orc::GridNeighborhood g(pts, cell_sizes, (size_t)divisions);
  orc::SpacePartitioning sp(&g);
  orc::Model m;
  std::memcpy(m.m, model, sizeof(m.m));
  std::vector<const std::vector<size_t>*> sets;
  size_t cnt = 0;
  sp.run(m, threshold, sets, cnt);
  *kept_points = (int64_t)cnt;
  int64_t c = 0;
  size_t s = 0;
  for (const auto& kv : g.getCells()) {
    if (c < max_cells) {
      keys_out[c] = (uint64_t)kv.first;
      npts_out[c] = (int64_t)kv.second.first.size();
      for (int d = 0; d < 4; ++d) idx_out[4 * c + d] = (int64_t)kv.second.second[(size_t)d];
      const bool kept = s < sets.size() && sets[s] == &kv.second.first;
      if (kept) ++s;
      kept_out[c] = kept ? 1 : 0;
    }
    ++c;
  }
  return c;
}
}
