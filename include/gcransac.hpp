// gcransac.hpp -- the reference's C++ template API over the MI355X engine.
//
// Header-only.  A program written against the reference's
//
//   gcransac::GCRANSAC<Estimator, NeighborhoodGraph[, ScoringFunction]>
//       (src/pygcransac/include/GCRANSAC.h:47-100)
//
// compiles against this header with the same names and call sequence:
//
//   gcransac::Points points(n, 4, data);                      // cv::Mat in the reference
//   gcransac::neighborhood::GridNeighborhoodGraph<4> grid(&points, {cw1, ch1, cw2, ch2}, 8);
//   gcransac::utils::DefaultHomographyEstimator estimator;
//   gcransac::sampler::UniformSampler main_sampler(&points);
//   gcransac::sampler::UniformSampler lo_sampler(&points);
//   gcransac::GCRANSAC<gcransac::utils::DefaultHomographyEstimator,
//                      gcransac::neighborhood::GridNeighborhoodGraph<4>> gcransac;
//   gcransac.settings.threshold = 3.0;
//   gcransac::Model model;
//   gcransac.run(points, estimator, &main_sampler, &lo_sampler, &grid, model);
//   const auto& stats = gcransac.getRansacStatistics();        // stats.inliers, .score, ...
//
// run() executes GCRANSAC::run (GCRANSAC.h:202-627) on the GPU through
// gcr_run() of gcr.h: sampling and control flow on the host, minimal solves,
// MSAC scoring, the grid neighbourhood, graph-cut labelling and the
// least-squares fits on the device.  The preemption and fast inlier
// selector slots (GCRANSAC.h:50-51) take preemption::SPRTPreemptiveVerfication
// and inlier_selector::SpacePartitioningRANSAC (homographies), run on the
// device too.  setFPS sets the engine's time limit.  A sampler class of the
// caller's own (derived from sampler::Sampler<Container, size_t>) draws the
// main loop's samples on the host through a callback.  What the engine does
// not implement is rejected with std::invalid_argument at run() instead of
// being silently replaced (a setting away from the reference default that
// the loop reads, a local-optimisation sampler other than the uniform one,
// space partitioning for fundamental or essential matrices).
//
// Any point container with the cv::Mat members `rows`, `cols`,
// `ptr<double>(row)` and `isContinuous()` is accepted, so a CV_64F cv::Mat
// works unchanged; gcransac::Points is the dependency-free container.
#pragma once

#include <cstddef>
#include <cstdint>
#include <exception>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "gcr.h"

namespace gcransac {

// Row-major FP64 matrix of `rows` points with `cols` coordinates each (the
// role cv::Mat plays in the reference).  For correspondences cols = 4:
// x1 y1 x2 y2.
class Points {
 public:
  int rows = 0, cols = 0;

  Points() = default;
  Points(int rows_, int cols_) : rows(rows_), cols(cols_), data_((size_t)rows_ * (size_t)cols_, 0.0) {}
  Points(int rows_, int cols_, const double* data) : Points(rows_, cols_) {
    for (size_t i = 0; i < data_.size(); ++i) data_[i] = data[i];
  }

  template <class T>
  T* ptr(int row = 0) {
    static_assert(std::is_same<T, double>::value, "gcransac::Points holds doubles");
    return data_.data() + (size_t)row * (size_t)cols;
  }
  template <class T>
  const T* ptr(int row = 0) const {
    static_assert(std::is_same<T, double>::value, "gcransac::Points holds doubles");
    return data_.data() + (size_t)row * (size_t)cols;
  }
  template <class T>
  T& at(int row, int col) {
    return ptr<T>(row)[col];
  }
  template <class T>
  const T& at(int row, int col) const {
    return ptr<T>(row)[col];
  }
  bool isContinuous() const { return true; }
  bool empty() const { return data_.empty(); }

 private:
  std::vector<double> data_;
};

// Dense row-major matrix for model parameters (Eigen::MatrixXd in the
// reference's Model, model.h:43-58).
class Matrix {
 public:
  Matrix() = default;
  Matrix(size_t rows, size_t cols) : rows_(rows), cols_(cols), v_(rows * cols, 0.0) {}
  double& operator()(size_t r, size_t c) { return v_[r * cols_ + c]; }
  double operator()(size_t r, size_t c) const { return v_[r * cols_ + c]; }
  size_t rows() const { return rows_; }
  size_t cols() const { return cols_; }
  size_t size() const { return v_.size(); }
  double* data() { return v_.data(); }
  const double* data() const { return v_.data(); }
  void resize(size_t rows, size_t cols) {
    rows_ = rows;
    cols_ = cols;
    v_.assign(rows * cols, 0.0);
  }

 private:
  size_t rows_ = 0, cols_ = 0;
  std::vector<double> v_;
};

// model.h:43-58: the estimated model; `descriptor` is 3 x 3 for the
// homography, fundamental and essential estimators and 0 x 0 when run() finds
// no model.
class Model {
 public:
  Matrix descriptor;
  Model() = default;
  explicit Model(const Matrix& descriptor_) : descriptor(descriptor_) {}
};

namespace utils {

// settings.h:38-88, same fields and defaults.  GCRANSAC::run reads
// threshold, confidence, spatial_coherence_weight, min/max_iteration_number
// and max_local_optimization_number (the inner-RANSAC trial count of the
// graph-cut local optimisation); the engine runs the remaining loop
// constants at the defaults below, and run() rejects other values of them.
struct Settings {
  bool do_final_iterated_least_squares = true, do_local_optimization = true, do_graph_cut = true,
       use_inlier_limit = false;
  double desired_fps = -1;
  size_t max_local_optimization_number = 10, min_iteration_number_before_lo = 20, min_iteration_number = 20,
         max_iteration_number = std::numeric_limits<size_t>::max(), max_unsuccessful_model_generations = 100,
         max_least_squares_iterations = 10, max_graph_cut_number = 10, core_number = 1;
  double confidence = 0.95, neighborhood_sphere_radius = 20, threshold = 2.0, spatial_coherence_weight = 0.14;
};

// statistics.h:43-71, same fields.  accepted_models, rejected_models and
// neighbor_number stay 0 as in the reference with no preemptive
// verification and no fast inlier selector (GCRANSAC.h:224,343-411).
struct RANSACStatistics {
  size_t graph_cut_number = 0, local_optimization_number = 0, iteration_number = 0, neighbor_number = 0,
         accepted_models = 0, rejected_models = 0;
  std::string main_sampler_name, local_optimizer_sampler_name;
  double processing_time = 0.0, score = 0.0;
  std::vector<size_t> inliers;
  // device counters of the last run (no reference counterpart)
  int64_t candidates_solved = 0, models_scored = 0;
};

}  // namespace utils

namespace estimator {

// Estimator tags.  The solvers run on the device; the host types carry the
// problem kind, the reference's sample sizes and the validity-test
// parameters the engine implements.

// homography_estimator.h:58-120 with the four-point solver for minimal and
// non-minimal samples (types.h:93-96).
class RobustHomographyEstimator {
 public:
  static constexpr int32_t kProblem = GCR_PROBLEM_HOMOGRAPHY;
  static constexpr size_t sampleSize() { return 4; }
  static constexpr size_t nonMinimalSampleSize() { return 4; }
  static constexpr size_t maximumMinimalSolutions() { return 1; }
  static constexpr size_t inlierLimit() { return 7 * sampleSize(); }
  void check() const {}
  int32_t runFlags() const { return 0; }
};

// fundamental_estimator.h:64-160 with the seven-point minimal and the
// bundle-adjustment non-minimal solver (types.h:78-81).  The validity test
// (oriented epipolar constraint, symmetric-error inlier ratio) and DEGENSAC
// (use_degensac: H-degenerate samples, nested plane-and-parallax GC-RANSAC)
// run on the device.
class FundamentalMatrixEstimator {
 public:
  static constexpr int32_t kProblem = GCR_PROBLEM_FUNDAMENTAL;
  explicit FundamentalMatrixEstimator(double minimum_inlier_ratio_in_validity_check = 0.5, bool use_degensac = true,
                                      double homography_threshold = 2.0)
      : min_ratio_(minimum_inlier_ratio_in_validity_check),
        use_degensac_(use_degensac),
        homography_threshold_(homography_threshold) {}
  static constexpr size_t sampleSize() { return 7; }
  static constexpr size_t nonMinimalSampleSize() { return 7; }
  static constexpr size_t maximumMinimalSolutions() { return 3; }
  static constexpr size_t inlierLimit() { return 7 * sampleSize(); }
  void check() const {
    if (min_ratio_ != 0.5)
      throw std::invalid_argument("FundamentalMatrixEstimator: only the validity ratio 0.5 is built");
    if (use_degensac_ && homography_threshold_ != 2.0)
      throw std::invalid_argument("FundamentalMatrixEstimator: only the DEGENSAC homography threshold 2.0 is built");
  }
  int32_t runFlags() const { return use_degensac_ ? 0 : GCR_RUN_NO_DEGENSAC; }
  bool useDegensac() const { return use_degensac_; }
  double homographyThreshold() const { return homography_threshold_; }

 private:
  double min_ratio_;
  bool use_degensac_;
  double homography_threshold_;
};

// essential_estimator.h:58-124 with the five-point Nister minimal and the
// Stewenius non-minimal solver (types.h:108-111).  run() takes the points
// already normalised by the intrinsics (x' = K^-1 x), the threshold in those
// units, as the reference's GCRANSAC does; the intrinsics are kept for the
// caller as the reference keeps them.
class EssentialMatrixEstimator {
 public:
  static constexpr int32_t kProblem = GCR_PROBLEM_ESSENTIAL;
  template <class Mat3>
  EssentialMatrixEstimator(const Mat3& intrinsics_src, const Mat3& intrinsics_dst,
                           double minimum_inlier_ratio_in_validity_check = 0.1,
                           double point_ratio_for_selecting_from_multiple_models = 0.05)
      : min_ratio_(minimum_inlier_ratio_in_validity_check),
        point_ratio_(point_ratio_for_selecting_from_multiple_models) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        K1_(r, c) = intrinsics_src(r, c);
        K2_(r, c) = intrinsics_dst(r, c);
      }
  }
  static constexpr size_t sampleSize() { return 5; }
  static constexpr size_t nonMinimalSampleSize() { return 5; }
  static constexpr size_t maximumMinimalSolutions() { return 10; }
  static constexpr size_t inlierLimit() { return 7 * sampleSize(); }
  void check() const {
    if (min_ratio_ != 0.1)
      throw std::invalid_argument("EssentialMatrixEstimator: only the validity ratio 0.1 is built");
  }
  int32_t runFlags() const { return 0; }
  const Matrix& intrinsicsSource() const { return K1_; }
  const Matrix& intrinsicsDestination() const { return K2_; }
  double pointRatioForSelectingFromMultipleModels() const { return point_ratio_; }

 private:
  double min_ratio_, point_ratio_;
  Matrix K1_{3, 3}, K2_{3, 3};
};

// A plain 3 x 3 row-major intrinsics matrix for EssentialMatrixEstimator when
// Eigen is not used.
struct Matrix3 {
  double v[9];
  double operator()(int r, int c) const { return v[3 * r + c]; }
};

}  // namespace estimator

namespace utils {
// types.h:78-111, the default estimators of the reference.
using DefaultHomographyEstimator = estimator::RobustHomographyEstimator;
using DefaultFundamentalMatrixEstimator = estimator::FundamentalMatrixEstimator;
using DefaultEssentialMatrixEstimator = estimator::EssentialMatrixEstimator;
}  // namespace utils

namespace sampler {

// samplers/sampler.h:44-87.  The built-in samplers (UniformSampler,
// ProsacSampler) carry their kind and seed; the engine draws their samples
// (on the device or the host).  The reference seeds them from a random
// device; the seed argument (no reference counterpart) makes runs
// repeatable.  The local-optimisation stream is derived from the main
// sampler's seed, so a run is fixed by that one value.
//
// A sampler class of the caller's own, derived from Sampler<Container,
// size_t> as in the reference, is a host plugin: run() passes it to the
// engine as a callback, and every minimal sample of the main loop comes from
// its sample(pool, subset, sample_size), in the reference's order.  The
// engine draws samples ahead in chunks, so update() is not called (a sampler
// that adapts to the loop's updates is not supported), and the local
// optimisation keeps the engine's uniform sampler, seeded then from the
// local-optimisation sampler's seed.
class SamplerBase {
 public:
  virtual ~SamplerBase() = default;
  virtual const std::string getName() const = 0;
  // 0 uniform, 1 PROSAC: drawn by the engine; -1: the caller's sampler
  virtual int32_t kind() const { return -1; }
  // sample_size indices drawn from pool_ into subset_; false: no sample
  virtual bool sample(const std::vector<size_t>& pool_, size_t* const subset_, size_t sample_size_) {
    (void)pool_;
    (void)subset_;
    (void)sample_size_;
    return false;
  }
  virtual void reset() {}
  virtual void update(const size_t* const subset_, const size_t& sample_size_, const size_t& iteration_number_,
                      const double& inlier_ratio_) {
    (void)subset_;
    (void)sample_size_;
    (void)iteration_number_;
    (void)inlier_ratio_;
  }
  uint64_t seed() const { return seed_; }
  bool isInitialized() const { return true; }

 protected:
  explicit SamplerBase(uint64_t seed) : seed_(seed) {}
  static uint64_t random_seed() { return ((uint64_t)std::random_device{}() << 32) ^ std::random_device{}(); }

 private:
  uint64_t seed_;
};

// The reference's base class of sampler plugins (samplers/sampler.h:44-87).
template <class _DataContainer = Points, class _IndexType = size_t>
class Sampler : public SamplerBase {
  static_assert(std::is_same<_IndexType, size_t>::value, "samplers draw size_t indices");

 public:
  explicit Sampler(const _DataContainer* const container_) : SamplerBase(0), container(container_) {}
  virtual bool initialize(const _DataContainer* const container_) = 0;

 protected:
  const _DataContainer* const container;
  bool initialized = false;
};

// samplers/uniform_sampler.h:50-101
class UniformSampler : public SamplerBase {
 public:
  template <class Mat>
  explicit UniformSampler(const Mat* container, uint64_t seed = random_seed()) : SamplerBase(seed) {
    (void)container;
  }
  const std::string getName() const override { return "Uniform Sampler"; }
  int32_t kind() const override { return 0; }
};

// samplers/prosac_sampler.h:50-200: the points must be ordered by
// decreasing quality.
class ProsacSampler : public SamplerBase {
 public:
  template <class Mat>
  ProsacSampler(const Mat* container, size_t sample_size, uint64_t seed = random_seed()) : SamplerBase(seed) {
    (void)container;
    (void)sample_size;
  }
  const std::string getName() const override { return "PROSAC Sampler"; }
  int32_t kind() const override { return 1; }
};

}  // namespace sampler

namespace neighborhood {

// neighborhood/grid_neighborhood_graph.h:150-301: the grid over the
// dimensions of the container's rows, cell sizes per axis and the number of
// cells along every axis.  The cells and edges are built on the device
// inside run(); the host object keeps the points and the cell geometry.
template <size_t _DimensionNumber>
class GridNeighborhoodGraph {
 public:
  static_assert(_DimensionNumber == 4, "the engine builds the 4D correspondence grid");

  GridNeighborhoodGraph() = default;
  template <class Mat>
  GridNeighborhoodGraph(const Mat* container, const std::vector<double>& cell_sizes, size_t cell_number_along_all_axes)
      : cell_sizes_(cell_sizes), divisions_(cell_number_along_all_axes) {
    if (cell_sizes_.size() != _DimensionNumber)
      throw std::invalid_argument("GridNeighborhoodGraph: one cell size per dimension is required");
    if (container == nullptr) throw std::invalid_argument("GridNeighborhoodGraph: null container");
    if (container->cols != (int)_DimensionNumber)
      throw std::invalid_argument("GridNeighborhoodGraph: container dimension differs from the graph's");
    rows_ = container->rows;
    if (rows_ > 0) base_ = container->template ptr<double>(0);
    continuous_ = container->isContinuous();
  }

  const std::vector<double>& getCellSizes() const { return cell_sizes_; }
  size_t getDivisionNumber() const { return divisions_; }
  bool isInitialized() const { return divisions_ > 0; }
  // the points the grid is built on (row-major, rows x 4)
  const double* points() const { return base_; }
  int rows() const { return rows_; }
  bool continuous() const { return continuous_; }

 private:
  std::vector<double> cell_sizes_;
  size_t divisions_ = 0;
  const double* base_ = nullptr;
  int rows_ = 0;
  bool continuous_ = true;
};

}  // namespace neighborhood

// scoring_function.h:137-233: the MSAC score, computed on the device.
template <class _ModelEstimator>
class MSACScoringFunction {};

namespace preemption {

// preemption/preemption_empty.h:44-66: every model is scored in full.
template <class _ModelEstimator>
class EmptyPreemptiveVerfication {
 public:
  static constexpr bool providesScore() { return false; }
  static constexpr const char* getName() { return "empty"; }
  static constexpr bool kSprt = false;
  double minimumInlierRatio() const { return 0.0; }
};

// preemption/preemption_sprt.h:141-190: Wald's sequential probability
// ratio test, run on the device.  The constructor takes the reference's
// arguments; the model-estimation time ratio t_M is 1 / m (the value the
// reference's timing code yields), and the random point pool is drawn from
// the rand() stream seeded with the main sampler's seed.
template <class _ModelEstimator>
class SPRTPreemptiveVerfication {
 public:
  static constexpr bool providesScore() { return true; }
  static constexpr const char* getName() { return "SPRT"; }
  static constexpr bool kSprt = true;
  template <class _Points>
  SPRTPreemptiveVerfication(const _Points& points_, const _ModelEstimator& estimator_,
                            const double& minimum_inlier_ratio_ = 0.01)
      : ratio_(minimum_inlier_ratio_) {
    (void)points_;
    (void)estimator_;
  }
  double minimumInlierRatio() const { return ratio_; }

 private:
  double ratio_;
};

}  // namespace preemption

namespace inlier_selector {

// inlier_selectors/empty_inlier_selector.h:42-70: all points are scored.
template <class _Estimator, class _NeighborhoodStructure>
class EmptyInlierSelector {
 public:
  static constexpr bool doesSomething() { return false; }
  static constexpr bool kSpacePartitioning = false;
  explicit EmptyInlierSelector(const _NeighborhoodStructure* kNeighborhoodGraph_ = nullptr) {
    (void)kNeighborhoodGraph_;
  }
};

// inlier_selectors/space_partitioning_ransac.h:46-110, 339-504: the cells of
// the neighbourhood grid a homography can map inliers between, selected on
// the device for every model before it is scored.  Built for homographies
// (the reference's selector keeps every cell for fundamental and essential
// matrices; run() rejects those here).
template <class _Estimator, class _NeighborhoodStructure>
class SpacePartitioningRANSAC {
 public:
  static constexpr bool doesSomething() { return true; }
  static constexpr bool kSpacePartitioning = true;
  explicit SpacePartitioningRANSAC(const _NeighborhoodStructure* kNeighborhoodGraph_)
      : graph_(kNeighborhoodGraph_) {
    if (!graph_) throw std::invalid_argument("SpacePartitioningRANSAC: null neighbourhood graph");
  }
  const _NeighborhoodStructure* graph() const { return graph_; }

 private:
  const _NeighborhoodStructure* graph_;
};

}  // namespace inlier_selector

namespace detail {
// A caller's sampler behind gcr_params_t::sample_fn: the pool as the
// reference passes it (a vector of size_t indices), the subset back as int32
// indices; an exception ends the sampler's draws (refused samples) and is
// rethrown by run().
struct SampleBridge {
  sampler::SamplerBase* s = nullptr;
  std::exception_ptr error;
  std::vector<size_t> pool, subset;
  bool identity = false;
  static int32_t call(void* ctx, const int32_t* pool_, int64_t pool_n, int32_t* out, int32_t m) {
    SampleBridge& b = *static_cast<SampleBridge*>(ctx);
    if (b.error || pool_n < 0 || m <= 0) return 0;
    try {
      if (pool_) {
        b.pool.assign(pool_, pool_ + pool_n);
        b.identity = false;
      } else if (!b.identity || (int64_t)b.pool.size() != pool_n) {
        b.pool.resize((size_t)pool_n);
        for (size_t i = 0; i < b.pool.size(); ++i) b.pool[i] = i;
        b.identity = true;
      }
      b.subset.assign((size_t)m, 0);
      if (!b.s->sample(b.pool, b.subset.data(), (size_t)m)) return 0;
      for (int32_t k = 0; k < m; ++k) {
        if (b.subset[(size_t)k] >= (size_t)pool_n) throw std::out_of_range("sampler index outside the point set");
        out[k] = (int32_t)b.subset[(size_t)k];
      }
      return 1;
    } catch (...) {
      b.error = std::current_exception();
      return 0;
    }
  }
};

inline void check_rc(int32_t rc) {
  if (rc >= 0) return;
  const std::string msg = gcr_last_error();
  if (rc == GCR_ERR_INVALID_ARGUMENT || rc == GCR_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}
}  // namespace detail

template <class _ModelEstimator, class _NeighborhoodGraph,
          class _ScoringFunction = MSACScoringFunction<_ModelEstimator>,
          class _PreemptiveModelVerification = preemption::EmptyPreemptiveVerfication<_ModelEstimator>,
          class _FastInlierSelector = inlier_selector::EmptyInlierSelector<_ModelEstimator, _NeighborhoodGraph>>
class GCRANSAC {
  static_assert(std::is_same<_ScoringFunction, MSACScoringFunction<_ModelEstimator>>::value,
                "the engine scores with MSAC");

 public:
  utils::Settings settings;

  GCRANSAC() = default;

  // GCRANSAC.h:69-84.  points_: rows x 4 correspondences; the samplers and
  // the neighbourhood graph as built by the caller.  obtained_model_ gets the
  // best model (empty descriptor when none is found); the statistics,
  // including the inlier indices, are in getRansacStatistics().
  template <class _Points>
  void run(const _Points& points_, const _ModelEstimator& estimator_, sampler::SamplerBase* main_sampler_,
           sampler::SamplerBase* local_optimization_sampler_, const _NeighborhoodGraph* neighborhood_graph_,
           Model& obtained_model_) {
    _PreemptiveModelVerification* none_p = nullptr;
    _FastInlierSelector* none_s = nullptr;
    run_impl(points_, estimator_, main_sampler_, local_optimization_sampler_, neighborhood_graph_, obtained_model_,
             none_p, none_s);
  }

  // GCRANSAC.h:69-78: the same with the preemptive verification and the fast
  // inlier selector the template names.
  template <class _Points>
  void run(const _Points& points_, const _ModelEstimator& estimator_, sampler::SamplerBase* main_sampler_,
           sampler::SamplerBase* local_optimization_sampler_, const _NeighborhoodGraph* neighborhood_graph_,
           Model& obtained_model_, _PreemptiveModelVerification& preemptive_verification_,
           _FastInlierSelector& fast_inlier_selector_) {
    run_impl(points_, estimator_, main_sampler_, local_optimization_sampler_, neighborhood_graph_, obtained_model_,
             &preemptive_verification_, &fast_inlier_selector_);
  }

  // GCRANSAC.h:86: the time limit 1 / fps, checked after every main-loop
  // iteration (GCRANSAC.h:503-516), timed from the engine run's start.
  void setFPS(double fps_) {
    settings.desired_fps = fps_;
    time_limit = 1.0 / fps_;
  }

  const utils::RANSACStatistics& getRansacStatistics() const { return statistics; }
  utils::RANSACStatistics& getMutableRansacStatistics() { return statistics; }

 protected:
  utils::RANSACStatistics statistics;

  template <class _Points>
  void run_impl(const _Points& points_, const _ModelEstimator& estimator_, sampler::SamplerBase* main_sampler_,
                sampler::SamplerBase* local_optimization_sampler_, const _NeighborhoodGraph* neighborhood_graph_,
                Model& obtained_model_, const _PreemptiveModelVerification* preemption_,
                const _FastInlierSelector* selector_) {
    estimator_.check();
    check_settings();
    if (!main_sampler_ || !local_optimization_sampler_)
      throw std::invalid_argument("GCRANSAC::run: null sampler");
    if (local_optimization_sampler_->kind() != 0)
      throw std::invalid_argument("GCRANSAC::run: the local optimisation sampler must be sampler::UniformSampler");
    if (!neighborhood_graph_ || !neighborhood_graph_->isInitialized())
      throw std::invalid_argument("GCRANSAC::run: an initialised grid neighbourhood graph is required");
    if (!neighborhood_graph_->continuous())
      throw std::invalid_argument("GCRANSAC::run: the neighbourhood points must be continuous");
    if (points_.cols != 4) throw std::invalid_argument("GCRANSAC::run: points must have 4 columns");
    if (!points_.isContinuous()) throw std::invalid_argument("GCRANSAC::run: points must be continuous");
    if (neighborhood_graph_->rows() != points_.rows)
      throw std::invalid_argument("GCRANSAC::run: the neighbourhood graph was built on a different point count");
    const int64_t n = points_.rows;
    const double* pts = n > 0 ? points_.template ptr<double>(0) : nullptr;

    gcr_params_t p;
    gcr_default_params(&p, _ModelEstimator::kProblem);
    p.threshold = settings.threshold;
    p.confidence = settings.confidence;
    p.spatial_coherence_weight = settings.spatial_coherence_weight;
    p.max_iters = clamp_i64(settings.max_iteration_number);
    p.min_iters = clamp_i64(settings.min_iteration_number);
    p.lo_number = (int32_t)std::min<size_t>(settings.max_local_optimization_number,
                                            (size_t)std::numeric_limits<int32_t>::max());
    detail::SampleBridge bridge;
    if (main_sampler_->kind() < 0) {
      // the caller's sampler class: a host plugin of the main loop
      bridge.s = main_sampler_;
      p.sampler = 0;
      p.sample_fn = &detail::SampleBridge::call;
      p.sample_ctx = &bridge;
    } else {
      p.sampler = main_sampler_->kind();
    }
    // the local-optimisation stream comes from this seed: the built-in main
    // sampler's, or with a caller's sampler the local sampler's own
    p.seed = main_sampler_->kind() < 0 ? local_optimization_sampler_->seed() : main_sampler_->seed();
    // (a limit of zero or less stops after the first iteration, as there)
    p.time_limit = settings.desired_fps > -1 ? (time_limit > 0.0 ? time_limit : 1e-300) : 0.0;
    p.neighborhood = 0;
    p.use_sprt = 0;
    p.use_space_partitioning = 0;
    if constexpr (_PreemptiveModelVerification::kSprt) {
      p.use_sprt = 1;
      p.min_inlier_ratio_for_sprt = preemption_ ? preemption_->minimumInlierRatio() : 0.01;
    }
    if constexpr (_FastInlierSelector::kSpacePartitioning) {
      if (_ModelEstimator::kProblem != GCR_PROBLEM_HOMOGRAPHY)
        throw std::invalid_argument("GCRANSAC::run: space partitioning is built for homographies only");
      if (selector_ && selector_->graph() != neighborhood_graph_)
        throw std::invalid_argument("GCRANSAC::run: the selector must be built on the run's neighbourhood graph");
      p.use_space_partitioning = 1;
    }

    const std::vector<double>& cs = neighborhood_graph_->getCellSizes();
    double model[9];
    std::vector<uint8_t> mask((size_t)(n > 0 ? n : 0));
    gcr_stats_t st{};
    const int32_t rc =
        gcr_run(_ModelEstimator::kProblem, pts, n, neighborhood_graph_->points(), cs.data(),
                (int32_t)std::min<size_t>(neighborhood_graph_->getDivisionNumber(), 1u << 30),
                estimator_.runFlags(), &p, model,
                mask.data(), &st);
    if (bridge.error) std::rethrow_exception(bridge.error);
    detail::check_rc(rc);

    statistics = utils::RANSACStatistics();
    statistics.main_sampler_name = main_sampler_->getName();
    statistics.local_optimizer_sampler_name = local_optimization_sampler_->getName();
    statistics.iteration_number = (size_t)st.iteration_number;
    statistics.local_optimization_number = (size_t)st.local_optimization_number;
    statistics.graph_cut_number = (size_t)st.graph_cut_number;
    statistics.processing_time = st.processing_time;
    statistics.score = st.score;
    statistics.candidates_solved = st.candidates_solved;
    statistics.models_scored = st.models_scored;
    for (int64_t i = 0; i < n; ++i)
      if (mask[(size_t)i]) statistics.inliers.push_back((size_t)i);
    if (rc > 0) {
      obtained_model_.descriptor.resize(3, 3);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) obtained_model_.descriptor(r, c) = model[3 * r + c];
    } else {
      obtained_model_.descriptor = Matrix();
    }
  }

  double time_limit = std::numeric_limits<double>::max();  // seconds (setFPS)

  static int64_t clamp_i64(size_t v) {
    return v > (size_t)std::numeric_limits<int64_t>::max() ? std::numeric_limits<int64_t>::max() : (int64_t)v;
  }

  // The loop constants the engine runs with (GCRANSAC.h:297,463,481,530,559,
  // 648,811 read them); a program that changes one gets an error instead of
  // a run with a different algorithm.
  void check_settings() const {
    const utils::Settings d;
    auto bad = [](const char* name) {
      throw std::invalid_argument(std::string("GCRANSAC: settings.") + name +
                                  " other than the reference default is not built");
    };
    if (settings.do_final_iterated_least_squares != d.do_final_iterated_least_squares)
      bad("do_final_iterated_least_squares");
    if (settings.do_local_optimization != d.do_local_optimization) bad("do_local_optimization");
    if (settings.min_iteration_number_before_lo != d.min_iteration_number_before_lo)
      bad("min_iteration_number_before_lo");
    if (settings.max_unsuccessful_model_generations != d.max_unsuccessful_model_generations)
      bad("max_unsuccessful_model_generations");
    if (settings.max_least_squares_iterations != d.max_least_squares_iterations) bad("max_least_squares_iterations");
    if (settings.max_graph_cut_number != d.max_graph_cut_number) bad("max_graph_cut_number");
  }
};

}  // namespace gcransac
