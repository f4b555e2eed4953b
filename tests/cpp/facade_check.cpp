// Test helper: runs include/gcransac.hpp's GCRANSAC template exactly as a
// reference program would, behind one C function the Python tests call with
// ctypes.  It uses only the public headers and libgcr_hip.so.
#include <cstring>
#include <random>
#include <exception>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "gcransac.hpp"

namespace {

thread_local std::string g_error;

// A sampler class written against the reference's plugin interface
// (samplers/sampler.h:44-87): the reference's uniform sampler
// (uniform_sampler.h:97-115 over uniform_random_generator.h:43-125) restated
// with std::mt19937 and std::uniform_int_distribution<size_t>, seeded with
// the low 32 bits of the seed as the engine seeds its own uniform sampler.
// throw_after > 0: the call after that many samples throws.
class UserUniformSampler : public gcransac::sampler::Sampler<gcransac::Points, size_t> {
 public:
  UserUniformSampler(const gcransac::Points* container_, uint64_t seed, int64_t throw_after)
      : Sampler(container_), gen_((uint32_t)(seed & 0xffffffffu)), throw_after_(throw_after) {}
  const std::string getName() const override { return "User Uniform Sampler"; }
  void reset() override {}
  bool initialize(const gcransac::Points* const) override { return true; }
  bool sample(const std::vector<size_t>& pool_, size_t* const subset_, size_t sample_size_) override {
    if (throw_after_ > 0 && calls_++ >= throw_after_) throw std::runtime_error("user sampler failure");
    if (sample_size_ > pool_.size() || pool_.empty()) return false;
    if (pool_.size() != range_) {
      dist_ = std::uniform_int_distribution<size_t>(0, pool_.size() - 1);
      range_ = pool_.size();
    }
    for (size_t k = 0; k < sample_size_;) {
      const size_t v = dist_(gen_);
      bool rep = false;
      for (size_t j = 0; j < k; ++j) rep |= subset_[j] == v;
      if (!rep) subset_[k++] = v;
    }
    for (size_t k = 0; k < sample_size_; ++k) subset_[k] = pool_[subset_[k]];
    return true;
  }

 private:
  std::mt19937 gen_;
  std::uniform_int_distribution<size_t> dist_;
  size_t range_ = 0;
  int64_t throw_after_ = 0, calls_ = 0;
};

struct FacadeArgs {
  double threshold, confidence, spatial_coherence_weight;
  int64_t max_iters, min_iters;
  int32_t sampler;  // 0 uniform, 1 PROSAC, 2 UserUniformSampler, 3 UserUniformSampler that throws
  int32_t lo_number;
  int32_t divisions;
  uint64_t seed;
  int32_t bad_setting;  // 0 none; 1.. set one unsupported setting (error-path tests)
  int32_t use_degensac;  // fundamental estimator's DEGENSAC switch
  int32_t use_sprt;      // preemption slot: SPRTPreemptiveVerfication
  int32_t use_sp;        // selector slot: SpacePartitioningRANSAC
  double sprt_ratio;     // the SPRT object's minimum inlier ratio
  double fps;            // > 0: setFPS(fps)
};

template <class Estimator, class Preemption, class Selector>
int run_typed(const Estimator& estimator, const double* pts_, int64_t n, const double* grid_pts_,
              const double* cell_sizes, const FacadeArgs& a, double* model_out, uint8_t* mask_out, int64_t* stats_out,
              double* score_out) {
  using namespace gcransac;
  Points points((int)n, 4, pts_);
  Points grid_points((int)n, 4, grid_pts_ ? grid_pts_ : pts_);
  neighborhood::GridNeighborhoodGraph<4> grid(&grid_points, {cell_sizes[0], cell_sizes[1], cell_sizes[2], cell_sizes[3]},
                                              (size_t)a.divisions);
  sampler::UniformSampler uniform(&points, a.seed);
  sampler::ProsacSampler prosac(&points, Estimator::sampleSize(), a.seed);
  sampler::UniformSampler lo_sampler(&points, a.seed);
  UserUniformSampler user(&points, a.seed, a.sampler == 3 ? 40 : 0);
  sampler::SamplerBase* main_sampler = &uniform;
  if (a.sampler == 1) main_sampler = &prosac;
  if (a.sampler >= 2) main_sampler = &user;

  GCRANSAC<Estimator, neighborhood::GridNeighborhoodGraph<4>, MSACScoringFunction<Estimator>, Preemption, Selector>
      gcransac;
  gcransac.settings.threshold = a.threshold;
  gcransac.settings.spatial_coherence_weight = a.spatial_coherence_weight;
  gcransac.settings.confidence = a.confidence;
  gcransac.settings.max_local_optimization_number = (size_t)a.lo_number;
  gcransac.settings.max_iteration_number = (size_t)a.max_iters;
  gcransac.settings.min_iteration_number = (size_t)a.min_iters;
  gcransac.settings.neighborhood_sphere_radius = a.divisions;
  if (a.bad_setting == 1) gcransac.settings.max_graph_cut_number = 5;
  if (a.fps > 0) gcransac.setFPS(a.fps);
  if (a.bad_setting == 3) gcransac.settings.do_local_optimization = false;

  Model model;
  if constexpr (std::is_same<Preemption, preemption::EmptyPreemptiveVerfication<Estimator>>::value &&
                std::is_same<Selector, inlier_selector::EmptyInlierSelector<Estimator,
                                                                            neighborhood::GridNeighborhoodGraph<4>>>::value) {
    gcransac.run(points, estimator, main_sampler, &lo_sampler, &grid, model);
  } else {
    Preemption preemption = [&]() {
      if constexpr (std::is_same<Preemption, preemption::SPRTPreemptiveVerfication<Estimator>>::value)
        return Preemption(points, estimator, a.sprt_ratio);
      else
        return Preemption();
    }();
    Selector selector(&grid);
    gcransac.run(points, estimator, main_sampler, &lo_sampler, &grid, model, preemption, selector);
  }
  const utils::RANSACStatistics& st = gcransac.getRansacStatistics();
  if (model.descriptor.size() == 9) std::memcpy(model_out, model.descriptor.data(), 9 * sizeof(double));
  std::memset(mask_out, 0, (size_t)n);
  for (size_t i : st.inliers) mask_out[i] = 1;
  stats_out[0] = (int64_t)st.iteration_number;
  stats_out[1] = (int64_t)st.local_optimization_number;
  stats_out[2] = (int64_t)st.graph_cut_number;
  stats_out[3] = (int64_t)st.inliers.size();
  stats_out[4] = st.main_sampler_name == "PROSAC Sampler" ? 1 : (st.main_sampler_name == "User Uniform Sampler" ? 2 : 0);
  *score_out = st.score;
  return (int)st.inliers.size();
}

template <class Estimator>
int run_one(const Estimator& estimator, const double* pts_, int64_t n, const double* grid_pts_,
            const double* cell_sizes, const FacadeArgs& a, double* model_out, uint8_t* mask_out, int64_t* stats_out,
            double* score_out) {
  using namespace gcransac;
  using Grid = neighborhood::GridNeighborhoodGraph<4>;
  using NoP = preemption::EmptyPreemptiveVerfication<Estimator>;
  using Sprt = preemption::SPRTPreemptiveVerfication<Estimator>;
  using NoS = inlier_selector::EmptyInlierSelector<Estimator, Grid>;
  using Sp = inlier_selector::SpacePartitioningRANSAC<Estimator, Grid>;
  if (a.use_sprt && a.use_sp)
    return run_typed<Estimator, Sprt, Sp>(estimator, pts_, n, grid_pts_, cell_sizes, a, model_out, mask_out, stats_out,
                                          score_out);
  if (a.use_sprt)
    return run_typed<Estimator, Sprt, NoS>(estimator, pts_, n, grid_pts_, cell_sizes, a, model_out, mask_out,
                                           stats_out, score_out);
  if (a.use_sp)
    return run_typed<Estimator, NoP, Sp>(estimator, pts_, n, grid_pts_, cell_sizes, a, model_out, mask_out, stats_out,
                                         score_out);
  return run_typed<Estimator, NoP, NoS>(estimator, pts_, n, grid_pts_, cell_sizes, a, model_out, mask_out, stats_out,
                                        score_out);
}

}  // namespace

extern "C" {

// kind 0 homography, 1 fundamental matrix, 2 essential matrix (pts already
// K-normalised; K: 18 doubles, K1 then K2, row-major).  Returns the inlier
// count, -1 for std::invalid_argument, -2 for other errors (message in
// facade_last_error()).
int facade_run(int32_t kind, const double* pts, int64_t n, const double* grid_pts, const double* cell_sizes,
               const FacadeArgs* args, const double* K, double* model_out, uint8_t* mask_out, int64_t* stats_out,
               double* score_out) {
  try {
    if (kind == 0)
      return run_one(gcransac::utils::DefaultHomographyEstimator(), pts, n, grid_pts, cell_sizes, *args, model_out,
                     mask_out, stats_out, score_out);
    if (kind == 1)
      return run_one(gcransac::utils::DefaultFundamentalMatrixEstimator(0.5, args->use_degensac != 0), pts, n,
                     grid_pts, cell_sizes, *args, model_out, mask_out, stats_out, score_out);
    if (kind == 2) {
      gcransac::estimator::Matrix3 K1{}, K2{};
      std::memcpy(K1.v, K, sizeof(K1.v));
      std::memcpy(K2.v, K + 9, sizeof(K2.v));
      return run_one(gcransac::utils::DefaultEssentialMatrixEstimator(K1, K2), pts, n, grid_pts, cell_sizes, *args,
                     model_out, mask_out, stats_out, score_out);
    }
    throw std::invalid_argument("unknown kind");
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return -1;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -2;
  }
}

const char* facade_last_error() { return g_error.c_str(); }
}
