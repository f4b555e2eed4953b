// MSAC scoring of many models, one wavefront per model.
//
// Restates MSACScoringFunction::getScore (scoring_function.h:137-233) with
// the reference's exact arithmetic: each lane evaluates the residual of one
// point (64 consecutive points per step), the inlier test is r2 < tau2, the
// inlier count is a popcount of the wave ballot, and the score value is
// accumulated in ascending point order (value -= r2 for every inlier) by
// walking the ballot's set bits through the scalar unit.  The result is
// therefore bit-identical to the sequential loop, then normalised as
// (value + count * tau2) / tau2, and (0, 0) when there is no inlier.
//
// HBM traffic: the pair's points (32 B each) are re-read from L2 by every
// model of the pair; the kernel is bound by FP64 VALU work (two divisions per
// homography residual, one per Sampson residual) and by the in-order
// accumulation chain for high-inlier models.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

// Scores model slot mdl on the calling wave (every lane takes part).
template <int KIND>
__device__ __forceinline__ void score_model(int64_t mdl, const double* __restrict__ pts,
                                            const int64_t* __restrict__ pair_pt_off,
                                            const int32_t* __restrict__ pair_n,
                                            const double* __restrict__ pair_sq_thr,
                                            const double* __restrict__ models,
                                            const int32_t* __restrict__ model_pair, int32_t* __restrict__ counts,
                                            double* __restrict__ values, uint8_t* __restrict__ masks,
                                            const int64_t* __restrict__ mask_off, int32_t* __restrict__ sym_counts,
                                            int slots_per, int32_t* q) {
  const int lane = lane_id();
  // model_pair holds one entry per group of slots_per model slots
  const int pair = model_pair[slots_per == 1 ? mdl : mdl / slots_per];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int n = pair_n[pair];
  const double thr = pair_sq_thr[pair];
  double M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = models[9 * mdl + k];
  uint8_t* mask = (masks && mask_off[mdl] >= 0) ? masks + mask_off[mdl] : nullptr;
  double acc = 0.0;
  int64_t cnt = 0, scnt = 0;
  if constexpr (KIND == 1) {
    // Sampson: the inlier test r2 < tau2 is decided without the division
    // when num = (x2^T F x1)^2 and tau2 * den differ by more than 1e-12
    // relative (far beyond the rounding of both sides); lanes inside that
    // band divide.  Inliers go to a per-wave queue of point indices (point
    // order); every 64 queued inliers are evaluated by the full wave (the
    // division of f_sq_sampson and the symmetric test, the same operations
    // as before) and accumulated in queue order.  Same bits, far fewer
    // divisions when most points are outliers.
    int qn = 0;
    auto flush = [&](int cnt_flush) {
      double r2 = 0.0;
      bool sym_in = false;
      if (lane < cnt_flush) {
        const double* s = P + 4 * (int64_t)q[lane];
        r2 = f_sq_sampson(s, M);
        if (sym_counts) sym_in = f_sq_symmetric(s, M) < thr;
      }
      if (sym_counts) scnt += __popcll(ballot64(sym_in));
      for (int l = 0; l < cnt_flush; ++l) acc = acc - readlane_f64(r2, l);
    };
    const double lo = 1.0 - 1e-12, hi = 1.0 + 1e-12;
    for (int base = 0; base < n; base += 64) {
      const int i = base + lane;
      bool in = false;
      if (i < n) {
        double num, den;
        f_sampson_parts(P + 4 * (int64_t)i, M, num, den);
        const double t = thr * den;
        if (num < t * lo) in = true;
        else if (!(num > t * hi)) in = (num / den) < thr;
        if (mask) mask[i] = in ? 1 : 0;
      }
      const uint64_t b = ballot64(in);
      if (in) q[qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = i;
      qn += __popcll(b);
      cnt += __popcll(b);
      if (qn >= 64) {
        flush(64);
        const int rest = qn - 64;
        const int32_t moved = lane < rest ? q[64 + lane] : 0;
        if (lane < rest) q[lane] = moved;
        qn = rest;
      }
    }
    flush(qn);
  } else {
    for (int base = 0; base < n; base += 64) {
      const int i = base + lane;
      double r2 = 0.0;
      bool in = false, sym_in = false;
      if (i < n) {
        r2 = sq_residual<KIND>(P + 4 * (int64_t)i, M);
        in = r2 < thr;
        if (mask) mask[i] = in ? 1 : 0;
        // fundamental_estimator.h:307-316: Sampson inliers that are also
        // symmetric-epipolar inliers at the same squared threshold
        if (KIND == 1 && sym_counts && in) sym_in = f_sq_symmetric(P + 4 * (int64_t)i, M) < thr;
      }
      uint64_t b = ballot64(in);
      cnt += __popcll(b);
      if (KIND == 1 && sym_counts) scnt += __popcll(ballot64(sym_in));
      while (b) {
        const int l = __ffsll((unsigned long long)b) - 1;
        acc = acc - readlane_f64(r2, l);
        b &= b - 1;
      }
    }
  }
  if (lane == 0) {
    counts[mdl] = (int32_t)cnt;
    values[mdl] = cnt == 0 ? 0.0 : (acc + (double)cnt * thr) / thr;
    if (sym_counts) sym_counts[mdl] = (int32_t)scnt;
  }
}

// One wave per model slot.  CANDIDATES: the launch scores the minimal-sample
// hypotheses of a candidate chunk (a separate instantiation, so profiles of
// the sampling loop's scoring are not mixed with the scoring of
// local-optimisation fits).
template <int KIND, bool CANDIDATES>
__global__ void __launch_bounds__(256) k_score(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                               const int32_t* __restrict__ pair_n,
                                               const double* __restrict__ pair_sq_thr,
                                               const double* __restrict__ models,
                                               const int32_t* __restrict__ model_pair,
                                               const int8_t* __restrict__ active, int64_t num_models,
                                               int32_t* __restrict__ counts, double* __restrict__ values,
                                               uint8_t* __restrict__ masks, const int64_t* __restrict__ mask_off,
                                               int32_t* __restrict__ sym_counts, int slots_per) {
  __shared__ int32_t queue[4][128];
  const int64_t mdl = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (mdl >= num_models) return;
  if (active && active[mdl] != 2) {
    if (lane_id() == 0) {
      counts[mdl] = -1;
      values[mdl] = 0.0;
      if (sym_counts) sym_counts[mdl] = 0;
    }
    return;
  }
  score_model<KIND>(mdl, pts, pair_pt_off, pair_n, pair_sq_thr, models, model_pair, counts, values, masks, mask_off,
                    sym_counts, slots_per, queue[threadIdx.x >> 6]);
}

// Candidate chunks: most model slots hold no model to score (a seven-point
// sample gives one or three models in three slots, a five-point sample up to
// ten, and samples or models rejected by the solver's tests are marked
// inactive).  k_score_gather gives every inactive slot its empty result and
// appends the active slots to a work list (one atomic per wave; the list
// order does not matter, every model's result goes to its own slot); a grid
// of at most SCORE_LIST_BLOCKS x 4 waves then walks the list.  XCD-aware:
// the list is cut into eight contiguous ranges, one per XCD (consecutive
// workgroups are dispatched round-robin over the eight XCDs), so the models
// of one pair -- adjacent in the list -- read its points through one L2.
constexpr int SCORE_LIST_BLOCKS = 2048;  // 8,192 waves: 8 per SIMD on 256 CUs
constexpr int NUM_XCD = 8;

__global__ void __launch_bounds__(256) k_score_gather(const int8_t* __restrict__ active, int64_t num_models,
                                                      int32_t* __restrict__ counts, double* __restrict__ values,
                                                      int32_t* __restrict__ sym_counts, int32_t* __restrict__ list,
                                                      uint32_t* __restrict__ list_count) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool act = false;
  if (s < num_models) {
    act = active[s] == 2;
    if (!act) {
      counts[s] = -1;
      values[s] = 0.0;
      if (sym_counts) sym_counts[s] = 0;
    }
  }
  const uint64_t b = ballot64(act);
  if (b == 0) return;
  const int lane = lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(list_count, (uint32_t)__popcll(b));
  base = (uint32_t)__shfl(base, 0, 64);
  if (act)
    list[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] =
        (int32_t)s;
}

__global__ void k_zero_count(uint32_t* __restrict__ c) {
  if (threadIdx.x == 0) *c = 0u;
}

template <int KIND>
__global__ void __launch_bounds__(256) k_score_list(const double* __restrict__ pts,
                                                    const int64_t* __restrict__ pair_pt_off,
                                                    const int32_t* __restrict__ pair_n,
                                                    const double* __restrict__ pair_sq_thr,
                                                    const double* __restrict__ models,
                                                    const int32_t* __restrict__ model_pair,
                                                    const int32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ list_count,
                                                    int32_t* __restrict__ counts, double* __restrict__ values,
                                                    int32_t* __restrict__ sym_counts, int slots_per) {
  __shared__ int32_t queue[4][128];
  const int64_t total = (int64_t)*list_count;
  // gridDim.x is a multiple of NUM_XCD: workgroup b runs on XCD b % NUM_XCD
  const int xcd = (int)(blockIdx.x % NUM_XCD);
  const int64_t waves_per_xcd = (int64_t)(gridDim.x / NUM_XCD) * 4;
  const int64_t w = (int64_t)(blockIdx.x / NUM_XCD) * 4 + (threadIdx.x >> 6);
  const int64_t lo = total * xcd / NUM_XCD, hi = total * (xcd + 1) / NUM_XCD;
  for (int64_t k = lo + w; k < hi; k += waves_per_xcd)
    score_model<KIND>((int64_t)list[k], pts, pair_pt_off, pair_n, pair_sq_thr, models, model_pair, counts, values,
                      nullptr, nullptr, sym_counts, slots_per, queue[threadIdx.x >> 6]);
}

void launch_score(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                  const double* pair_sq_thr, const double* models, const int32_t* model_pair, const int8_t* active,
                  int64_t num_models, int32_t* counts, double* values, uint8_t* masks, const int64_t* mask_off,
                  int32_t* sym_counts, bool candidates, int slots_per, int32_t* work_list, uint32_t* work_count) {
  if (num_models <= 0) return;
  if (work_list && work_count && active && !masks && num_models <= (int64_t)INT32_MAX) {
    // work_list: num_models int32 entries; work_count: one zeroed counter
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, st, work_count);
    hipLaunchKernelGGL(k_score_gather, dim3((unsigned)((num_models + 255) / 256)), dim3(256), 0, st, active,
                       num_models, counts, values, sym_counts, work_list, work_count);
    int64_t blocks = (num_models + 3) / 4;
    if (blocks > SCORE_LIST_BLOCKS) blocks = SCORE_LIST_BLOCKS;
    blocks = (blocks + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    if (kind == 0)
      hipLaunchKernelGGL(k_score_list<0>, dim3((unsigned)blocks), dim3(256), 0, st, pts, pair_pt_off, pair_n,
                         pair_sq_thr, models, model_pair, work_list, work_count, counts, values, sym_counts,
                         slots_per);
    else
      hipLaunchKernelGGL(k_score_list<1>, dim3((unsigned)blocks), dim3(256), 0, st, pts, pair_pt_off, pair_n,
                         pair_sq_thr, models, model_pair, work_list, work_count, counts, values, sym_counts,
                         slots_per);
    return;
  }
  const int wpb = 4;
  const int64_t grid = (num_models + wpb - 1) / wpb;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * wpb), 0, st, pts, pair_pt_off, pair_n, pair_sq_thr,
                       models, model_pair, active, num_models, counts, values, masks, mask_off, sym_counts,
                       slots_per);
  };
  if (kind == 0) {
    if (candidates) go(k_score<0, true>);
    else go(k_score<0, false>);
  } else {
    if (candidates) go(k_score<1, true>);
    else go(k_score<1, false>);
  }
}

}  // namespace gcr
