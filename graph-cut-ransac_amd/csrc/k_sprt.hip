// SPRT preemptive verification (preemption_sprt.h:218-296) of a pair's
// candidate models, in the order the main loop meets them.
//
// The test is sequential twice over: each model walks the pair's points in
// the order of a random pool, starting where the previous model stopped,
// and multiplies a likelihood ratio point by point until it exceeds the
// decision threshold A; and the test's parameters (delta, A) change after a
// rejection that moves the estimated delta by more than 5 %.  One wavefront
// per pair runs the whole sequence: 64 pool positions at a time it computes
// residuals and inlier flags in parallel, then steps through them in order
// (every lane carries the same scalar chain) with the reference's
// operations: lambda * (delta / epsilon) or lambda * ((1 - delta) /
// (1 - epsilon)), the test against A, and the score sum of
// 1 - r^2 / tau^2 over the tested inliers.
//
// A after a delta change comes from a table the host fills for the current
// epsilon (A for delta = k / N, k = 0..N), so every logarithm is the host C
// library's.  The walk stops at the first accepted model whose score beats
// the best value given: the host then makes the reference's epsilon update,
// the validity check and the best-model update, and continues the walk.
#include "dev_common.hpp"
#include "dev_sp.hpp"
#include "kernels.hpp"

namespace gcr {

// The same test after the space partitioning selector
// (preemption_sprt.h:241-258): a model whose kept cells hold fewer points
// than the best model's inliers is skipped untested; the others are tested
// over their kept cells' points in cell order, without the pool.  A
// rejection leaves only the current cell: the following cells are still
// tested from the last accepted ratio (the reference's break exits the inner
// loop only), and their tested inliers count.  Each tested model's inliers
// are written in test order to list (the last tested model's list is the
// loop's inlier container).
__device__ void sprt_walk_sp(SprtItem& it, const double* __restrict__ pts, const SpPair& sp,
                             const SpCell* __restrict__ cells, const int32_t* __restrict__ perm,
                             const int32_t* __restrict__ cell_of, const double* __restrict__ a_table, int n,
                             double sq_thr, uint8_t* keep, int32_t* __restrict__ list) {
  const int lane = threadIdx.x;
  double delta = it.delta, A = it.A;
  const double eps = it.eps;
  double f_in = delta / eps, f_out = (1 - delta) / (1 - eps);
  it.stop = -1;
  it.last = -1;
  for (int s = 0; s < it.num_slots; ++s) {
    const double* mp = it.models + 9 * (int64_t)it.slots[s];
    double M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = mp[k];
    const int64_t kept = sp_select_wave(M, sp, cells, keep);
    if (kept < it.best_count) continue;
    double lambda = 1.0, value = 0.0;
    int64_t count = 0;
    bool rejected = false;
    int broken = -1;  // the cell a rejection left
    for (int b = 0; b < n; b += 64) {
      const int e = b + lane;
      int c = -1;
      bool kp = false, inl = false;
      double term = 0.0;
      int32_t p = 0;
      if (e < n) {
        c = cell_of[e];
        kp = keep[c] != 0;
        if (kp) {
          p = perm[e];
          const double r2 = h_sq_residual(pts + 4 * (int64_t)p, M);
          inl = r2 < sq_thr;
          if (inl) term = 1.0 - r2 / sq_thr;
        }
      }
      uint64_t mk = __ballot(kp);
      const uint64_t mi = __ballot(inl);
      while (mk) {
        const int l = __ffsll((unsigned long long)mk) - 1;
        mk &= mk - 1;
        const int cl = __shfl(c, l, 64);
        const double tl = __shfl(term, l, 64);
        const int32_t pl = __shfl(p, l, 64);
        if (cl == broken) continue;
        double lambda_new;
        if ((mi >> l) & 1ull) {
          lambda_new = lambda * f_in;
          value += tl;
          if (lane == 0) list[count] = pl;
          ++count;
        } else {
          lambda_new = lambda * f_out;
        }
        if (lambda_new > A) {
          rejected = true;
          broken = cl;
          continue;
        }
        lambda = lambda_new;
      }
    }
    it.last = s;
    it.last_start = -2;
    it.last_tested = 0;
    it.last_count = count;
    it.last_value = value;
    it.last_accepted = rejected ? 0 : 1;
    if (!rejected) {
      if (value > it.best_value) {
        it.stop = s;
        break;
      }
    } else {
      const double de = (double)count / (double)n;
      if (de > 0 && fabs(delta - de) / delta > 0.05) {
        delta = de;
        A = a_table[count];
        f_in = delta / eps;
        f_out = (1 - delta) / (1 - eps);
        ++it.delta_changes;
      }
    }
  }
  it.delta = delta;
  it.A = A;
}

template <int KIND>
__device__ void sprt_walk_one(SprtItem& it, const double* __restrict__ pts, const int32_t* __restrict__ pool,
                              const double* __restrict__ a_table, int n, double sq_thr) {
  const int lane = threadIdx.x;
  int64_t idx = it.pool_idx;
  double delta = it.delta, A = it.A;
  const double eps = it.eps;
  double f_in = delta / eps, f_out = (1 - delta) / (1 - eps);
  it.stop = -1;
  it.last = -1;
  for (int s = 0; s < it.num_slots; ++s) {
    const double* mp = it.models + 9 * (int64_t)it.slots[s];
    double M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = mp[k];
    const int64_t start = idx;
    double lambda = 1.0, value = 0.0;
    int64_t count = 0, tested = 0;
    bool rejected = false;
    for (int b = 0; b < n && !rejected; b += 64) {
      const int t = b + lane;
      bool inl = false;
      double term = 0.0;
      if (t < n) {
        int64_t pos = start + t;
        if (pos >= n) pos -= n;
        const double r2 = sq_residual<KIND>(pts + 4 * (int64_t)pool[pos], M);
        inl = r2 < sq_thr;
        if (inl) term = 1.0 - r2 / sq_thr;
      }
      const uint64_t mask = __ballot(inl);
      const int cnt = n - b < 64 ? n - b : 64;
      for (int l = 0; l < cnt; ++l) {
        const bool in = (mask >> l) & 1ull;
        const double tl = __shfl(term, l, 64);
        double lambda_new;
        if (in) {
          lambda_new = lambda * f_in;
          ++count;
          value += tl;
        } else {
          lambda_new = lambda * f_out;
        }
        ++tested;
        if (lambda_new > A) {
          rejected = true;
          break;
        }
        lambda = lambda_new;
      }
    }
    idx = (start + tested) % n;
    it.last = s;
    it.last_start = start;
    it.last_tested = tested;
    it.last_count = count;
    it.last_value = value;
    it.last_accepted = rejected ? 0 : 1;
    if (!rejected) {
      if (value > it.best_value) {
        it.stop = s;
        break;
      }
    } else {
      const double de = (double)count / (double)n;
      if (de > 0 && fabs(delta - de) / delta > 0.05) {
        delta = de;
        A = a_table[count];
        f_in = delta / eps;
        f_out = (1 - delta) / (1 - eps);
        ++it.delta_changes;
      }
    }
  }
  it.pool_idx = idx;
  it.delta = delta;
  it.A = A;
}

__global__ void __launch_bounds__(64) k_sprt_walk(SprtItem* items, const double* __restrict__ pts,
                                                  const int64_t* __restrict__ pair_pt_off,
                                                  const int32_t* __restrict__ pair_n,
                                                  const double* __restrict__ pair_sq_thr,
                                                  const int32_t* __restrict__ pools,
                                                  const SpPair* __restrict__ sp_pairs,
                                                  const SpCell* __restrict__ sp_cells,
                                                  const int32_t* __restrict__ sp_perm,
                                                  const int32_t* __restrict__ sp_cell_of, int32_t* sp_list) {
  extern __shared__ uint8_t keep[];
  SprtItem it = items[blockIdx.x];
  const int64_t base = pair_pt_off[it.pair];
  const int n = pair_n[it.pair];
  const double sq = pair_sq_thr[it.pair];
  it.delta_changes = 0;
  if (it.sp) {
    const SpPair sp = sp_pairs[it.pair];
    sprt_walk_sp(it, pts + 4 * base, sp, sp_cells + sp.cell_off, sp_perm + base, sp_cell_of + base, it.a_table, n,
                 sq, keep, sp_list + base);
  } else if (it.kind == 0) {
    sprt_walk_one<0>(it, pts + 4 * base, pools + base, it.a_table, n, sq);
  } else {
    sprt_walk_one<1>(it, pts + 4 * base, pools + base, it.a_table, n, sq);
  }
  if (threadIdx.x == 0) items[blockIdx.x] = it;
}

void launch_sprt_walk(hipStream_t st, SprtItem* items, int num_items, const double* pts, const int64_t* pair_pt_off,
                      const int32_t* pair_n, const double* pair_sq_thr, const int32_t* pools, const SpPair* sp_pairs,
                      const SpCell* sp_cells, const int32_t* sp_perm, const int32_t* sp_cell_of, int32_t* sp_list,
                      int max_cells) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_sprt_walk, dim3(num_items), dim3(64), (size_t)std::max(1, max_cells), st, items, pts,
                     pair_pt_off, pair_n, pair_sq_thr, pools, sp_pairs, sp_cells, sp_perm, sp_cell_of, sp_list);
}

}  // namespace gcr
