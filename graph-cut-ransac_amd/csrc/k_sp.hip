// Space partitioning inlier selector (space_partitioning_ransac.h:339-504)
// and the MSAC score over the cells it keeps (scoring_function.h:186-214).
//
// One wavefront per model: the cell selection goes into LDS (one byte per
// cell), then the pair's points are walked in cell order 64 at a time; the
// kept inliers' squared residuals are subtracted from the score one by one
// in that order (the reference's sum), so counts and values are the
// reference's bits.
#include "dev_common.hpp"
#include "dev_sp.hpp"
#include "kernels.hpp"

namespace gcr {

__global__ void __launch_bounds__(64)
    k_score_sp(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
               const int32_t* __restrict__ pair_n, const double* __restrict__ pair_sq_thr,
               const int8_t* __restrict__ pair_sp, const SpPair* __restrict__ sp_pairs,
               const SpCell* __restrict__ sp_cells, const int32_t* __restrict__ sp_perm,
               const int32_t* __restrict__ sp_cell_of, const double* __restrict__ models,
               const int32_t* __restrict__ cpair, const int8_t* __restrict__ status, int32_t* __restrict__ counts,
               double* __restrict__ values, int32_t* __restrict__ sel) {
  extern __shared__ uint8_t keep[];
  const int64_t slot = blockIdx.x;
  const int pair = cpair[slot];
  if (!pair_sp[pair] || status[slot] != 2) return;
  const int lane = threadIdx.x;
  double H[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) H[k] = models[9 * slot + k];
  const SpPair sp = sp_pairs[pair];
  const int64_t kept = sp_select_wave(H, sp, sp_cells + sp.cell_off, keep);
  const int64_t base = pair_pt_off[pair];
  const int n = pair_n[pair];
  const double sq = pair_sq_thr[pair];
  double value = 0.0;
  int count = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    bool inl = false;
    double r2 = 0.0;
    if (e < n && keep[sp_cell_of[base + e]]) {
      r2 = h_sq_residual(pts + 4 * (base + sp_perm[base + e]), H);
      inl = r2 < sq;
    }
    uint64_t m = __ballot(inl);
    while (m) {
      const int l = __ffsll((unsigned long long)m) - 1;
      value = value - __shfl(r2, l, 64);
      ++count;
      m &= m - 1;
    }
  }
  if (lane == 0) {
    counts[slot] = count;
    values[slot] = count == 0 ? 0.0 : (value + (double)count * sq) / sq;
    sel[slot] = (int32_t)kept;
  }
}

__global__ void __launch_bounds__(64)
    k_sp_list(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
              const int32_t* __restrict__ pair_n, const double* __restrict__ thr, const SpPair* __restrict__ sp_pairs,
              const SpCell* __restrict__ sp_cells, const int32_t* __restrict__ sp_perm,
              const int32_t* __restrict__ sp_cell_of, const double* __restrict__ models,
              const int32_t* __restrict__ item_pair, const int8_t* __restrict__ item_sp,
              const int64_t* __restrict__ list_off, int32_t* __restrict__ lists, int32_t* __restrict__ counts) {
  extern __shared__ uint8_t keep[];
  const int i = blockIdx.x;
  if (!item_sp[i]) return;
  const int pair = item_pair[i];
  const int lane = threadIdx.x;
  double H[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) H[k] = models[9 * (int64_t)i + k];
  const SpPair sp = sp_pairs[pair];
  sp_select_wave(H, sp, sp_cells + sp.cell_off, keep);
  const int64_t base = pair_pt_off[pair];
  const int n = pair_n[pair];
  const double sq = thr[pair];
  int32_t* out = lists + list_off[i];
  int cnt = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    bool inl = false;
    int32_t p = 0;
    if (e < n && keep[sp_cell_of[base + e]]) {
      p = sp_perm[base + e];
      inl = h_sq_residual(pts + 4 * (base + p), H) < sq;
    }
    const uint64_t m = __ballot(inl);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (inl) out[cnt + before] = p;
    cnt += __popcll(m);
  }
  if (lane == 0) counts[i] = cnt;
}

void launch_score_sp(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                     const double* pair_sq_thr, const int8_t* pair_sp, const SpPair* sp_pairs, const SpCell* sp_cells,
                     const int32_t* sp_perm, const int32_t* sp_cell_of, const double* models, const int32_t* cpair,
                     const int8_t* status, int64_t num_slots, int32_t* counts, double* values, int32_t* sel,
                     int max_cells) {
  if (num_slots <= 0) return;
  hipLaunchKernelGGL(k_score_sp, dim3((unsigned)num_slots), dim3(64), (size_t)std::max(1, max_cells), st, pts,
                     pair_pt_off, pair_n, pair_sq_thr, pair_sp, sp_pairs, sp_cells, sp_perm, sp_cell_of, models, cpair,
                     status, counts, values, sel);
}

void launch_sp_list(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                    const double* thr, const SpPair* sp_pairs, const SpCell* sp_cells, const int32_t* sp_perm,
                    const int32_t* sp_cell_of, const double* models, const int32_t* item_pair, const int8_t* item_sp,
                    const int64_t* list_off, int num_items, int32_t* lists, int32_t* counts, int max_cells) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_sp_list, dim3((unsigned)num_items), dim3(64), (size_t)std::max(1, max_cells), st, pts,
                     pair_pt_off, pair_n, thr, sp_pairs, sp_cells, sp_perm, sp_cell_of, models, item_pair, item_sp,
                     list_off, lists, counts);
}

}  // namespace gcr
