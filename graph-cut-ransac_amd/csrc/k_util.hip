// Small utility kernels: model-slot copies, range zeroing and IRLS weights.
#include <algorithm>
#include <atomic>

#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

__global__ void k_copy_models(const double* const* __restrict__ src, double* const* __restrict__ dst, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = t / 9, k = t % 9;
  if (c >= count) return;
  dst[c][k] = src[c][k];
}

int lds_limit_bytes() {
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 64 * 1024;
  int v = cached[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 64 * 1024;
    cached[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

void launch_copy_models(hipStream_t st, const double* const* src, double* const* dst, int count) {
  if (count <= 0) return;
  const int th = count * 9;
  hipLaunchKernelGGL(k_copy_models, dim3((th + 255) / 256), dim3(256), 0, st, src, dst, count);
}

// Zeroes the ranges [off[r], off[r] + len[r]) of a; blockIdx.y = range.
__global__ void k_zero_ranges(double* __restrict__ a, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ len) {
  const int r = blockIdx.y;
  const int64_t n = len[r];
  double* const dst = a + off[r];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = 0.0;
}

void launch_zero_ranges(hipStream_t st, double* a, const int64_t* off, const int32_t* len, int count, int max_len) {
  if (count <= 0 || max_len <= 0) return;
  const dim3 grid((unsigned)std::min(16, (max_len + 255) / 256), (unsigned)count);
  hipLaunchKernelGGL(k_zero_ranges, grid, dim3(256), 0, st, a, off, len);
}

// Tukey bisquare weights of the current inliers, stored by point index
// (GCRANSAC.h:655-667): w = max(0, 1 - r2 / tau2)^2.  zero != 0 resets the
// entries of the listed inliers to 0 instead (GCRANSAC.h:677-679).
template <int KIND>
__global__ void k_irls_weights(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                               const double* __restrict__ pair_sq_thr, const int32_t* __restrict__ item_pair,
                               const double* const* __restrict__ item_model, const int32_t* __restrict__ inliers,
                               const int64_t* __restrict__ inl_off, const int32_t* __restrict__ inl_n,
                               double* __restrict__ weights, int zero) {
  const int it = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= inl_n[it]) return;
  const int pair = item_pair[it];
  const int64_t base = pair_pt_off[pair];
  const int idx = inliers[inl_off[it] + k];
  if (zero) {
    weights[base + idx] = 0.0;
    return;
  }
  double M[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) M[j] = item_model[it][j];
  const double r2 = sq_residual<KIND>(pts + 4 * (base + idx), M);
  const double w0 = 1.0 - r2 / pair_sq_thr[pair];
  const double w = 0.0 > w0 ? 0.0 : w0;
  weights[base + idx] = w * w;
}

void launch_irls_weights(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off,
                         const double* pair_sq_thr, const int32_t* item_pair, const double* const* item_model,
                         const int32_t* inliers, const int64_t* inl_off, const int32_t* inl_n, double* weights,
                         int num_items, int zero, int max_n) {
  if (num_items <= 0 || max_n <= 0) return;
  const dim3 grid((max_n + 255) / 256, num_items);
  if (kind == 0)
    hipLaunchKernelGGL(k_irls_weights<0>, grid, dim3(256), 0, st, pts, pair_pt_off, pair_sq_thr, item_pair, item_model,
                       inliers, inl_off, inl_n, weights, zero);
  else
    hipLaunchKernelGGL(k_irls_weights<1>, grid, dim3(256), 0, st, pts, pair_pt_off, pair_sq_thr, item_pair, item_model,
                       inliers, inl_off, inl_n, weights, zero);
}

// ---- candidate results filtered on the device --------------------------------
// One 256-thread block per item walks its slots in order, 1,024 at a time
// (thread t takes the 4 consecutive slots 4t .. 4t + 3 of the round): a slot
// (candidate c, model j) holds a model when j < nmodels[c]; the last model's
// count goes to last_count[c]; the slots whose value beats the item's best
// are compacted in slot order (block prefix sum of the per-thread counts)
// into the item's region of recs.
__global__ void __launch_bounds__(256)
    k_chunk_filter(int h, int S, const int64_t* __restrict__ item_base, const int32_t* __restrict__ item_ncand,
                   const double* __restrict__ item_best, const int8_t* __restrict__ status, int8_t* nmodels,
                   const int32_t* __restrict__ counts, const double* __restrict__ values,
                   const int32_t* __restrict__ sym, const int8_t* __restrict__ valid, const int8_t* __restrict__ degen,
                   const int32_t* __restrict__ samples, int m, int32_t* __restrict__ last_count,
                   SlotRec* __restrict__ recs, int32_t* __restrict__ rec_count) {
  __shared__ int wsum[4];
  const int it = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int64_t b = item_base[it];
  const int64_t nslots = (int64_t)item_ncand[it] * S;
  const double best = item_best[it];
  SlotRec* out = recs + b * S;
  int cnt = 0;
  for (int64_t q0 = 0; q0 < nslots; q0 += 1024) {
    SlotRec r[4];
    bool keep[4];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t q = q0 + 4 * t + k;
      keep[k] = false;
      if (q < nslots) {
        const int64_t c = b + q / S;
        const int j = (int)(q % S);
        const int64_t g = b * S + q;
        int nm;
        if (h) {
          nm = status[g] == 2 ? 1 : 0;
          nmodels[c] = (int8_t)nm;
        } else {
          nm = nmodels[c];
        }
        if (j < nm) {
          if (j == nm - 1) last_count[c] = counts[g];
          const double v = values[g];
          if (v > best) {
            keep[k] = true;
            r[k].slot = (int32_t)q;
            r[k].count = counts[g];
            r[k].value = v;
            r[k].sym = h ? 0 : sym[g];
            r[k].valid = h ? valid[c] : 1;
            r[k].degen = degen ? degen[g] : 0;
            r[k].pad[0] = r[k].pad[1] = 0;
            r[k].sample0 = samples[c * m];
            r[k].pad2 = 0;
          }
        }
      }
      mine += keep[k] ? 1 : 0;
    }
    // inclusive prefix of the per-thread counts inside the wave, then across
    // the four waves
    int incl = mine;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      const int o = __shfl_up(incl, s, 64);
      if (lane >= s) incl += o;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int woff = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (w < wv) woff += wsum[w];
      total += wsum[w];
    }
    int pos = cnt + woff + incl - mine;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (keep[k]) out[pos++] = r[k];
    cnt += total;
    __syncthreads();
  }
  if (t == 0) rec_count[it] = cnt;
}

// cand_pair[base[i] + c] = pair[i] for c < count[i] (one block per item)
__global__ void __launch_bounds__(256) k_fill_cand_pair(const int64_t* __restrict__ base,
                                                        const int32_t* __restrict__ count,
                                                        const int32_t* __restrict__ pair, int32_t* __restrict__ cand_pair) {
  const int it = blockIdx.x;
  const int64_t b = base[it];
  const int32_t n = count[it], p = pair[it];
  for (int32_t c = threadIdx.x; c < n; c += 256) cand_pair[b + c] = p;
}

void launch_fill_cand_pair(hipStream_t st, int num_items, const int64_t* base, const int32_t* count,
                           const int32_t* pair, int32_t* cand_pair) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_fill_cand_pair, dim3(num_items), dim3(256), 0, st, base, count, pair, cand_pair);
}

// Item i's records to packed + (sum of the earlier items' counts).
__global__ void __launch_bounds__(64) k_chunk_pack(int S, const int64_t* __restrict__ item_base,
                                                   const int32_t* __restrict__ rec_count,
                                                   const SlotRec* __restrict__ recs, SlotRec* __restrict__ packed) {
  const int it = blockIdx.x;
  const int lane = threadIdx.x;
  int64_t off = 0;
  for (int i = lane; i < it; i += 64) off += rec_count[i];
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) off += __shfl_xor(off, s, 64);
  const SlotRec* src = recs + item_base[it] * S;
  const int n = rec_count[it];
  for (int k = lane; k < n; k += 64) packed[off + k] = src[k];
}

void launch_chunk_filter(hipStream_t st, int h, int S, int num_items, const int64_t* item_base,
                         const int32_t* item_ncand, const double* item_best, const int8_t* status, int8_t* nmodels,
                         const int32_t* counts, const double* values, const int32_t* sym, const int8_t* valid,
                         const int8_t* degen, const int32_t* samples, int m, int32_t* last_count, SlotRec* recs,
                         int32_t* rec_count, SlotRec* packed) {
  if (num_items <= 0) return;
  hipLaunchKernelGGL(k_chunk_filter, dim3(num_items), dim3(256), 0, st, h, S, item_base, item_ncand, item_best, status,
                     nmodels, counts, values, sym, valid, degen, samples, m, last_count, recs, rec_count);
  hipLaunchKernelGGL(k_chunk_pack, dim3(num_items), dim3(64), 0, st, S, item_base, rec_count, recs, packed);
}

}  // namespace gcr
