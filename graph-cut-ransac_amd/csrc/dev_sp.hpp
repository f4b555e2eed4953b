// Space partitioning inlier selector for homographies
// (space_partitioning_ransac.h:339-504), device side.
//
// A pair's grid cells are kept in the order the reference's grid container
// iterates them (the host builds the same container); the points of a cell
// are contiguous in the pair's "cell order" permutation.  Per model, a cell
// of fewer than 4 points is kept; otherwise its first-image corners are
// projected by H and the cell is kept when a projected corner's
// second-image cell (x86 truncating conversions) is the cell's, or when the
// rectangle of its projected (0,0) and (1,1) corners, grown by the
// threshold, passes the reference's overlap test.  Cells with a first-image
// axis index at or beyond the division count (points outside the image, for
// which the reference reads corner projections it did not compute) are kept.
// A model with |det H| < 1e-3 (Eigen's fixed-size cofactor determinant)
// keeps no cell.
#pragma once

#include <climits>

#include "kernels.hpp"

namespace gcr {

// x86 cvttsd2si: truncation, INT_MIN when out of range or NaN
__device__ __forceinline__ int x86_cvtt_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
  return (int)v;
}

__device__ __forceinline__ double sp_det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

struct SpCorner {
  int ix, iy;
  double x, y;
};

__device__ __forceinline__ SpCorner sp_project(const double* H, uint64_t h, uint64_t v, const SpPair& sp) {
  const double kX1 = (double)h * sp.cs[0], kY1 = (double)v * sp.cs[1];
  double x2p = kX1 * H[0] + kY1 * H[1] + H[2], y2p = kX1 * H[3] + kY1 * H[4] + H[5],
         h2p = kX1 * H[6] + kY1 * H[7] + H[8];
  x2p /= h2p;
  y2p /= h2p;
  return SpCorner{x86_cvtt_i32(x2p / sp.cs[2]), x86_cvtt_i32(y2p / sp.cs[3]), x2p, y2p};
}

__device__ __forceinline__ bool sp_keep(const double* H, const SpCell& c, const SpPair& sp) {
  if (c.len < 4) return true;
  const double P = sp.P;
  if ((double)c.idx[0] >= P || (double)c.idx[1] >= P) return true;
  for (int s = 0; s < 4; ++s) {
    const uint64_t h = c.idx[0] + (uint64_t)(s >> 1);
    if ((double)h >= P) continue;
    const uint64_t v = c.idx[1] + (uint64_t)(s & 1);
    if ((double)v >= P) continue;
    const SpCorner k = sp_project(H, h, v, sp);
    if ((uint64_t)(int64_t)k.ix == c.idx[2] && (uint64_t)(int64_t)k.iy == c.idx[3]) return true;
  }
  const SpCorner c00 = sp_project(H, c.idx[0], c.idx[1], sp);
  const SpCorner c11 = sp_project(H, c.idx[0] + 1, c.idx[1] + 1, sp);
  const double l1x = c00.x - sp.tau, l1y = c00.y - sp.tau;
  const double r1x = c11.x + sp.tau, r1y = c11.y + sp.tau;
  const double l2x = sp.cs[2] * (double)c.idx[2], l2y = sp.cs[3] * (double)c.idx[3];
  const double r2x = l2x + sp.cs[2], r2y = l2y + sp.cs[3];
  return (l1x <= r2x && l2x <= r1x) || (r1y <= l2y && r2y <= l1y);
}

// Per cell kept flag into keep[] (LDS, one byte per cell) and the kept
// point count, for one model on one wavefront.  Returns the count (the same
// on every lane).
__device__ __forceinline__ int64_t sp_select_wave(const double* H, const SpPair& sp,
                                                  const SpCell* __restrict__ cells, uint8_t* keep) {
  const int lane = threadIdx.x;
  const bool any = !(fabs(sp_det3(H)) < 1e-3);
  int64_t cnt = 0;
  for (int c = lane; c < sp.ncells; c += 64) {
    const SpCell cell = cells[c];
    const bool k = any && sp_keep(H, cell, sp);
    keep[c] = k ? 1 : 0;
    if (k) cnt += cell.len;
  }
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) cnt += __shfl_xor(cnt, s, 64);
  __syncthreads();
  return cnt;
}

}  // namespace gcr
