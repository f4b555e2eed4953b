// Device helpers shared by the engine's kernels (gfx950, wave64).
//
// Every FP64 routine here is written so that, compiled with
// -ffp-contract=off, it performs the same IEEE operations in the same order
// as the sequential code it restates: no fused multiply-add, divisions and
// square roots correctly rounded.  Parallel reductions use one fixed order
// (see wave_sum64) that host code can reproduce exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gcr {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Broadcast a double held by lane `l` (uniform) to the whole wave through the
// scalar unit.
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
  return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// Reduction order used for every sum the engine evaluates in parallel:
// lane l owns the terms whose index is congruent to l mod 64 (added in
// ascending order into a partial that starts at 0.0); the 64 partials are
// then combined as p[l] += p[l + s] for s = 32, 16, 8, 4, 2, 1.  Lane 0 ends
// with the total, which is returned to every lane.
//
// The lane moves stay in the vector ALU: v_permlane32_swap / v_permlane16_swap
// (gfx950) bring lanes 32..63 to 0..31 and 16..31 to 0..15, DPP row_ror:8,
// row_ror:12 and the quad permutations [2,3,0,1] and [1,0,3,2] bring lane
// l + s to lane l for s = 8, 4, 2, 1 (l < s).  At each step only lanes l < s
// carry a partial the total depends on, and each of them adds exactly
// p[l] + p[l + s], so the total is that of the tree above.
template <int CTRL>
__device__ __forceinline__ double dpp_mov_f64(double v) {
  const uint64_t u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// lanes_down32: lane l < 32 gets the value of lane l + 32.  lanes_down16:
// lane l < 16 (and 32 <= l < 48) gets the value of lane l + 16.  Called with
// the whole wave active.
__device__ __forceinline__ double lanes_down32(double v) {
  const uint64_t u = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)u, (uint32_t)u, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
  return __longlong_as_double(((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ double lanes_down16(double v) {
  const uint64_t u = __double_as_longlong(v);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)u, (uint32_t)u, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
  return __longlong_as_double(((uint64_t)hi[1] << 32) | lo[1]);
}

__device__ __forceinline__ double wave_tree64(double partial) {
  partial = partial + lanes_down32(partial);
  partial = partial + lanes_down16(partial);
  partial = partial + dpp_mov_f64<0x128>(partial);  // row_ror:8
  partial = partial + dpp_mov_f64<0x12C>(partial);  // row_ror:12
  partial = partial + dpp_mov_f64<0x4E>(partial);   // quad_perm [2,3,0,1]
  partial = partial + dpp_mov_f64<0xB1>(partial);   // quad_perm [1,0,3,2]
  return readlane_f64(partial, 0);
}

// Four sums of the wave order in one register.  permlane32_swap(a, b) leaves
// [a lanes 0..31, b lanes 0..31] and [a lanes 32..63, b lanes 32..63], so
// their sum holds the s = 32 step of a on lanes 0..31 and of b on lanes
// 32..63; permlane16_swap of two such registers (a|b, c|d) does the s = 16
// step of all four, leaving rows 0..3 (16 lanes each) = a, c, b, d; the DPP
// steps s = 8..1 then run inside every row.  Each total is the same sum of
// the same partials as wave_tree64's.  Returns the register: the total of a
// on lane 0, b on lane 32, c on lane 16, d on lane 48.
__device__ __forceinline__ double swap32_add(double a, double b) {
  const uint64_t ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
  return __longlong_as_double(((uint64_t)hi[0] << 32) | lo[0]) +
         __longlong_as_double(((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ double swap16_add(double x, double y) {
  const uint64_t ux = __double_as_longlong(x), uy = __double_as_longlong(y);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ux, (uint32_t)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(ux >> 32), (uint32_t)(uy >> 32), false, false);
  return __longlong_as_double(((uint64_t)hi[0] << 32) | lo[0]) +
         __longlong_as_double(((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ __forceinline__ double wave_tree64x4(double a, double b, double c, double d) {
  double z = swap16_add(swap32_add(a, b), swap32_add(c, d));
  z = z + dpp_mov_f64<0x128>(z);
  z = z + dpp_mov_f64<0x12C>(z);
  z = z + dpp_mov_f64<0x4E>(z);
  z = z + dpp_mov_f64<0xB1>(z);
  return z;
}
// lane that holds the total of sum 4g + r after wave_tree64x4
__device__ __forceinline__ constexpr int x4_lane(int r) { return r == 0 ? 0 : r == 1 ? 32 : r == 2 ? 16 : 48; }

// wave_tree64 of v[0..K) (every lane's partials), four sums at a time; every
// lane gets every total.
template <int K>
__device__ __forceinline__ void wave_tree64_many(double* v) {
#pragma unroll
  for (int g = 0; g < K; g += 4) {
    const double z = wave_tree64x4(v[g], g + 1 < K ? v[g + 1] : 0.0, g + 2 < K ? v[g + 2] : 0.0,
                                   g + 3 < K ? v[g + 3] : 0.0);
#pragma unroll
    for (int r = 0; r < 4 && g + r < K; ++r) v[g + r] = readlane_f64(z, x4_lane(r));
  }
}

// Sums of a fit's per-point terms over a block of NT threads (NT = 64: one
// wave; NT = 256: four waves): thread t holds the partial of the terms
// i = t mod NT (ascending i); the partials combine as p[t] += p[t + s] for
// s = NT/2 .. 1, the s >= 64 steps through LDS (red: K x 128 doubles, tot: K
// doubles), the rest inside wave 0.  Every thread returns the K totals in
// acc.  For NT = 64 this is the wave order of wave_tree64.
template <int NT, int K>
__device__ __forceinline__ void block_tree(double* acc, double* red, double* tot) {
  if constexpr (NT == 64) {
    wave_tree64_many<K>(acc);
  } else {
    static_assert(NT == 256, "fits run on 64 or 256 threads");
    const int t = threadIdx.x;
    __syncthreads();
    if (t >= 128) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k * 128 + (t - 128)] = acc[k];
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = acc[k] + red[k * 128 + t];
    }
    __syncthreads();
    if (t >= 64 && t < 128) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k * 128 + (t - 64)] = acc[k];
    }
    __syncthreads();
    if (t < 64) {
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = acc[k] + red[k * 128 + t];
      // four sums at a time; the lane holding a total writes it
#pragma unroll
      for (int g = 0; g < K; g += 4) {
        const double z = wave_tree64x4(acc[g], g + 1 < K ? acc[g + 1] : 0.0, g + 2 < K ? acc[g + 2] : 0.0,
                                       g + 3 < K ? acc[g + 3] : 0.0);
#pragma unroll
        for (int r = 0; r < 4 && g + r < K; ++r)
          if (t == x4_lane(r)) tot[g + r] = z;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = tot[k];
  }
}

// Reduction staging of a 256-thread fit (dynamic LDS after the rows).
struct BlockRed {
  double* red;  // 45 x 128
  double* tot;  // 45
};
constexpr int kRedValues = 45;

template <class F>
__device__ __forceinline__ double wave_sum64(int64_t begin, int64_t end, F term) {
  const int l = lane_id();
  double p = 0.0;
  // first index >= begin congruent to l mod 64
  int64_t i = begin + ((l - (int)(begin & 63)) & 63);
  for (; i < end; i += 64) p = p + term(i);
  return wave_tree64(p);
}

// Sequential sum term(0) + term(1) + ... + term(n-1), left to right, computed
// by the whole wave: each chunk of 64 terms is evaluated in parallel and then
// added in index order through the scalar unit.  Same result as the plain
// loop `s = 0; for i: s += term(i)`; every lane returns it.
template <class F>
__device__ __forceinline__ double wave_seq_sum(int n, F term) {
  const int lane = lane_id();
  double acc = 0.0;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    const double v = i < n ? term(i) : 0.0;
    const int cnt = n - base < 64 ? n - base : 64;
    for (int l = 0; l < cnt; ++l) acc = acc + readlane_f64(v, l);
  }
  return acc;
}

// Four sequential sums at once (independent accumulators, same order each).
template <class F>
__device__ __forceinline__ void wave_seq_sum4(int n, F term, double* out) {
  const int lane = lane_id();
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) term(i, v);
    const int cnt = n - base < 64 ? n - base : 64;
    for (int l = 0; l < cnt; ++l) {
      a0 = a0 + readlane_f64(v[0], l);
      a1 = a1 + readlane_f64(v[1], l);
      a2 = a2 + readlane_f64(v[2], l);
      a3 = a3 + readlane_f64(v[3], l);
    }
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
  out[3] = a3;
}

__device__ __forceinline__ double clamp01(double d) { return d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d); }

// One-way transfer error of a homography (row-major h), evaluated left to
// right exactly as RobustHomographyEstimator::squaredResidual does.
__device__ __forceinline__ double h_sq_residual(const double* __restrict__ s, const double* h) {
  const double x1 = s[0], y1 = s[1], x2 = s[2], y2 = s[3];
  const double t1 = h[0] * x1 + h[1] * y1 + h[2];
  const double t2 = h[3] * x1 + h[4] * y1 + h[5];
  const double t3 = h[6] * x1 + h[7] * y1 + h[8];
  const double d1 = x2 - (t1 / t3);
  const double d2 = y2 - (t2 / t3);
  return d1 * d1 + d2 * d2;
}

// Squared Sampson distance of a fundamental / essential matrix (row-major).
__device__ __forceinline__ double f_sq_sampson(const double* __restrict__ s, const double* F) {
  const double x1 = s[0], y1 = s[1], x2 = s[2], y2 = s[3];
  const double e11 = F[0], e12 = F[1], e13 = F[2], e21 = F[3], e22 = F[4], e23 = F[5];
  const double rxc = e11 * x2 + e21 * y2 + F[6];
  const double ryc = e12 * x2 + e22 * y2 + F[7];
  const double rwc = e13 * x2 + e23 * y2 + F[8];
  const double r = (x1 * rxc + y1 * ryc + rwc);
  const double rx = e11 * x1 + e12 * y1 + e13;
  const double ry = e21 * x1 + e22 * y1 + e23;
  return r * r / (rxc * rxc + ryc * ryc + rx * rx + ry * ry);
}

// The numerator (x2^T F x1)^2 and denominator of f_sq_sampson, with the
// same operations (f_sq_sampson = num / den).
__device__ __forceinline__ void f_sampson_parts(const double* __restrict__ s, const double* F, double& num,
                                                double& den) {
  const double x1 = s[0], y1 = s[1], x2 = s[2], y2 = s[3];
  const double e11 = F[0], e12 = F[1], e13 = F[2], e21 = F[3], e22 = F[4], e23 = F[5];
  const double rxc = e11 * x2 + e21 * y2 + F[6];
  const double ryc = e12 * x2 + e22 * y2 + F[7];
  const double rwc = e13 * x2 + e23 * y2 + F[8];
  const double r = (x1 * rxc + y1 * ryc + rwc);
  const double rx = e11 * x1 + e12 * y1 + e13;
  const double ry = e21 * x1 + e22 * y1 + e23;
  num = r * r;
  den = rxc * rxc + ryc * ryc + rx * rx + ry * ry;
}

// Squared symmetric epipolar distance (row-major F).
__device__ __forceinline__ double f_sq_symmetric(const double* __restrict__ s, const double* F) {
  const double x1 = s[0], y1 = s[1], x2 = s[2], y2 = s[3];
  const double e11 = F[0], e12 = F[1], e13 = F[2], e21 = F[3], e22 = F[4], e23 = F[5];
  const double rxc = e11 * x2 + e21 * y2 + F[6];
  const double ryc = e12 * x2 + e22 * y2 + F[7];
  const double rwc = e13 * x2 + e23 * y2 + F[8];
  const double r = (x1 * rxc + y1 * ryc + rwc);
  const double rx = e11 * x1 + e12 * y1 + e13;
  const double ry = e21 * x1 + e22 * y1 + e23;
  const double a = rxc * rxc + ryc * ryc;
  const double b = rx * rx + ry * ry;
  return r * r * (a + b) / (a * b);
}

template <int KIND>
__device__ __forceinline__ double sq_residual(const double* __restrict__ s, const double* m) {
  if constexpr (KIND == 0) return h_sq_residual(s, m);
  else return f_sq_sampson(s, m);
}

// 3x3 helpers (row-major), same formulas as the host restatement.
__device__ __forceinline__ void mat3_mul(const double* a, const double* b, double* c) {
  double t[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      t[3 * i + j] = a[3 * i + 0] * b[0 + j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
#pragma unroll
  for (int i = 0; i < 9; ++i) c[i] = t[i];
}

__device__ __forceinline__ void mat3_inverse(const double* a, double* inv) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / det;
  double t[9];
  t[0] = c00 * id;
  t[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  t[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  t[3] = c01 * id;
  t[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  t[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  t[6] = c02 * id;
  t[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  t[8] = (a[0] * a[4] - a[1] * a[3]) * id;
#pragma unroll
  for (int i = 0; i < 9; ++i) inv[i] = t[i];
}

// Determinant of a 3x3 by partial-pivot LU.
__device__ __forceinline__ double det3_partial_pivot(const double* a_in) {
  double a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = a_in[i];
  double det = 1.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int p = k;
    double best = fabs(a[k * 3 + k]);
#pragma unroll
    for (int i = k + 1; i < 3; ++i) {
      const double v = fabs(a[i * 3 + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (best == 0.0) return 0.0;
    if (p != k) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        // select-based swap keeps the array in registers
        const double ak = a[k * 3 + j];
        const double ap1 = a[(k + 1 < 3 ? k + 1 : 2) * 3 + j];
        const double ap2 = a[2 * 3 + j];
        const double ap = (p == k + 1) ? ap1 : ap2;
        a[k * 3 + j] = ap;
        if (p == k + 1) a[(k + 1 < 3 ? k + 1 : 2) * 3 + j] = ak;
        else a[2 * 3 + j] = ak;
      }
      det = -det;
    }
    const double piv = a[k * 3 + k];
    det *= piv;
#pragma unroll
    for (int i = k + 1; i < 3; ++i) {
      const double f = a[i * 3 + k] / piv;
#pragma unroll
      for (int j = k + 1; j < 3; ++j) a[i * 3 + j] = a[i * 3 + j] - f * a[k * 3 + j];
    }
  }
  return det;
}

}  // namespace gcr
