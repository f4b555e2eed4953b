// Device numerics of the fundamental-matrix path (gfx950, wave64).
//
// Wave-cooperative pieces (seven-point elimination, sums over points) spread
// data over lanes; the small dense pieces (cubic roots, 3x3 SVD, 7x7
// Cholesky, LM updates) run redundantly on every lane with identical
// inputs, so every lane holds the same result without broadcasts.  All of it
// uses only +, -, *, /, sqrt and the fixed reduction order of wave_sum64, so
// results are bit-identical to a sequential host restatement of the same
// operations compiled without fused multiply-add.
#pragma once

#include "dev_common.hpp"

namespace gcr {

// sin and cos by Cody-Waite reduction (pi/2 in two parts) and Taylor series
// to r^17 / r^18, in three parts for callers that form the 17 series
// quotients r^2 / ((n - 1) n) in parallel: the reduction, the denominators in
// the order the series use them (m = 0..7: sine, n = 17, 15, .., 3; m =
// 8..16: cosine, n = 18, 16, .., 2), and the series from the quotients.
__device__ __forceinline__ void sct_reduce(double x, double& k, double& r, double& r2) {
  const double pio2_hi = 1.5707963267948966;
  const double pio2_lo = 6.123233995736766e-17;
  k = rint(x / pio2_hi);
  r = (x - k * pio2_hi) - k * pio2_lo;
  r2 = r * r;
}

__device__ __forceinline__ double sct_denominator(int m) {
  const int n = m < 8 ? 17 - 2 * m : 18 - 2 * (m - 8);
  return (double)((n - 1) * n);
}

__device__ __forceinline__ void sct_finish(double k, double r, const double* q, double& s, double& c) {
  double ps = 1.0, pc = 1.0;
#pragma unroll
  for (int m = 0; m < 8; ++m) ps = 1.0 - q[m] * ps;
#pragma unroll
  for (int m = 8; m < 17; ++m) pc = 1.0 - q[m] * pc;
  const double sr = r * ps, cr = pc;
  long qd = (long)k;
  qd &= 3;
  if (qd < 0) qd += 4;
  if (qd == 0) {
    s = sr;
    c = cr;
  } else if (qd == 1) {
    s = cr;
    c = -sr;
  } else if (qd == 2) {
    s = -sr;
    c = -cr;
  } else {
    s = -cr;
    c = sr;
  }
}

__device__ __forceinline__ void sin_cos_taylor(double x, double& s, double& c) {
  double k, r, r2, q[17];
  sct_reduce(x, k, r, r2);
#pragma unroll
  for (int m = 0; m < 17; ++m) q[m] = r2 / sct_denominator(m);
  sct_finish(k, r, q, s, c);
}

__device__ __forceinline__ double cubic_eval(double a, double b, double c, double x) {
  return ((x + a) * x + b) * x + c;
}

__device__ __forceinline__ double cubic_root_in(double a, double b, double c, double lo, double hi) {
  double flo = cubic_eval(a, b, c, lo);
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    const double fx = cubic_eval(a, b, c, x);
    if (fx == 0.0) return x;
    if ((fx < 0.0) == (flo < 0.0)) {
      lo = x;
      flo = fx;
    } else
      hi = x;
    const double d = (3.0 * x + 2.0 * a) * x + b;
    double xn = d != 0.0 ? x - fx / d : 0.5 * (lo + hi);
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (fabs(xn - x) <= 1e-15 * fabs(x) || hi - lo <= 1e-15 * fabs(x)) return xn;
    x = xn;
  }
  return x;
}

// Real roots of c0 + c1 x + c2 x^2 + c3 x^3, ascending (c3 != 0), into
// r0, r1, r2; returns their number.  The critical points split [-bound,
// bound] into at most three monotone pieces, taken left to right; a piece
// whose left end is a root contributes that end (unless it repeats the last
// root), otherwise a sign change is searched by cubic_root_in; the right end
// of the last piece is a root when the cubic vanishes there.  Written with
// named slots instead of indexed arrays so everything stays in registers.
__device__ __forceinline__ int cubic_real_roots3(const double* c, double& r0, double& r1, double& r2) {
  const double a = c[2] / c[3], b = c[1] / c[3], cc = c[0] / c[3];
  double bound = 1.0;
  bound = fmax(bound, 1.0 + fabs(a));
  bound = fmax(bound, 1.0 + fabs(b));
  bound = fmax(bound, 1.0 + fabs(cc));
  const double D = a * a - 3.0 * b;
  double s1 = 0.0, s2 = 0.0;
  bool v1 = false, v2 = false;
  if (D > 0.0) {
    const double sq = sqrt(D);
    s1 = (-a - sq) / 3.0;
    s2 = (-a + sq) / 3.0;
    v1 = s1 > -bound && s1 < bound;
    v2 = s2 > -bound && s2 < bound && s2 > s1;
  }
  int nr = 0;
  double last = 0.0;
  r0 = r1 = r2 = 0.0;
  double lo = -bound;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bool use = k == 0 ? v1 : (k == 1 ? v2 : true);
    const double hi = k == 0 ? s1 : (k == 1 ? s2 : bound);
    if (use) {
      const double flo = cubic_eval(a, b, cc, lo), fhi = cubic_eval(a, b, cc, hi);
      bool add = false;
      double v = 0.0;
      if (flo == 0.0) {
        if (nr == 0 || last != lo) {
          add = true;
          v = lo;
        }
      } else if ((flo < 0.0) != (fhi < 0.0) && fhi != 0.0) {
        add = true;
        v = cubic_root_in(a, b, cc, lo, hi);
      }
      // value selects (not stores through a chosen reference) keep the
      // roots in registers
      r0 = (add && nr == 0) ? v : r0;
      r1 = (add && nr == 1) ? v : r1;
      r2 = (add && nr >= 2) ? v : r2;
      last = add ? v : last;
      nr += add ? 1 : 0;
      lo = hi;
    }
  }
  const double fb = cubic_eval(a, b, cc, bound);
  const bool addb = fb == 0.0 && nr < 3;
  r0 = (addb && nr == 0) ? bound : r0;
  r1 = (addb && nr == 1) ? bound : r1;
  r2 = (addb && nr == 2) ? bound : r2;
  nr += addb ? 1 : 0;
  return nr;
}

// Wave argmax of v (ties: smallest index); returns the winning lane.
__device__ __forceinline__ int wave_argmax(double v, int idx, double& best) {
  for (int s = 32; s >= 1; s >>= 1) {
    const double ov = __shfl_xor(v, s, 64);
    const int oi = __shfl_xor(idx, s, 64);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  best = readlane_f64(v, 0);
  return __builtin_amdgcn_readlane(idx, 0);
}

// The same argmax for values that are never NaN, reduced towards lane 0
// through permlane swaps and DPP (the lane moves of wave_tree64): at each
// step lane l < s keeps the larger of its own (value, index) and lane
// l + s's, ties to the smaller index.  The order is total, so lane 0 ends
// with the argmax wave_argmax returns.
template <int CTRL>
__device__ __forceinline__ int dpp_mov_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int lanes_down32_i32(int v) {
  return (int)__builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false)[1];
}
__device__ __forceinline__ int lanes_down16_i32(int v) {
  return (int)__builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false)[1];
}
__device__ __forceinline__ void argmax_step(double& v, int& idx, double ov, int oi) {
  if (ov > v || (ov == v && oi < idx)) {
    v = ov;
    idx = oi;
  }
}
__device__ __forceinline__ int wave_argmax_dpp(double v, int idx, double& best) {
  argmax_step(v, idx, lanes_down32(v), lanes_down32_i32(idx));
  argmax_step(v, idx, lanes_down16(v), lanes_down16_i32(idx));
  argmax_step(v, idx, dpp_mov_f64<0x128>(v), dpp_mov_i32<0x128>(idx));
  argmax_step(v, idx, dpp_mov_f64<0x12C>(v), dpp_mov_i32<0x12C>(idx));
  argmax_step(v, idx, dpp_mov_f64<0x4E>(v), dpp_mov_i32<0x4E>(idx));
  argmax_step(v, idx, dpp_mov_f64<0xB1>(v), dpp_mov_i32<0xB1>(idx));
  best = readlane_f64(v, 0);
  return __builtin_amdgcn_readlane(idx, 0);
}

// Models of the seven-point solver from the two null vectors f1, f2 of the
// 7 x 9 system (solver_fundamental_matrix_seven_point.h:150-291): the cubic
// det(l f1 + (1 - l) f2) = 0, its real roots and F33 = 1 scaling.  Returns
// the bitmask of models written (root order), -1 on failure.
__device__ __forceinline__ int f7_models_from_null_space(double* f1, double* f2, double models[3][9]) {
#pragma unroll
  for (int i = 0; i < 9; ++i) f1[i] -= f2[i];
  double t0, t1, t2, c[4];
  t0 = f2[4] * f2[8] - f2[5] * f2[7];
  t1 = f2[3] * f2[8] - f2[5] * f2[6];
  t2 = f2[3] * f2[7] - f2[4] * f2[6];
  c[0] = f2[0] * t0 - f2[1] * t1 + f2[2] * t2;
  c[1] = f1[0] * t0 - f1[1] * t1 + f1[2] * t2 - f1[3] * (f2[1] * f2[8] - f2[2] * f2[7]) +
         f1[4] * (f2[0] * f2[8] - f2[2] * f2[6]) - f1[5] * (f2[0] * f2[7] - f2[1] * f2[6]) +
         f1[6] * (f2[1] * f2[5] - f2[2] * f2[4]) - f1[7] * (f2[0] * f2[5] - f2[2] * f2[3]) +
         f1[8] * (f2[0] * f2[4] - f2[1] * f2[3]);
  t0 = f1[4] * f1[8] - f1[5] * f1[7];
  t1 = f1[3] * f1[8] - f1[5] * f1[6];
  t2 = f1[3] * f1[7] - f1[4] * f1[6];
  c[2] = f2[0] * t0 - f2[1] * t1 + f2[2] * t2 - f2[3] * (f1[1] * f1[8] - f1[2] * f1[7]) +
         f2[4] * (f1[0] * f1[8] - f1[2] * f1[6]) - f2[5] * (f1[0] * f1[7] - f1[1] * f1[6]) +
         f2[6] * (f1[1] * f1[5] - f1[2] * f1[4]) - f2[7] * (f1[0] * f1[5] - f1[2] * f1[3]) +
         f2[8] * (f1[0] * f1[4] - f1[1] * f1[3]);
  c[3] = f1[0] * t0 - f1[1] * t1 + f1[2] * t2;
  const double eps = 2.220446049250313e-16;
  if (fabs(c[0] + c[1] + c[2] + c[3]) < 1e-9 || fabs(c[0]) < eps || fabs(c[1]) < eps || fabs(c[2]) < eps ||
      fabs(c[3]) < eps)
    return -1;
  double rt0, rt1, rt2;
  const int n = cubic_real_roots3(c, rt0, rt1, rt2);
  if (n < 1 || n > 3) return -1;
  int mask = 0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double root = r == 0 ? rt0 : (r == 1 ? rt1 : rt2);
    double lambda = root, mu = 1.;
    const double s = f1[8] * root + f2[8];
    if (r < n && fabs(s) > eps) {
      mu = 1.0 / s;
      lambda *= mu;
#pragma unroll
      for (int i = 0; i < 8; ++i) models[r][i] = f1[i] * lambda + f2[i] * mu;
      models[r][8] = 1.0;
      mask |= 1 << r;
    }
  }
  return mask;
}

// Seven-point solver on one wave.  Lane L < 63 holds entry (L / 9, L % 9) of
// the 7 x 9 system; the caller passes the lane's own row (x0 y0 x1 y1 of
// sample point L / 9; ignored on lane 63).  Gauss-Jordan with full pivoting:
// the pivot is the largest |a_ij| over rows >= r and unused columns (first in
// row-major order), the rank cut-off 7 * eps * (first pivot).  Model r
// (F33 = 1) belongs to the r-th real root of the cubic (ascending); returns
// the bit mask of the roots that give a model (their order is the model
// order), or -1 when the kernel is not two-dimensional or the cubic is
// degenerate.  Pivot rows and free columns are kept in bit fields so no
// per-lane array is indexed at run time.
__device__ __forceinline__ int wave_f7(double x0, double y0, double x1, double y1, double models[3][9]) {
  const int lane = lane_id();
  const int ri = lane / 9, rj = lane % 9;
  const bool live = lane < 63;
  double a = 0.0;
  if (live) {
    switch (rj) {
      case 0: a = x1 * x0; break;
      case 1: a = x1 * y0; break;
      case 2: a = x1; break;
      case 3: a = y1 * x0; break;
      case 4: a = y1 * y0; break;
      case 5: a = y1; break;
      case 6: a = x0; break;
      case 7: a = y0; break;
      default: a = 1; break;
    }
  }
  uint32_t used = 0;
  uint32_t row_of_col = 0;  // 3 bits per pivot column: its pivot row
  int rank = 0;
  double max_pivot = 0.0;
  for (int r = 0; r < 7; ++r) {
    const bool elig = live && ri >= r && !((used >> rj) & 1u);
    double best;
    const int win = wave_argmax(elig ? fabs(a) : -1.0, lane, best);
    if (r == 0) max_pivot = best;
    if (best < 0.0 || best <= 7.0 * 2.220446049250313e-16 * max_pivot) break;
    const int pr_ = win / 9, pc = win % 9;
    // swap rows r and pr_
    const double from_pr = __shfl(a, pr_ * 9 + rj, 64);
    const double from_r = __shfl(a, r * 9 + rj, 64);
    if (live && pr_ != r) {
      if (ri == r) a = from_pr;
      else if (ri == pr_) a = from_r;
    }
    const double piv = readlane_f64(a, r * 9 + pc);
    if (live && ri == r) a = a / piv;
    const double arj = __shfl(a, r * 9 + rj, 64);
    const double f = __shfl(a, ri * 9 + pc, 64);
    if (live && ri != r && f != 0.0) a = a - f * arj;
    used |= 1u << pc;
    row_of_col |= (uint32_t)r << (3 * pc);
    ++rank;
  }
  if (9 - rank != 2) return -1;
  const uint32_t freem = ~used & 0x1ffu;
  const int fc0 = __builtin_ctz(freem);
  const int fc1 = __builtin_ctz(freem & ~(1u << fc0));
  double f1[9], f2[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    if ((used >> j) & 1u) {
      const int r = (int)((row_of_col >> (3 * j)) & 7u);
      f1[j] = -readlane_f64(a, r * 9 + fc0);
      f2[j] = -readlane_f64(a, r * 9 + fc1);
    } else {
      f1[j] = j == fc0 ? 1.0 : 0.0;
      f2[j] = j == fc1 ? 1.0 : 0.0;
    }
  }
  return f7_models_from_null_space(f1, f2, models);
}

// Seven-point solve of one sample by one thread: the same Gauss-Jordan
// elimination with complete pivoting (first maximum in row-major scan order
// over unused columns, 7 eps relative rank cut), the same null vectors, cubic
// and models as wave_f7.  The 7 x 9 system stays in registers: every index is
// a compile-time constant and the pivot row / column are applied by selects.
// P: the pair's points, s: the 7 sample indices.  Returns the bitmask of
// models written to models[0..2] (root order), -1 on failure.
__device__ __forceinline__ int thread_f7(const double* __restrict__ P, const int32_t* __restrict__ s,
                                         double models[3][9]) {
  double a[7][9];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const double* d = P + 4 * (int64_t)s[i];
    const double x0 = d[0], y0 = d[1], x1 = d[2], y1 = d[3];
    a[i][0] = x1 * x0;
    a[i][1] = x1 * y0;
    a[i][2] = x1;
    a[i][3] = y1 * x0;
    a[i][4] = y1 * y0;
    a[i][5] = y1;
    a[i][6] = x0;
    a[i][7] = y0;
    a[i][8] = 1;
  }
  uint32_t used = 0;
  uint32_t row_of_col = 0;  // 3 bits per pivot column: its pivot row
  int rank = 0;
  double max_pivot = 0.0;
  bool active = true;
#pragma unroll
  for (int r = 0; r < 7; ++r) {
    if (active) {
      double best = 0.0;
      int pr = -1, pc = -1;
#pragma unroll
      for (int i = r; i < 7; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) {
          const double v = fabs(a[i][j]);
          const bool take = !((used >> j) & 1u) && v > best;
          best = take ? v : best;
          pr = take ? i : pr;
          pc = take ? j : pc;
        }
      if (r == 0) max_pivot = best;
      if (pr < 0 || best <= 7.0 * 2.220446049250313e-16 * max_pivot) {
        active = false;
      } else {
#pragma unroll
        for (int i = r + 1; i < 7; ++i) {
          const bool sw = i == pr;
#pragma unroll
          for (int j = 0; j < 9; ++j) {
            const double t = a[r][j];
            a[r][j] = sw ? a[i][j] : t;
            a[i][j] = sw ? t : a[i][j];
          }
        }
        double piv = a[r][0];
#pragma unroll
        for (int j = 1; j < 9; ++j) piv = j == pc ? a[r][j] : piv;
#pragma unroll
        for (int j = 0; j < 9; ++j) a[r][j] = a[r][j] / piv;
#pragma unroll
        for (int i = 0; i < 7; ++i) {
          if (i == r) continue;
          double f = a[i][0];
#pragma unroll
          for (int j = 1; j < 9; ++j) f = j == pc ? a[i][j] : f;
          const bool upd = f != 0.0;
#pragma unroll
          for (int j = 0; j < 9; ++j) a[i][j] = upd ? a[i][j] - f * a[r][j] : a[i][j];
        }
        used |= 1u << pc;
        row_of_col |= (uint32_t)r << (3 * pc);
        ++rank;
      }
    }
  }
  if (9 - rank != 2) return -1;
  const uint32_t freem = ~used & 0x1ffu;
  const int fc0 = __builtin_ctz(freem);
  const int fc1 = __builtin_ctz(freem & ~(1u << fc0));
  // the free columns of every row, then the null vectors
  double v0[7], v1[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    double x = a[i][0], y = a[i][0];
#pragma unroll
    for (int j = 1; j < 9; ++j) {
      x = j == fc0 ? a[i][j] : x;
      y = j == fc1 ? a[i][j] : y;
    }
    v0[i] = x;
    v1[i] = y;
  }
  double f1[9], f2[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    if ((used >> j) & 1u) {
      const int r = (int)((row_of_col >> (3 * j)) & 7u);
      double x = v0[0], y = v1[0];
#pragma unroll
      for (int i = 1; i < 7; ++i) {
        x = i == r ? v0[i] : x;
        y = i == r ? v1[i] : y;
      }
      f1[j] = -x;
      f2[j] = -y;
    } else {
      f1[j] = j == fc0 ? 1.0 : 0.0;
      f2[j] = j == fc1 ? 1.0 : 0.0;
    }
  }
  return f7_models_from_null_space(f1, f2, models);
}

// Oriented epipolar constraint over n rows (x1 y1 x2 y2 each).
__device__ inline bool f_orientation_valid(const double* F, const double* rows, int n) {
  const double eps = 1.9984e-15;
  double e0 = F[1] * F[8] - F[2] * F[7];
  double e1 = F[2] * F[6] - F[0] * F[8];
  double e2 = F[0] * F[7] - F[1] * F[6];
  if (!(e0 > eps || e0 < -eps || e1 > eps || e1 < -eps || e2 > eps || e2 < -eps)) {
    e0 = F[4] * F[8] - F[5] * F[7];
    e1 = F[5] * F[6] - F[3] * F[8];
    e2 = F[3] * F[7] - F[4] * F[6];
  }
  (void)e0;
  const double* p0 = rows;
  const double s2 = (F[0] * p0[2] + F[3] * p0[3] + F[6]) * (e1 - e2 * p0[1]);
  for (int i = 1; i < n; i++) {
    const double* p = rows + 4 * i;
    const double s1 = (F[0] * p[2] + F[3] * p[3] + F[6]) * (e1 - e2 * p[1]);
    if (s2 * s1 < 0) return false;
  }
  return true;
}

// The same test on points P + 4 * s[i] (sample indices into a point array).
__device__ __forceinline__ bool f_orientation_valid_idx(const double* F, const double* P, const int32_t* s, int n) {
  const double eps = 1.9984e-15;
  double e0 = F[1] * F[8] - F[2] * F[7];
  double e1 = F[2] * F[6] - F[0] * F[8];
  double e2 = F[0] * F[7] - F[1] * F[6];
  if (!(e0 > eps || e0 < -eps || e1 > eps || e1 < -eps || e2 > eps || e2 < -eps)) {
    e0 = F[4] * F[8] - F[5] * F[7];
    e1 = F[5] * F[6] - F[3] * F[8];
    e2 = F[3] * F[7] - F[4] * F[6];
  }
  (void)e0;
  const double* p0 = P + 4 * (int64_t)s[0];
  const double s2 = (F[0] * p0[2] + F[3] * p0[3] + F[6]) * (e1 - e2 * p0[1]);
  for (int i = 1; i < n; i++) {
    const double* p = P + 4 * (int64_t)s[i];
    const double s1 = (F[0] * p[2] + F[3] * p[3] + F[6]) * (e1 - e2 * p[1]);
    if (s2 * s1 < 0) return false;
  }
  return true;
}

// 3x3 SVD by one-sided Jacobi (same sweep order and stopping rule as the
// host restatement).
__device__ inline void svd3(const double* A, double* U, double* s, double* V) {
  double B[9], Vm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int i = 0; i < 9; ++i) B[i] = A[i];
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < 3; ++i) {
          alpha += B[3 * i + p] * B[3 * i + p];
          beta += B[3 * i + q] * B[3 * i + q];
          gamma += B[3 * i + p] * B[3 * i + q];
        }
        if (gamma == 0.0) continue;
        const double rel = fabs(gamma) / sqrt(alpha * beta);
        if (rel > off) off = rel;
        if (rel < 1e-15) continue;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < 3; ++i) {
          const double bp = B[3 * i + p], bq = B[3 * i + q];
          B[3 * i + p] = c * bp - sn * bq;
          B[3 * i + q] = sn * bp + c * bq;
          const double vp = Vm[3 * i + p], vq = Vm[3 * i + q];
          Vm[3 * i + p] = c * vp - sn * vq;
          Vm[3 * i + q] = sn * vp + c * vq;
        }
      }
    if (off < 1e-15) break;
  }
  double nrm[3];
  for (int j = 0; j < 3; ++j) nrm[j] = sqrt(B[j] * B[j] + B[3 + j] * B[3 + j] + B[6 + j] * B[6 + j]);
  int ord[3] = {0, 1, 2};
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 3; ++b)
      if (nrm[ord[b]] > nrm[ord[a]]) {
        const int t = ord[a];
        ord[a] = ord[b];
        ord[b] = t;
      }
  for (int k = 0; k < 3; ++k) {
    const int j = ord[k];
    s[k] = nrm[j];
    for (int i = 0; i < 3; ++i) V[3 * i + k] = Vm[3 * i + j];
    if (k < 2)
      for (int i = 0; i < 3; ++i) U[3 * i + k] = nrm[j] > 0 ? B[3 * i + j] / nrm[j] : (i == k ? 1.0 : 0.0);
  }
  U[2] = U[3] * U[7] - U[6] * U[4];
  U[5] = U[6] * U[1] - U[0] * U[7];
  U[8] = U[0] * U[4] - U[3] * U[1];
}

}  // namespace gcr
