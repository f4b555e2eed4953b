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
// to r^17 / r^18.
__device__ __forceinline__ void sin_cos_taylor(double x, double& s, double& c) {
  const double pio2_hi = 1.5707963267948966;
  const double pio2_lo = 6.123233995736766e-17;
  const double k = rint(x / pio2_hi);
  const double r = (x - k * pio2_hi) - k * pio2_lo;
  const double r2 = r * r;
  double ps = 1.0, pc = 1.0;
#pragma unroll
  for (int n = 17; n >= 3; n -= 2) ps = 1.0 - r2 / (double)((n - 1) * n) * ps;
#pragma unroll
  for (int n = 18; n >= 2; n -= 2) pc = 1.0 - r2 / (double)((n - 1) * n) * pc;
  const double sr = r * ps, cr = pc;
  long q = (long)k;
  q &= 3;
  if (q < 0) q += 4;
  if (q == 0) {
    s = sr;
    c = cr;
  } else if (q == 1) {
    s = cr;
    c = -sr;
  } else if (q == 2) {
    s = -sr;
    c = -cr;
  } else {
    s = -cr;
    c = sr;
  }
}

__device__ __forceinline__ double cubic_eval(double a, double b, double c, double x) {
  return ((x + a) * x + b) * x + c;
}

__device__ inline double cubic_root_in(double a, double b, double c, double lo, double hi) {
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

// Real roots of c0 + c1 x + c2 x^2 + c3 x^3, ascending (c3 != 0).
__device__ inline int cubic_real_roots(const double* c, double* roots) {
  const double a = c[2] / c[3], b = c[1] / c[3], cc = c[0] / c[3];
  double bound = 1.0;
  bound = fmax(bound, 1.0 + fabs(a));
  bound = fmax(bound, 1.0 + fabs(b));
  bound = fmax(bound, 1.0 + fabs(cc));
  const double D = a * a - 3.0 * b;
  double pts[4];
  int np = 0;
  pts[np++] = -bound;
  if (D > 0.0) {
    const double sq = sqrt(D);
    const double x1 = (-a - sq) / 3.0, x2 = (-a + sq) / 3.0;
    if (x1 > -bound && x1 < bound) pts[np++] = x1;
    if (x2 > -bound && x2 < bound && x2 > x1) pts[np++] = x2;
  }
  pts[np++] = bound;
  int nr = 0;
  for (int i = 0; i + 1 < np; ++i) {
    const double flo = cubic_eval(a, b, cc, pts[i]), fhi = cubic_eval(a, b, cc, pts[i + 1]);
    if (flo == 0.0) {
      if (nr == 0 || roots[nr - 1] != pts[i]) roots[nr++] = pts[i];
      continue;
    }
    if ((flo < 0.0) != (fhi < 0.0) && fhi != 0.0) roots[nr++] = cubic_root_in(a, b, cc, pts[i], pts[i + 1]);
  }
  const double fb = cubic_eval(a, b, cc, pts[np - 1]);
  if (fb == 0.0 && nr < 3) roots[nr++] = pts[np - 1];
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

// Seven-point solver on one wave.  Lane L < 63 holds entry (L / 9, L % 9) of
// the 7 x 9 system built from the seven rows given as x1 y1 x2 y2 (pr holds
// the 28 values, uniform).  Gauss-Jordan with full pivoting: the pivot is the
// largest |a_ij| over rows >= r and unused columns (first in row-major
// order), the rank cut-off 7 * eps * (first pivot).  Writes up to three
// models (F33 = 1) and returns their number; -1 when the kernel is not
// two-dimensional or the cubic is degenerate.
__device__ inline int wave_f7(const double* pr, double models[3][9]) {
  const int lane = lane_id();
  const int ri = lane / 9, rj = lane % 9;
  const bool live = lane < 63;
  double a = 0.0;
  if (live) {
    const double x0 = pr[4 * ri], y0 = pr[4 * ri + 1], x1 = pr[4 * ri + 2], y1 = pr[4 * ri + 3];
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
  int col_of_row[7];
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
    col_of_row[r] = pc;
    ++rank;
  }
  if (9 - rank != 2) return -1;
  int freec[2], nf = 0;
  for (int j = 0; j < 9; ++j)
    if (!((used >> j) & 1u)) freec[nf++] = j;
  double f1[9], f2[9];
  for (int j = 0; j < 9; ++j) {
    f1[j] = 0.0;
    f2[j] = 0.0;
  }
  f1[freec[0]] = 1.0;
  f2[freec[1]] = 1.0;
  for (int r = 0; r < rank; ++r) {
    f1[col_of_row[r]] = -readlane_f64(a, r * 9 + freec[0]);
    f2[col_of_row[r]] = -readlane_f64(a, r * 9 + freec[1]);
  }
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
  double roots[3];
  const int n = cubic_real_roots(c, roots);
  if (n < 1 || n > 3) return -1;
  int nm = 0;
  for (int r = 0; r < n; ++r) {
    double lambda = roots[r], mu = 1.;
    const double s = f1[8] * roots[r] + f2[8];
    if (fabs(s) > eps) {
      mu = 1.0 / s;
      lambda *= mu;
      for (int i = 0; i < 8; ++i) models[nm][i] = f1[i] * lambda + f2[i] * mu;
      models[nm][8] = 1.0;
      ++nm;
    }
  }
  return nm;
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
