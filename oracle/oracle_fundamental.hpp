// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Fundamental-matrix estimator restatement:
//   inc/estimators/fundamental_estimator.h (FundamentalMatrixEstimator)
//   inc/estimators/solver_fundamental_matrix_seven_point.h (7-point)
//   inc/estimators/solver_fundamental_matrix_eight_point.h (8-point)
//   inc/estimators/solver_fundamental_matrix_bundle_adjustment.h
//   inc/relative_pose/bundle.cpp:252-339,454-500 (lm_F_impl, refine_fundamental)
//   inc/relative_pose/jacobian_impl.h:468-614 (FactorizedFundamentalMatrix,
//     FundamentalJacobianAccumulator), robust_loss.h:19-31 (TruncatedLoss)
//   py/gcransac_python.cpp:1832-1887 (post-hoc weighted bundle adjustment)
//
// Eigen's decompositions inside these functions (FullPivLU kernel,
// PolynomialSolver, FullPivHouseholderQR, JacobiSVD, LLT) are restated with
// textbook algorithms written here; their results agree with Eigen's to
// rounding, not bit for bit (parity with the reference binary is unpinned
// there).  Reductions the restatement adds (A^T A, the LM sums) use the
// sum256 order of oracle_core.hpp.  sin/cos are evaluated by sin_cos below
// (Cody-Waite reduction and Taylor series) so the result depends on no libm.
#pragma once

#include "oracle_core.hpp"
#include "oracle_gcransac.hpp"
#include "oracle_homography.hpp"

namespace orc {

// sin and cos: reduce x by k*pi/2 (pi/2 split in two parts) to |r| <= pi/4,
// Taylor series to r^17 / r^18, then the quadrant.
inline void sin_cos(double x, double& s, double& c) {
  const double pio2_hi = 1.5707963267948966;      // nearest double to pi/2
  const double pio2_lo = 6.123233995736766e-17;   // pi/2 - pio2_hi
  const double k = std::nearbyint(x / pio2_hi);
  const double r = (x - k * pio2_hi) - k * pio2_lo;
  const double r2 = r * r;
  // sin r = r (1 - r^2/3! (1 - r^2/(4*5) (1 - ...)))
  double ps = 1.0, pc = 1.0;
  for (int n = 17; n >= 3; n -= 2) ps = 1.0 - r2 / (double)((n - 1) * n) * ps;
  for (int n = 18; n >= 2; n -= 2) pc = 1.0 - r2 / (double)((n - 1) * n) * pc;
  const double sr = r * ps, cr = pc;
  long q = (long)k;
  q &= 3;
  if (q < 0) q += 4;
  switch (q) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
  }
}

// ---- seven-point solver -----------------------------------------------------
// Null space of the 7 x 9 system by Gauss-Jordan elimination with full
// pivoting (the role of FullPivLU::kernel, seven_point.h:161-166).  Returns
// the kernel dimension (2 for a generic sample); the basis vectors set one
// free unknown to 1 and the other to 0.
inline int null_space_7x9(double a[7][9], double f1[9], double f2[9]) {
  int col_of_row[7];
  int used_col[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int rank = 0;
  double max_pivot = 0.0;
  for (int r = 0; r < 7; ++r) {
    int pr = -1, pc = -1;
    double best = 0.0;
    for (int i = r; i < 7; ++i)
      for (int j = 0; j < 9; ++j)
        if (!used_col[j] && std::fabs(a[i][j]) > best) {
          best = std::fabs(a[i][j]);
          pr = i;
          pc = j;
        }
    if (r == 0) max_pivot = best;
    if (pr < 0 || best <= 7.0 * std::numeric_limits<double>::epsilon() * max_pivot) break;
    if (pr != r)
      for (int j = 0; j < 9; ++j) std::swap(a[r][j], a[pr][j]);
    const double piv = a[r][pc];
    for (int j = 0; j < 9; ++j) a[r][j] = a[r][j] / piv;
    for (int i = 0; i < 7; ++i) {
      if (i == r) continue;
      const double f = a[i][pc];
      if (f != 0.0)
        for (int j = 0; j < 9; ++j) a[i][j] = a[i][j] - f * a[r][j];
    }
    used_col[pc] = 1;
    col_of_row[r] = pc;
    ++rank;
  }
  const int dim = 9 - rank;
  if (dim != 2) return dim;
  int freec[2], nf = 0;
  for (int j = 0; j < 9; ++j)
    if (!used_col[j]) freec[nf++] = j;
  double* out[2] = {f1, f2};
  for (int k = 0; k < 2; ++k) {
    for (int j = 0; j < 9; ++j) out[k][j] = 0.0;
    out[k][freec[k]] = 1.0;
    for (int r = 0; r < rank; ++r) out[k][col_of_row[r]] = -a[r][freec[k]];
  }
  return 2;
}

// Real roots of c0 + c1 x + c2 x^2 + c3 x^3 (c3 != 0), ascending.  The
// critical points of the monic cubic split the line into monotone pieces;
// each piece with a sign change is searched by safeguarded Newton steps
// (bisection when a step leaves the bracket).  Only +, -, *, / and sqrt.
inline double cubic_eval(double a, double b, double c, double x) { return ((x + a) * x + b) * x + c; }

inline double cubic_root_in(double a, double b, double c, double lo, double hi) {
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
    if (std::fabs(xn - x) <= 1e-15 * std::fabs(x) || hi - lo <= 1e-15 * std::fabs(x)) return xn;
    x = xn;
  }
  return x;
}

inline int cubic_real_roots(const double c[4], double roots[3]) {
  const double a = c[2] / c[3], b = c[1] / c[3], cc = c[0] / c[3];
  double bound = 1.0;
  bound = std::max(bound, 1.0 + std::fabs(a));
  bound = std::max(bound, 1.0 + std::fabs(b));
  bound = std::max(bound, 1.0 + std::fabs(cc));
  const double D = a * a - 3.0 * b;
  double pts[4];
  int np = 0;
  pts[np++] = -bound;
  if (D > 0.0) {
    const double sq = std::sqrt(D);
    double x1 = (-a - sq) / 3.0, x2 = (-a + sq) / 3.0;
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

// seven_point.h:91-291 on a sample of 7 points (raw coordinates).
inline bool f7_solve(const Points& pts, const size_t* sample, std::vector<Model>& models) {
  double a[7][9];
  for (int i = 0; i < 7; i++) {
    const double* p = pts.row(sample[i]);
    const double x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
    a[i][0] = x1 * x0; a[i][1] = x1 * y0; a[i][2] = x1;
    a[i][3] = y1 * x0; a[i][4] = y1 * y0; a[i][5] = y1;
    a[i][6] = x0; a[i][7] = y0; a[i][8] = 1;
  }
  double f1[9], f2[9];
  if (null_space_7x9(a, f1, f2) != 2) return false;
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
  const double eps = std::numeric_limits<double>::epsilon();
  if (std::fabs(c[0] + c[1] + c[2] + c[3]) < 1e-9 || std::fabs(c[0]) < eps || std::fabs(c[1]) < eps ||
      std::fabs(c[2]) < eps || std::fabs(c[3]) < eps)
    return false;
  double roots[3];
  const int n = cubic_real_roots(c, roots);
  if (n < 1 || n > 3) return false;
  for (int r = 0; r < n; ++r) {
    double lambda = roots[r], mu = 1.;
    const double s = f1[8] * roots[r] + f2[8];
    if (std::fabs(s) > eps) {
      mu = 1.0 / s;
      lambda *= mu;
      Model m;
      for (int i = 0; i < 8; ++i) m.m[i] = f1[i] * lambda + f2[i] * mu;
      m.m[8] = 1.0;
      models.push_back(m);
    }
  }
  return true;
}

// fundamental_estimator.h:737-800 (oriented epipolar constraint)
inline bool f_orientation_valid(const double* F, const Points& pts, const size_t* sample, size_t n) {
  const double eps = 1.9984e-15;
  double e[3];
  // epipole = row0 x row2, else row1 x row2
  e[0] = F[1] * F[8] - F[2] * F[7];
  e[1] = F[2] * F[6] - F[0] * F[8];
  e[2] = F[0] * F[7] - F[1] * F[6];
  if (!(e[0] > eps || e[0] < -eps || e[1] > eps || e[1] < -eps || e[2] > eps || e[2] < -eps)) {
    e[0] = F[4] * F[8] - F[5] * F[7];
    e[1] = F[5] * F[6] - F[3] * F[8];
    e[2] = F[3] * F[7] - F[4] * F[6];
  }
  auto signum = [&](const double* p) {
    const double s1 = F[0] * p[2] + F[3] * p[3] + F[6];
    const double s2 = e[1] - e[2] * p[1];
    return s1 * s2;
  };
  const double s2 = signum(sample ? pts.row(sample[0]) : pts.row(0));
  for (size_t i = 1; i < n; i++) {
    const double s1 = signum(sample ? pts.row(sample[i]) : pts.row(i));
    if (s2 * s1 < 0) return false;
  }
  return true;
}

// fundamental_estimator.h:224-246
inline double f_squared_symmetric(const double* s, const double* F) {
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

// inc/estimators/fundamental_estimator.h:195-222
inline double f_squared_sampson(const double* s, const double* F) {
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

// ---- eight-point solver -----------------------------------------------------
// eight_point.h:93-182: null vector of A^T A as the last column of Q of a
// Householder QR with full pivoting (P A P' = Q R).  rows: points addressed
// through row(i) (i < n); weights[i] multiply row i when given.
template <class RowF, class WF>
inline void f8_solve(size_t n, RowF row, WF weight, bool weighted, double F[9]) {
  // A^T A (9 x 9), entries by sum256 over the rows
  double M[9][9];
  auto coef = [&](size_t i, int k) {
    const double* p = row(i);
    const double x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
    if (!weighted) {
      switch (k) {
        case 0: return x1 * x0; case 1: return x1 * y0; case 2: return x1;
        case 3: return y1 * x0; case 4: return y1 * y0; case 5: return y1;
        case 6: return x0; case 7: return y0; default: return 1.0;
      }
    }
    const double w = weight(i);
    const double wx0 = w * x0, wy0 = w * y0, wx1 = w * x1, wy1 = w * y1;
    switch (k) {
      case 0: return wx1 * x0; case 1: return wx1 * y0; case 2: return wx1;
      case 3: return wy1 * x0; case 4: return wy1 * y0; case 5: return wy1;
      case 6: return wx0; case 7: return wy0; default: return w;
    }
  };
  double acc[45];
  sum256_multi<45>(0, n, [&](size_t i, double* t) {
    double c[9];
    for (int k = 0; k < 9; ++k) c[k] = coef(i, k);
    int o = 0;
    for (int a = 0; a < 9; ++a)
      for (int b = a; b < 9; ++b) t[o++] = c[a] * c[b];
  }, acc);
  {
    int o = 0;
    for (int a = 0; a < 9; ++a)
      for (int b = a; b < 9; ++b) {
        M[a][b] = acc[o];
        M[b][a] = acc[o];
        ++o;
      }
  }
  // Householder QR with full pivoting; Q = T0 H0 T1 H1 ... T7 H7
  int rowt[9];
  double hv[9][9], tau[9];
  for (int k = 0; k < 9; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int j = k; j < 9; ++j)
      for (int i = k; i < 9; ++i)
        if (std::fabs(M[i][j]) > best) {
          best = std::fabs(M[i][j]);
          pr = i;
          pc = j;
        }
    rowt[k] = pr;
    if (pr != k)
      for (int j = 0; j < 9; ++j) std::swap(M[k][j], M[pr][j]);
    if (pc != k)
      for (int i = 0; i < 9; ++i) std::swap(M[i][k], M[i][pc]);
    const double alpha = M[k][k];
    double sigma = 0.0;
    for (int i = k + 1; i < 9; ++i) sigma += M[i][k] * M[i][k];
    tau[k] = 0.0;
    for (int i = 0; i < 9; ++i) hv[k][i] = 0.0;
    hv[k][k] = 1.0;
    if (sigma != 0.0) {
      const double nrm = std::sqrt(alpha * alpha + sigma);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      tau[k] = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
      for (int i = k + 1; i < 9; ++i) hv[k][i] = M[i][k] * scale;
      M[k][k] = beta;
      for (int i = k + 1; i < 9; ++i) M[i][k] = 0.0;
      for (int j = k + 1; j < 9; ++j) {
        double w = M[k][j];
        for (int i = k + 1; i < 9; ++i) w += hv[k][i] * M[i][j];
        const double tw = tau[k] * w;
        M[k][j] = M[k][j] - tw;
        for (int i = k + 1; i < 9; ++i) M[i][j] = M[i][j] - tw * hv[k][i];
      }
    }
  }
  // q = Q e_8 = T0 H0 T1 H1 ... T8 H8 e_8
  double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 1.0};
  for (int k = 8; k >= 0; --k) {
    if (tau[k] != 0.0) {
      double w = 0.0;
      for (int i = k; i < 9; ++i) w += hv[k][i] * q[i];
      const double tw = tau[k] * w;
      for (int i = k; i < 9; ++i) q[i] = q[i] - tw * hv[k][i];
    }
    if (rowt[k] != k) std::swap(q[k], q[rowt[k]]);
  }
  for (int i = 0; i < 9; ++i) F[i] = q[i];
}

// ---- 3x3 SVD (one-sided Jacobi) -----------------------------------------------
// A = U diag(s) V^T with s descending; U's third column completes a
// right-handed frame when the third singular value vanishes.
inline void svd3(const double* A, double* U, double* s, double* V) {
  double B[9];
  std::memcpy(B, A, sizeof(B));  // row-major; columns are rotated
  double Vm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
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
        const double rel = std::fabs(gamma) / std::sqrt(alpha * beta);
        if (rel > off) off = rel;
        if (rel < 1e-15) continue;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
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
  for (int j = 0; j < 3; ++j) nrm[j] = std::sqrt(B[j] * B[j] + B[3 + j] * B[3 + j] + B[6 + j] * B[6 + j]);
  int ord[3] = {0, 1, 2};
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 3; ++b)
      if (nrm[ord[b]] > nrm[ord[a]]) std::swap(ord[a], ord[b]);
  for (int k = 0; k < 3; ++k) {
    const int j = ord[k];
    s[k] = nrm[j];
    for (int i = 0; i < 3; ++i) V[3 * i + k] = Vm[3 * i + j];
    if (k < 2)
      for (int i = 0; i < 3; ++i) U[3 * i + k] = nrm[j] > 0 ? B[3 * i + j] / nrm[j] : (i == k ? 1.0 : 0.0);
  }
  // third column of U: u0 x u1
  U[2] = U[3] * U[7] - U[6] * U[4];
  U[5] = U[6] * U[1] - U[0] * U[7];
  U[8] = U[0] * U[4] - U[3] * U[1];
}

// ---- Levenberg-Marquardt refinement of F (bundle.cpp:252-339) ----------------
struct FactorizedF {
  double U[9], V[9], sigma;
  void from(const double* F) {
    double s[3];
    svd3(F, U, s, V);
    sigma = s[1] / s[0];
  }
  // U.col(0) V.col(0)^T + sigma U.col(1) V.col(1)^T
  void F(double* out) const {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) out[3 * i + j] = U[3 * i] * V[3 * j] + sigma * U[3 * i + 1] * V[3 * j + 1];
  }
};

// Truncated loss with scale 1 (robust_loss.h:19-31): loss min(r2, 1),
// weight r2 < 1.
template <class RowF, class WF>
inline double lm_f_cost(size_t n, RowF row, WF wt, bool weighted, const double* F) {
  return sum256(0, n, [&](size_t k) {
    const double* p = row(k);
    const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
    const double f0 = F[0] * x1 + F[1] * y1 + F[2];
    const double f1 = F[3] * x1 + F[4] * y1 + F[5];
    const double f2 = F[6] * x1 + F[7] * y1 + F[8];
    const double C = x2 * f0 + y2 * f1 + f2;
    const double g0 = F[0] * x2 + F[3] * y2 + F[6];
    const double g1 = F[1] * x2 + F[4] * y2 + F[7];
    const double nJc = (f0 * f0 + f1 * f1) + (g0 * g0 + g1 * g1);
    const double r2 = (C * C) / nJc;
    const double l = r2 < 1.0 ? r2 : 1.0;
    return weighted ? wt(k) * l : l;
  });
}

// Normal equations of one LM step (jacobian_impl.h:490-614): the gradient
// Jtr = sum_k w_k r_k J_k and the lower triangle of JtJ = sum_k w_k J_k J_k^T,
// where r_k is the signed Sampson residual of row k, J_k its derivative with
// respect to the 7 step parameters (rotations of U and V, then sigma) and w_k
// = [r_k^2 < 1] / n (times the row weight when weighted).  Returns the
// squared norm of Jtr.
template <class RowF, class WF>
inline double lm_f_normal_equations(size_t n, RowF row, WF wt, bool weighted, const FactorizedF& FF,
                                    double Jtr[7], double JtJ[7][7]) {
  double Fm[9];
  FF.F(Fm);
  // column-major view: Fc[c] with c = 3*col + row
  double ds[9];  // d_sigma = U.col(1) V.col(1)^T, row-major
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ds[3 * i + j] = FF.U[3 * i + 1] * FF.V[3 * j + 1];
  auto Fe = [&](int i, int j) { return Fm[3 * i + j]; };
  auto D = [&](int i, int j) { return ds[3 * i + j]; };
  // dF_dparams (jacobian_impl.h:511-520), rows in column-major F order
  const double P[9][7] = {
      {0, Fe(2, 0), -Fe(1, 0), 0, Fe(0, 2), -Fe(0, 1), D(0, 0)},
      {-Fe(2, 0), 0, Fe(0, 0), 0, Fe(1, 2), -Fe(1, 1), D(1, 0)},
      {Fe(1, 0), -Fe(0, 0), 0, 0, Fe(2, 2), -Fe(2, 1), D(2, 0)},
      {0, Fe(2, 1), -Fe(1, 1), -Fe(0, 2), 0, Fe(0, 0), D(0, 1)},
      {-Fe(2, 1), 0, Fe(0, 1), -Fe(1, 2), 0, Fe(1, 0), D(1, 1)},
      {Fe(1, 1), -Fe(0, 1), 0, -Fe(2, 2), 0, Fe(2, 0), D(2, 1)},
      {0, Fe(2, 2), -Fe(1, 2), Fe(0, 1), -Fe(0, 0), 0, D(0, 2)},
      {-Fe(2, 2), 0, Fe(0, 2), Fe(1, 1), -Fe(1, 0), 0, D(1, 2)},
      {Fe(1, 2), -Fe(0, 2), 0, Fe(2, 1), -Fe(2, 0), 0, D(2, 2)}};
  // per-point weighted J (7) and residual term; reductions by sum256
  auto jac = [&](size_t k, double* J, double& wres) -> double {
    const double* p = row(k);
    const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
    const double f0 = Fe(0, 0) * x1 + Fe(0, 1) * y1 + Fe(0, 2);
    const double f1 = Fe(1, 0) * x1 + Fe(1, 1) * y1 + Fe(1, 2);
    const double f2 = Fe(2, 0) * x1 + Fe(2, 1) * y1 + Fe(2, 2);
    const double C = x2 * f0 + y2 * f1 + f2;
    const double J0 = Fe(0, 0) * x2 + Fe(1, 0) * y2 + Fe(2, 0);
    const double J1 = Fe(0, 1) * x2 + Fe(1, 1) * y2 + Fe(2, 1);
    const double J2 = f0, J3 = f1;
    const double nJ = std::sqrt(J0 * J0 + J1 * J1 + J2 * J2 + J3 * J3);
    const double inv = 1.0 / nJ;
    const double r = C * inv;
    double w = ((r * r) < 1.0 ? 1.0 : 0.0) / (double)n;
    if (weighted) w = wt(k) * w;
    if (w == 0.0) return 0.0;
    double dF[9] = {x1 * x2, x1 * y2, x1, y1 * x2, y1 * y2, y1, x2, y2, 1.0};
    const double s = C * inv * inv;
    dF[0] -= s * (J2 * x1 + J0 * x2);
    dF[1] -= s * (J3 * x1 + J0 * y2);
    dF[2] -= s * (J0);
    dF[3] -= s * (J2 * y1 + J1 * x2);
    dF[4] -= s * (J3 * y1 + J1 * y2);
    dF[5] -= s * (J1);
    dF[6] -= s * (J2);
    dF[7] -= s * (J3);
    for (int t = 0; t < 9; ++t) dF[t] *= inv;
    for (int c = 0; c < 7; ++c) {
      double v = 0.0;
      for (int t = 0; t < 9; ++t) v += dF[t] * P[t][c];
      J[c] = v;
    }
    wres = w * C * inv;
    return w;
  };
  // 7 Jtr entries then the 28 lower-triangle JtJ entries, row by row
  double acc[35];
  sum256_multi<35>(0, n, [&](size_t k, double* t) {
    double J[7], wres = 0.0;
    const double w = jac(k, J, wres);
    int o = 0;
    for (int i = 0; i < 7; ++i) t[o++] = w == 0.0 ? 0.0 : wres * J[i];
    for (int i = 0; i < 7; ++i)
      for (int j = 0; j <= i; ++j) t[o++] = w == 0.0 ? 0.0 : w * (J[i] * J[j]);
  }, acc);
  {
    int o = 0;
    for (int i = 0; i < 7; ++i) Jtr[i] = acc[o++];
    for (int i = 0; i < 7; ++i)
      for (int j = 0; j <= i; ++j) JtJ[i][j] = acc[o++];
  }
  double gn = 0.0;
  for (int i = 0; i < 7; ++i) gn += Jtr[i] * Jtr[i];
  return gn;
}

// trace (nullable): the cost before the first step, then the cost of every
// trial step; accepted (nullable): 1 per trial step that was kept.
template <class RowF, class WF>
inline void lm_refine_f(size_t n, RowF row, WF wt, bool weighted, double* Fio, int max_iterations = 25,
                        std::vector<double>* trace = nullptr, std::vector<char>* accepted = nullptr) {
  FactorizedF FF;
  FF.from(Fio);
  double lambda = 1e-3;
  double Fm[9];
  FF.F(Fm);
  double cost = lm_f_cost(n, row, wt, weighted, Fm);
  if (trace) trace->push_back(cost);
  double JtJ[7][7], Jtr[7];
  bool recompute = true;
  for (int iter = 0; iter < max_iterations; ++iter) {
    if (recompute) {
      const double gn = lm_f_normal_equations(n, row, wt, weighted, FF, Jtr, JtJ);
      if (std::sqrt(gn) < 1e-8) break;
    }
    for (int k = 0; k < 7; ++k) JtJ[k][k] += lambda;
    // LLT on the lower triangle, solve JtJ sol = -Jtr
    double L[7][7];
    bool ok = true;
    for (int j = 0; j < 7; ++j) {
      double d = JtJ[j][j];
      for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
      if (!(d > 0.0)) ok = false;
      L[j][j] = std::sqrt(d);
      for (int i = j + 1; i < 7; ++i) {
        double v = JtJ[i][j];
        for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
        L[i][j] = v / L[j][j];
      }
    }
    double y[7], sol[7];
    for (int i = 0; i < 7; ++i) {
      double v = -Jtr[i];
      for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
      y[i] = v / L[i][i];
    }
    for (int i = 6; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < 7; ++k) v -= L[k][i] * sol[k];
      sol[i] = v / L[i][i];
    }
    (void)ok;
    double sn = 0.0;
    for (int i = 0; i < 7; ++i) sn += sol[i] * sol[i];
    if (std::sqrt(sn) < 1e-8) break;
    FactorizedF Fn;
    double w1[3] = {sol[0], sol[1], sol[2]}, w2[3] = {sol[3], sol[4], sol[5]};
    const double th1 = std::sqrt(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
    const double th2 = std::sqrt(w2[0] * w2[0] + w2[1] * w2[1] + w2[2] * w2[2]);
    for (int i = 0; i < 3; ++i) {
      w1[i] /= th1;
      w2[i] /= th2;
    }
    double a1, b1, a2, b2;
    sin_cos(th1, a1, b1);
    sin_cos(th2, a2, b2);
    auto update = [](const double* w, double a, double b, const double* Uin, double* Uout) {
      const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
      double sw2[9], R[9];
      mat3_mul(sw, sw, sw2);
      for (int i = 0; i < 9; ++i) R[i] = a * sw[i] + (1 - b) * sw2[i];
      double RU[9];
      mat3_mul(R, Uin, RU);
      for (int i = 0; i < 9; ++i) Uout[i] = Uin[i] + RU[i];
    };
    update(w1, a1, b1, FF.U, Fn.U);
    update(w2, a2, b2, FF.V, Fn.V);
    Fn.sigma = FF.sigma + sol[6];
    double Fnm[9];
    Fn.F(Fnm);
    const double cost_new = lm_f_cost(n, row, wt, weighted, Fnm);
    if (trace) trace->push_back(cost_new);
    if (cost_new < cost) {
      if (accepted) accepted->push_back(1);
      FF = Fn;
      lambda /= 10;
      cost = cost_new;
      recompute = true;
    } else {
      if (accepted) accepted->push_back(0);
      for (int k = 0; k < 7; ++k) JtJ[k][k] -= lambda;
      lambda *= 10;
      recompute = false;
    }
  }
  FF.F(Fio);
}

// FundamentalMatrixBundleAdjustmentSolver::estimateModel
// (solver_fundamental_matrix_bundle_adjustment.h:114-178) on rows(0..n-1):
// 8-point (n >= 8) or 7-point initialisation, then LM per model.
template <class RowF, class WF>
inline bool f_bundle_adjustment(size_t n, RowF row, WF wt, bool weighted, std::vector<Model>& models) {
  if (n < 7) return false;
  if (n >= 8) {
    Model m;
    f8_solve(n, row, wt, weighted, m.m);
    models.push_back(m);
  } else {
    // 7 rows: the seven-point solver on rows 0..6 (weights do not enter)
    std::vector<double> buf(28);
    for (size_t i = 0; i < 7; ++i)
      for (int k = 0; k < 4; ++k) buf[4 * i + k] = row(i)[k];
    Points p7{buf.data(), 7, 4};
    const size_t s[7] = {0, 1, 2, 3, 4, 5, 6};
    f7_solve(p7, s, models);
  }
  for (auto& m : models) lm_refine_f(n, row, wt, weighted, m.m);
  return !models.empty();
}

class FundamentalEstimator {
 public:
  explicit FundamentalEstimator(uint64_t seed, double min_ratio = 0.5, bool degensac = true)
      : seed_(seed), min_ratio_(min_ratio), use_degensac_(degensac) {}
  static constexpr size_t sampleSize() { return 7; }
  static constexpr size_t nonMinimalSampleSize() { return 7; }
  static constexpr size_t maximumMinimalSolutions() { return 3; }
  static constexpr bool isWeightingApplicable() { return true; }
  size_t inlierLimit() const { return 7 * sampleSize(); }

  // fundamental_estimator.h:161-184
  bool estimateModel(const Points& pts, const size_t* sample, std::vector<Model>* models) const {
    f7_solve(pts, sample, *models);
    for (int k = (int)models->size() - 1; k >= 0; --k)
      if (!f_orientation_valid((*models)[k].m, pts, sample, 7)) models->erase(models->begin() + k);
    return !models->empty();
  }

  // fundamental_estimator.h:574-618
  bool estimateModelNonminimal(const Points& pts, const size_t* sample, size_t n, std::vector<Model>* models,
                               const double* weights) const {
    if (n < nonMinimalSampleSize()) return false;
    std::vector<double> np(4 * n);
    double T1[9], T2[9];
    normalize_points(pts, sample, n, np.data(), T1, T2);
    auto row = [&](size_t i) { return np.data() + 4 * i; };
    auto wt = [&](size_t i) { return weights[i]; };
    std::vector<Model> out;
    f_bundle_adjustment(n, row, wt, weights != nullptr, out);
    double T2t[9];
    mat3_transpose(T2, T2t);
    for (auto& m : out) {
      double tmp[9];
      mat3_mul(T2t, m.m, tmp);
      mat3_mul(tmp, T1, m.m);
      double s = 0.0;
      for (int i = 0; i < 9; ++i) s += m.m[i] * m.m[i];
      s = std::sqrt(s);
      for (int i = 0; i < 9; ++i) m.m[i] /= s;
      if (m.m[8] < 0)
        for (int i = 0; i < 9; ++i) m.m[i] = -m.m[i];
      models->push_back(m);
    }
    return true;
  }

  double squaredResidual(const double* pt, const Model& m) const { return f_squared_sampson(pt, m.m); }
  bool isValidSample(const Points&, const size_t*) const { return true; }

  // fundamental_estimator.h:287-338.  threshold is the truncated threshold.
  bool isValidModel(Model& model, const Points& pts, const std::vector<size_t>& inliers, const size_t* sample,
                    double threshold, bool& updated) const {
    updated = false;
    const size_t minimum = (size_t)std::max((double)sampleSize(), (double)inliers.size() * min_ratio_);
    const double sq = threshold * threshold;
    size_t cnt = 0;
    bool passed = false;
    for (size_t idx : inliers)
      if (f_squared_symmetric(pts.row(idx), model.m) < sq)
        if (++cnt >= minimum) {
          passed = true;
          break;
        }
    if (!passed) return false;
    if (use_degensac_) return degensac(model, pts, inliers, sample, threshold, updated);
    return true;
  }

  // fundamental_estimator.h:341-572.  The epipole is the left null vector of
  // F (the third left singular vector in the reference), normalised to
  // e[2] = 1.
  bool degensac(Model& model, const Points& pts, const std::vector<size_t>& inliers, const size_t* sample,
                double threshold, bool& updated) const;

 private:
  uint64_t seed_;
  double min_ratio_;
  bool use_degensac_;
};

// Plane-and-parallax minimal solver (solver_fundamental_matrix_plane_and_parallax.h)
// with the 8-point non-minimal solver: the estimator DEGENSAC's nested
// GC-RANSAC uses (fundamental_estimator.h:530-556).
class PlaneParallaxEstimator {
 public:
  explicit PlaneParallaxEstimator(const double* H) { std::memcpy(H_, H, sizeof(H_)); }
  static constexpr size_t sampleSize() { return 2; }
  static constexpr size_t nonMinimalSampleSize() { return 8; }
  static constexpr size_t maximumMinimalSolutions() { return 1; }
  static constexpr bool isWeightingApplicable() { return true; }
  size_t inlierLimit() const { return 7 * sampleSize(); }
  bool estimateModel(const Points& pts, const size_t* sample, std::vector<Model>* models) const {
    const double* p1 = pts.row(sample[0]);
    const double* p2 = pts.row(sample[1]);
    double a1[3], a2[3];
    for (int i = 0; i < 3; ++i) {
      a1[i] = H_[3 * i] * p1[0] + H_[3 * i + 1] * p1[1] + H_[3 * i + 2] * 1.0;
      a2[i] = H_[3 * i] * p2[0] + H_[3 * i + 1] * p2[1] + H_[3 * i + 2] * 1.0;
    }
    const double d1[3] = {p1[2], p1[3], 1.0}, d2[3] = {p2[2], p2[3], 1.0};
    auto cross = [](const double* u, const double* v, double* r) {
      r[0] = u[1] * v[2] - u[2] * v[1];
      r[1] = u[2] * v[0] - u[0] * v[2];
      r[2] = u[0] * v[1] - u[1] * v[0];
    };
    double l1[3], l2[3], e[3];
    cross(a1, d1, l1);
    cross(a2, d2, l2);
    cross(l1, l2, e);
    if (std::abs(e[2]) < std::numeric_limits<double>::epsilon()) return false;
    const double ex[9] = {0, -e[2], e[1], e[2], 0, -e[0], -e[1], e[0], 0};
    Model m;
    mat3_mul(ex, H_, m.m);
    models->push_back(m);
    // orientation filter of the estimator (fundamental_estimator.h:170-180)
    if (!f_orientation_valid(m.m, pts, sample, 2)) models->pop_back();
    return !models->empty();
  }
  // fundamental_estimator.h:574-618 with the plain 8-point solver
  bool estimateModelNonminimal(const Points& pts, const size_t* sample, size_t n, std::vector<Model>* models,
                               const double* weights) const {
    if (n < nonMinimalSampleSize()) return false;
    std::vector<double> np(4 * n);
    double T1[9], T2[9];
    normalize_points(pts, sample, n, np.data(), T1, T2);
    auto row = [&](size_t i) { return np.data() + 4 * i; };
    auto wt = [&](size_t i) { return weights[i]; };
    Model m;
    f8_solve(n, row, wt, weights != nullptr, m.m);
    double T2t[9], tmp[9];
    mat3_transpose(T2, T2t);
    mat3_mul(T2t, m.m, tmp);
    mat3_mul(tmp, T1, m.m);
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += m.m[i] * m.m[i];
    s = std::sqrt(s);
    for (int i = 0; i < 9; ++i) m.m[i] /= s;
    if (m.m[8] < 0)
      for (int i = 0; i < 9; ++i) m.m[i] = -m.m[i];
    models->push_back(m);
    return true;
  }
  double squaredResidual(const double* pt, const Model& m) const { return f_squared_sampson(pt, m.m); }
  bool isValidSample(const Points&, const size_t*) const { return true; }
  // minimum ratio 0: the symmetric-epipolar test needs max(2, 0) = 2 inliers
  bool isValidModel(Model& model, const Points& pts, const std::vector<size_t>& inliers, const size_t*,
                    double threshold, bool& updated) const {
    updated = false;
    const double sq = threshold * threshold;
    size_t cnt = 0;
    for (size_t idx : inliers)
      if (f_squared_symmetric(pts.row(idx), model.m) < sq)
        if (++cnt >= 2) return true;
    return false;
  }

 private:
  double H_[9];
};

inline bool FundamentalEstimator::degensac(Model& model, const Points& pts, const std::vector<size_t>& inliers,
                                           const size_t* sample, double threshold, bool& updated) const {
  updated = false;
  static const size_t triplets[15] = {0, 1, 2, 3, 4, 5, 0, 1, 6, 3, 4, 6, 2, 5, 6};
  const double* F = model.m;
  // left null vector of F: the largest of the three column cross products
  const double c0[3] = {F[0], F[3], F[6]}, c1[3] = {F[1], F[4], F[7]}, c2[3] = {F[2], F[5], F[8]};
  auto cross = [](const double* u, const double* v, double* r) {
    r[0] = u[1] * v[2] - u[2] * v[1];
    r[1] = u[2] * v[0] - u[0] * v[2];
    r[2] = u[0] * v[1] - u[1] * v[0];
  };
  double e01[3], e02[3], e12[3];
  cross(c0, c1, e01);
  cross(c0, c2, e02);
  cross(c1, c2, e12);
  auto n2 = [](const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; };
  const double* eb = e01;
  if (n2(e02) > n2(eb)) eb = e02;
  if (n2(e12) > n2(eb)) eb = e12;
  const double e[3] = {eb[0] / eb[2], eb[1] / eb[2], 1.0};
  const double ex[9] = {0, -e[2], e[1], e[2], 0, -e[0], -e[1], e[0], 0};
  double A[9];
  mat3_mul(ex, F, A);
  bool degenerate = false;
  double best_h[9];
  for (int t = 0; t < 5 && !degenerate; ++t) {
    const size_t i1 = sample[triplets[3 * t]], i2 = sample[triplets[3 * t + 1]], i3 = sample[triplets[3 * t + 2]];
    const double* P[3] = {pts.row(i1), pts.row(i2), pts.row(i3)};
    double b[3], M[9];
    for (int k = 0; k < 3; ++k) {
      const double x1[3] = {P[k][0], P[k][1], 1.0}, x2[3] = {P[k][2], P[k][3], 1.0};
      double ce[3], Ax[3], c2x[3];
      cross(x2, e, ce);
      for (int r = 0; r < 3; ++r) Ax[r] = A[3 * r] * x1[0] + A[3 * r + 1] * x1[1] + A[3 * r + 2] * x1[2];
      cross(x2, Ax, c2x);
      b[k] = (c2x[0] * ce[0] + c2x[1] * ce[1] + c2x[2] * ce[2]) / n2(ce);
      M[3 * k] = x1[0];
      M[3 * k + 1] = x1[1];
      M[3 * k + 2] = x1[2];
    }
    double Mi[9];
    mat3_inverse(M, Mi);
    double v[3];
    for (int r = 0; r < 3; ++r) v[r] = Mi[3 * r] * b[0] + Mi[3 * r + 1] * b[1] + Mi[3 * r + 2] * b[2];
    double Hm[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Hm[3 * r + c] = A[3 * r + c] - e[r] * v[c];
    size_t cnt = 3;
    for (int k = 0; k < 7; ++k) {
      const size_t idx = sample[k];
      if (idx == i1 || idx == i2 || idx == i3) continue;
      if (h_squared_residual(pts.row(idx), Hm) < 4.0) ++cnt;
    }
    if (cnt >= 5) {
      std::memcpy(best_h, Hm, sizeof(best_h));
      degenerate = true;
    }
  }
  if (!degenerate) return true;
  const bool trace = getenv("ORACLE_TRACE_DEGENSAC") != nullptr;
  if (trace) fprintf(stderr, "degensac: degenerate sample\n");
  std::vector<size_t> hin;
  for (size_t idx : inliers)
    if (h_squared_residual(pts.row(idx), best_h) < 4.0) hin.push_back(idx);
  if (trace)
    fprintf(stderr, "degensac: sample0 %zu model inliers %zu, H inliers %zu\n", sample[0], inliers.size(), hin.size());
  if (hin.size() < 4) return false;
  HomographyEstimator he;
  std::vector<Model> hs;
  he.estimateModelNonminimal(pts, hin.data(), hin.size(), &hs, nullptr);
  if (hs.size() != 1) return false;
  // nested GC-RANSAC with the plane-and-parallax solver (:530-556); its
  // sampler is seeded from the run's seed and the sample's first index
  PlaneParallaxEstimator pe(hs[0].m);
  GCRANSAC<PlaneParallaxEstimator> gc;
  gc.settings.threshold = threshold;
  gc.settings.spatial_coherence_weight = 0;
  gc.settings.confidence = 0.99;
  const uint32_t s = (uint32_t)((seed_ * 2654435761ull) ^ (uint64_t)sample[0]);
  UniformSampler sampler(s, pts.rows);
  Model m;
  gc.run(pts, pe, &sampler, &sampler, nullptr, m);
  if (trace)
    fprintf(stderr, "degensac: nested run %zu inliers, %zu iterations\n", gc.getRansacStatistics().inliers.size(),
            gc.getRansacStatistics().iteration_number);
  if (gc.getRansacStatistics().inliers.size() > inliers.size()) {
    updated = true;
    model = m;
  }
  return true;
}

// py/gcransac_python.cpp:1832-1887: weighted bundle adjustment on the final
// inliers in pixel coordinates, kept when its MSAC value beats the run's.
// The reference's 8-point initialiser reads the inlier-length weight vector
// by point index (undefined behaviour past its end); here it reads it by
// position in the sample, as the LM does.
inline void fundamental_post_bundle_adjustment(const Points& pts, const FundamentalEstimator& est, double threshold,
                                               Model& model, std::vector<size_t>& inliers, double& score) {
  const size_t n = inliers.size();
  if (n > 7) {
    const double tt = 3.0 / 2.0 * threshold;
    const double sq = tt * tt;
    std::vector<double> w(n, 0.0);
    for (size_t k = 0; k < n; ++k) {
      const double v = 1.0 - est.squaredResidual(pts.row(inliers[k]), model) / sq;
      w[k] = 0 > v ? 0 : v;  // the reference's MAX macro (NaN stays NaN)
    }
    std::vector<Model> models;
    auto row = [&](size_t k) { return pts.row(inliers[k]); };
    auto wt = [&](size_t k) { return w[k]; };
    f_bundle_adjustment(n, row, wt, true, models);
    std::vector<size_t> cur;
    for (auto& m : models) {
      const Score s = msac_score(pts, m, est, sq, &cur);
      if (score < s.value) {
        model = m;
        score = s.value;
        inliers.swap(cur);
      }
    }
  } else {
    Model I;
    I.m[0] = I.m[4] = I.m[8] = 1.0;
    model = I;
  }
}

}  // namespace orc
