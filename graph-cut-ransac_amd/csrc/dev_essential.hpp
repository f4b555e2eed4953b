// Device numerics of the essential-matrix path (gfx950).
//
// Thread-level routines (one thread runs each call): the five-point solver
// (Nister: null space, trace and determinant constraints, elimination,
// degree-10 polynomial, Sturm-sequence roots, back-substitution), the dense
// steps of the Stewenius non-minimal solver (column-pivoted QR, Gauss-Jordan,
// real Schur eigenvalues, null vectors) and the relative-pose helpers of the
// bundle adjustment.  They use only +, -, *, / and sqrt in a fixed order and
// are compiled without fused multiply-add, so results are reproducible
// bit for bit by a sequential host restatement of the same steps.
//
// Reference: inc/estimators/solver_essential_matrix_five_point_nister.h,
// solver_essential_matrix_five_point_stewenius.h,
// inc/maths/sturm_polynomial_solver.h, inc/relative_pose/{essential.cpp,
// bundle.cpp,bundle.h,jacobian_impl.h}.
#pragma once

#include "dev_common.hpp"
#include "dev_fundamental.hpp"

namespace gcr {

template <class T>
__device__ __forceinline__ void gswap(T& a, T& b) {
  T t = a;
  a = b;
  b = t;
}

// ---- monomials of x, y, z ----------------------------------------------------
// Nister order: degree 2 [x^2, xy, xz, x, y^2, yz, y, z^2, z, 1], degree 3
// [x^3, y^3, x^2y, xy^2, x^2z, x^2, y^2z, y^2, xyz, xy, xz^2, xz, x, yz^2,
// yz, y, z^3, z^2, z, 1]; Stewenius order: degree 2 [x^2, xy, y^2, xz, yz,
// z^2, x, y, z, 1], degree 3 [x^3, x^2y, xy^2, y^3, x^2z, xyz, y^2z, xz^2,
// yz^2, z^3, x^2, xy, y^2, xz, yz, z^2, x, y, z, 1].  Degree 1 is [x, y, z, 1].
// A monomial is encoded as 16 a + 4 b + c (exponents of x, y, z).
struct MonoOrder {
  int d2[10];
  int d3[20];
};
constexpr int mono_code(int a, int b, int c) { return 16 * a + 4 * b + c; }
constexpr int kDeg1Code[4] = {mono_code(1, 0, 0), mono_code(0, 1, 0), mono_code(0, 0, 1), mono_code(0, 0, 0)};
constexpr MonoOrder kNisterOrder = {
    {mono_code(2, 0, 0), mono_code(1, 1, 0), mono_code(1, 0, 1), mono_code(1, 0, 0), mono_code(0, 2, 0),
     mono_code(0, 1, 1), mono_code(0, 1, 0), mono_code(0, 0, 2), mono_code(0, 0, 1), mono_code(0, 0, 0)},
    {mono_code(3, 0, 0), mono_code(0, 3, 0), mono_code(2, 1, 0), mono_code(1, 2, 0), mono_code(2, 0, 1),
     mono_code(2, 0, 0), mono_code(0, 2, 1), mono_code(0, 2, 0), mono_code(1, 1, 1), mono_code(1, 1, 0),
     mono_code(1, 0, 2), mono_code(1, 0, 1), mono_code(1, 0, 0), mono_code(0, 1, 2), mono_code(0, 1, 1),
     mono_code(0, 1, 0), mono_code(0, 0, 3), mono_code(0, 0, 2), mono_code(0, 0, 1), mono_code(0, 0, 0)}};
constexpr MonoOrder kSteweniusOrder = {
    {mono_code(2, 0, 0), mono_code(1, 1, 0), mono_code(0, 2, 0), mono_code(1, 0, 1), mono_code(0, 1, 1),
     mono_code(0, 0, 2), mono_code(1, 0, 0), mono_code(0, 1, 0), mono_code(0, 0, 1), mono_code(0, 0, 0)},
    {mono_code(3, 0, 0), mono_code(2, 1, 0), mono_code(1, 2, 0), mono_code(0, 3, 0), mono_code(2, 0, 1),
     mono_code(1, 1, 1), mono_code(0, 2, 1), mono_code(1, 0, 2), mono_code(0, 1, 2), mono_code(0, 0, 3),
     mono_code(2, 0, 0), mono_code(1, 1, 0), mono_code(0, 2, 0), mono_code(1, 0, 1), mono_code(0, 1, 1),
     mono_code(0, 0, 2), mono_code(1, 0, 0), mono_code(0, 1, 0), mono_code(0, 0, 1), mono_code(0, 0, 0)}};

// Product index tables: out index of (degree-1 i) x (degree-1 j) and of
// (degree-2 i) x (degree-1 j) in each order (compile-time).
struct MulTable {
  int m11[16];
  int m21[40];
};
constexpr int find_code(const int* t, int n, int code) {
  for (int i = 0; i < n; ++i)
    if (t[i] == code) return i;
  return -1;
}
constexpr MulTable make_mul_table(const MonoOrder& o) {
  MulTable t{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t.m11[4 * i + j] = find_code(o.d2, 10, kDeg1Code[i] + kDeg1Code[j]);
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 4; ++j) t.m21[4 * i + j] = find_code(o.d3, 20, o.d2[i] + kDeg1Code[j]);
  return t;
}
__constant__ MulTable kNisterMul = make_mul_table(kNisterOrder);
__constant__ MulTable kSteweniusMul = make_mul_table(kSteweniusOrder);

// out += s * (p * q), products added in (i, j) order.
__device__ inline void mul11_acc(const MulTable& t, const double* p, const double* q, double s, double* out) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int k = t.m11[4 * i + j];
      out[k] = out[k] + s * (p[i] * q[j]);
    }
}
__device__ inline void mul21_acc(const MulTable& t, const double* p, const double* q, double s, double* out) {
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 4; ++j) {
      const int k = t.m21[4 * i + j];
      out[k] = out[k] + s * (p[i] * q[j]);
    }
}

// Cubic constraints on E = x N0 + y N1 + z N2 + N3 (ee[i][j]: coefficients
// of E(i, j)): rows 3 i + j = entry (i, j) of eet_scale E E^T E -
// 0.5 tr(eet_scale E E^T) E, row 9 = det E along the third row; C is 10 x 20.
__device__ inline void essential_constraints(const double (*ee)[3][4], const MulTable& t, double eet_scale,
                                             double* C) {
  for (int i = 0; i < 200; ++i) C[i] = 0.0;
  const int cof[3][4] = {{0, 1, 1, 2}, {0, 2, 1, 0}, {0, 0, 1, 1}};
  const int cofn[3][4] = {{0, 2, 1, 1}, {0, 0, 1, 2}, {0, 1, 1, 0}};
  for (int k = 0; k < 3; ++k) {
    double d[10];
    for (int i = 0; i < 10; ++i) d[i] = 0.0;
    mul11_acc(t, ee[cof[k][0]][cof[k][1]], ee[cof[k][2]][cof[k][3]], 1.0, d);
    mul11_acc(t, ee[cofn[k][0]][cofn[k][1]], ee[cofn[k][2]][cofn[k][3]], -1.0, d);
    mul21_acc(t, d, ee[2][k], 1.0, C + 20 * 9);
  }
  double eet[3][3][10];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      for (int k = 0; k < 10; ++k) eet[i][j][k] = 0.0;
      for (int k = 0; k < 3; ++k) mul11_acc(t, ee[i][k], ee[j][k], eet_scale, eet[i][j]);
    }
  double tr[10];
  for (int k = 0; k < 10; ++k) tr[k] = eet[0][0][k] + eet[1][1][k] + eet[2][2][k];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double* row = C + 20 * (3 * i + j);
      for (int k = 0; k < 3; ++k) mul21_acc(t, eet[i][k], ee[k][j], 1.0, row);
      mul21_acc(t, tr, ee[i][j], -0.5, row);
    }
}

// Coefficient row of one correspondence for vec(E) stacked column by column:
// row . vec(E) = x2^T E x1.
__device__ __forceinline__ void e_row(const double* p, double w, double* r) {
  const double x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
  const double wx0 = w * x0, wx1 = w * x1, wy0 = w * y0, wy1 = w * y1;
  r[0] = wx0 * x1;
  r[1] = wx0 * y1;
  r[2] = wx0;
  r[3] = wy0 * x1;
  r[4] = wy0 * y1;
  r[5] = wy0;
  r[6] = wx1;
  r[7] = wy1;
  r[8] = w;
}

struct RelPose {
  double q[4];  // unit quaternion (w, x, y, z)
  double t[3];
};

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ inline int null_space_rx9(double* a, int r, double* basis /* (9-r) x 9 */) {
  int col_of_row[9];
  int used[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int rank = 0;
  double maxp = 0.0;
  for (int s = 0; s < r; ++s) {
    int pr = -1, pc = -1;
    double best = 0.0;
    for (int i = s; i < r; ++i)
      for (int j = 0; j < 9; ++j)
        if (!used[j] && fabs(a[9 * i + j]) > best) {
          best = fabs(a[9 * i + j]);
          pr = i;
          pc = j;
        }
    if (s == 0) maxp = best;
    if (pr < 0 || best <= (double)r * 2.220446049250313e-16 * maxp) break;
    if (pr != s)
      for (int j = 0; j < 9; ++j) gswap(a[9 * s + j], a[9 * pr + j]);
    const double piv = a[9 * s + pc];
    for (int j = 0; j < 9; ++j) a[9 * s + j] = a[9 * s + j] / piv;
    for (int i = 0; i < r; ++i) {
      if (i == s) continue;
      const double f = a[9 * i + pc];
      if (f != 0.0)
        for (int j = 0; j < 9; ++j) a[9 * i + j] = a[9 * i + j] - f * a[9 * s + j];
    }
    used[pc] = 1;
    col_of_row[s] = pc;
    ++rank;
  }
  const int dim = 9 - rank;
  if (dim != 9 - r) return dim;
  int freec[9], nf = 0;
  for (int j = 0; j < 9; ++j)
    if (!used[j]) freec[nf++] = j;
  for (int k = 0; k < nf; ++k) {
    double* v = basis + 9 * k;
    for (int j = 0; j < 9; ++j) v[j] = 0.0;
    v[freec[k]] = 1.0;
    for (int s = 0; s < rank; ++s) v[col_of_row[s]] = -a[9 * s + freec[k]];
  }
  return dim;
}

__device__ inline void lu_partial_solve(double* A, double* B, int n, int m) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[n * k + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[n * i + k]) > best) {
        best = fabs(A[n * i + k]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < n; ++j) gswap(A[n * k + j], A[n * p + j]);
      for (int j = 0; j < m; ++j) gswap(B[m * k + j], B[m * p + j]);
    }
    const double piv = A[n * k + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = A[n * i + k] / piv;
      A[n * i + k] = f;
      for (int j = k + 1; j < n; ++j) A[n * i + j] = A[n * i + j] - f * A[n * k + j];
      for (int j = 0; j < m; ++j) B[m * i + j] = B[m * i + j] - f * B[m * k + j];
    }
  }
  for (int j = 0; j < m; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double v = B[m * i + j];
      for (int k = i + 1; k < n; ++k) v = v - A[n * i + k] * B[m * k + j];
      B[m * i + j] = v / A[n * i + i];
    }
}

__device__ inline void lu_full_solve(double* A, double* B, int n, int m) {
  int colp[16];
  for (int j = 0; j < n; ++j) colp[j] = j;
  for (int k = 0; k < n; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int i = k; i < n; ++i)
      for (int j = k; j < n; ++j)
        if (fabs(A[n * i + j]) > best) {
          best = fabs(A[n * i + j]);
          pr = i;
          pc = j;
        }
    if (pr != k) {
      for (int j = 0; j < n; ++j) gswap(A[n * k + j], A[n * pr + j]);
      for (int j = 0; j < m; ++j) gswap(B[m * k + j], B[m * pr + j]);
    }
    if (pc != k) {
      for (int i = 0; i < n; ++i) gswap(A[n * i + k], A[n * i + pc]);
      gswap(colp[k], colp[pc]);
    }
    const double piv = A[n * k + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = A[n * i + k] / piv;
      for (int j = k + 1; j < n; ++j) A[n * i + j] = A[n * i + j] - f * A[n * k + j];
      for (int j = 0; j < m; ++j) B[m * i + j] = B[m * i + j] - f * B[m * k + j];
    }
  }
  double X[16 * 16];
  for (int j = 0; j < m; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double v = B[m * i + j];
      for (int k = i + 1; k < n; ++k) v = v - A[n * i + k] * X[m * k + j];
      X[m * i + j] = v / A[n * i + i];
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) B[m * colp[i] + j] = X[m * i + j];
}

// Real eigenvalues of a real n x n matrix (n <= 10): Householder reduction to
// Hessenberg form, then Francis double-shift QR steps -- the published
// "hqr" method (Martin, Peters and Wilkinson, "The QR algorithm for real
// Hessenberg matrices", Numer. Math. 14, 1970), written here.
__device__ inline int real_eigenvalues(const double* M, int n, double* ev) {
  int pos[10];
  double a[10][10];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i][j] = M[n * i + j];
  // Hessenberg reduction
  for (int k = 0; k < n - 2; ++k) {
    double sigma = 0.0;
    for (int i = k + 2; i < n; ++i) sigma += a[i][k] * a[i][k];
    if (sigma == 0.0) continue;
    const double alpha = a[k + 1][k];
    const double nrm = sqrt(alpha * alpha + sigma);
    const double beta = alpha >= 0.0 ? -nrm : nrm;
    double v[10];
    v[k + 1] = alpha - beta;
    for (int i = k + 2; i < n; ++i) v[i] = a[i][k];
    double vv = 0.0;
    for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
    // A <- (I - 2 v v^T / vv) A (I - 2 v v^T / vv)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * a[i][j];
      s = 2.0 * s / vv;
      for (int i = k + 1; i < n; ++i) a[i][j] -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += a[i][j] * v[j];
      s = 2.0 * s / vv;
      for (int j = k + 1; j < n; ++j) a[i][j] -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i) a[i][k] = 0.0;
  }
  // shifted QR (hqr)
  int nr = 0;
  int nn = n - 1;
  double t = 0.0;
  double anorm = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = (i - 1 > 0 ? i - 1 : 0); j < n; ++j) anorm += fabs(a[i][j]);
  int its = 0;
  const double eps = 2.220446049250313e-16;
  while (nn >= 0) {
    int l;
    for (l = nn; l >= 1; --l) {
      const double s = fabs(a[l - 1][l - 1]) + fabs(a[l][l]);
      const double sc = s == 0.0 ? anorm : s;
      if (fabs(a[l][l - 1]) <= eps * sc) {
        a[l][l - 1] = 0.0;
        break;
      }
    }
    const double x = a[nn][nn];
    if (l == nn) {  // one root
      pos[nr] = nn;
      ev[nr++] = x + t;
      --nn;
      its = 0;
      continue;
    }
    const double y = a[nn - 1][nn - 1];
    const double w = a[nn][nn - 1] * a[nn - 1][nn];
    if (l == nn - 1) {  // two roots
      const double p = 0.5 * (y - x);
      const double q = p * p + w;
      const double z = sqrt(fabs(q));
      const double xx = x + t;
      if (q >= 0.0) {
        const double zz = p + (p >= 0.0 ? fabs(z) : -fabs(z));
        const double r1 = xx + zz;
        const double r2 = zz != 0.0 ? xx - w / zz : xx + zz;
        pos[nr] = nn - 1;
        ev[nr++] = r1;
        pos[nr] = nn;
        ev[nr++] = r2;
      }
      nn -= 2;
      its = 0;
      continue;
    }
    if (its == 30 * n) return -1;  // no convergence
    double xs = x, ys = y, ws = w;
    if (its == 10 || its == 20) {  // exceptional shift
      t += x;
      for (int i = 0; i <= nn; ++i) a[i][i] -= x;
      const double s = fabs(a[nn][nn - 1]) + fabs(a[nn - 1][nn - 2]);
      xs = ys = 0.75 * s;
      ws = -0.4375 * s * s;
    }
    ++its;
    int m;
    double p = 0, q = 0, r = 0, z;
    for (m = nn - 2; m >= l; --m) {
      z = a[m][m];
      const double rr = xs - z, ss = ys - z;
      p = (rr * ss - ws) / a[m + 1][m] + a[m][m + 1];
      q = a[m + 1][m + 1] - z - rr - ss;
      r = a[m + 2][m + 1];
      const double s = fabs(p) + fabs(q) + fabs(r);
      p /= s;
      q /= s;
      r /= s;
      if (m == l) break;
      const double u = fabs(a[m][m - 1]) * (fabs(q) + fabs(r));
      const double v = fabs(p) * (fabs(a[m - 1][m - 1]) + fabs(z) + fabs(a[m + 1][m + 1]));
      if (u <= eps * v) break;
    }
    for (int i = m + 2; i <= nn; ++i) {
      a[i][i - 2] = 0.0;
      if (i != m + 2) a[i][i - 3] = 0.0;
    }
    for (int k = m; k <= nn - 1; ++k) {
      double xx = 0.0;
      if (k != m) {
        p = a[k][k - 1];
        q = a[k + 1][k - 1];
        r = 0.0;
        if (k != nn - 1) r = a[k + 2][k - 1];
        xx = fabs(p) + fabs(q) + fabs(r);
        if (xx != 0.0) {
          p /= xx;
          q /= xx;
          r /= xx;
        }
      }
      const double sq = sqrt(p * p + q * q + r * r);
      const double s = p >= 0.0 ? sq : -sq;
      if (s != 0.0) {
        if (k == m) {
          if (l != m) a[k][k - 1] = -a[k][k - 1];
        } else
          a[k][k - 1] = -s * xx;
        p += s;
        const double xr = p / s, yr = q / s, zr = r / s;
        q /= p;
        r /= p;
        for (int j = k; j <= nn; ++j) {
          double pp = a[k][j] + q * a[k + 1][j];
          if (k != nn - 1) {
            pp += r * a[k + 2][j];
            a[k + 2][j] -= pp * zr;
          }
          a[k + 1][j] -= pp * yr;
          a[k][j] -= pp * xr;
        }
        const int mmin = nn < k + 3 ? nn : k + 3;
        for (int i = l; i <= mmin; ++i) {
          double pp = xr * a[i][k] + yr * a[i][k + 1];
          if (k != nn - 1) {
            pp += zr * a[i][k + 2];
            a[i][k + 2] -= pp * r;
          }
          a[i][k + 1] -= pp * q;
          a[i][k] -= pp;
        }
      }
    }
  }
  // report in diagonal order (top to bottom), as the real Schur form lists them
  for (int a1 = 1; a1 < nr; ++a1)
    for (int b1 = a1; b1 > 0 && pos[b1 - 1] > pos[b1]; --b1) {
      gswap(pos[b1 - 1], pos[b1]);
      gswap(ev[b1 - 1], ev[b1]);
    }
  return nr;
}

__device__ inline void solve3x2_fullpiv(const double* b0, const double* b1, const double* rhs, double* xz) {
  double a[3][2] = {{b0[0], b1[0]}, {b0[1], b1[1]}, {b0[2], b1[2]}};
  double r[3] = {rhs[0], rhs[1], rhs[2]};
  int col[2] = {0, 1};
  int rank = 0;
  double first = 0.0;
  for (int k = 0; k < 2; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int i = k; i < 3; ++i)
      for (int j = k; j < 2; ++j)
        if (fabs(a[i][j]) > best) {
          best = fabs(a[i][j]);
          pr = i;
          pc = j;
        }
    if (k == 0) first = best;
    if (best <= 2.0 * 2.220446049250313e-16 * first || best == 0.0) break;
    if (pr != k) {
      for (int j = 0; j < 2; ++j) gswap(a[k][j], a[pr][j]);
      gswap(r[k], r[pr]);
    }
    if (pc != k) {
      for (int i = 0; i < 3; ++i) gswap(a[i][k], a[i][pc]);
      gswap(col[k], col[pc]);
    }
    for (int i = k + 1; i < 3; ++i) {
      const double f = a[i][k] / a[k][k];
      for (int j = k + 1; j < 2; ++j) a[i][j] = a[i][j] - f * a[k][j];
      r[i] = r[i] - f * r[k];
    }
    ++rank;
  }
  double z[2] = {0.0, 0.0};
  for (int k = rank - 1; k >= 0; --k) {
    double v = r[k];
    for (int j = k + 1; j < rank; ++j) v = v - a[k][j] * z[j];
    z[k] = v / a[k][k];
  }
  xz[col[0]] = z[0];
  xz[col[1]] = z[1];
}

__device__ inline void null_vector_fullpiv(double* A, int n, double* v) {
  int col[16];
  for (int j = 0; j < n; ++j) col[j] = j;
  int rank = 0;
  double first = 0.0;
  for (int k = 0; k < n; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int i = k; i < n; ++i)
      for (int j = k; j < n; ++j)
        if (fabs(A[n * i + j]) > best) {
          best = fabs(A[n * i + j]);
          pr = i;
          pc = j;
        }
    if (k == 0) first = best;
    if (best <= (double)n * 2.220446049250313e-16 * first || best == 0.0) break;
    if (pr != k)
      for (int j = 0; j < n; ++j) gswap(A[n * k + j], A[n * pr + j]);
    if (pc != k) {
      for (int i = 0; i < n; ++i) gswap(A[n * i + k], A[n * i + pc]);
      gswap(col[k], col[pc]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[n * i + k] / A[n * k + k];
      for (int j = k + 1; j < n; ++j) A[n * i + j] = A[n * i + j] - f * A[n * k + j];
      A[n * i + k] = 0.0;
    }
    ++rank;
  }
  double z[16];
  for (int j = 0; j < n; ++j) z[j] = 0.0;
  const int fr = rank < n ? rank : n - 1;  // the free unknown (in pivoted order)
  z[fr] = 1.0;
  for (int k = fr - 1; k >= 0; --k) {
    double s = 0.0;
    for (int j = k + 1; j <= fr; ++j) s = s + A[n * k + j] * z[j];
    z[k] = -s / A[n * k + k];
  }
  for (int j = 0; j < n; ++j) v[col[j]] = z[j];
}

__device__ inline void frobenius_normalise_sign(double* m) {
  double z = 0.0;
  for (int i = 0; i < 9; ++i) z += m[i] * m[i];
  if (z > 0.0) {
    const double s = sqrt(z);
    for (int i = 0; i < 9; ++i) m[i] /= s;
  }
  if (m[8] < 0)
    for (int i = 0; i < 9; ++i) m[i] = -m[i];
}

__device__ inline void quat_to_R(const double* q, double* R) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x;
  const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R[0] = 1 - (tyy + tzz);
  R[1] = txy - twz;
  R[2] = txz + twy;
  R[3] = txy + twz;
  R[4] = 1 - (txx + tzz);
  R[5] = tyz - twx;
  R[6] = txz - twy;
  R[7] = tyz + twx;
  R[8] = 1 - (txx + tyy);
}

__device__ inline void R_to_quat(const double* R, double* q) {
  auto m = [&](int i, int j) { return R[3 * i + j]; };
  double w, v[3];
  double t = m(0, 0) + m(1, 1) + m(2, 2);
  if (t > 0) {
    t = sqrt(t + 1.0);
    w = 0.5 * t;
    t = 0.5 / t;
    v[0] = (m(2, 1) - m(1, 2)) * t;
    v[1] = (m(0, 2) - m(2, 0)) * t;
    v[2] = (m(1, 0) - m(0, 1)) * t;
  } else {
    int i = 0;
    if (m(1, 1) > m(0, 0)) i = 1;
    if (m(2, 2) > m(i, i)) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    t = sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
    v[i] = 0.5 * t;
    t = 0.5 / t;
    w = (m(k, j) - m(j, k)) * t;
    v[j] = (m(j, i) + m(i, j)) * t;
    v[k] = (m(k, i) + m(i, k)) * t;
  }
  const double nrm = sqrt(w * w + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  q[0] = w / nrm;
  q[1] = v[0] / nrm;
  q[2] = v[1] / nrm;
  q[3] = v[2] / nrm;
}

__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ inline void normalise3(double* a) {
  const double n = sqrt(dot3(a, a));
  for (int i = 0; i < 3; ++i) a[i] /= n;
}

__device__ inline void essential_from_motion(const RelPose& p, double* E) {
  double R[9];
  quat_to_R(p.q, R);
  const double tx[9] = {0.0, -p.t[2], p.t[1], p.t[2], 0.0, -p.t[0], -p.t[1], p.t[0], 0.0};
  mat3_mul(tx, R, E);
}

__device__ inline bool check_cheirality(const RelPose& p, const double* x1, const double* x2) {
  double R[9], Rx1[3];
  quat_to_R(p.q, R);
  for (int i = 0; i < 3; ++i) Rx1[i] = R[3 * i] * x1[0] + R[3 * i + 1] * x1[1] + R[3 * i + 2] * x1[2];
  const double a = -dot3(Rx1, x2);
  const double b1 = -dot3(Rx1, p.t);
  const double b2 = dot3(x2, p.t);
  const double l1 = b1 - a * b2;
  const double l2 = -a * b1 + b2;
  const double md = 0.0 * (1 - a * a);
  return l1 > md && l2 > md;
}

__device__ inline double relpose_r2(const double* E, const double* p, double* JC, double* Cout) {
  const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
  const double e0 = E[0] * x1 + E[1] * y1 + E[2];
  const double e1 = E[3] * x1 + E[4] * y1 + E[5];
  const double e2 = E[6] * x1 + E[7] * y1 + E[8];
  const double C = x2 * e0 + y2 * e1 + e2;
  const double g0 = E[0] * x2 + E[3] * y2 + E[6];
  const double g1 = E[1] * x2 + E[4] * y2 + E[7];
  if (JC) {
    JC[0] = g0;
    JC[1] = g1;
    JC[2] = e0;
    JC[3] = e1;
  }
  if (Cout) *Cout = C;
  return (C * C) / ((e0 * e0 + e1 * e1) + (g0 * g0 + g1 * g1));
}

// The four (R, t) of E from cross products of its columns (essential.cpp:25-80),
// kept when the correspondence (x1, x2) is in front of both cameras, in the
// order (R_a, t), (R_a, -t), (R_b, t), (R_b, -t).  Returns their number.
__device__ inline int motion_from_essential_dev(const double* Erm, const double* x1, const double* x2,
                                                RelPose* out) {
  double c[3][3];  // columns of E
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) c[j][i] = Erm[3 * i + j];
  double u12[3], u13[3], u23[3];
  cross3(c[0], c[1], u12);
  cross3(c[0], c[2], u13);
  cross3(c[1], c[2], u23);
  const double n12 = dot3(u12, u12), n13 = dot3(u13, u13), n23 = dot3(u23, u23);
  double UW[3][3];  // columns
  const double* col1;
  const double* u;
  double nu;
  if (n12 > n13) {
    if (n12 > n23) {
      col1 = c[0];
      u = u12;
      nu = n12;
    } else {
      col1 = c[1];
      u = u23;
      nu = n23;
    }
  } else {
    if (n13 > n23) {
      col1 = c[0];
      u = u13;
      nu = n13;
    } else {
      col1 = c[1];
      u = u23;
      nu = n23;
    }
  }
  {
    double v[3] = {col1[0], col1[1], col1[2]};
    normalise3(v);
    const double sq = sqrt(nu);
    for (int i = 0; i < 3; ++i) {
      UW[1][i] = v[i];
      UW[2][i] = u[i] / sq;
    }
  }
  double tmp[3];
  cross3(UW[2], UW[1], tmp);
  for (int i = 0; i < 3; ++i) UW[0][i] = -tmp[i];
  double Vt[3][3];
  for (int j = 0; j < 3; ++j) {
    Vt[0][j] = UW[1][0] * c[j][0] + UW[1][1] * c[j][1] + UW[1][2] * c[j][2];
    Vt[1][j] = -(UW[0][0] * c[j][0] + UW[0][1] * c[j][1] + UW[0][2] * c[j][2]);
  }
  normalise3(Vt[0]);
  const double d01 = dot3(Vt[0], Vt[1]);
  for (int j = 0; j < 3; ++j) Vt[1][j] -= d01 * Vt[0][j];
  normalise3(Vt[1]);
  cross3(Vt[0], Vt[1], Vt[2]);
  int cnt = 0;
  RelPose p;
  double R[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[3 * i + j] = UW[0][i] * Vt[0][j] + UW[1][i] * Vt[1][j] + UW[2][i] * Vt[2][j];
  R_to_quat(R, p.q);
  for (int i = 0; i < 3; ++i) p.t[i] = UW[2][i];
  if (check_cheirality(p, x1, x2)) out[cnt++] = p;
  for (int i = 0; i < 3; ++i) p.t[i] = -p.t[i];
  if (check_cheirality(p, x1, x2)) out[cnt++] = p;
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 3; ++i) UW[k][i] = -UW[k][i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[3 * i + j] = UW[0][i] * Vt[0][j] + UW[1][i] * Vt[1][j] + UW[2][i] * Vt[2][j];
  R_to_quat(R, p.q);
  if (check_cheirality(p, x1, x2)) out[cnt++] = p;
  for (int i = 0; i < 3; ++i) p.t[i] = -p.t[i];
  if (check_cheirality(p, x1, x2)) out[cnt++] = p;
  return cnt;
}

// ---- Sturm-sequence roots of a degree-10 polynomial ----------------------------
// The steps of this solver follow PoseLib's Sturm-sequence solver as the
// reference ships it (inc/maths/sturm_polynomial_solver.h), which carries
// this licence:
//   Copyright (c) 2020, Viktor Larsson.  All rights reserved.
//   Redistribution and use in source and binary forms, with or without
//   modification, are permitted provided that the following conditions are
//   met: (1) Redistributions of source code must retain the above copyright
//   notice, this list of conditions and the following disclaimer.
//   (2) Redistributions in binary form must reproduce the above copyright
//   notice, this list of conditions and the following disclaimer in the
//   documentation and/or other materials provided with the distribution.
//   (3) Neither the name of the copyright holder nor the names of its
//   contributors may be used to endorse or promote products derived from
//   this software without specific prior written permission.
//   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS
//   IS" AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO,
//   THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR
//   PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER BE LIABLE
//   FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL
//   DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR
//   SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER
//   CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT
//   LIABILITY, OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY
//   OUT OF THE USE OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF
//   SUCH DAMAGE.
//
// Monic polynomial and normalised derivative, the Sturm sequence by
// pseudo-division with -|leading| scaling, sign changes by the three-term
// recurrence, bisection isolation (depth <= 30, left half first: an explicit
// stack replaces recursion), Ridders' method to width 1e-3, then up to ten
// Newton steps (tolerance 1e-10).
struct DevSturm10 {
  static constexpr int N = 10;
  double f[2 * N + 1];
  double s[3 * N];

  __device__ double eval(const double* p, int deg, double x) const {
    double v = x + p[deg - 1];
    for (int i = deg - 2; i >= 0; --i) v = x * v + p[i];
    return v;
  }

  __device__ void build() {
    double buf[3 * N];
    double* f1 = buf;
    double* f2 = f1 + N + 1;
    double* f3 = f2 + N;
    for (int i = 0; i < 2 * N + 1; ++i) buf[i] = f[i];
    for (int i = 0; i < N - 1; ++i) {
      const double q1 = f1[N - i] * f2[N - 1 - i];
      const double q0 = f1[N - 1 - i] * f2[N - 1 - i] - f1[N - i] * f2[N - 2 - i];
      f3[0] = f1[0] - q0 * f2[0];
      for (int j = 1; j < N - 1 - i; ++j) f3[j] = f1[j] - q1 * f2[j - 1] - q0 * f2[j];
      const double c = -fabs(f3[N - 2 - i]);
      const double ci = 1.0 / c;
      for (int j = 0; j < N - 1 - i; ++j) f3[j] = f3[j] * ci;
      double* t = f1;
      f1 = f2;
      f2 = f3;
      f3 = t;
      s[3 * i] = q0;
      s[3 * i + 1] = q1;
      s[3 * i + 2] = c;
    }
    s[3 * N - 3] = f1[0];
    s[3 * N - 2] = f1[1];
    s[3 * N - 1] = f2[0];
  }

  __device__ int sign_changes(double x) const {
    double g[N + 1];
    g[N] = s[3 * N - 1];
    g[N - 1] = s[3 * N - 3] + x * s[3 * N - 2];
    for (int i = N - 2; i >= 0; --i) g[i] = (s[3 * i] + x * s[3 * i + 1]) * g[i + 1] + s[3 * i + 2] * g[i + 2];
    int cnt = 0;
    for (int i = 0; i < N; ++i)
      if ((g[i] < 0) != (g[i + 1] < 0)) ++cnt;
    return cnt;
  }

  __device__ void refine(double a, double b, double* roots, int& nr, double tol) const {
    double fa = eval(f, N, a), fb = eval(f, N, b);
    if (!((fa < 0) ^ (fb < 0))) return;
    for (int it = 0; it < 30; ++it) {
      if (fabs(a - b) < 1e-3) break;
      const double c = (a + b) * 0.5;
      const double fc = eval(f, N, c);
      const double sq = sqrt(fc * fc - fa * fb);
      if (!sq) break;
      const double d = (fa < fb) ? c + (a - c) * fc / sq : c + (c - a) * fc / sq;
      const double fd = eval(f, N, d);
      if (fd >= 0 ? (fc < 0) : (fc > 0)) {
        a = c;
        fa = fc;
        b = d;
        fb = fd;
      } else if (fd >= 0 ? (fa < 0) : (fa > 0)) {
        b = d;
        fb = fd;
      } else {
        a = d;
        fa = fd;
      }
    }
    double x = (a + b) * 0.5;
    for (int it = 0; it < 10; ++it) {
      const double fx = eval(f, N, x);
      if (fabs(fx) < tol) break;
      const double fpx = static_cast<double>(N) * eval(f + N + 1, N - 1, x);
      const double dx = fx / fpx;
      x = x - dx;
      if (fabs(dx) < tol) break;
    }
    roots[nr++] = x;
  }

  // coefficients ascending (11); returns the number of roots found
  __device__ int solve(const double* coeffs, double* roots, double tol = 1e-10) {
    if (coeffs[N] == 0.0) return 0;
    for (int i = 0; i <= N; ++i) f[i] = coeffs[i];
    const double ci = 1.0 / f[N];
    for (int i = 0; i < N; ++i) f[i] *= ci;
    f[N] = 1.0;
    for (int i = 0; i < N - 1; ++i) f[N + 1 + i] = f[i + 1] * ((i + 1) / static_cast<double>(N));
    f[2 * N] = 1.0;
    build();
    double mx = 0.0;
    for (int i = 0; i < N; ++i) mx = fmax(mx, fabs(f[i]));
    const double r0 = 1.0 + mx;
    const int sa0 = sign_changes(-r0), sb0 = sign_changes(r0);
    if (sa0 - sb0 == 0) return 0;
    int nr = 0;
    // depth-first, left half first
    // at most depth + 2 <= 32 pending intervals (one right sibling per level)
    constexpr int kStack = 40;
    double sta[kStack], stb[kStack];
    int ssa[kStack], ssb[kStack], sdep[kStack];
    int top = 0;
    sta[0] = -r0;
    stb[0] = r0;
    ssa[0] = sa0;
    ssb[0] = sb0;
    sdep[0] = 0;
    top = 1;
    while (top > 0) {
      --top;
      const double a = sta[top], b = stb[top];
      const int sa = ssa[top], sb = ssb[top], depth = sdep[top];
      if (depth > 30) continue;
      const int n = sa - sb;
      if (n > 1 && top + 2 <= kStack) {
        const double c = (a + b) * 0.5;
        const int sc = sign_changes(c);
        // push the right half, then the left half (popped first)
        sta[top] = c;
        stb[top] = b;
        ssa[top] = sc;
        ssb[top] = sb;
        sdep[top] = depth + 1;
        ++top;
        sta[top] = a;
        stb[top] = c;
        ssa[top] = sa;
        ssb[top] = sc;
        sdep[top] = depth + 1;
        ++top;
      } else if (n == 1) {
        refine(a, b, roots, nr, tol);
      }
    }
    return nr;
  }
};

}  // namespace gcr
