// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Essential-matrix estimator restatement:
//   inc/estimators/essential_estimator.h (EssentialMatrixEstimator)
//   inc/estimators/solver_essential_matrix_five_point_nister.h (minimal)
//   inc/estimators/solver_essential_matrix_five_point_stewenius.h (non-minimal)
//   inc/maths/sturm_polynomial_solver.h (Sturm-sequence root bracketing)
//   inc/utils.h:65-110 (normalizeCorrespondences)
//   py/gcransac_python.cpp:2519-2855 (findEssentialMatrix_)
//
// Polynomial algebra is written here from the mathematics: E = x N0 + y N1 +
// z N2 + N3 over a null-space basis, the cubic constraints det(E) = 0 and
// 2 E E^T E - tr(E E^T) E = 0, Gauss-Jordan elimination of the ten leading
// monomials, and the degree-10 polynomial det[p q r](z) of the remaining
// 3 x 3 system.  Eigen's decompositions inside the reference (FullPivLU
// kernel / solve, PartialPivLU, ColPivHouseholderQR, EigenSolver) are
// restated with textbook algorithms; parity with the reference binary is
// unpinned there.  The Sturm solver follows the reference's steps and is
// pinned against the reference's own header (tests/golden).
#pragma once

#include "oracle_core.hpp"
#include "oracle_fundamental.hpp"
#include "oracle_gcransac.hpp"

namespace orc {

// ---- small polynomial helpers (monomials of x, y, z) -------------------------
// degree 1: [x, y, z, 1]
// degree 2: [x^2, xy, xz, x, y^2, yz, y, z^2, z, 1]
// degree 3: [x^3, y^3, x^2y, xy^2, x^2z, x^2, y^2z, y^2, xyz, xy,
//            xz^2, xz, x, yz^2, yz, y, z^3, z^2, z, 1]
struct Mono {
  int a, b, c;  // exponents of x, y, z
};
static const Mono kDeg2[10] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 0}, {0, 2, 0},
                               {0, 1, 1}, {0, 1, 0}, {0, 0, 2}, {0, 0, 1}, {0, 0, 0}};
static const Mono kDeg3[20] = {{3, 0, 0}, {0, 3, 0}, {2, 1, 0}, {1, 2, 0}, {2, 0, 1}, {2, 0, 0}, {0, 2, 1},
                               {0, 2, 0}, {1, 1, 1}, {1, 1, 0}, {1, 0, 2}, {1, 0, 1}, {1, 0, 0}, {0, 1, 2},
                               {0, 1, 1}, {0, 1, 0}, {0, 0, 3}, {0, 0, 2}, {0, 0, 1}, {0, 0, 0}};
static const Mono kDeg1[4] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

inline int deg2_index(int a, int b, int c) {
  for (int i = 0; i < 10; ++i)
    if (kDeg2[i].a == a && kDeg2[i].b == b && kDeg2[i].c == c) return i;
  return -1;
}
inline int deg3_index(int a, int b, int c) {
  for (int i = 0; i < 20; ++i)
    if (kDeg3[i].a == a && kDeg3[i].b == b && kDeg3[i].c == c) return i;
  return -1;
}

// out (deg 2) = p * q (deg 1 each); terms added in (i, j) order
inline void poly11(const double* p, const double* q, double* out) {
  for (int k = 0; k < 10; ++k) out[k] = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int k = deg2_index(kDeg1[i].a + kDeg1[j].a, kDeg1[i].b + kDeg1[j].b, kDeg1[i].c + kDeg1[j].c);
      out[k] = out[k] + p[i] * q[j];
    }
}
// out (deg 3) += s * p (deg 2) * q (deg 1)
inline void poly21_acc(const double* p, const double* q, double s, double* out) {
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 4; ++j) {
      const int k = deg3_index(kDeg2[i].a + kDeg1[j].a, kDeg2[i].b + kDeg1[j].b, kDeg2[i].c + kDeg1[j].c);
      out[k] = out[k] + s * (p[i] * q[j]);
    }
}

// ---- Gauss-Jordan null space with full pivoting (r x 9, kernel dim 9 - r) ----
// Pivot: largest |a_ij| over rows >= step and unused columns (first in
// row-major order); rank cut-off r * eps * first pivot.  Basis vector k sets
// free column k to 1 and the other free columns to 0.
inline int null_space_rx9(double* a, int r, double* basis /* (9-r) x 9 */) {
  int col_of_row[9];
  int used[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int rank = 0;
  double maxp = 0.0;
  for (int s = 0; s < r; ++s) {
    int pr = -1, pc = -1;
    double best = 0.0;
    for (int i = s; i < r; ++i)
      for (int j = 0; j < 9; ++j)
        if (!used[j] && std::fabs(a[9 * i + j]) > best) {
          best = std::fabs(a[9 * i + j]);
          pr = i;
          pc = j;
        }
    if (s == 0) maxp = best;
    if (pr < 0 || best <= (double)r * std::numeric_limits<double>::epsilon() * maxp) break;
    if (pr != s)
      for (int j = 0; j < 9; ++j) std::swap(a[9 * s + j], a[9 * pr + j]);
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

// ---- Sturm-sequence root isolation (inc/maths/sturm_polynomial_solver.h) ----
// The steps of the reference's solver: monic polynomial plus normalised
// derivative, Sturm sequence by pseudo-division with -|leading| scaling,
// sign-change counts by the three-term recurrence, bisection isolation (depth
// <= 30), then Ridders' method down to width 1e-3 and up to ten Newton steps.
template <int N>
struct Sturm {
  double f[2 * N + 1];  // monic polynomial (f[N] = 1) then derivative / N
  double s[3 * N];      // recurrence coefficients

  double eval(const double* p, int deg, double x) const {
    double v = x + p[deg - 1];
    for (int i = deg - 2; i >= 0; --i) v = x * v + p[i];
    return v;
  }

  void build() {
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
      const double c = -std::abs(f3[N - 2 - i]);
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

  int sign_changes(double x) const {
    double g[N + 1];
    g[N] = s[3 * N - 1];
    g[N - 1] = s[3 * N - 3] + x * s[3 * N - 2];
    for (int i = N - 2; i >= 0; --i) g[i] = (s[3 * i] + x * s[3 * i + 1]) * g[i + 1] + s[3 * i + 2] * g[i + 2];
    int cnt = 0;
    for (int i = 0; i < N; ++i)
      if ((g[i] < 0) != (g[i + 1] < 0)) ++cnt;
    return cnt;
  }

  void refine(double a, double b, double* roots, int& nr, double tol) const {
    double fa = eval(f, N, a), fb = eval(f, N, b);
    if (!((fa < 0) ^ (fb < 0))) return;
    for (int it = 0; it < 30; ++it) {
      if (std::abs(a - b) < 1e-3) break;
      const double c = (a + b) * 0.5;
      const double fc = eval(f, N, c);
      const double sq = std::sqrt(fc * fc - fa * fb);
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
      if (std::abs(fx) < tol) break;
      const double fpx = static_cast<double>(N) * eval(f + N + 1, N - 1, x);
      const double dx = fx / fpx;
      x = x - dx;
      if (std::abs(dx) < tol) break;
    }
    roots[nr++] = x;
  }

  void isolate(double a, double b, int sa, int sb, double* roots, int& nr, double tol, int depth) const {
    if (depth > 30) return;
    const int n = sa - sb;
    if (n > 1) {
      const double c = (a + b) * 0.5;
      const int sc = sign_changes(c);
      isolate(a, c, sa, sc, roots, nr, tol, depth + 1);
      isolate(c, b, sc, sb, roots, nr, tol, depth + 1);
    } else if (n == 1) {
      refine(a, b, roots, nr, tol);
    }
  }

  // coeffs ascending, degree N; returns the number of roots found
  int solve(const double* coeffs, double* roots, double tol = 1e-10) {
    if (coeffs[N] == 0.0) return 0;
    for (int i = 0; i <= N; ++i) f[i] = coeffs[i];
    const double ci = 1.0 / f[N];
    for (int i = 0; i < N; ++i) f[i] *= ci;
    f[N] = 1.0;
    for (int i = 0; i < N - 1; ++i) f[N + 1 + i] = f[i + 1] * ((i + 1) / static_cast<double>(N));
    f[2 * N] = 1.0;
    build();
    double mx = 0.0;
    for (int i = 0; i < N; ++i) mx = std::max(mx, std::abs(f[i]));
    const double r0 = 1.0 + mx;
    const int sa = sign_changes(-r0), sb = sign_changes(r0);
    if (sa - sb == 0) return 0;
    int nr = 0;
    isolate(-r0, r0, sa, sb, roots, nr, tol, 0);
    return nr;
  }
};

// ---- dense helpers -------------------------------------------------------------
// X = A^-1 B for A n x n (row-major), B n x m, by LU with partial pivoting
// (max |a_ik| in the column, first on ties).
inline void lu_partial_solve(double* A, double* B, int n, int m) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[n * k + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[n * i + k]) > best) {
        best = std::fabs(A[n * i + k]);
        p = i;
      }
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[n * k + j], A[n * p + j]);
      for (int j = 0; j < m; ++j) std::swap(B[m * k + j], B[m * p + j]);
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

// X = A^-1 B by Gaussian elimination with full pivoting (the role of
// FullPivLU::solve for a square system).
inline void lu_full_solve(double* A, double* B, int n, int m) {
  int colp[16];
  for (int j = 0; j < n; ++j) colp[j] = j;
  for (int k = 0; k < n; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int i = k; i < n; ++i)
      for (int j = k; j < n; ++j)
        if (std::fabs(A[n * i + j]) > best) {
          best = std::fabs(A[n * i + j]);
          pr = i;
          pc = j;
        }
    if (pr != k) {
      for (int j = 0; j < n; ++j) std::swap(A[n * k + j], A[n * pr + j]);
      for (int j = 0; j < m; ++j) std::swap(B[m * k + j], B[m * pr + j]);
    }
    if (pc != k) {
      for (int i = 0; i < n; ++i) std::swap(A[n * i + k], A[n * i + pc]);
      std::swap(colp[k], colp[pc]);
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
// Hessenberg form, then Francis double-shift QR steps (exceptional shifts at
// iterations 10 and 20, at most 30 per eigenvalue) until the matrix is quasi
// upper triangular.  1 x 1 blocks and 2 x 2 blocks with real eigenvalues give
// real eigenvalues; complex pairs are skipped.  Returns the count, in the
// order the blocks deflate (last row first).
inline int real_eigenvalues(const double* M, int n, double* ev) {
  double a[10][10];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[i][j] = M[n * i + j];
  // Hessenberg reduction
  for (int k = 0; k < n - 2; ++k) {
    double sigma = 0.0;
    for (int i = k + 2; i < n; ++i) sigma += a[i][k] * a[i][k];
    if (sigma == 0.0) continue;
    const double alpha = a[k + 1][k];
    const double nrm = std::sqrt(alpha * alpha + sigma);
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
    for (int j = std::max(i - 1, 0); j < n; ++j) anorm += std::fabs(a[i][j]);
  int its = 0;
  const double eps = std::numeric_limits<double>::epsilon();
  while (nn >= 0) {
    int l;
    for (l = nn; l >= 1; --l) {
      const double s = std::fabs(a[l - 1][l - 1]) + std::fabs(a[l][l]);
      const double sc = s == 0.0 ? anorm : s;
      if (std::fabs(a[l][l - 1]) <= eps * sc) {
        a[l][l - 1] = 0.0;
        break;
      }
    }
    const double x = a[nn][nn];
    if (l == nn) {  // one root
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
      const double z = std::sqrt(std::fabs(q));
      const double xx = x + t;
      if (q >= 0.0) {
        const double zz = p + (p >= 0.0 ? std::fabs(z) : -std::fabs(z));
        const double r1 = xx + zz;
        const double r2 = zz != 0.0 ? xx - w / zz : xx + zz;
        ev[nr++] = r1;
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
      const double s = std::fabs(a[nn][nn - 1]) + std::fabs(a[nn - 1][nn - 2]);
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
      const double s = std::fabs(p) + std::fabs(q) + std::fabs(r);
      p /= s;
      q /= s;
      r /= s;
      if (m == l) break;
      const double u = std::fabs(a[m][m - 1]) * (std::fabs(q) + std::fabs(r));
      const double v = std::fabs(p) * (std::fabs(a[m - 1][m - 1]) + std::fabs(z) + std::fabs(a[m + 1][m + 1]));
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
        xx = std::fabs(p) + std::fabs(q) + std::fabs(r);
        if (xx != 0.0) {
          p /= xx;
          q /= xx;
          r /= xx;
        }
      }
      const double sq = std::sqrt(p * p + q * q + r * r);
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
  return nr;
}

// Essential-matrix driver (py/gcransac_python.cpp:2519-2855).  The five-point
// estimator above is being assembled; until it is complete the driver
// reports "unsupported" (-2) rather than a wrong answer.
inline int find_essential(const double*, size_t, const double*, const double*, int, int, int, int, double, double,
                          double, size_t, size_t, int, double, size_t, uint64_t, Model&, std::vector<size_t>&,
                          Statistics&, double&) {
  return -2;
}

}  // namespace orc
