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
// A polynomial is a coefficient array over a monomial table.  The Nister
// solver orders its monomials as
//   degree 1: [x, y, z, 1]
//   degree 2: [x^2, xy, xz, x, y^2, yz, y, z^2, z, 1]
//   degree 3: [x^3, y^3, x^2y, xy^2, x^2z, x^2, y^2z, y^2, xyz, xy,
//              xz^2, xz, x, yz^2, yz, y, z^3, z^2, z, 1]
// (nister.h:100-199); the Stewenius solver as
//   degree 2: [x^2, xy, y^2, xz, yz, z^2, x, y, z, 1]
//   degree 3: [x^3, x^2y, xy^2, y^3, x^2z, xyz, y^2z, xz^2, yz^2, z^3,
//              x^2, xy, y^2, xz, yz, z^2, x, y, z, 1]
// (stewenius.h:234-283).  The column order decides the elimination, so each
// solver builds its constraint matrix over its own table.
struct Mono {
  int a, b, c;  // exponents of x, y, z
};
static const Mono kDeg1[4] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};
static const Mono kDeg2[10] = {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 0}, {0, 2, 0},
                               {0, 1, 1}, {0, 1, 0}, {0, 0, 2}, {0, 0, 1}, {0, 0, 0}};
static const Mono kDeg3[20] = {{3, 0, 0}, {0, 3, 0}, {2, 1, 0}, {1, 2, 0}, {2, 0, 1}, {2, 0, 0}, {0, 2, 1},
                               {0, 2, 0}, {1, 1, 1}, {1, 1, 0}, {1, 0, 2}, {1, 0, 1}, {1, 0, 0}, {0, 1, 2},
                               {0, 1, 1}, {0, 1, 0}, {0, 0, 3}, {0, 0, 2}, {0, 0, 1}, {0, 0, 0}};
static const Mono kSDeg2[10] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1},
                                {0, 0, 2}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};
static const Mono kSDeg3[20] = {{3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}, {2, 0, 1}, {1, 1, 1}, {0, 2, 1},
                                {1, 0, 2}, {0, 1, 2}, {0, 0, 3}, {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1},
                                {0, 1, 1}, {0, 0, 2}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

inline int mono_index(const Mono* t, int n, int a, int b, int c) {
  for (int i = 0; i < n; ++i)
    if (t[i].a == a && t[i].b == b && t[i].c == c) return i;
  return -1;
}

// out += s * (p * q): p over table tp (np terms), q over tq, out over tout;
// products are added in (i, j) order.
inline void poly_mul_acc(const Mono* tp, int np, const double* p, const Mono* tq, int nq, const double* q,
                         const Mono* tout, int nout, double s, double* out) {
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < nq; ++j) {
      const int k = mono_index(tout, nout, tp[i].a + tq[j].a, tp[i].b + tq[j].b, tp[i].c + tq[j].c);
      out[k] = out[k] + s * (p[i] * q[j]);
    }
}

// The ten cubic constraints on E = x N0 + y N1 + z N2 + N3 (ee[i][j] holds the
// [x, y, z, 1] coefficients of E(i, j)): rows 3 i + j are entry (i, j) of
// (E E^T - trace_scale * tr(E E^T) I) E scaled by eet_scale, row 9 is det(E)
// by cofactors along the third row.  C is 10 x 20 over t3 (t2 for E E^T).
inline void essential_constraints(const double ee[3][3][4], const Mono* t2, const Mono* t3, double eet_scale,
                                  double* C) {
  for (int i = 0; i < 200; ++i) C[i] = 0.0;
  // det(E) = E20 (E01 E12 - E02 E11) + E21 (E02 E10 - E00 E12) + E22 (E00 E11 - E01 E10)
  {
    const int cof[3][4] = {{0, 1, 1, 2}, {0, 2, 1, 0}, {0, 0, 1, 1}};  // (r, c) of the positive product
    const int cofn[3][4] = {{0, 2, 1, 1}, {0, 0, 1, 2}, {0, 1, 1, 0}};
    for (int k = 0; k < 3; ++k) {
      double d[10] = {0};
      poly_mul_acc(kDeg1, 4, ee[cof[k][0]][cof[k][1]], kDeg1, 4, ee[cof[k][2]][cof[k][3]], t2, 10, 1.0, d);
      poly_mul_acc(kDeg1, 4, ee[cofn[k][0]][cofn[k][1]], kDeg1, 4, ee[cofn[k][2]][cofn[k][3]], t2, 10, -1.0, d);
      poly_mul_acc(t2, 10, d, kDeg1, 4, ee[2][k], t3, 20, 1.0, C + 20 * 9);
    }
  }
  // EE^T (scaled), its trace, and the nine trace rows
  double eet[3][3][10];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      for (int k = 0; k < 10; ++k) eet[i][j][k] = 0.0;
      for (int k = 0; k < 3; ++k) poly_mul_acc(kDeg1, 4, ee[i][k], kDeg1, 4, ee[j][k], t2, 10, eet_scale, eet[i][j]);
    }
  double tr[10];
  for (int k = 0; k < 10; ++k) tr[k] = eet[0][0][k] + eet[1][1][k] + eet[2][2][k];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double* row = C + 20 * (3 * i + j);
      for (int k = 0; k < 3; ++k) poly_mul_acc(t2, 10, eet[i][k], kDeg1, 4, ee[k][j], t3, 20, 1.0, row);
      poly_mul_acc(t2, 10, tr, kDeg1, 4, ee[i][j], t3, 20, -0.5, row);
    }
}

// Coefficient row of one correspondence for vec(E) stacked column by column
// (nister.h:264-303, stewenius.h:137-176): row . vec(E) = x2^T E x1.
inline void e_row(const double* p, double w, double* r) {
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
// Hessenberg form, then Francis double-shift QR steps -- the "hqr" method of
// Martin, Peters and Wilkinson (Numer. Math. 14, 1970) -- (exceptional shifts at
// iterations 10 and 20, at most 30 per eigenvalue) until the matrix is quasi
// upper triangular.  1 x 1 blocks and 2 x 2 blocks with real eigenvalues give
// real eigenvalues; complex pairs are skipped.  Returns the count; the
// eigenvalues are listed by their position on the diagonal (top first).
inline int real_eigenvalues(const double* M, int n, double* ev) {
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
      const double z = std::sqrt(std::fabs(q));
      const double xx = x + t;
      if (q >= 0.0) {
        const double zz = p + (p >= 0.0 ? std::fabs(z) : -std::fabs(z));
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
  // report in diagonal order (top to bottom), as the real Schur form lists them
  for (int a1 = 1; a1 < nr; ++a1)
    for (int b1 = a1; b1 > 0 && pos[b1 - 1] > pos[b1]; --b1) {
      std::swap(pos[b1 - 1], pos[b1]);
      std::swap(ev[b1 - 1], ev[b1]);
    }
  return nr;
}

// ---- more dense helpers ----------------------------------------------------------
// Solve the 3 x 2 system [b0 b1] xz = rhs by Gaussian elimination with full
// pivoting (the role of FullPivLU::solve, nister.h:600): pivots are taken
// largest-first, a pivot below 2 * eps * (first pivot) ends the rank, and a
// non-pivot unknown is 0.
inline void solve3x2_fullpiv(const double* b0, const double* b1, const double* rhs, double* xz) {
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
        if (std::fabs(a[i][j]) > best) {
          best = std::fabs(a[i][j]);
          pr = i;
          pc = j;
        }
    if (k == 0) first = best;
    if (best <= 2.0 * std::numeric_limits<double>::epsilon() * first || best == 0.0) break;
    if (pr != k) {
      for (int j = 0; j < 2; ++j) std::swap(a[k][j], a[pr][j]);
      std::swap(r[k], r[pr]);
    }
    if (pc != k) {
      for (int i = 0; i < 3; ++i) std::swap(a[i][k], a[i][pc]);
      std::swap(col[k], col[pc]);
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

// A null vector of the n x n matrix A (rank n - 1 expected): Gaussian
// elimination with full pivoting; the last pivot-free unknown is set to 1.
inline void null_vector_fullpiv(double* A, int n, double* v) {
  int col[16];
  for (int j = 0; j < n; ++j) col[j] = j;
  int rank = 0;
  double first = 0.0;
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
    if (k == 0) first = best;
    if (best <= (double)n * std::numeric_limits<double>::epsilon() * first || best == 0.0) break;
    if (pr != k)
      for (int j = 0; j < n; ++j) std::swap(A[n * k + j], A[n * pr + j]);
    if (pc != k) {
      for (int i = 0; i < n; ++i) std::swap(A[n * i + k], A[n * i + pc]);
      std::swap(col[k], col[pc]);
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

// Q of the Householder QR with column pivoting of the n x n matrix A
// (row-major, n <= 9), M P = Q R (the role of ColPivHouseholderQR::matrixQ,
// stewenius.h:183-186).  Column norms use the sum64 order.
inline void colpiv_qr_Q(const double* Ain, int n, double* Q) {
  double A[81], hv[9][9], tau[9];
  for (int i = 0; i < n * n; ++i) A[i] = Ain[i];
  auto at = [&](int i, int j) -> double& { return A[n * i + j]; };
  double norms[9];
  for (int j = 0; j < n; ++j) norms[j] = sum64(0, (size_t)n, [&](size_t i) { return at((int)i, j) * at((int)i, j); });
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int j = k + 1; j < n; ++j)
      if (norms[j] > norms[p]) p = j;
    if (p != k) {
      for (int i = 0; i < n; ++i) std::swap(at(i, k), at(i, p));
      std::swap(norms[k], norms[p]);
    }
    const double alpha = at(k, k);
    const double sigma = sum64((size_t)k + 1, (size_t)n, [&](size_t i) { return at((int)i, k) * at((int)i, k); });
    tau[k] = 0.0;
    for (int i = 0; i < n; ++i) hv[k][i] = 0.0;
    hv[k][k] = 1.0;
    if (sigma != 0.0) {
      const double nrm = std::sqrt(alpha * alpha + sigma);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      tau[k] = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
      for (int i = k + 1; i < n; ++i) hv[k][i] = at(i, k) * scale;
      at(k, k) = beta;
      for (int i = k + 1; i < n; ++i) at(i, k) = 0.0;
      for (int j = k + 1; j < n; ++j) {
        const double w = at(k, j) + sum64((size_t)k + 1, (size_t)n, [&](size_t i) { return hv[k][i] * at((int)i, j); });
        const double tw = tau[k] * w;
        at(k, j) = at(k, j) - tw;
        for (int i = k + 1; i < n; ++i) at(i, j) = at(i, j) - tw * hv[k][i];
      }
    }
    for (int j = k + 1; j < n; ++j)
      norms[j] = sum64((size_t)k + 1, (size_t)n, [&](size_t i) { return at((int)i, j) * at((int)i, j); });
  }
  // Q = H0 H1 ... H(n-1): apply the reflectors to the identity, last first
  for (int i = 0; i < n * n; ++i) Q[i] = 0.0;
  for (int i = 0; i < n; ++i) Q[n * i + i] = 1.0;
  for (int k = n - 1; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    for (int j = 0; j < n; ++j) {
      double w = 0.0;
      for (int i = k; i < n; ++i) w += hv[k][i] * Q[n * i + j];
      const double tw = tau[k] * w;
      for (int i = k; i < n; ++i) Q[n * i + j] = Q[n * i + j] - tw * hv[k][i];
    }
  }
}

// ---- five-point minimal solver (nister.h:204-615) ---------------------------------
// On a sample of 5 (normalised) correspondences; up to 10 models.
inline bool nister_solve(const Points& pts, const size_t* sample, std::vector<Model>& models) {
  double a[5 * 9];
  for (int i = 0; i < 5; ++i) e_row(pts.row(sample[i]), 1.0, a + 9 * i);
  double N[4 * 9];  // row k: k-th kernel vector (vec(E) column by column)
  if (null_space_rx9(a, 5, N) != 4) return false;
  double ee[3][3][4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 4; ++k) ee[i][j][k] = N[9 * k + 3 * j + i];
  double C[200];
  essential_constraints(ee, kDeg2, kDeg3, 1.0, C);
  // Gauss-Jordan on the ten leading monomials: B = C[:, :10]^-1 C[:, 10:]
  double L[100], B[100];
  for (int r = 0; r < 10; ++r)
    for (int c = 0; c < 10; ++c) {
      L[10 * r + c] = C[20 * r + c];
      B[10 * r + c] = C[20 * r + 10 + c];
    }
  lu_partial_solve(L, B, 10, 10);
  // rows (x^2 z, x^2), (y^2 z, y^2), (x y z, x y) of B: <r1> - z <r2> leaves
  // x p(z) + y q(z) + r(z) = 0 over the columns [xz^2, xz, x, yz^2, yz, y,
  // z^3, z^2, z, 1]
  double A[3][13];
  for (int i = 0; i < 3; ++i) {
    const double* r1 = B + 10 * (4 + 2 * i);
    const double* r2 = B + 10 * (5 + 2 * i);
    A[i][0] = 0.0;
    for (int c = 0; c < 3; ++c) A[i][1 + c] = r1[c];
    for (int c = 0; c < 3; ++c) A[i][c] = A[i][c] - r2[c];
    A[i][4] = 0.0;
    for (int c = 0; c < 3; ++c) A[i][5 + c] = r1[3 + c];
    for (int c = 0; c < 3; ++c) A[i][4 + c] = A[i][4 + c] - r2[3 + c];
    A[i][8] = 0.0;
    for (int c = 0; c < 4; ++c) A[i][9 + c] = r1[6 + c];
    for (int c = 0; c < 4; ++c) A[i][8 + c] = A[i][8 + c] - r2[6 + c];
  }
  // det [p q r](z), degree 10, ascending coefficients
  double P[3][4], Qp[3][4], R[3][5];
  for (int i = 0; i < 3; ++i) {
    for (int d = 0; d < 4; ++d) {
      P[i][d] = A[i][3 - d];
      Qp[i][d] = A[i][7 - d];
    }
    for (int d = 0; d < 5; ++d) R[i][d] = A[i][12 - d];
  }
  auto mul = [](const double* u, int du, const double* v, int dv, double* out) {
    for (int k = 0; k <= du + dv; ++k) out[k] = 0.0;
    for (int i = 0; i <= du; ++i)
      for (int j = 0; j <= dv; ++j) out[i + j] = out[i + j] + u[i] * v[j];
  };
  double c[11];
  for (int k = 0; k < 11; ++k) c[k] = 0.0;
  {
    double t1[8], t2[8], m7[8], full[11];
    // p0 (q1 r2 - q2 r1)
    mul(Qp[1], 3, R[2], 4, t1);
    mul(Qp[2], 3, R[1], 4, t2);
    for (int k = 0; k < 8; ++k) m7[k] = t1[k] - t2[k];
    mul(P[0], 3, m7, 7, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] + full[k];
    // - q0 (p1 r2 - p2 r1)
    mul(P[1], 3, R[2], 4, t1);
    mul(P[2], 3, R[1], 4, t2);
    for (int k = 0; k < 8; ++k) m7[k] = t1[k] - t2[k];
    mul(Qp[0], 3, m7, 7, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] - full[k];
    // + r0 (p1 q2 - p2 q1)
    double s1[7], s2[7], m6[7];
    mul(P[1], 3, Qp[2], 3, s1);
    mul(P[2], 3, Qp[1], 3, s2);
    for (int k = 0; k < 7; ++k) m6[k] = s1[k] - s2[k];
    mul(R[0], 4, m6, 6, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] + full[k];
  }
  double roots[10];
  Sturm<10> st;
  const int nr = st.solve(c, roots);
  for (int s = 0; s < nr; ++s) {
    const double z = roots[s];
    const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
    double b0[3], b1[3], bb[3];
    for (int i = 0; i < 3; ++i) {
      b0[i] = A[i][0] * z3 + A[i][1] * z2 + A[i][2] * z + A[i][3];
      b1[i] = A[i][4] * z3 + A[i][5] * z2 + A[i][6] * z + A[i][7];
      bb[i] = A[i][8] * z4 + A[i][9] * z3 + A[i][10] * z2 + A[i][11] * z + A[i][12];
    }
    // 2 x 2 inverse on the first two rows; full-pivot solve of all three
    // when the third row disagrees (nister.h:590-601)
    const double det = b0[0] * b1[1] - b1[0] * b0[1];
    const double id = 1.0 / det;
    double xz[2];
    xz[0] = (b1[1] * id) * bb[0] + (-b1[0] * id) * bb[1];
    xz[1] = (-b0[1] * id) * bb[0] + (b0[0] * id) * bb[1];
    if (std::fabs(b0[2] * xz[0] + b1[2] * xz[1] - bb[2]) > 1e-6) solve3x2_fullpiv(b0, b1, bb, xz);
    const double x = -xz[0], y = -xz[1];
    double e[9];
    const double sc = 1.0 / std::sqrt(x * x + y * y + z * z + 1.0);
    for (int k = 0; k < 9; ++k) e[k] = (N[k] * x + N[9 + k] * y + N[18 + k] * z + N[27 + k]) * sc;
    Model m;
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) m.m[3 * r + cc] = e[3 * cc + r];
    models.push_back(m);
  }
  return !models.empty();
}

// ---- five-point non-minimal solver (stewenius.h:114-231) --------------------------
// weight_of(i) gives the weight of sample position i (nullptr-free callers
// pass 1.0).  Up to 10 models, not normalised.
template <class WF>
inline bool stewenius_solve(const Points& pts, const size_t* sample, size_t n, WF weight_of,
                            std::vector<Model>& models) {
  double NS[9][4];  // columns: the four null-space vectors (x, y, z, 1)
  if (n == 5) {
    double a[5 * 9];
    for (size_t i = 0; i < 5; ++i) e_row(pts.row(sample[i]), weight_of(i), a + 9 * i);
    double N[4 * 9];
    if (null_space_rx9(a, 5, N) != 4) return false;
    for (int r = 0; r < 9; ++r)
      for (int k = 0; k < 4; ++k) NS[r][k] = N[9 * k + r];
  } else {
    // A^T A of the n x 9 system in the sum64 order, then Q of its
    // column-pivoted QR; the last four columns span the null space
    std::vector<double> rows(9 * n);
    for (size_t i = 0; i < n; ++i) e_row(pts.row(sample[i]), weight_of(i), rows.data() + 9 * i);
    double M[81];
    double acc[45];
    sum64_multi<45>(0, n, [&](size_t i, double* out) {
      const double* r = rows.data() + 9 * i;
      int o = 0;
      for (int a1 = 0; a1 < 9; ++a1)
        for (int b1 = a1; b1 < 9; ++b1) out[o++] = r[a1] * r[b1];
    }, acc);
    int o = 0;
    for (int a1 = 0; a1 < 9; ++a1)
      for (int b1 = a1; b1 < 9; ++b1) {
        M[9 * a1 + b1] = acc[o];
        M[9 * b1 + a1] = acc[o];
        ++o;
      }
    double Q[81];
    colpiv_qr_Q(M, 9, Q);
    for (int r = 0; r < 9; ++r)
      for (int k = 0; k < 4; ++k) NS[r][k] = Q[9 * r + 5 + k];
  }
  double ee[3][3][4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 4; ++k) ee[i][j][k] = NS[3 * j + i][k];
  double C[200];
  essential_constraints(ee, kSDeg2, kSDeg3, 2.0, C);
  double L[100], B[100];
  for (int r = 0; r < 10; ++r)
    for (int c = 0; c < 10; ++c) {
      L[10 * r + c] = C[20 * r + c];
      B[10 * r + c] = C[20 * r + 10 + c];
    }
  lu_full_solve(L, B, 10, 10);
  // action matrix of multiplication by x on [x^2, xy, y^2, xz, yz, z^2, x, y, z, 1]
  double Act[100];
  for (int i = 0; i < 100; ++i) Act[i] = 0.0;
  const int from[6] = {0, 1, 2, 4, 5, 7};
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 10; ++c) Act[10 * r + c] = B[10 * from[r] + c];
  Act[10 * 6 + 0] = -1.0;
  Act[10 * 7 + 1] = -1.0;
  Act[10 * 8 + 3] = -1.0;
  Act[10 * 9 + 6] = -1.0;
  double ev[10];
  const int ne = real_eigenvalues(Act, 10, ev);
  for (int q = 0; q < ne; ++q) {
    double Ms[100], v[10];
    for (int i = 0; i < 100; ++i) Ms[i] = Act[i];
    for (int i = 0; i < 10; ++i) Ms[11 * i] = Ms[11 * i] - ev[q];
    null_vector_fullpiv(Ms, 10, v);
    double e[9];
    for (int r = 0; r < 9; ++r) e[r] = NS[r][0] * v[6] + NS[r][1] * v[7] + NS[r][2] * v[8] + NS[r][3] * v[9];
    Model m;
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) m.m[3 * r + cc] = e[3 * cc + r];
    models.push_back(m);
  }
  return !models.empty();
}

inline void frobenius_normalise_sign(Model& m) {
  double z = 0.0;
  for (int i = 0; i < 9; ++i) z += m.m[i] * m.m[i];
  if (z > 0.0) {
    const double s = std::sqrt(z);
    for (int i = 0; i < 9; ++i) m.m[i] /= s;
  }
  if (m.m[8] < 0)
    for (int i = 0; i < 9; ++i) m.m[i] = -m.m[i];
}

// ---- relative pose (inc/relative_pose/) ----------------------------------------
// Camera pose: unit quaternion q = (w, x, y, z) and translation t.
struct RelPose {
  double q[4];
  double t[3];
};

// Rotation matrix of a unit quaternion (row-major).
inline void quat_to_R(const double* q, double* R) {
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

// Unit quaternion of a rotation matrix (Shepperd's method: the largest of
// w, x, y, z is taken from the diagonal), then normalised (bundle.h:35-41).
inline void R_to_quat(const double* R, double* q) {
  auto m = [&](int i, int j) { return R[3 * i + j]; };
  double w, v[3];
  double t = m(0, 0) + m(1, 1) + m(2, 2);
  if (t > 0) {
    t = std::sqrt(t + 1.0);
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
    t = std::sqrt(m(i, i) - m(j, j) - m(k, k) + 1.0);
    v[i] = 0.5 * t;
    t = 0.5 / t;
    w = (m(k, j) - m(j, k)) * t;
    v[j] = (m(j, i) + m(i, j)) * t;
    v[k] = (m(k, i) + m(i, k)) * t;
  }
  const double nrm = std::sqrt(w * w + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  q[0] = w / nrm;
  q[1] = v[0] / nrm;
  q[2] = v[1] / nrm;
  q[3] = v[2] / nrm;
}

inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void normalise3(double* a) {
  const double n = std::sqrt(dot3(a, a));
  for (int i = 0; i < 3; ++i) a[i] /= n;
}

// essential.cpp:5-11: E = [t]_x R.
inline void essential_from_motion(const RelPose& p, double* E) {
  double R[9];
  quat_to_R(p.q, R);
  const double tx[9] = {0.0, -p.t[2], p.t[1], p.t[2], 0.0, -p.t[0], -p.t[1], p.t[0], 0.0};
  mat3_mul(tx, R, E);
}

// essential.cpp:13-23 (min_depth 0).
inline bool check_cheirality(const RelPose& p, const double* x1, const double* x2) {
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

// essential.cpp:25-80: the four (R, t) of E from cross products of its
// columns, kept when the correspondence (x1, x2) is in front of both cameras.
inline void motion_from_essential(const double* Erm, const double* x1, const double* x2, std::vector<RelPose>& out) {
  double c[3][3];  // columns of E
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) c[j][i] = Erm[3 * i + j];
  double u12[3], u13[3], u23[3];
  cross3(c[0], c[1], u12);
  cross3(c[0], c[2], u13);
  cross3(c[1], c[2], u23);
  const double n12 = dot3(u12, u12), n13 = dot3(u13, u13), n23 = dot3(u23, u23);
  double UW[3][3];  // columns
  auto set = [&](const double* col1, const double* u, double nu) {
    double v[3] = {col1[0], col1[1], col1[2]};
    normalise3(v);
    const double s = std::sqrt(nu);
    for (int i = 0; i < 3; ++i) {
      UW[1][i] = v[i];
      UW[2][i] = u[i] / s;
    }
  };
  if (n12 > n13) {
    if (n12 > n23) set(c[0], u12, n12);
    else set(c[1], u23, n23);
  } else {
    if (n13 > n23) set(c[0], u13, n13);
    else set(c[1], u23, n23);
  }
  double tmp[3];
  cross3(UW[2], UW[1], tmp);
  for (int i = 0; i < 3; ++i) UW[0][i] = -tmp[i];
  // Vt rows: UW1^T E, -UW0^T E, then Gram-Schmidt and a cross product
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
  auto rot = [&](double* R) {  // R = UW * Vt (row-major)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[3 * i + j] = UW[0][i] * Vt[0][j] + UW[1][i] * Vt[1][j] + UW[2][i] * Vt[2][j];
  };
  RelPose p;
  double R[9];
  rot(R);
  R_to_quat(R, p.q);
  for (int i = 0; i < 3; ++i) p.t[i] = UW[2][i];
  if (check_cheirality(p, x1, x2)) out.push_back(p);
  for (int i = 0; i < 3; ++i) p.t[i] = -p.t[i];
  if (check_cheirality(p, x1, x2)) out.push_back(p);
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < 3; ++i) UW[k][i] = -UW[k][i];
  rot(R);
  R_to_quat(R, p.q);
  if (check_cheirality(p, x1, x2)) out.push_back(p);
  for (int i = 0; i < 3; ++i) p.t[i] = -p.t[i];
  if (check_cheirality(p, x1, x2)) out.push_back(p);
}

// Sampson residual pieces of one correspondence for E (jacobian_impl.h:321-346).
inline double relpose_r2(const double* E, const double* p, double* JC, double* Cout) {
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

// refine_relpose / lm_5dof_impl (bundle.cpp:176-251,412-452) with the
// truncated loss of scale 1 (robust_loss.h:19-31), at most 25 iterations.
// Point k of the sample has weight w(k); sums over points use sum64.
template <class RowF, class WF>
inline void refine_relpose(size_t n, RowF row, WF wt, RelPose& pose, int max_iterations = 25) {
  auto cost_of = [&](const RelPose& p) {
    double E[9];
    essential_from_motion(p, E);
    return sum64(0, n, [&](size_t k) {
      const double r2 = relpose_r2(E, row(k), nullptr, nullptr);
      return wt(k) * (r2 < 1.0 ? r2 : 1.0);
    });
  };
  double lambda = 1e-3;
  double cost = cost_of(pose);
  double JtJ[5][5], Jtr[5], tb[3][2];
  bool recompute = true;
  for (int iter = 0; iter < max_iterations; ++iter) {
    if (recompute) {
      // tangent basis of t (jacobian_impl.h:350-367)
      const double* t = pose.t;
      double b0[3];
      const double ux[3] = {1, 0, 0}, uy[3] = {0, 1, 0}, uz[3] = {0, 0, 1};
      if (std::fabs(t[0]) < std::fabs(t[1])) {
        if (std::fabs(t[0]) < std::fabs(t[2])) cross3(t, ux, b0);
        else cross3(t, uz, b0);
      } else {
        if (std::fabs(t[1]) < std::fabs(t[2])) cross3(t, uy, b0);
        else cross3(t, uz, b0);
      }
      normalise3(b0);
      double b1[3];
      cross3(b0, t, b1);
      normalise3(b1);
      for (int i = 0; i < 3; ++i) {
        tb[i][0] = b0[i];
        tb[i][1] = b1[i];
      }
      double E[9], R[9];
      essential_from_motion(pose, E);
      quat_to_R(pose.q, R);
      // d vec(E) (column by column) / d(rotation, translation)
      double dR[9][3], dt[9][2];
      auto Ecol = [&](int j, int i) { return E[3 * i + j]; };
      for (int i = 0; i < 3; ++i) {
        dR[i][0] = 0.0;
        dR[i][1] = -Ecol(2, i);
        dR[i][2] = Ecol(1, i);
        dR[3 + i][0] = Ecol(2, i);
        dR[3 + i][1] = 0.0;
        dR[3 + i][2] = -Ecol(0, i);
        dR[6 + i][0] = -Ecol(1, i);
        dR[6 + i][1] = Ecol(0, i);
        dR[6 + i][2] = 0.0;
      }
      for (int col = 0; col < 3; ++col) {
        const double rc[3] = {R[col], R[3 + col], R[6 + col]};
        double c0[3], c1[3];
        cross3(b0, rc, c0);
        cross3(b1, rc, c1);
        for (int i = 0; i < 3; ++i) {
          dt[3 * col + i][0] = c0[i];
          dt[3 * col + i][1] = c1[i];
        }
      }
      double acc[20];
      sum64_multi<20>(0, n, [&](size_t k, double* out) {
        for (int e = 0; e < 20; ++e) out[e] = 0.0;
        const double* p = row(k);
        double JC[4], C;
        relpose_r2(E, p, JC, &C);
        const double nJ = std::sqrt(JC[0] * JC[0] + JC[1] * JC[1] + JC[2] * JC[2] + JC[3] * JC[3]);
        const double inv = 1.0 / nJ;
        const double r = C * inv;
        double w = ((r * r) < 1.0 ? 1.0 : 0.0) / (double)n;
        w = wt(k) * w;
        if (w == 0.0) return;
        const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
        double dF[9] = {x1 * x2, x1 * y2, x1, y1 * x2, y1 * y2, y1, x2, y2, 1.0};
        const double s = C * inv * inv;
        dF[0] -= s * (JC[2] * x1 + JC[0] * x2);
        dF[1] -= s * (JC[3] * x1 + JC[0] * y2);
        dF[2] -= s * (JC[0]);
        dF[3] -= s * (JC[2] * y1 + JC[1] * x2);
        dF[4] -= s * (JC[3] * y1 + JC[1] * y2);
        dF[5] -= s * (JC[1]);
        dF[6] -= s * (JC[2]);
        dF[7] -= s * (JC[3]);
        for (int q = 0; q < 9; ++q) dF[q] *= inv;
        double J[5];
        for (int cc = 0; cc < 3; ++cc) {
          double v = 0.0;
          for (int q = 0; q < 9; ++q) v += dF[q] * dR[q][cc];
          J[cc] = v;
        }
        for (int cc = 0; cc < 2; ++cc) {
          double v = 0.0;
          for (int q = 0; q < 9; ++q) v += dF[q] * dt[q][cc];
          J[3 + cc] = v;
        }
        const double wres = w * C * inv;
        int o = 0;
        for (int i = 0; i < 5; ++i) out[o++] = wres * J[i];
        for (int i = 0; i < 5; ++i)
          for (int j = 0; j <= i; ++j) out[o++] = w * (J[i] * J[j]);
      }, acc);
      int o = 0;
      for (int i = 0; i < 5; ++i) Jtr[i] = acc[o++];
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j <= i; ++j) JtJ[i][j] = acc[o++];
      double gn = 0.0;
      for (int i = 0; i < 5; ++i) gn += Jtr[i] * Jtr[i];
      if (std::sqrt(gn) < 1e-8) break;
    }
    for (int k = 0; k < 5; ++k) JtJ[k][k] += lambda;
    // Cholesky of the lower triangle, then -JtJ^-1 Jtr
    double L[5][5], y[5], sol[5];
    for (int j = 0; j < 5; ++j) {
      double d = JtJ[j][j];
      for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
      L[j][j] = std::sqrt(d);
      for (int i = j + 1; i < 5; ++i) {
        double v = JtJ[i][j];
        for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
        L[i][j] = v / L[j][j];
      }
    }
    for (int i = 0; i < 5; ++i) {
      double v = -Jtr[i];
      for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
      y[i] = v / L[i][i];
    }
    for (int i = 4; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < 5; ++k) v -= L[k][i] * sol[k];
      sol[i] = v / L[i][i];
    }
    double sn = 0.0;
    for (int i = 0; i < 5; ++i) sn += sol[i] * sol[i];
    if (std::sqrt(sn) < 1e-8) break;
    double w[3] = {sol[0], sol[1], sol[2]};
    const double theta = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int i = 0; i < 3; ++i) w[i] /= theta;
    double a, b;
    sin_cos(theta, a, b);
    const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double sw2[9], Rm[9], R[9], RR[9], Rn[9];
    mat3_mul(sw, sw, sw2);
    for (int i = 0; i < 9; ++i) Rm[i] = a * sw[i] + (1 - b) * sw2[i];
    quat_to_R(pose.q, R);
    mat3_mul(R, Rm, RR);
    for (int i = 0; i < 9; ++i) Rn[i] = R[i] + RR[i];
    RelPose np;
    R_to_quat(Rn, np.q);
    for (int i = 0; i < 3; ++i) np.t[i] = pose.t[i] + (tb[i][0] * sol[3] + tb[i][1] * sol[4]);
    const double cost_new = cost_of(np);
    if (cost_new < cost) {
      pose = np;
      lambda /= 10;
      cost = cost_new;
      recompute = true;
    } else {
      for (int k = 0; k < 5; ++k) JtJ[k][k] -= lambda;
      lambda *= 10;
      recompute = false;
    }
  }
}

// ---- estimator (essential_estimator.h) ----------------------------------------------
class EssentialEstimator {
 public:
  explicit EssentialEstimator(double min_ratio = 0.1) : min_ratio_(min_ratio) {}
  static constexpr size_t sampleSize() { return 5; }
  static constexpr size_t nonMinimalSampleSize() { return 5; }
  static constexpr size_t maximumMinimalSolutions() { return 10; }
  static constexpr bool isWeightingApplicable() { return true; }
  size_t inlierLimit() const { return 7 * sampleSize(); }

  // essential_estimator.h:152-175
  bool estimateModel(const Points& pts, const size_t* sample, std::vector<Model>* models) const {
    if (!nister_solve(pts, sample, *models)) return false;
    for (int k = (int)models->size() - 1; k >= 0; --k)
      if (!f_orientation_valid((*models)[k].m, pts, sample, 5)) models->erase(models->begin() + k);
    return !models->empty();
  }

  // essential_estimator.h:308-335; weights are indexed by point index
  // (stewenius.h:150-158 reads weights_[sample_[i]]).
  bool estimateModelNonminimal(const Points& pts, const size_t* sample, size_t n, std::vector<Model>* models,
                               const double* weights) const {
    if (n < nonMinimalSampleSize()) return false;
    std::vector<Model> out;
    auto wf = [&](size_t i) { return weights ? weights[sample[i]] : 1.0; };
    if (!stewenius_solve(pts, sample, n, wf, out)) return false;
    for (auto& m : out) {
      frobenius_normalise_sign(m);
      models->push_back(m);
    }
    return true;
  }

  double squaredResidual(const double* pt, const Model& m) const { return f_squared_sampson(pt, m.m); }
  bool isValidSample(const Points&, const size_t*) const { return true; }

  // essential_estimator.h:279-305.  threshold is the truncated threshold.
  bool isValidModel(Model& model, const Points& pts, const std::vector<size_t>& inliers, const size_t*,
                    double threshold, bool& updated) const {
    updated = false;
    const double need = (double)inliers.size() * min_ratio_;
    const size_t minimum = (size_t)((double)sampleSize() > need ? (double)sampleSize() : need);
    const double sq = threshold * threshold;
    size_t cnt = 0;
    for (size_t idx : inliers)
      if (f_squared_symmetric(pts.row(idx), model.m) < sq)
        if (++cnt >= minimum) return true;
    return false;
  }

 private:
  double min_ratio_;
};

// py/gcransac_python.cpp:2766-2823: weighted pose bundle adjustment of the
// inliers in normalised coordinates, kept when its MSAC value is higher.
// trunc_thr is 3 / 2 * threshold / threshold_normaliser as the driver forms it.
inline void essential_post_bundle_adjustment(const Points& pts, const EssentialEstimator& est, double trunc_thr,
                                             Model& model, std::vector<size_t>& inliers, double& score) {
  const size_t n = inliers.size();
  if (n <= 5) {
    Model I;
    I.m[0] = I.m[4] = I.m[8] = 1.0;
    model = I;
    return;
  }
  const double sq = trunc_thr * trunc_thr;
  std::vector<double> w(n, 0.0);
  for (size_t k = 0; k < n; ++k) {
    const double v = 1.0 - est.squaredResidual(pts.row(inliers[k]), model) / sq;
    w[k] = 0 > v ? 0 : v;
  }
  // EssentialMatrixBundleAdjustmentSolver::estimateModel (…bundle_adjustment.h:114-188):
  // five-point initialisation on the weighted inliers (weights by inlier
  // position: the reference indexes this inlier-sized array by point index,
  // which is undefined; position is the defined reading), the poses of each
  // model, each pose refined.
  std::vector<Model> init;
  stewenius_solve(pts, inliers.data(), n, [&](size_t i) { return w[i]; }, init);
  const double* p0 = pts.row(inliers[0]);
  double x1[3] = {p0[0], p0[1], 1.0}, x2[3] = {p0[2], p0[3], 1.0};
  normalise3(x1);
  normalise3(x2);
  std::vector<Model> models;
  for (auto& m0 : init) {
    std::vector<RelPose> poses;
    motion_from_essential(m0.m, x1, x2, poses);
    for (auto& pose : poses) {
      refine_relpose(n, [&](size_t k) { return pts.row(inliers[k]); }, [&](size_t k) { return w[k]; }, pose);
      Model m;
      essential_from_motion(pose, m.m);
      models.push_back(m);
    }
  }
  std::vector<size_t> cur;
  for (auto& m : models) {
    const Score s = msac_score(pts, m, est, sq, &cur);
    if (score < s.value) {
      model = m;
      score = s.value;
      inliers.swap(cur);
    }
  }
}

// K^-1 [x y 1]^T for both images (utils.h:65-110); rows x1 y1 x2 y2.
inline void normalise_correspondences(const double* corr, size_t n, const double* K1, const double* K2, double* out) {
  double i1[9], i2[9];
  mat3_inverse(K1, i1);
  mat3_inverse(K2, i2);
  for (size_t r = 0; r < n; ++r) {
    const double* p = corr + 4 * r;
    out[4 * r + 0] = i1[0] * p[0] + i1[1] * p[1] + i1[2] * 1.0;
    out[4 * r + 1] = i1[3] * p[0] + i1[4] * p[1] + i1[5] * 1.0;
    out[4 * r + 2] = i2[0] * p[2] + i2[1] * p[3] + i2[2] * 1.0;
    out[4 * r + 3] = i2[3] * p[2] + i2[4] * p[3] + i2[5] * 1.0;
  }
}

}  // namespace orc
