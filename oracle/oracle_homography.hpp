// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
// Homography estimator restatement:
//   inc/estimators/homography_estimator.h:98-381 (RobustHomographyEstimator)
//   inc/estimators/solver_homography_four_point.h:109-264 (4-point solver)
#pragma once

#include "oracle_core.hpp"

namespace orc {

// 4-point DLT with h33 = 1 solved by inc/math_utils.h:45-87.
// solver_homography_four_point.h:109-190.  Returns false on NaN.
inline bool h4_minimal(const Points& pts, const size_t* sample, Model& out) {
  double a[8][9];
  int r = 0;
  for (int i = 0; i < 4; ++i) {
    const double* p = pts.row(sample[i]);
    const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
    const double weight = 1.0;
    const double mwx1 = -weight * x1, mwy1 = -weight * y1, wx2 = weight * x2, wy2 = weight * y2;
    a[r][0] = mwx1; a[r][1] = mwy1; a[r][2] = -weight; a[r][3] = 0; a[r][4] = 0; a[r][5] = 0;
    a[r][6] = wx2 * x1; a[r][7] = wx2 * y1; a[r][8] = -wx2;
    ++r;
    a[r][0] = 0; a[r][1] = 0; a[r][2] = 0; a[r][3] = mwx1; a[r][4] = mwy1; a[r][5] = -weight;
    a[r][6] = wy2 * x1; a[r][7] = wy2 * y1; a[r][8] = -wy2;
    ++r;
  }
  double h[8];
  gaussElimination<8>(a, h);
  for (int i = 0; i < 8; ++i)
    if (std::isnan(h[i])) return false;
  for (int i = 0; i < 8; ++i) out.m[i] = h[i];
  out.m[8] = 1.0;
  return true;
}

// Weighted least squares on (already normalised) points, sample == nullptr
// in the reference: row i uses point i and weight weights[i]
// (solver_homography_four_point.h:192-264).
inline void h_nonminimal_ls(const double* npts, size_t n, const double* weights, Model& out) {
  const int m = (int)(2 * n);
  std::vector<double> A((size_t)m * 8), b(m);
  for (size_t i = 0; i < n; ++i) {
    const double* p = npts + 4 * i;
    const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
    const double weight = weights ? weights[i] : 1.0;
    const double mwx1 = -weight * x1, mwy1 = -weight * y1, wx2 = weight * x2, wy2 = weight * y2;
    const int r0 = (int)(2 * i), r1 = r0 + 1;
    A[0 * m + r0] = mwx1; A[1 * m + r0] = mwy1; A[2 * m + r0] = -weight;
    A[3 * m + r0] = 0; A[4 * m + r0] = 0; A[5 * m + r0] = 0;
    A[6 * m + r0] = wx2 * x1; A[7 * m + r0] = wx2 * y1; b[r0] = -wx2;
    A[0 * m + r1] = 0; A[1 * m + r1] = 0; A[2 * m + r1] = 0;
    A[3 * m + r1] = mwx1; A[4 * m + r1] = mwy1; A[5 * m + r1] = -weight;
    A[6 * m + r1] = wy2 * x1; A[7 * m + r1] = wy2 * y1; b[r1] = -wy2;
  }
  double h[8];
  colpiv_qr_solve(A.data(), b.data(), m, 8, h);
  for (int i = 0; i < 8; ++i) out.m[i] = h[i];
  out.m[8] = 1.0;
}

// Hartley normalisation, homography_estimator.h:213-310 (identical code in
// fundamental_estimator.h:637-734).  Writes n x 4 normalised coordinates.
inline void normalize_points(const Points& pts, const size_t* sample, size_t n, double* out, double T1[9],
                             double T2[9]) {
  double ms0 = 0, ms1 = 0, md0 = 0, md1 = 0;
  for (size_t i = 0; i < n; ++i) {
    const double* d = pts.row(sample[i]);
    ms0 += d[0]; ms1 += d[1]; md0 += d[2]; md1 += d[3];
  }
  ms0 /= n; ms1 /= n; md0 /= n; md1 /= n;
  double as = 0.0, ad = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double* d = pts.row(sample[i]);
    const double dx1 = ms0 - d[0], dy1 = ms1 - d[1], dx2 = md0 - d[2], dy2 = md1 - d[3];
    as += std::sqrt(dx1 * dx1 + dy1 * dy1);
    ad += std::sqrt(dx2 * dx2 + dy2 * dy2);
  }
  as /= n;
  ad /= n;
  const double rs = M_SQRT2 / as, rd = M_SQRT2 / ad;
  for (size_t i = 0; i < n; ++i) {
    const double* d = pts.row(sample[i]);
    out[4 * i + 0] = (d[0] - ms0) * rs;
    out[4 * i + 1] = (d[1] - ms1) * rs;
    out[4 * i + 2] = (d[2] - md0) * rd;
    out[4 * i + 3] = (d[3] - md1) * rd;
  }
  const double t1[9] = {rs, 0, -rs * ms0, 0, rs, -rs * ms1, 0, 0, 1};
  const double t2[9] = {rd, 0, -rd * md0, 0, rd, -rd * md1, 0, 0, 1};
  std::memcpy(T1, t1, sizeof(t1));
  std::memcpy(T2, t2, sizeof(t2));
}

// homography_estimator.h:185-199 (left-to-right evaluation, no fused ops).
inline double h_squared_residual(const double* s, const double* h) {
  const double x1 = s[0], y1 = s[1], x2 = s[2], y2 = s[3];
  const double t1 = h[0] * x1 + h[1] * y1 + h[2];
  const double t2 = h[3] * x1 + h[4] * y1 + h[5];
  const double t3 = h[6] * x1 + h[7] * y1 + h[8];
  const double d1 = x2 - (t1 / t3);
  const double d2 = y2 - (t2 / t3);
  return d1 * d1 + d2 * d2;
}

// homography_estimator.h:346-381
inline bool h_is_valid_sample(const Points& pts, const size_t* sample) {
  const double *a = pts.row(sample[0]), *b = pts.row(sample[1]), *c = pts.row(sample[2]), *d = pts.row(sample[3]);
  auto cross = [](double* r, const double* v1, const double* v2) {
    r[0] = v1[1] - v2[1];
    r[1] = v2[0] - v1[0];
    r[2] = v1[0] * v2[1] - v1[1] * v2[0];
  };
  double p[3], q[3];
  cross(p, a, b);
  cross(q, a + 2, b + 2);
  if ((p[0] * c[0] + p[1] * c[1] + p[2]) * (q[0] * c[2] + q[1] * c[3] + q[2]) < 0) return false;
  if ((p[0] * d[0] + p[1] * d[1] + p[2]) * (q[0] * d[2] + q[1] * d[3] + q[2]) < 0) return false;
  cross(p, c, d);
  cross(q, c + 2, d + 2);
  if ((p[0] * a[0] + p[1] * a[1] + p[2]) * (q[0] * a[2] + q[1] * a[3] + q[2]) < 0) return false;
  if ((p[0] * b[0] + p[1] * b[1] + p[2]) * (q[0] * b[2] + q[1] * b[3] + q[2]) < 0) return false;
  return true;
}

class HomographyEstimator {
 public:
  static constexpr size_t sampleSize() { return 4; }
  static constexpr size_t nonMinimalSampleSize() { return 4; }
  static constexpr size_t maximumMinimalSolutions() { return 1; }
  static constexpr bool isWeightingApplicable() { return true; }
  size_t inlierLimit() const { return 7 * sampleSize(); }

  bool estimateModel(const Points& pts, const size_t* sample, std::vector<Model>* models) const {
    Model h;
    if (!h4_minimal(pts, sample, h)) return false;
    models->push_back(h);
    return true;
  }

  // homography_estimator.h:136-173
  bool estimateModelNonminimal(const Points& pts, const size_t* sample, size_t n, std::vector<Model>* models,
                               const double* weights) const {
    if (n < nonMinimalSampleSize()) return false;
    std::vector<double> np(4 * n);
    double T1[9], T2[9];
    normalize_points(pts, sample, n, np.data(), T1, T2);
    Model h;
    h_nonminimal_ls(np.data(), n, weights, h);
    double T2inv[9], tmp[9];
    mat3_inverse(T2, T2inv);
    mat3_mul(T2inv, h.m, tmp);
    mat3_mul(tmp, T1, h.m);
    models->push_back(h);
    return true;
  }

  double squaredResidual(const double* pt, const Model& m) const { return h_squared_residual(pt, m.m); }
  bool isValidSample(const Points& pts, const size_t* sample) const { return h_is_valid_sample(pts, sample); }

  // homography_estimator.h:326-342
  bool isValidModel(Model& model, const Points&, const std::vector<size_t>&, const size_t*, double,
                    bool& updated) const {
    updated = false;
    const double det = det_partial_pivot(model.m, 3);
    return !(std::abs(det) < 1e-2);
  }
};

}  // namespace orc
