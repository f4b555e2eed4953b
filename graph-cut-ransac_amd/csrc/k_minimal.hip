// Minimal solvers, one candidate sample per lane.
//
// Homography: oriented-sample test (homography_estimator.h:346-381), 8x9 DLT
// with h33 = 1 solved by the reference's pivoting Gauss elimination
// (solver_homography_four_point.h:109-190, math_utils.h:45-87), NaN check,
// and the |det H| >= 1e-2 validity test (homography_estimator.h:326-342).
// Loops over the matrix are fully unrolled so the 8x9 system stays in VGPRs;
// row swaps are selects, never dynamic indexing.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

__device__ __forceinline__ void h_cross(double* r, const double* v1, const double* v2) {
  r[0] = v1[1] - v2[1];
  r[1] = v2[0] - v1[0];
  r[2] = v1[0] * v2[1] - v1[1] * v2[0];
}

__device__ __forceinline__ bool h_valid_sample(const double* a, const double* b, const double* c, const double* d) {
  double p[3], q[3];
  h_cross(p, a, b);
  h_cross(q, a + 2, b + 2);
  if ((p[0] * c[0] + p[1] * c[1] + p[2]) * (q[0] * c[2] + q[1] * c[3] + q[2]) < 0) return false;
  if ((p[0] * d[0] + p[1] * d[1] + p[2]) * (q[0] * d[2] + q[1] * d[3] + q[2]) < 0) return false;
  h_cross(p, c, d);
  h_cross(q, c + 2, d + 2);
  if ((p[0] * a[0] + p[1] * a[1] + p[2]) * (q[0] * a[2] + q[1] * a[3] + q[2]) < 0) return false;
  if ((p[0] * b[0] + p[1] * b[1] + p[2]) * (q[0] * b[2] + q[1] * b[3] + q[2]) < 0) return false;
  return true;
}

template <int S>
__device__ __forceinline__ void gauss_elimination(double (&m)[S][S + 1], double (&x)[S]) {
#pragma unroll
  for (int i = 0; i < S; i++)
#pragma unroll
    for (int k = i + 1; k < S; k++) {
      const bool sw = fabs(m[i][i]) < fabs(m[k][i]);
#pragma unroll
      for (int j = 0; j <= S; j++) {
        const double a = m[i][j], b = m[k][j];
        m[i][j] = sw ? b : a;
        m[k][j] = sw ? a : b;
      }
    }
#pragma unroll
  for (int i = 0; i < S - 1; i++)
#pragma unroll
    for (int k = i + 1; k < S; k++) {
      const double t = m[k][i] / m[i][i];
#pragma unroll
      for (int j = 0; j <= S; j++) m[k][j] = m[k][j] - t * m[i][j];
    }
#pragma unroll
  for (int i = S - 1; i >= 0; i--) {
    double v = m[i][S];
#pragma unroll
    for (int j = i + 1; j < S; j++) v = v - m[i][j] * x[j];
    x[i] = v / m[i][i];
  }
}

// status: 0 rejected by the sample test, 1 solver failure, 2 model written.
// valid: isValidModel outcome for written models.
__global__ void __launch_bounds__(256) k_h4_solve(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                                  const int32_t* __restrict__ cand_pair,
                                                  const int32_t* __restrict__ samples, int64_t num_cand,
                                                  double* __restrict__ models, int8_t* __restrict__ status,
                                                  int8_t* __restrict__ valid) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= num_cand) return;
  const double* P = pts + 4 * pair_pt_off[cand_pair[c]];
  const int32_t* s = samples + 4 * c;
  double pa[4], pb[4], pc[4], pd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    pa[k] = P[4 * (int64_t)s[0] + k];
    pb[k] = P[4 * (int64_t)s[1] + k];
    pc[k] = P[4 * (int64_t)s[2] + k];
    pd[k] = P[4 * (int64_t)s[3] + k];
  }
  if (!h_valid_sample(pa, pb, pc, pd)) {
    status[c] = 0;
    valid[c] = 0;
    return;
  }
  double a[8][9];
  const double* rows[4] = {pa, pb, pc, pd};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double x1 = rows[i][0], y1 = rows[i][1], x2 = rows[i][2], y2 = rows[i][3];
    const double weight = 1.0;
    const double mwx1 = -weight * x1, mwy1 = -weight * y1, wx2 = weight * x2, wy2 = weight * y2;
    const int r = 2 * i;
    a[r][0] = mwx1; a[r][1] = mwy1; a[r][2] = -weight; a[r][3] = 0; a[r][4] = 0; a[r][5] = 0;
    a[r][6] = wx2 * x1; a[r][7] = wx2 * y1; a[r][8] = -wx2;
    a[r + 1][0] = 0; a[r + 1][1] = 0; a[r + 1][2] = 0; a[r + 1][3] = mwx1; a[r + 1][4] = mwy1; a[r + 1][5] = -weight;
    a[r + 1][6] = wy2 * x1; a[r + 1][7] = wy2 * y1; a[r + 1][8] = -wy2;
  }
  double h[8];
  gauss_elimination<8>(a, h);
  bool nan = false;
#pragma unroll
  for (int i = 0; i < 8; ++i) nan |= (h[i] != h[i]);
  if (nan) {
    status[c] = 1;
    valid[c] = 0;
    return;
  }
  double* out = models + 9 * c;
  double H[9];
#pragma unroll
  for (int i = 0; i < 8; ++i) H[i] = h[i];
  H[8] = 1.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = H[i];
  status[c] = 2;
  const double det = det3_partial_pivot(H);
  valid[c] = (fabs(det) < 1e-2) ? 0 : 1;
}

void launch_h4_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* status, int8_t* valid) {
  if (num_cand <= 0) return;
  const int bs = 256;
  const int64_t grid = (num_cand + bs - 1) / bs;
  hipLaunchKernelGGL(k_h4_solve, dim3((unsigned)grid), dim3(bs), 0, st, pts, pair_pt_off, cand_pair, samples, num_cand,
                     models, status, valid);
}

}  // namespace gcr
