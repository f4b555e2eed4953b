// Non-minimal (least-squares) homography fits, one wavefront per fit.
//
// Restates RobustHomographyEstimator::estimateModelNonminimal
// (homography_estimator.h:136-173): Hartley normalisation of the sample
// (:213-310, mass point and mean distance accumulated sequentially in sample
// order, as the reference loops do), the weighted 2n x 8 system of
// HomographyFourPointSolver::estimateNonMinimalModel
// (solver_homography_four_point.h:192-264) solved by Householder QR with
// column pivoting, and denormalisation T2^-1 * H * T1.
//
// The reference calls the solver with sample == nullptr on the compacted
// normalised points, so row i reads weights[i] of the point array, i.e. the
// weight stored for point index i (not for the i-th sampled point); the
// kernel reproduces that indexing.
//
// Reductions inside the QR (column norms, Householder dot products) use the
// fixed 64-lane order of wave_sum64.  Row i of every column is only ever
// written and read by lane i mod 64; values other lanes need are broadcast
// after a barrier.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

// Sequential sum over the sample in sample order, run by a single lane.
template <class F>
__device__ __forceinline__ double seq_sum(int n, F term) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += term(i);
  return s;
}

// Wave-cooperative Householder QR with column pivoting solving
// min ||A x - b|| (A m x n column-major in global scratch, n <= 8).
__device__ void wave_colpiv_qr_solve(double* __restrict__ A, double* __restrict__ b, int m, int n, double* x) {
  const int lane = lane_id();
  int perm[8];
  double norms[8];
  for (int j = 0; j < n; ++j) {
    perm[j] = j;
    const double* c = A + (int64_t)j * m;
    norms[j] = wave_sum64(0, m, [&](int64_t i) { return c[i] * c[i]; });
  }
  const int kmax = m < n ? m : n;
  for (int k = 0; k < kmax; ++k) {
    int p = k;
    for (int j = k + 1; j < n; ++j)
      if (norms[j] > norms[p]) p = j;
    if (p != k) {
      double* ck = A + (int64_t)k * m;
      double* cp = A + (int64_t)p * m;
      for (int i = lane; i < m; i += 64) {
        const double t = ck[i];
        ck[i] = cp[i];
        cp[i] = t;
      }
      const double tn = norms[k];
      norms[k] = norms[p];
      norms[p] = tn;
      const int tp = perm[k];
      perm[k] = perm[p];
      perm[p] = tp;
    }
    __syncthreads();
    double* col = A + (int64_t)k * m;
    const double alpha = col[k];
    const double sigma = wave_sum64(k + 1, m, [&](int64_t i) { return col[i] * col[i]; });
    double tau = 0.0, beta = alpha;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
      for (int i = lane; i < m; i += 64)
        if (i > k) col[i] = col[i] * scale;
    }
    __syncthreads();
    if (sigma != 0.0 && lane == (k & 63)) col[k] = beta;
    __syncthreads();
    if (tau != 0.0) {
      for (int j = k + 1; j < n; ++j) {
        double* cj = A + (int64_t)j * m;
        const double w = cj[k] + wave_sum64(k + 1, m, [&](int64_t i) { return col[i] * cj[i]; });
        const double tw = tau * w;
        __syncthreads();
        for (int i = lane; i < m; i += 64) {
          if (i == k) cj[i] = cj[i] - tw;
          else if (i > k) cj[i] = cj[i] - tw * col[i];
        }
        __syncthreads();
      }
      const double w = b[k] + wave_sum64(k + 1, m, [&](int64_t i) { return col[i] * b[i]; });
      const double tw = tau * w;
      __syncthreads();
      for (int i = lane; i < m; i += 64) {
        if (i == k) b[i] = b[i] - tw;
        else if (i > k) b[i] = b[i] - tw * col[i];
      }
      __syncthreads();
    }
    for (int j = k + 1; j < n; ++j) {
      const double* cj = A + (int64_t)j * m;
      norms[j] = wave_sum64(k + 1, m, [&](int64_t i) { return cj[i] * cj[i]; });
    }
  }
  const double thr = 2.220446049250313e-16 * (double)kmax;
  const double r00 = fabs(A[0]);
  int rank = 0;
  for (int k = 0; k < kmax; ++k) {
    if (fabs(A[(int64_t)k * m + k]) > thr * r00) ++rank;
    else break;
  }
  double z[8];
  for (int k = 0; k < n; ++k) z[k] = 0.0;
  for (int k = rank - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < rank; ++j) s = s - A[(int64_t)j * m + k] * z[j];
    z[k] = s / A[(int64_t)k * m + k];
  }
  for (int k = 0; k < n; ++k) x[perm[k]] = z[k];
}

// One block (one wave) per fit.  fit_samples[fit_sample_off[f] ...] holds
// fit_ns[f] pair-local point indices.  weights: per-pair array indexed by
// pair-local point index (nullptr = unweighted).
__global__ void __launch_bounds__(64) k_fit_h(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                              const int32_t* __restrict__ fit_pair,
                                              const int64_t* __restrict__ fit_sample_off,
                                              const int32_t* __restrict__ fit_ns,
                                              const int32_t* __restrict__ samples,
                                              const double* __restrict__ weights,
                                              const int64_t* __restrict__ weight_off,
                                              double* __restrict__ scratch,
                                              const int64_t* __restrict__ scratch_off,
                                              double* __restrict__ models_out, int8_t* __restrict__ ok_out) {
  const int f = blockIdx.x;
  const int lane = lane_id();
  const int pair = fit_pair[f];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int32_t* __restrict__ S = samples + fit_sample_off[f];
  const int ns = fit_ns[f];
  const double* __restrict__ W = weights ? weights + weight_off[pair] : nullptr;
  if (ns < 4) {
    if (lane == 0) ok_out[f] = 0;
    return;
  }
  const int m = 2 * ns;
  double* A = scratch + scratch_off[f];
  double* b = A + (int64_t)8 * m;

  // Mass points: four independent sequential sums in sample order.
  double msum = 0.0;
  if (lane < 4) msum = seq_sum(ns, [&](int i) { return P[4 * (int64_t)S[i] + lane]; });
  const double ms0 = readlane_f64(msum, 0) / ns, ms1 = readlane_f64(msum, 1) / ns;
  const double md0 = readlane_f64(msum, 2) / ns, md1 = readlane_f64(msum, 3) / ns;
  // Mean distances: two sequential sums.
  double dsum = 0.0;
  if (lane < 2) {
    const double c0 = lane == 0 ? ms0 : md0, c1 = lane == 0 ? ms1 : md1;
    const int o = lane == 0 ? 0 : 2;
    dsum = seq_sum(ns, [&](int i) {
      const double* d = P + 4 * (int64_t)S[i];
      const double dx = c0 - d[o], dy = c1 - d[o + 1];
      return sqrt(dx * dx + dy * dy);
    });
  }
  const double as = readlane_f64(dsum, 0) / ns, ad = readlane_f64(dsum, 1) / ns;
  const double rs = 1.4142135623730951 / as, rd = 1.4142135623730951 / ad;

  // Build the weighted system (rows 2i, 2i+1 owned by lanes (2i) mod 64 and
  // (2i+1) mod 64).
  for (int r = lane; r < m; r += 64) {
    const int i = r >> 1;
    const double* d = P + 4 * (int64_t)S[i];
    const double x1 = (d[0] - ms0) * rs, y1 = (d[1] - ms1) * rs;
    const double x2 = (d[2] - md0) * rd, y2 = (d[3] - md1) * rd;
    const double weight = W ? W[i] : 1.0;
    const double mwx1 = -weight * x1, mwy1 = -weight * y1, wx2 = weight * x2, wy2 = weight * y2;
    if ((r & 1) == 0) {
      A[0 * (int64_t)m + r] = mwx1; A[1 * (int64_t)m + r] = mwy1; A[2 * (int64_t)m + r] = -weight;
      A[3 * (int64_t)m + r] = 0; A[4 * (int64_t)m + r] = 0; A[5 * (int64_t)m + r] = 0;
      A[6 * (int64_t)m + r] = wx2 * x1; A[7 * (int64_t)m + r] = wx2 * y1; b[r] = -wx2;
    } else {
      A[0 * (int64_t)m + r] = 0; A[1 * (int64_t)m + r] = 0; A[2 * (int64_t)m + r] = 0;
      A[3 * (int64_t)m + r] = mwx1; A[4 * (int64_t)m + r] = mwy1; A[5 * (int64_t)m + r] = -weight;
      A[6 * (int64_t)m + r] = wy2 * x1; A[7 * (int64_t)m + r] = wy2 * y1; b[r] = -wy2;
    }
  }
  __syncthreads();
  double h[8];
  wave_colpiv_qr_solve(A, b, m, 8, h);
  if (lane == 0) {
    double H[9];
    for (int i = 0; i < 8; ++i) H[i] = h[i];
    H[8] = 1.0;
    const double T1[9] = {rs, 0, -rs * ms0, 0, rs, -rs * ms1, 0, 0, 1};
    const double T2[9] = {rd, 0, -rd * md0, 0, rd, -rd * md1, 0, 0, 1};
    double T2inv[9], tmp[9];
    mat3_inverse(T2, T2inv);
    mat3_mul(T2inv, H, tmp);
    mat3_mul(tmp, T1, H);
    for (int i = 0; i < 9; ++i) models_out[9 * (int64_t)f + i] = H[i];
    ok_out[f] = 2;
  }
}

void launch_fit_h(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, double* scratch, const int64_t* scratch_off, double* models_out,
                  int8_t* ok_out, int num_fits) {
  if (num_fits <= 0) return;
  hipLaunchKernelGGL(k_fit_h, dim3(num_fits), dim3(64), 0, st, pts, pair_pt_off, fit_pair, fit_sample_off, fit_ns,
                     samples, weights, weight_off, scratch, scratch_off, models_out, ok_out);
}

}  // namespace gcr
