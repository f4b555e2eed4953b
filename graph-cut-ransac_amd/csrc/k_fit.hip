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
// The system lives in LDS (rows beyond the LDS capacity of one workgroup in
// the fit's global scratch), so the QR passes wait on LDS, not HBM.
// Reductions inside the QR (column norms, Householder dot products) use a
// fixed lane order: partials of rows i = t mod NT, then the block tree of
// block_tree (64 lanes for small systems, 256 for large ones; the same bits
// up to 128 rows).  Row i of every column is only ever written and read by
// thread i mod NT; values other threads need are read after a barrier.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

// Storage of the m x 9 system: rows [0, cap) in LDS, rows [cap, m) in the
// fit's global scratch.  cap is a multiple of 64 (or >= m); every row is
// always read and written by the same thread.
struct Rows9 {
  double* lds;
  double* glob;
  int cap;
  __device__ __forceinline__ double* row(int i) const { return i < cap ? lds + 9 * i : glob + 9 * (int64_t)i; }
};

// Sequential sum over the sample in sample order, run by a single lane.
template <class F>
__device__ __forceinline__ double seq_sum(int n, F term) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += term(i);
  return s;
}

// Block-cooperative Householder QR with column pivoting solving
// min ||A x - b|| for the homography system (n = 8 unknowns).  A is stored
// row-major with b as a ninth column (row i at A.row(i)); row i is only ever
// touched by thread i mod NT, so each thread streams its own rows and every
// column reduction keeps the fixed order of block_tree (thread partials over
// ascending rows, then the tree).  One pass over the rows accumulates every
// dot product a step needs.
template <int NT>
__device__ void wave_colpiv_qr_solve9(const Rows9& A, int m, double* x, const BlockRed& B) {
  constexpr int n = 8;
  const int tid = threadIdx.x;
  int perm[8];
  double norms[8];
  {
    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = 0.0;
    for (int i = tid; i < m; i += NT) {
      const double* r = A.row(i);
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[j] = acc[j] + r[j] * r[j];
    }
    block_tree<NT, 8>(acc, B.red, B.tot);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      perm[j] = j;
      norms[j] = acc[j];
    }
  }
  // logical column c lives in physical column phys[c]
  int phys[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) phys[j] = j;
  const int kmax = m < n ? m : n;
  double rkk[8];
  for (int k = 0; k < kmax; ++k) {
    int p = k;
    for (int j = k + 1; j < n; ++j)
      if (norms[j] > norms[p]) p = j;
    if (p != k) {
      const double tn = norms[k];
      norms[k] = norms[p];
      norms[p] = tn;
      const int tp = perm[k];
      perm[k] = perm[p];
      perm[p] = tp;
      const int tq = phys[k];
      phys[k] = phys[p];
      phys[p] = tq;
    }
    const int ck = phys[k];
    __syncthreads();
    const double alpha = A.row(k)[ck];
    double sacc[1] = {0.0};
    for (int i = tid; i < m; i += NT)
      if (i > k) {
        const double v = A.row(i)[ck];
        sacc[0] = sacc[0] + v * v;
      }
    block_tree<NT, 1>(sacc, B.red, B.tot);
    const double sigma = sacc[0];
    double tau = 0.0, beta = alpha, scale = 1.0;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    rkk[k] = beta;
    // physical columns still to be reduced (logical k+1 .. n-1)
    uint32_t amask = 0;
    for (int c = k + 1; c < n; ++c) amask |= 1u << phys[c];
    if (tau != 0.0) {
      // Householder vector entries v_i = A[i][ck] * scale (i > k) are formed
      // where they are used and never stored: only R and Q^T b are read
      // later.  w_j = A[k][j] + sum_{i>k} v_i A[i][j] for the remaining
      // columns and b.
      double wacc[9];
#pragma unroll
      for (int j = 0; j < 9; ++j) wacc[j] = 0.0;
      for (int i = tid; i < m; i += NT)
        if (i > k) {
          const double* r = A.row(i);
          const double v = r[ck] * scale;
#pragma unroll
          for (int j = 0; j < 9; ++j) wacc[j] = wacc[j] + v * r[j];
        }
      block_tree<NT, 9>(wacc, B.red, B.tot);
      double twv[9];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        twv[j] = 0.0;
        if ((amask >> j) & 1u) {
          const double w = A.row(k)[j] + wacc[j];
          twv[j] = tau * w;
        }
      }
      {
        const double w = A.row(k)[8] + wacc[8];
        twv[8] = tau * w;
      }
      __syncthreads();
      // update rows k.. (row k without the vector factor) and accumulate the
      // new column norms over rows > k; the next row is loaded before this
      // row is stored so the loads do not wait on the stores
      double nphys[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) nphys[j] = 0.0;
      double cur[9];
      int i = tid;
      if (i < m) {
#pragma unroll
        for (int c = 0; c < 9; ++c) cur[c] = A.row(i)[c];
      }
      for (; i < m; i += NT) {
        double nxt[9];
        const bool more = i + NT < m;
        if (more) {
#pragma unroll
          for (int c = 0; c < 9; ++c) nxt[c] = A.row(i + NT)[c];
        }
        if (i >= k) {
          double vck = 0.0;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            if (c == ck) vck = cur[c];
          if (i == k) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if ((amask >> j) & 1u) cur[j] = cur[j] - twv[j];
            cur[8] = cur[8] - twv[8];
          } else {
            const double v = vck * scale;
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if ((amask >> j) & 1u) {
                const double u = cur[j] - twv[j] * v;
                cur[j] = u;
                nphys[j] = nphys[j] + u * u;
              }
            cur[8] = cur[8] - twv[8] * v;
          }
          double* r = A.row(i);
#pragma unroll
          for (int c = 0; c < 9; ++c) r[c] = cur[c];
        }
        if (more) {
#pragma unroll
          for (int c = 0; c < 9; ++c) cur[c] = nxt[c];
        }
      }
      block_tree<NT, 8>(nphys, B.red, B.tot);
      double tn[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) tn[j] = ((amask >> j) & 1u) ? nphys[j] : 0.0;
      for (int c = k + 1; c < n; ++c) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j == phys[c]) v = tn[j];
        norms[c] = v;
      }
    } else {
      double nphys[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) nphys[j] = 0.0;
      for (int i = tid; i < m; i += NT)
        if (i > k) {
          const double* r = A.row(i);
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if ((amask >> j) & 1u) nphys[j] = nphys[j] + r[j] * r[j];
        }
      block_tree<NT, 8>(nphys, B.red, B.tot);
      double tn[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) tn[j] = ((amask >> j) & 1u) ? nphys[j] : 0.0;
      for (int c = k + 1; c < n; ++c) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j == phys[c]) v = tn[j];
        norms[c] = v;
      }
    }
    __syncthreads();
  }
  // R (upper part: row k, logical column c at phys[c]; diagonal rkk) and Q^T b
  const double thr = 2.220446049250313e-16 * (double)kmax;
  const double r00 = fabs(rkk[0]);
  int rank = 0;
  for (int k = 0; k < kmax; ++k) {
    if (fabs(rkk[k]) > thr * r00) ++rank;
    else break;
  }
  double z[8];
  for (int k = 0; k < n; ++k) z[k] = 0.0;
  for (int k = rank - 1; k >= 0; --k) {
    const double* rk = A.row(k);
    double s = rk[8];
    for (int j = k + 1; j < rank; ++j) s = s - rk[phys[j]] * z[j];
    z[k] = s / rkk[k];
  }
  for (int k = 0; k < n; ++k) x[perm[k]] = z[k];
}

// One block per fit (one wave, or four for systems of more than 64 rows).
// fit_samples[fit_sample_off[f] ...] holds fit_ns[f] pair-local point
// indices.  weights: per-pair array indexed by pair-local point index
// (nullptr = unweighted).
template <int NT>
__global__ void __launch_bounds__(NT) k_fit_h(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                              const int32_t* __restrict__ fit_pair,
                                              const int64_t* __restrict__ fit_sample_off,
                                              const int32_t* __restrict__ fit_ns,
                                              const int32_t* __restrict__ samples,
                                              const double* __restrict__ weights,
                                              const int64_t* __restrict__ weight_off,
                                              double* __restrict__ scratch,
                                              const int64_t* __restrict__ scratch_off,
                                              double* __restrict__ models_out, int8_t* __restrict__ ok_out,
                                              int lds_rows) {
  const int f = blockIdx.x;
  const bool t0 = threadIdx.x == 0;
  const int pair = fit_pair[f];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int32_t* __restrict__ S = samples + fit_sample_off[f];
  const int ns = fit_ns[f];
  const double* __restrict__ W = weights ? weights + weight_off[pair] : nullptr;
  if (ns < 4) {
    if (t0) ok_out[f] = 0;
    return;
  }
  const int m = 2 * ns;
  // m x 9, row-major: the first lds_rows rows in LDS, the rest in scratch;
  // the reduction staging of a four-wave block after the rows
  extern __shared__ double lds_sys[];
  const Rows9 A{lds_sys, scratch + scratch_off[f], lds_rows};
  const BlockRed B{lds_sys + 9 * (int64_t)lds_rows, lds_sys + 9 * (int64_t)lds_rows + 9 * 128};

  // Mass points and mean distances: sequential sums in sample order
  // (homography_estimator.h:236-270), evaluated by the whole wave.
  double msum[4];
  wave_seq_sum4(ns, [&](int i, double* v) {
    const double* d = P + 4 * (int64_t)S[i];
    v[0] = d[0];
    v[1] = d[1];
    v[2] = d[2];
    v[3] = d[3];
  }, msum);
  const double ms0 = msum[0] / ns, ms1 = msum[1] / ns;
  const double md0 = msum[2] / ns, md1 = msum[3] / ns;
  double dsum[4];
  wave_seq_sum4(ns, [&](int i, double* v) {
    const double* d = P + 4 * (int64_t)S[i];
    const double dx1 = ms0 - d[0], dy1 = ms1 - d[1], dx2 = md0 - d[2], dy2 = md1 - d[3];
    v[0] = sqrt(dx1 * dx1 + dy1 * dy1);
    v[1] = sqrt(dx2 * dx2 + dy2 * dy2);
    v[2] = 0.0;
    v[3] = 0.0;
  }, dsum);
  const double as = dsum[0] / ns, ad = dsum[1] / ns;
  const double rs = 1.4142135623730951 / as, rd = 1.4142135623730951 / ad;

  // Build the weighted system, row-major with b as column 8 (row r owned by
  // thread r mod NT).
  for (int r = threadIdx.x; r < m; r += NT) {
    const int i = r >> 1;
    const double* d = P + 4 * (int64_t)S[i];
    const double x1 = (d[0] - ms0) * rs, y1 = (d[1] - ms1) * rs;
    const double x2 = (d[2] - md0) * rd, y2 = (d[3] - md1) * rd;
    const double weight = W ? W[i] : 1.0;
    const double mwx1 = -weight * x1, mwy1 = -weight * y1, wx2 = weight * x2, wy2 = weight * y2;
    double* row = A.row(r);
    if ((r & 1) == 0) {
      row[0] = mwx1; row[1] = mwy1; row[2] = -weight; row[3] = 0; row[4] = 0; row[5] = 0;
      row[6] = wx2 * x1; row[7] = wx2 * y1; row[8] = -wx2;
    } else {
      row[0] = 0; row[1] = 0; row[2] = 0; row[3] = mwx1; row[4] = mwy1; row[5] = -weight;
      row[6] = wy2 * x1; row[7] = wy2 * y1; row[8] = -wy2;
    }
  }
  __syncthreads();
  double h[8];
  wave_colpiv_qr_solve9<NT>(A, m, h, B);
  if (t0) {
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
                  int8_t* ok_out, int num_fits, int max_ns) {
  if (num_fits <= 0) return;
  // systems of more than 64 rows run on four waves (256-lane sums, the same
  // bits as 64 lanes up to 128 rows); the largest system of the launch in LDS
  // when it fits (next to the reduction staging), else its first rows (a
  // multiple of 64) with the rest in scratch
  const int m = 2 * max_ns;
  const bool wide = m > 64;
  const int red_bytes = wide ? (9 * 128 + 9) * (int)sizeof(double) : 0;
  const int limit_rows = ((lds_limit_bytes() - red_bytes - 256) / (9 * (int)sizeof(double))) & ~63;
  const int lds_rows = m <= limit_rows ? m : limit_rows;
  const size_t shmem = (size_t)(lds_rows > 0 ? lds_rows : 1) * 9 * sizeof(double) + (size_t)red_bytes;
  if (wide)
    hipLaunchKernelGGL(k_fit_h<256>, dim3(num_fits), dim3(256), shmem, st, pts, pair_pt_off, fit_pair,
                       fit_sample_off, fit_ns, samples, weights, weight_off, scratch, scratch_off, models_out, ok_out,
                       lds_rows);
  else
    hipLaunchKernelGGL(k_fit_h<64>, dim3(num_fits), dim3(64), shmem, st, pts, pair_pt_off, fit_pair,
                       fit_sample_off, fit_ns, samples, weights, weight_off, scratch, scratch_off, models_out, ok_out,
                       lds_rows);
}

}  // namespace gcr
