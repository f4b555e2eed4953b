// Essential-matrix kernels: the five-point minimal solver (one thread per
// candidate sample), the Stewenius non-minimal fit (one wavefront per fit)
// and the driver's pose bundle adjustment (one wavefront per pair).
//
// Minimal: EssentialMatrixEstimator::estimateModel
// (essential_estimator.h:152-175) -> EssentialMatrixFivePointNisterSolver
// (solver_essential_matrix_five_point_nister.h:204-615) on K-normalised
// coordinates, then the oriented epipolar filter (essential_estimator.h:338-398).
//
// Non-minimal: estimateModelNonminimal (essential_estimator.h:308-335) ->
// EssentialMatrixFivePointSteweniusSolver (…stewenius.h:114-231), then
// Frobenius normalisation and E33 >= 0.
//
// Bundle adjustment: py/gcransac_python.cpp:2766-2823 ->
// EssentialMatrixBundleAdjustmentSolver (…bundle_adjustment.h:114-188):
// Stewenius initialisation, motion_from_essential (essential.cpp:25-80),
// refine_relpose / lm_5dof_impl (bundle.cpp:176-251,412-452) with the
// truncated loss, essential_from_motion.
//
// Sums over points use the fixed 64-lane order of wave_sum64; the small
// dense steps run on lane 0 (their state in LDS) so every lane sees the same
// values after a barrier.
#include "dev_essential.hpp"
#include "kernels.hpp"

namespace gcr {

// ---------------------------------------------------------------------------
// Five-point minimal solver
// ---------------------------------------------------------------------------
// Up to 10 models of one sample (row-major 3x3 each, in root order); returns
// their number.
__device__ int nister_solve_thread(const double* P, const int32_t* s, double* models) {
  double a[45];
  for (int i = 0; i < 5; ++i) e_row(P + 4 * (int64_t)s[i], 1.0, a + 9 * i);
  double N[36];
  if (null_space_rx9(a, 5, N) != 4) return 0;
  double ee[3][3][4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 4; ++k) ee[i][j][k] = N[9 * k + 3 * j + i];
  double C[200];
  essential_constraints(ee, kNisterMul, 1.0, C);
  double L[100], B[100];
  for (int r = 0; r < 10; ++r)
    for (int c = 0; c < 10; ++c) {
      L[10 * r + c] = C[20 * r + c];
      B[10 * r + c] = C[20 * r + 10 + c];
    }
  lu_partial_solve(L, B, 10, 10);
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
  double Pp[3][4], Qp[3][4], Rp[3][5];
  for (int i = 0; i < 3; ++i) {
    for (int d = 0; d < 4; ++d) {
      Pp[i][d] = A[i][3 - d];
      Qp[i][d] = A[i][7 - d];
    }
    for (int d = 0; d < 5; ++d) Rp[i][d] = A[i][12 - d];
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
    mul(Qp[1], 3, Rp[2], 4, t1);
    mul(Qp[2], 3, Rp[1], 4, t2);
    for (int k = 0; k < 8; ++k) m7[k] = t1[k] - t2[k];
    mul(Pp[0], 3, m7, 7, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] + full[k];
    mul(Pp[1], 3, Rp[2], 4, t1);
    mul(Pp[2], 3, Rp[1], 4, t2);
    for (int k = 0; k < 8; ++k) m7[k] = t1[k] - t2[k];
    mul(Qp[0], 3, m7, 7, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] - full[k];
    double s1[7], s2[7], m6[7];
    mul(Pp[1], 3, Qp[2], 3, s1);
    mul(Pp[2], 3, Qp[1], 3, s2);
    for (int k = 0; k < 7; ++k) m6[k] = s1[k] - s2[k];
    mul(Rp[0], 4, m6, 6, full);
    for (int k = 0; k < 11; ++k) c[k] = c[k] + full[k];
  }
  double roots[10];
  DevSturm10 st;
  const int nr = st.solve(c, roots);
  for (int q = 0; q < nr; ++q) {
    const double z = roots[q];
    const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
    double b0[3], b1[3], bb[3];
    for (int i = 0; i < 3; ++i) {
      b0[i] = A[i][0] * z3 + A[i][1] * z2 + A[i][2] * z + A[i][3];
      b1[i] = A[i][4] * z3 + A[i][5] * z2 + A[i][6] * z + A[i][7];
      bb[i] = A[i][8] * z4 + A[i][9] * z3 + A[i][10] * z2 + A[i][11] * z + A[i][12];
    }
    const double det = b0[0] * b1[1] - b1[0] * b0[1];
    const double id = 1.0 / det;
    double xz[2];
    xz[0] = (b1[1] * id) * bb[0] + (-b1[0] * id) * bb[1];
    xz[1] = (-b0[1] * id) * bb[0] + (b0[0] * id) * bb[1];
    if (fabs(b0[2] * xz[0] + b1[2] * xz[1] - bb[2]) > 1e-6) solve3x2_fullpiv(b0, b1, bb, xz);
    const double x = -xz[0], y = -xz[1];
    double e[9];
    const double sc = 1.0 / sqrt(x * x + y * y + z * z + 1.0);
    for (int k = 0; k < 9; ++k) e[k] = (N[k] * x + N[9 + k] * y + N[18 + k] * z + N[27 + k]) * sc;
    double* m = models + 9 * q;
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) m[3 * r + cc] = e[3 * cc + r];
  }
  return nr;
}

// One thread per candidate; 10 model slots per candidate (status 2 = kept,
// compacted in root order after the oriented filter).
__global__ void __launch_bounds__(64) k_e5_solve(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                                 const int32_t* __restrict__ cand_pair,
                                                 const int32_t* __restrict__ samples, int64_t num_cand,
                                                 double* __restrict__ models, int8_t* __restrict__ mstatus,
                                                 int8_t* __restrict__ nmodels) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= num_cand) return;
  const double* P = pts + 4 * pair_pt_off[cand_pair[c]];
  const int32_t* s = samples + 5 * c;
  double* out = models + 90 * c;
  int kept = 0;
  if (s[0] >= 0) {
    const int nr = nister_solve_thread(P, s, out);
    for (int q = 0; q < nr; ++q) {
      const double* m = out + 9 * q;
      if (f_orientation_valid_idx(m, P, s, 5)) {
        if (kept != q)
          for (int i = 0; i < 9; ++i) out[9 * kept + i] = m[i];
        ++kept;
      }
    }
  }
  for (int k = 0; k < 10; ++k) mstatus[10 * c + k] = k < kept ? 2 : 0;
  nmodels[c] = (int8_t)kept;
}

void launch_e5_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* mstatus, int8_t* nmodels) {
  if (num_cand <= 0) return;
  const int64_t grid = (num_cand + 63) / 64;
  hipLaunchKernelGGL(k_e5_solve, dim3((unsigned)grid), dim3(64), 0, st, pts, pair_pt_off, cand_pair, samples,
                     num_cand, models, mstatus, nmodels);
}

// ---------------------------------------------------------------------------
// Non-minimal fit and bundle adjustment (one wave each)
// ---------------------------------------------------------------------------
// sum64 order on one lane: partial i & 63 per term, then the 64-lane tree.
template <class F>
__device__ double lane_sum64(int begin, int end, F term) {
  double p[64];
  for (int l = 0; l < 64; ++l) p[l] = 0.0;
  for (int i = begin; i < end; ++i) p[i & 63] += term(i);
  for (int s = 32; s >= 1; s >>= 1)
    for (int l = 0; l < s; ++l) p[l] += p[l + s];
  return p[0];
}

// Q of the Householder QR with column pivoting of the 9 x 9 matrix A (the
// role of ColPivHouseholderQR::matrixQ); column norms in the sum64 order.
__device__ void colpiv_qr_Q9(double* A, double* Q) {
  const int n = 9;
  double hv[9][9], tau[9], norms[9];
  for (int j = 0; j < n; ++j) norms[j] = lane_sum64(0, n, [&](int i) { return A[n * i + j] * A[n * i + j]; });
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int j = k + 1; j < n; ++j)
      if (norms[j] > norms[p]) p = j;
    if (p != k) {
      for (int i = 0; i < n; ++i) gswap(A[n * i + k], A[n * i + p]);
      gswap(norms[k], norms[p]);
    }
    const double alpha = A[n * k + k];
    const double sigma = lane_sum64(k + 1, n, [&](int i) { return A[n * i + k] * A[n * i + k]; });
    tau[k] = 0.0;
    for (int i = 0; i < n; ++i) hv[k][i] = 0.0;
    hv[k][k] = 1.0;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      tau[k] = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
      for (int i = k + 1; i < n; ++i) hv[k][i] = A[n * i + k] * scale;
      A[n * k + k] = beta;
      for (int i = k + 1; i < n; ++i) A[n * i + k] = 0.0;
      for (int j = k + 1; j < n; ++j) {
        const double w = A[n * k + j] + lane_sum64(k + 1, n, [&](int i) { return hv[k][i] * A[n * i + j]; });
        const double tw = tau[k] * w;
        A[n * k + j] = A[n * k + j] - tw;
        for (int i = k + 1; i < n; ++i) A[n * i + j] = A[n * i + j] - tw * hv[k][i];
      }
    }
    for (int j = k + 1; j < n; ++j)
      norms[j] = lane_sum64(k + 1, n, [&](int i) { return A[n * i + j] * A[n * i + j]; });
  }
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

// Shared state of one fit (LDS).
struct EssShared {
  double M[81];        // A^T A, then scratch of the QR
  double Q[81];
  double NS[9][4];     // null-space basis (columns: x, y, z, 1)
  double models[10][9];
  int nmodels;
  // bundle adjustment
  double dR[9][3], dt[9][2], tb[3][2];
  double pose_q[4], pose_t[3], new_q[4], new_t[3];
  double sol[5];
  double JtJ[5][5];
  double E[9];
};

// Stewenius steps after the null space (lane 0): constraints, elimination,
// action matrix, real eigenvalues, null vectors -> S.models / S.nmodels.
__device__ void stewenius_tail(EssShared& S) {
  double ee[3][3][4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 4; ++k) ee[i][j][k] = S.NS[3 * j + i][k];
  double C[200];
  essential_constraints(ee, kSteweniusMul, 2.0, C);
  double L[100], B[100];
  for (int r = 0; r < 10; ++r)
    for (int c = 0; c < 10; ++c) {
      L[10 * r + c] = C[20 * r + c];
      B[10 * r + c] = C[20 * r + 10 + c];
    }
  lu_full_solve(L, B, 10, 10);
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
    for (int r = 0; r < 9; ++r)
      e[r] = S.NS[r][0] * v[6] + S.NS[r][1] * v[7] + S.NS[r][2] * v[8] + S.NS[r][3] * v[9];
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) S.models[q][3 * r + cc] = e[3 * cc + r];
  }
  S.nmodels = ne;
}

// Stewenius solve of the n rows written to `rows` (9 doubles each, row i by
// lane i mod 64).  Leaves the models in S (all lanes after the call).
__device__ void wave_stewenius(const double* rows, int n, EssShared& S) {
  const int lane = lane_id();
  __syncthreads();
  if (n == 5) {
    if (lane == 0) {
      double a[45], N[36];
      for (int i = 0; i < 45; ++i) a[i] = rows[i];
      S.nmodels = 0;
      if (null_space_rx9(a, 5, N) == 4) {
        for (int r = 0; r < 9; ++r)
          for (int k = 0; k < 4; ++k) S.NS[r][k] = N[9 * k + r];
        S.nmodels = -1;  // null space ready
      }
    }
  } else {
    double acc[45];
#pragma unroll
    for (int e = 0; e < 45; ++e) acc[e] = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double* r = rows + 9 * (int64_t)i;
      int o = 0;
#pragma unroll
      for (int a1 = 0; a1 < 9; ++a1)
#pragma unroll
        for (int b1 = a1; b1 < 9; ++b1) {
          acc[o] = acc[o] + r[a1] * r[b1];
          ++o;
        }
    }
    wave_tree64_many<45>(acc);
    int o = 0;
#pragma unroll
    for (int a1 = 0; a1 < 9; ++a1)
#pragma unroll
      for (int b1 = a1; b1 < 9; ++b1) {
        const double v = acc[o];
        if (lane == 0) {
          S.M[9 * a1 + b1] = v;
          S.M[9 * b1 + a1] = v;
        }
        ++o;
      }
    if (lane == 0) {
      colpiv_qr_Q9(S.M, S.Q);
      for (int r = 0; r < 9; ++r)
        for (int k = 0; k < 4; ++k) S.NS[r][k] = S.Q[9 * r + 5 + k];
      S.nmodels = -1;
    }
  }
  if (lane == 0 && S.nmodels == -1) stewenius_tail(S);
  __syncthreads();
}

// One wave per fit: the estimator's non-minimal fit.  weight_mode 0:
// unweighted; 1: W[point index] (the Stewenius solver reads weights by the
// sample's point index).  Up to 10 models per fit at models_out[(10 f + j) 9],
// Frobenius-normalised with E33 >= 0.
__global__ void __launch_bounds__(64) k_fit_e(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                              const int32_t* __restrict__ fit_pair,
                                              const int64_t* __restrict__ fit_sample_off,
                                              const int32_t* __restrict__ fit_ns,
                                              const int32_t* __restrict__ samples,
                                              const double* __restrict__ weights,
                                              const int64_t* __restrict__ weight_off, int weight_mode,
                                              double* __restrict__ scratch, const int64_t* __restrict__ scratch_off,
                                              double* __restrict__ models_out, int8_t* __restrict__ ok_out) {
  __shared__ EssShared SH;
  const int f = blockIdx.x;
  const int lane = lane_id();
  const int pair = fit_pair[f];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int32_t* __restrict__ S = samples + fit_sample_off[f];
  const int ns = fit_ns[f];
  const double* __restrict__ W = weight_mode == 1 ? weights + weight_off[pair] : nullptr;
  if (ns < 5) {
    if (lane == 0)
      for (int k = 0; k < 10; ++k) ok_out[10 * f + k] = 0;
    return;
  }
  double* rows = scratch + scratch_off[f];
  for (int i = lane; i < ns; i += 64) e_row(P + 4 * (int64_t)S[i], W ? W[S[i]] : 1.0, rows + 9 * (int64_t)i);
  wave_stewenius(rows, ns, SH);
  if (lane == 0) {
    const int nm = SH.nmodels > 0 ? SH.nmodels : 0;
    for (int k = 0; k < 10; ++k) {
      ok_out[10 * f + k] = k < nm ? 2 : 0;
      if (k < nm) {
        double m[9];
        for (int i = 0; i < 9; ++i) m[i] = SH.models[k][i];
        frobenius_normalise_sign(m);
        for (int i = 0; i < 9; ++i) models_out[(10 * (int64_t)f + k) * 9 + i] = m[i];
      }
    }
  }
}

void launch_fit_e(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, int weight_mode, double* scratch, const int64_t* scratch_off,
                  double* models_out, int8_t* ok_out, int num_fits) {
  if (num_fits <= 0) return;
  hipLaunchKernelGGL(k_fit_e, dim3(num_fits), dim3(64), 0, st, pts, pair_pt_off, fit_pair, fit_sample_off, fit_ns,
                     samples, weights, weight_off, weight_mode, scratch, scratch_off, models_out, ok_out);
}

// ---- pose bundle adjustment --------------------------------------------------
__device__ __forceinline__ double relpose_cost_term(const double* E, const double* p, double w) {
  const double r2 = relpose_r2(E, p, nullptr, nullptr);
  return w * (r2 < 1.0 ? r2 : 1.0);
}

// refine_relpose on the pose in S.pose_q / S.pose_t (LDS) over the sample S
// of n points with position weights wpos (scratch, entry k by lane k mod 64).
__device__ void wave_refine_relpose(const double* P, const int32_t* Sm, int n, const double* wpos, EssShared& S) {
  const int lane = lane_id();
  auto cost_of = [&](const double* q, const double* t) {
    RelPose p;
    for (int i = 0; i < 4; ++i) p.q[i] = q[i];
    for (int i = 0; i < 3; ++i) p.t[i] = t[i];
    double E[9];
    essential_from_motion(p, E);
    return wave_sum64(0, n, [&](int64_t k) { return relpose_cost_term(E, P + 4 * (int64_t)Sm[k], wpos[k]); });
  };
  double lambda = 1e-3;
  double cost = cost_of(S.pose_q, S.pose_t);
  double Jtr[5];
  bool recompute = true;
  for (int iter = 0; iter < 25; ++iter) {
    if (recompute) {
      if (lane == 0) {
        const double* t = S.pose_t;
        double b0[3];
        const double ux[3] = {1, 0, 0}, uy[3] = {0, 1, 0}, uz[3] = {0, 0, 1};
        if (fabs(t[0]) < fabs(t[1])) {
          if (fabs(t[0]) < fabs(t[2])) cross3(t, ux, b0);
          else cross3(t, uz, b0);
        } else {
          if (fabs(t[1]) < fabs(t[2])) cross3(t, uy, b0);
          else cross3(t, uz, b0);
        }
        normalise3(b0);
        double b1[3];
        cross3(b0, t, b1);
        normalise3(b1);
        for (int i = 0; i < 3; ++i) {
          S.tb[i][0] = b0[i];
          S.tb[i][1] = b1[i];
        }
        RelPose p;
        for (int i = 0; i < 4; ++i) p.q[i] = S.pose_q[i];
        for (int i = 0; i < 3; ++i) p.t[i] = S.pose_t[i];
        double R[9];
        essential_from_motion(p, S.E);
        quat_to_R(p.q, R);
        const double* E = S.E;
        for (int i = 0; i < 3; ++i) {
          S.dR[i][0] = 0.0;
          S.dR[i][1] = -E[3 * i + 2];
          S.dR[i][2] = E[3 * i + 1];
          S.dR[3 + i][0] = E[3 * i + 2];
          S.dR[3 + i][1] = 0.0;
          S.dR[3 + i][2] = -E[3 * i + 0];
          S.dR[6 + i][0] = -E[3 * i + 1];
          S.dR[6 + i][1] = E[3 * i + 0];
          S.dR[6 + i][2] = 0.0;
        }
        for (int col = 0; col < 3; ++col) {
          const double rc[3] = {R[col], R[3 + col], R[6 + col]};
          double c0[3], c1[3];
          cross3(b0, rc, c0);
          cross3(b1, rc, c1);
          for (int i = 0; i < 3; ++i) {
            S.dt[3 * col + i][0] = c0[i];
            S.dt[3 * col + i][1] = c1[i];
          }
        }
      }
      __syncthreads();
      double acc[20];
#pragma unroll
      for (int e = 0; e < 20; ++e) acc[e] = 0.0;
      for (int k = lane; k < n; k += 64) {
        asm volatile("" ::: "memory");
        const double* p = P + 4 * (int64_t)Sm[k];
        double JC[4], C;
        relpose_r2(S.E, p, JC, &C);
        const double nJ = sqrt(JC[0] * JC[0] + JC[1] * JC[1] + JC[2] * JC[2] + JC[3] * JC[3]);
        const double inv = 1.0 / nJ;
        const double r = C * inv;
        double w = ((r * r) < 1.0 ? 1.0 : 0.0) / (double)n;
        w = wpos[k] * w;
        if (w != 0.0) {
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
#pragma unroll
          for (int q = 0; q < 9; ++q) dF[q] *= inv;
          double J[5];
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) v += dF[q] * S.dR[q][cc];
            J[cc] = v;
          }
#pragma unroll
          for (int cc = 0; cc < 2; ++cc) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) v += dF[q] * S.dt[q][cc];
            J[3 + cc] = v;
          }
          const double wres = w * C * inv;
          int o = 0;
#pragma unroll
          for (int i = 0; i < 5; ++i) {
            acc[o] = acc[o] + wres * J[i];
            ++o;
          }
#pragma unroll
          for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
              acc[o] = acc[o] + w * (J[i] * J[j]);
              ++o;
            }
        }
      }
      {
        wave_tree64_many<20>(acc);
        int o = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i) Jtr[i] = acc[o++];
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
          for (int j = 0; j <= i; ++j) {
            const double v = acc[o++];
            if (lane == 0) S.JtJ[i][j] = v;
          }
      }
      double gn = 0.0;
      for (int i = 0; i < 5; ++i) gn += Jtr[i] * Jtr[i];
      if (sqrt(gn) < 1e-8) break;
    }
    if (lane == 0) {
      auto& JtJ = S.JtJ;
      for (int k = 0; k < 5; ++k) JtJ[k][k] += lambda;
      double L[5][5], y[5];
      for (int j = 0; j < 5; ++j) {
        double d = JtJ[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        L[j][j] = sqrt(d);
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
        for (int k = i + 1; k < 5; ++k) v -= L[k][i] * S.sol[k];
        S.sol[i] = v / L[i][i];
      }
    }
    __syncthreads();
    double sn = 0.0;
    for (int i = 0; i < 5; ++i) sn += S.sol[i] * S.sol[i];
    if (sqrt(sn) < 1e-8) break;
    if (lane == 0) {
      double w[3] = {S.sol[0], S.sol[1], S.sol[2]};
      const double theta = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      for (int i = 0; i < 3; ++i) w[i] /= theta;
      double a, b;
      sin_cos_taylor(theta, a, b);
      const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
      double sw2[9], Rm[9], R[9], RR[9], Rn[9];
      mat3_mul(sw, sw, sw2);
      for (int i = 0; i < 9; ++i) Rm[i] = a * sw[i] + (1 - b) * sw2[i];
      quat_to_R(S.pose_q, R);
      mat3_mul(R, Rm, RR);
      for (int i = 0; i < 9; ++i) Rn[i] = R[i] + RR[i];
      R_to_quat(Rn, S.new_q);
      for (int i = 0; i < 3; ++i) S.new_t[i] = S.pose_t[i] + (S.tb[i][0] * S.sol[3] + S.tb[i][1] * S.sol[4]);
    }
    __syncthreads();
    const double cost_new = cost_of(S.new_q, S.new_t);
    const bool accept = cost_new < cost;
    if (lane == 0) {
      if (accept) {
        for (int i = 0; i < 4; ++i) S.pose_q[i] = S.new_q[i];
        for (int i = 0; i < 3; ++i) S.pose_t[i] = S.new_t[i];
      } else {
        for (int k = 0; k < 5; ++k) S.JtJ[k][k] -= lambda;
      }
    }
    if (accept) {
      lambda /= 10;
      cost = cost_new;
      recompute = true;
    } else {
      lambda *= 10;
      recompute = false;
    }
    __syncthreads();
  }
  __syncthreads();
}

// One wave per pair: the driver's bundle adjustment of the inlier sample
// (fit_ns[f] >= 6 points) with weights max(0, 1 - r2 / post_sq[pair]) from
// the model wmodels[f] by inlier position.  Up to 40 models per fit at
// models_out[(40 f + j) 9] (the refined poses of every initial model).
__global__ void __launch_bounds__(64) k_ba_e(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                             const int32_t* __restrict__ fit_pair,
                                             const int64_t* __restrict__ fit_sample_off,
                                             const int32_t* __restrict__ fit_ns, const int32_t* __restrict__ samples,
                                             const double* const* __restrict__ wmodels,
                                             const double* __restrict__ post_sq, double* __restrict__ scratch,
                                             const int64_t* __restrict__ scratch_off, double* __restrict__ models_out,
                                             int8_t* __restrict__ ok_out) {
  __shared__ EssShared SH;
  __shared__ double init[10][9];
  __shared__ double poses[4][7];
  __shared__ int npose;
  const int f = blockIdx.x;
  const int lane = lane_id();
  const int pair = fit_pair[f];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int32_t* __restrict__ Sm = samples + fit_sample_off[f];
  const int ns = fit_ns[f];
  const double sq = post_sq[pair];
  const double* wm = wmodels[f];
  if (ns < 6) {
    if (lane == 0)
      for (int k = 0; k < 40; ++k) ok_out[40 * f + k] = 0;
    return;
  }
  double* wpos = scratch + scratch_off[f];
  double* rows = wpos + ns;
  for (int k = lane; k < ns; k += 64) {
    const double* p = P + 4 * (int64_t)Sm[k];
    const double v = 1.0 - f_sq_sampson(p, wm) / sq;
    const double w = 0 > v ? 0 : v;
    wpos[k] = w;
    e_row(p, w, rows + 9 * (int64_t)k);
  }
  wave_stewenius(rows, ns, SH);
  const int ninit = SH.nmodels > 0 ? SH.nmodels : 0;
  if (lane == 0)
    for (int k = 0; k < ninit; ++k)
      for (int i = 0; i < 9; ++i) init[k][i] = SH.models[k][i];
  __syncthreads();
  int out = 0;
  for (int k = 0; k < ninit; ++k) {
    if (lane == 0) {
      const double* p0 = P + 4 * (int64_t)Sm[0];
      double x1[3] = {p0[0], p0[1], 1.0}, x2[3] = {p0[2], p0[3], 1.0};
      normalise3(x1);
      normalise3(x2);
      RelPose ps[4];
      npose = motion_from_essential_dev(init[k], x1, x2, ps);
      for (int j = 0; j < npose; ++j) {
        for (int i = 0; i < 4; ++i) poses[j][i] = ps[j].q[i];
        for (int i = 0; i < 3; ++i) poses[j][4 + i] = ps[j].t[i];
      }
    }
    __syncthreads();
    const int np = npose;
    for (int j = 0; j < np; ++j) {
      if (lane == 0) {
        for (int i = 0; i < 4; ++i) SH.pose_q[i] = poses[j][i];
        for (int i = 0; i < 3; ++i) SH.pose_t[i] = poses[j][4 + i];
      }
      __syncthreads();
      wave_refine_relpose(P, Sm, ns, wpos, SH);
      if (lane == 0) {
        RelPose p;
        for (int i = 0; i < 4; ++i) p.q[i] = SH.pose_q[i];
        for (int i = 0; i < 3; ++i) p.t[i] = SH.pose_t[i];
        double E[9];
        essential_from_motion(p, E);
        if (out < 40)
          for (int i = 0; i < 9; ++i) models_out[(40 * (int64_t)f + out) * 9 + i] = E[i];
      }
      ++out;
      __syncthreads();
    }
  }
  if (lane == 0)
    for (int k = 0; k < 40; ++k) ok_out[40 * f + k] = k < out ? 2 : 0;
}

void launch_ba_e(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                 const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples,
                 const double* const* wmodels, const double* post_sq, double* scratch, const int64_t* scratch_off,
                 double* models_out, int8_t* ok_out, int num_fits) {
  if (num_fits <= 0) return;
  hipLaunchKernelGGL(k_ba_e, dim3(num_fits), dim3(64), 0, st, pts, pair_pt_off, fit_pair, fit_sample_off, fit_ns,
                     samples, wmodels, post_sq, scratch, scratch_off, models_out, ok_out);
}

// K^-1-normalised copy of the correspondences of the pairs flagged in
// norm_pair (two 3x3 inverses per pair in kinv, row-major); other pairs are
// copied as they are.  One thread per point.
__global__ void k_prepare_points(const double* __restrict__ pix, const int32_t* __restrict__ pt_pair,
                                 const int8_t* __restrict__ norm_pair, const double* __restrict__ kinv,
                                 int64_t total, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double* p = pix + 4 * i;
  double* o = out + 4 * i;
  const int pr = pt_pair[i];
  if (!norm_pair[pr]) {
    for (int k = 0; k < 4; ++k) o[k] = p[k];
    return;
  }
  const double* i1 = kinv + 18 * (int64_t)pr;
  const double* i2 = i1 + 9;
  o[0] = i1[0] * p[0] + i1[1] * p[1] + i1[2] * 1.0;
  o[1] = i1[3] * p[0] + i1[4] * p[1] + i1[5] * 1.0;
  o[2] = i2[0] * p[2] + i2[1] * p[3] + i2[2] * 1.0;
  o[3] = i2[3] * p[2] + i2[4] * p[3] + i2[5] * 1.0;
}

void launch_prepare_points(hipStream_t st, const double* pix, const int32_t* pt_pair, const int8_t* norm_pair,
                           const double* kinv, int64_t total, double* out) {
  if (total <= 0) return;
  const int64_t grid = (total + 255) / 256;
  hipLaunchKernelGGL(k_prepare_points, dim3((unsigned)grid), dim3(256), 0, st, pix, pt_pair, norm_pair, kinv, total,
                     out);
}

}  // namespace gcr
