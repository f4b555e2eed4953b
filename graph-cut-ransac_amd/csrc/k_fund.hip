// Fundamental-matrix kernels: seven-point minimal solver (one thread per
// candidate sample) and the non-minimal fit (one wavefront per fit).
//
// Minimal: FundamentalMatrixEstimator::estimateModel
// (fundamental_estimator.h:161-184) -> seven-point solver
// (solver_fundamental_matrix_seven_point.h:91-291) on raw pixel coordinates,
// then the oriented epipolar filter (fundamental_estimator.h:737-800).
//
// Non-minimal: estimateModelNonminimal (fundamental_estimator.h:574-618):
// Hartley normalisation, FundamentalMatrixBundleAdjustmentSolver
// (solver_fundamental_matrix_bundle_adjustment.h:114-178): eight-point
// initialisation (eight_point.h:93-182) and Levenberg-Marquardt on the
// factorised F with the truncated loss (bundle.cpp:252-339,
// jacobian_impl.h:468-614), then T2^T F T1, Frobenius normalisation and
// F33 >= 0.  The same kernel runs the driver's post-hoc bundle adjustment on
// raw coordinates (py/gcransac_python.cpp:1832-1887).
#include "dev_fundamental.hpp"
#include "kernels.hpp"

namespace gcr {

// status per model slot (3 per candidate): 2 = model kept, 0 = none.
// nmodels per candidate: number of kept models (compacted to slots 0..k-1).
__global__ void __launch_bounds__(256) k_f7_solve(const double* __restrict__ pts,
                                                  const int64_t* __restrict__ pair_pt_off,
                                                  const int32_t* __restrict__ cand_pair,
                                                  const int32_t* __restrict__ samples, int64_t num_cand,
                                                  double* __restrict__ models, int8_t* __restrict__ mstatus,
                                                  int8_t* __restrict__ nmodels) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= num_cand) return;
  const double* P = pts + 4 * pair_pt_off[cand_pair[c]];
  const int32_t* s = samples + 7 * c;
  double F[3][9];
  const int mask = thread_f7(P, s, F);
  // orientation filter, keeping the order of the remaining models
  int kept = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (mask > 0 && ((mask >> k) & 1) && f_orientation_valid_idx(F[k], P, s, 7)) {
      for (int i = 0; i < 9; ++i) models[(3 * c + kept) * 9 + i] = F[k][i];
      ++kept;
    }
  }
  for (int k = 0; k < 3; ++k) mstatus[3 * c + k] = k < kept ? 2 : 0;
  nmodels[c] = (int8_t)kept;
}

void launch_f7_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* mstatus, int8_t* nmodels) {
  if (num_cand <= 0) return;
  hipLaunchKernelGGL(k_f7_solve, dim3((unsigned)((num_cand + 255) / 256)), dim3(256), 0, st, pts, pair_pt_off,
                     cand_pair, samples, num_cand, models, mstatus, nmodels);
}

// ---------------------------------------------------------------------------
// Non-minimal fit
// ---------------------------------------------------------------------------
struct FitRows {
  // ns x 4 rows (lane i mod 64 wrote row i): rows [0, cap) in LDS, the rest
  // in the fit's scratch; cap is a multiple of 64 or >= ns
  const double* lds;
  const double* glob;
  int cap;
  int n;
  int weight_mode;       // 0 none, 1 W[k] (array by row position), 2 from wmodel
  const double* W;       // mode 1
  const double* wmodel;  // mode 2: w_k = max(0, 1 - sampson(raw row k) / wthr)
  double wthr;
  __device__ __forceinline__ const double* row(int64_t k) const { return k < cap ? lds + 4 * k : glob + 4 * k; }
  __device__ __forceinline__ double weight(int k) const {
    if (weight_mode == 1) return W[k];
    const double r2 = f_sq_sampson(row(k), wmodel);
    const double v = 1.0 - r2 / wthr;
    return 0 > v ? 0 : v;
  }
};

// Work arrays of the dense steps of one fit, in LDS (one wave per block):
// run-time indexed small matrices stay out of per-lane scratch, and the
// wave-uniform state of the Levenberg-Marquardt loop is held once.
struct FitShared {
  double hv[9][9];   // Householder vectors
  double tau[9];
  int rowt[9];
  double q[9];       // null vector
  double JtJ[7][7];  // lower triangle
  double sol[7];
  double U[9], V[9], sigma;     // current factorisation
  double Un[9], Vn[9], sigman;  // candidate step
  double F[3][9];    // models of the fit
};


// Householder QR with full pivoting of the 9 x 9 matrix A^T A (P A P' = Q R,
// Q = T0 H0 ... T8 H8; the last column of Q is the eight-point null vector)
// on one wave: lane j < 9 holds column j of the matrix in registers and
// keeps its column position (the column swaps of the pivoting exchange
// positions, not data).  Pivot search per column (rows in ascending order,
// strict >) then a wave argmax over the unprocessed columns (ties: smaller
// position), row swaps inside each lane, the reflector formed by the pivot
// column's lane and broadcast, every other unprocessed column updated by its
// own lane with the same operation order as the sequential version.  The
// reflectors go to S.hv / S.tau; thread 0 accumulates the last column of Q
// into S.q.  In a multi-wave block every wave runs it on the same values.
__device__ void wave_fullpiv_qr_nullvec(FitShared& S, double* col) {
  const int lane = lane_id();
  const bool live = lane < 9;
  int cpos = live ? lane : 64;  // this lane's column position
  int rowt[9];
  double tauv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    // pivot: first maximum in column-major scan over rows >= k, cols >= k
    double best = -1.0;
    int besti = k;
#pragma unroll
    for (int i = k; i < 9; ++i)
      if (fabs(col[i]) > best) {
        best = fabs(col[i]);
        besti = i;
      }
    double gbest;
    const bool open = live && cpos >= k;
    const int pc = wave_argmax_dpp(open ? best : -2.0, open ? cpos : 64 + lane, gbest);
    // the lane holding position pc becomes position k, the one at k takes pc
    const int wl = __ffsll((unsigned long long)__ballot(live && cpos == pc)) - 1;
    const int pr = __builtin_amdgcn_readlane(besti, wl);
    rowt[k] = pr;
    if (live && cpos == k) cpos = pc;
    if (lane == wl) cpos = k;
    // swap rows k and pr in every column
    if (pr != k) {
      double vpr = 0.0;
#pragma unroll
      for (int i = k + 1; i < 9; ++i)
        if (i == pr) vpr = col[i];
      const double vk = col[k];
#pragma unroll
      for (int i = k + 1; i < 9; ++i)
        if (i == pr) col[i] = vk;
      col[k] = vpr;
    }
    // reflector of the pivot column (formed by lane wl)
    double sig = 0.0;
#pragma unroll
    for (int i = k + 1; i < 9; ++i) sig += col[i] * col[i];
    const double alpha = readlane_f64(col[k], wl);
    const double sigma = readlane_f64(sig, wl);
    double tau = 0.0;
    double hv[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) hv[i] = 0.0;
    hv[k] = 1.0;
    if (sigma != 0.0) {
      const double nrm = sqrt(alpha * alpha + sigma);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      const double scale = 1.0 / (alpha - beta);
#pragma unroll
      for (int i = k + 1; i < 9; ++i) hv[i] = readlane_f64(col[i] * scale, wl);
      if (lane == wl) {
        col[k] = beta;
#pragma unroll
        for (int i = k + 1; i < 9; ++i) col[i] = 0.0;
      } else if (live && cpos > k) {
        double w = col[k];
#pragma unroll
        for (int i = k + 1; i < 9; ++i) w += hv[i] * col[i];
        const double tw = tau * w;
        col[k] = col[k] - tw;
#pragma unroll
        for (int i = k + 1; i < 9; ++i) col[i] = col[i] - tw * hv[i];
      }
    }
    tauv[k] = tau;
    if (threadIdx.x == 0) {
#pragma unroll
      for (int i = 0; i < 9; ++i) S.hv[k][i] = hv[i];
    }
  }
  if (threadIdx.x == 0) {
    double* q = S.q;
    for (int i = 0; i < 9; ++i) q[i] = 0.0;
    q[8] = 1.0;
#pragma unroll
    for (int k = 8; k >= 0; --k) {
      if (tauv[k] != 0.0) {
        double w = 0.0;
        for (int i = k; i < 9; ++i) w += S.hv[k][i] * q[i];
        const double tw = tauv[k] * w;
        for (int i = k; i < 9; ++i) q[i] = q[i] - tw * S.hv[k][i];
      }
      if (rowt[k] != k) {
        const double t = q[k];
        q[k] = q[rowt[k]];
        q[rowt[k]] = t;
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ double f8_coef(const double* p, int k, bool weighted, double w) {
  const double x0 = p[0], y0 = p[1], x1 = p[2], y1 = p[3];
  if (!weighted) {
    switch (k) {
      case 0: return x1 * x0;
      case 1: return x1 * y0;
      case 2: return x1;
      case 3: return y1 * x0;
      case 4: return y1 * y0;
      case 5: return y1;
      case 6: return x0;
      case 7: return y0;
      default: return 1.0;
    }
  }
  const double wx0 = w * x0, wy0 = w * y0, wx1 = w * x1, wy1 = w * y1;
  switch (k) {
    case 0: return wx1 * x0;
    case 1: return wx1 * y0;
    case 2: return wx1;
    case 3: return wy1 * x0;
    case 4: return wy1 * y0;
    case 5: return wy1;
    case 6: return wx0;
    case 7: return wy0;
    default: return w;
  }
}

// Eight-point: A^T A by the 64-lane order, then the QR null vector.
template <int NT>
__device__ void wave_f8(const FitRows& R, FitShared& S, double* Fout, const BlockRed& B) {
  const int lane = lane_id();
  const bool weighted = R.weight_mode != 0;
  // lane j < 9 keeps column j of the symmetric A^T A in registers
  double col[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) col[i] = 0.0;
  if constexpr (NT == 64) {
    // at most 64 rows: one row per lane; each of the 45 sums is this lane's
    // 0.0 + c_a c_b reduced on its own (the general path's operations and
    // order), one row a of the matrix at a time
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = 0.0;
    const bool has = lane < R.n;
    if (has) {
      const double* p = R.row(lane);
      const double w = weighted ? R.weight(lane) : 1.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) c[k] = f8_coef(p, k, weighted, w);
    }
    // the 45 sums (a, b >= a) in row order, four at a time
    constexpr int kA[45] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2,
                            3, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 7, 7, 8};
    constexpr int kB[45] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 1, 2, 3, 4, 5, 6, 7, 8, 2, 3, 4, 5, 6, 7, 8,
                            3, 4, 5, 6, 7, 8, 4, 5, 6, 7, 8, 5, 6, 7, 8, 6, 7, 8, 7, 8, 8};
    auto term = [&](int e) -> double { return (e < 45 && has) ? 0.0 + c[kA[e]] * c[kB[e]] : 0.0; };
#pragma unroll
    for (int g = 0; g < 48; g += 4) {
      const double z = wave_tree64x4(term(g), term(g + 1), term(g + 2), term(g + 3));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = g + r;
        if (e < 45) {
          const double v = readlane_f64(z, x4_lane(r));
          if (lane == kB[e]) col[kA[e]] = v;
          if (lane == kA[e]) col[kB[e]] = v;
        }
      }
    }
  } else {
  double acc[45];
#pragma unroll
  for (int e = 0; e < 45; ++e) acc[e] = 0.0;
  for (int i = threadIdx.x; i < R.n; i += NT) {
    const double* p = R.row(i);
    const double w = weighted ? R.weight(i) : 1.0;
    double c[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = f8_coef(p, k, weighted, w);
    int o = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
      for (int b = a; b < 9; ++b) {
        acc[o] = acc[o] + c[a] * c[b];
        ++o;
      }
  }
  block_tree<NT, 45>(acc, B.red, B.tot);
  {
    int o = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
      for (int b = a; b < 9; ++b) {
        const double v = acc[o];
        if (lane == b) col[a] = v;
        if (lane == a) col[b] = v;
        ++o;
      }
  }
  }
  wave_fullpiv_qr_nullvec(S, col);
  if (threadIdx.x == 0)
    for (int i = 0; i < 9; ++i) Fout[i] = S.q[i];
  __syncthreads();
}

// F = U diag(1, sigma, 0) V^T of the factorisation (U, V row-major).
__device__ __forceinline__ void fact_to_F(const double* U, const double* V, double sigma, double* out) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) out[3 * i + j] = U[3 * i] * V[3 * j] + sigma * U[3 * i + 1] * V[3 * j + 1];
}

template <int NT>
__device__ double lm_cost(const FitRows& R, const double* F, const BlockRed& B) {
  const bool weighted = R.weight_mode != 0;
  double acc[1] = {0.0};
  for (int k = threadIdx.x; k < R.n; k += NT) {
    const double* p = R.row(k);
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
    acc[0] = acc[0] + (weighted ? R.weight(k) * l : l);
  }
  block_tree<NT, 1>(acc, B.red, B.tot);
  return acc[0];
}

// U_out = U_in + (a [w]x + (1 - b) [w]x^2) U_in (the rotation update of the
// factorised F for the step w with |w| = theta, a = sin theta, b = cos theta).
__device__ void rot_update(const double* w, double a, double b, const double* Uin, double* Uout) {
  const double sw[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
  double sw2[9], Rm[9], RU[9];
  mat3_mul(sw, sw, sw2);
  for (int i = 0; i < 9; ++i) Rm[i] = a * sw[i] + (1 - b) * sw2[i];
  mat3_mul(Rm, Uin, RU);
  for (int i = 0; i < 9; ++i) Uout[i] = Uin[i] + RU[i];
}

// Levenberg-Marquardt refinement (25 iterations, truncated loss, scale 1) of
// the model in Fio (LDS).  The sums over points are wave-parallel; the dense
// 7 x 7 steps run on lane 0 with their state in S, and every value the whole
// wave needs is read back from S after a barrier.
template <int NT>
__device__ void wave_lm_f(const FitRows& R, FitShared& S, double* Fio, int& jac_passes, int& cost_passes,
                          const BlockRed& B) {
  const bool weighted = R.weight_mode != 0;
  const bool t0 = threadIdx.x == 0;
  if (t0) {
    double sv[3];
    svd3(Fio, S.U, sv, S.V);
    S.sigma = sv[1] / sv[0];
  }
  __syncthreads();
  double lambda = 1e-3;
  double Fm[9];
  fact_to_F(S.U, S.V, S.sigma, Fm);
  double cost = lm_cost<NT>(R, Fm, B);
  ++cost_passes;
  double Jtr[7];
  bool recompute = true;
  for (int iter = 0; iter < 25; ++iter) {
    if (recompute) {
      ++jac_passes;
      fact_to_F(S.U, S.V, S.sigma, Fm);
#define FE(i, j) Fm[3 * (i) + (j)]
#define DS(i, j) ds[3 * (i) + (j)]
// rows: dF/dvec(F) -> (rotation of U (3), rotation of V (3), sigma)
#define PROWS                                                                            \
  {{0, FE(2, 0), -FE(1, 0), 0, FE(0, 2), -FE(0, 1), DS(0, 0)},                           \
   {-FE(2, 0), 0, FE(0, 0), 0, FE(1, 2), -FE(1, 1), DS(1, 0)},                           \
   {FE(1, 0), -FE(0, 0), 0, 0, FE(2, 2), -FE(2, 1), DS(2, 0)},                           \
   {0, FE(2, 1), -FE(1, 1), -FE(0, 2), 0, FE(0, 0), DS(0, 1)},                           \
   {-FE(2, 1), 0, FE(0, 1), -FE(1, 2), 0, FE(1, 0), DS(1, 1)},                           \
   {FE(1, 1), -FE(0, 1), 0, -FE(2, 2), 0, FE(2, 0), DS(2, 1)},                           \
   {0, FE(2, 2), -FE(1, 2), FE(0, 1), -FE(0, 0), 0, DS(0, 2)},                           \
   {-FE(2, 2), 0, FE(0, 2), FE(1, 1), -FE(1, 0), 0, DS(1, 2)},                           \
   {FE(1, 2), -FE(0, 2), 0, FE(2, 1), -FE(2, 0), 0, DS(2, 2)}}
      if constexpr (NT == 64) {
        // fits of at most 64 points: one point per lane, so each of the 35
        // sums is this lane's term (0.0 + term, as the running partial of the
        // general path holds it) reduced on its own: the same operations and
        // reduction order as accumulating all 35 partials first.  Every lane
        // forms the map dF / d(parameters) in registers (its zero entries
        // multiply as in the dense product), and the sums are reduced four at
        // a time.
        double ds[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) ds[3 * i + j] = S.U[3 * i + 1] * S.V[3 * j + 1];
        const double Pm[9][7] = PROWS;
        double J[7], wres = 0.0, w = 0.0;
#pragma unroll
        for (int cc = 0; cc < 7; ++cc) J[cc] = 0.0;
        const int k = threadIdx.x;
        if (k < R.n) {
          const double* p = R.row(k);
          const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
          const double f0 = FE(0, 0) * x1 + FE(0, 1) * y1 + FE(0, 2);
          const double f1 = FE(1, 0) * x1 + FE(1, 1) * y1 + FE(1, 2);
          const double f2 = FE(2, 0) * x1 + FE(2, 1) * y1 + FE(2, 2);
          const double C = x2 * f0 + y2 * f1 + f2;
          const double J0 = FE(0, 0) * x2 + FE(1, 0) * y2 + FE(2, 0);
          const double J1 = FE(0, 1) * x2 + FE(1, 1) * y2 + FE(2, 1);
          const double J2 = f0, J3 = f1;
          const double nJ = sqrt(J0 * J0 + J1 * J1 + J2 * J2 + J3 * J3);
          const double inv = 1.0 / nJ;
          const double r = C * inv;
          w = ((r * r) < 1.0 ? 1.0 : 0.0) / (double)R.n;
          if (weighted) w = R.weight(k) * w;
          if (w != 0.0) {
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
#pragma unroll
            for (int q = 0; q < 9; ++q) dF[q] *= inv;
#pragma unroll
            for (int cc = 0; cc < 7; ++cc) {
              double v = 0.0;
#pragma unroll
              for (int q = 0; q < 9; ++q) v += dF[q] * Pm[q][cc];
              J[cc] = v;
            }
            wres = w * C * inv;
          }
        }
        // term e: Jtr (sum_k wres_k J_k[e]) for e < 7, then JtJ row by row
        // (sum_k w_k J_k[i] J_k[j], j <= i)
        constexpr int kTi[35] = {0, 1, 2, 3, 4, 5, 6, 0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4,
                                 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 6};
        constexpr int kTj[35] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0,
                                 1, 2, 3, 4, 0, 1, 2, 3, 4, 5, 0, 1, 2, 3, 4, 5, 6};
        auto term = [&](int e) -> double {
          if (e >= 35 || w == 0.0) return 0.0;
          if (e < 7) return 0.0 + wres * J[e];
          return 0.0 + w * (J[kTi[e]] * J[kTj[e]]);
        };
        const int lane = lane_id();
#pragma unroll
        for (int g = 0; g < 36; g += 4) {
          const double z = wave_tree64x4(term(g), term(g + 1), term(g + 2), term(g + 3));
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int e = g + r;
            if (e < 7) {
              Jtr[e] = readlane_f64(z, x4_lane(r));
            } else if (e < 35) {
              if (lane == x4_lane(r)) S.JtJ[kTi[e]][kTj[e]] = z;
            }
          }
        }
      } else {
        // the map dF / d(parameters) in registers on every thread, as above
        double ds[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) ds[3 * i + j] = S.U[3 * i + 1] * S.V[3 * j + 1];
        const double Pm[9][7] = PROWS;
        double acc[35];
  #pragma unroll
        for (int e = 0; e < 35; ++e) acc[e] = 0.0;
        for (int k = threadIdx.x; k < R.n; k += NT) {
          const double* p = R.row(k);
          const double x1 = p[0], y1 = p[1], x2 = p[2], y2 = p[3];
          const double f0 = FE(0, 0) * x1 + FE(0, 1) * y1 + FE(0, 2);
          const double f1 = FE(1, 0) * x1 + FE(1, 1) * y1 + FE(1, 2);
          const double f2 = FE(2, 0) * x1 + FE(2, 1) * y1 + FE(2, 2);
          const double C = x2 * f0 + y2 * f1 + f2;
          const double J0 = FE(0, 0) * x2 + FE(1, 0) * y2 + FE(2, 0);
          const double J1 = FE(0, 1) * x2 + FE(1, 1) * y2 + FE(2, 1);
          const double J2 = f0, J3 = f1;
          const double nJ = sqrt(J0 * J0 + J1 * J1 + J2 * J2 + J3 * J3);
          const double inv = 1.0 / nJ;
          const double r = C * inv;
          double w = ((r * r) < 1.0 ? 1.0 : 0.0) / (double)R.n;
          if (weighted) w = R.weight(k) * w;
          // a zero weight adds +0.0 to partials that are never -0.0: skipping
          // the point leaves every partial unchanged
          if (w != 0.0) {
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
  #pragma unroll
            for (int q = 0; q < 9; ++q) dF[q] *= inv;
            double J[7];
  #pragma unroll
            for (int cc = 0; cc < 7; ++cc) {
              double v = 0.0;
  #pragma unroll
              for (int q = 0; q < 9; ++q) v += dF[q] * Pm[q][cc];
              J[cc] = v;
            }
            const double wres = w * C * inv;
            int o = 0;
  #pragma unroll
            for (int i = 0; i < 7; ++i) {
              acc[o] = acc[o] + wres * J[i];
              ++o;
            }
  #pragma unroll
            for (int i = 0; i < 7; ++i)
  #pragma unroll
              for (int j = 0; j <= i; ++j) {
                acc[o] = acc[o] + w * (J[i] * J[j]);
                ++o;
              }
          }
        }
        block_tree<NT, 35>(acc, B.red, B.tot);
        {
          int o = 0;
  #pragma unroll
          for (int i = 0; i < 7; ++i) Jtr[i] = acc[o++];
  #pragma unroll
          for (int i = 0; i < 7; ++i)
  #pragma unroll
            for (int j = 0; j <= i; ++j) {
              const double v = acc[o++];
              if (t0) S.JtJ[i][j] = v;
            }
        }
      }
#undef PROWS
#undef DS
#undef FE
      double gn = 0.0;
      for (int i = 0; i < 7; ++i) gn += Jtr[i] * Jtr[i];
      if (sqrt(gn) < 1e-8) break;
    }
    __syncthreads();  // JtJ written by the lanes that hold its totals
    if (t0)
      for (int k = 0; k < 7; ++k) S.JtJ[k][k] += lambda;  // the LDS copy keeps the reject bookkeeping
    __syncthreads();
    if (threadIdx.x < 64) {
      // Cholesky of the damped lower triangle, column by column: every lane
      // forms the diagonal (the same value everywhere) from row j, lane i > j
      // forms L[i][j] from its own row i, and the column is broadcast.  The
      // forward substitution runs beside the factorisation (y[j] once row j
      // is complete), the back substitution after it, both on every lane
      // without branches.  Every element keeps the sequential order of its
      // sums.
      const int l = threadIdx.x;
      double a[7], Lr[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        a[j] = (l < 7 && j <= l) ? S.JtJ[l < 7 ? l : 0][j] : 0.0;
        Lr[j] = 0.0;
      }
      double L[7][7], y[7], sol[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        double d = S.JtJ[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        L[j][j] = sqrt(d);
        double v = a[j];
#pragma unroll
        for (int k = 0; k < j; ++k) v -= Lr[k] * L[j][k];
        const double lij = v / L[j][j];
        if (l > j) Lr[j] = lij;
#pragma unroll
        for (int i = j + 1; i < 7; ++i) L[i][j] = readlane_f64(Lr[j], i);
        double u = -Jtr[j];
#pragma unroll
        for (int k = 0; k < j; ++k) u -= L[j][k] * y[k];
        y[j] = u / L[j][j];
      }
#pragma unroll
      for (int i = 6; i >= 0; --i) {
        double v = y[i];
#pragma unroll
        for (int k = i + 1; k < 7; ++k) v -= L[k][i] * sol[k];
        sol[i] = v / L[i][i];
      }
      if (l == 0) {
#pragma unroll
        for (int i = 0; i < 7; ++i) S.sol[i] = sol[i];
      }
    }
    __syncthreads();
    double sn = 0.0;
    for (int i = 0; i < 7; ++i) sn += S.sol[i] * S.sol[i];
    if (sqrt(sn) < 1e-8) break;
    if (threadIdx.x < 64) {
      // the rotation steps of U (lane 0) and V (lane 1) side by side.  The
      // divisions they need, per side 3 axis components w_i / theta and the
      // 17 series quotients of sin / cos theta, are independent: lane
      // 20 side + j forms division j of that side, and lanes 0 and 1 fetch
      // their side's 20.
      const int l = threadIdx.x;
      double thv[2], kv[2], rv[2], r2v[2];
#pragma unroll
      for (int sd = 0; sd < 2; ++sd) {
        const double w0 = S.sol[3 * sd], w1 = S.sol[3 * sd + 1], w2 = S.sol[3 * sd + 2];
        thv[sd] = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        sct_reduce(thv[sd], kv[sd], rv[sd], r2v[sd]);
      }
      const int sd = l >= 20 ? 1 : 0;
      const int j = l - 20 * sd;
      const int jc = j < 20 ? j : 19;
      const double num = jc < 3 ? S.sol[3 * sd + jc] : (sd ? r2v[1] : r2v[0]);
      const double den = jc < 3 ? (sd ? thv[1] : thv[0]) : sct_denominator(jc - 3);
      const double qt = num / den;
      const int src = l < 2 ? 20 * l : 0;
      double w[3], q[17];
#pragma unroll
      for (int i = 0; i < 3; ++i) w[i] = __shfl(qt, src + i, 64);
#pragma unroll
      for (int m = 0; m < 17; ++m) q[m] = __shfl(qt, src + 3 + m, 64);
      if (l < 2) {
        double a, b;
        sct_finish(l == 0 ? kv[0] : kv[1], l == 0 ? rv[0] : rv[1], q, a, b);
        rot_update(w, a, b, l == 0 ? S.U : S.V, l == 0 ? S.Un : S.Vn);
        if (l == 0) S.sigman = S.sigma + S.sol[6];
      }
    }
    __syncthreads();
    double Fnm[9];
    fact_to_F(S.Un, S.Vn, S.sigman, Fnm);
    const double cost_new = lm_cost<NT>(R, Fnm, B);
    ++cost_passes;
    const bool accept = cost_new < cost;
    if (t0) {
      if (accept) {
        for (int i = 0; i < 9; ++i) {
          S.U[i] = S.Un[i];
          S.V[i] = S.Vn[i];
        }
        S.sigma = S.sigman;
      } else {
        for (int k = 0; k < 7; ++k) S.JtJ[k][k] -= lambda;
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
  fact_to_F(S.U, S.V, S.sigma, Fm);
  __syncthreads();
  if (t0)
    for (int i = 0; i < 9; ++i) Fio[i] = Fm[i];
  __syncthreads();
}

// One wave per fit; up to 3 models per fit at models_out[(3 f + j) * 9].
// flags bit 0: Hartley-normalise (and denormalise, Frobenius-normalise,
// F33 >= 0 afterwards); bit 1: the plain eight-point solver (no seven-point
// branch, no LM refinement; fewer than 8 points fail); bits 8-9: weight mode.
template <int NT>
__global__ void __launch_bounds__(NT) k_fit_f(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                                              const int32_t* __restrict__ fit_pair,
                                              const int64_t* __restrict__ fit_sample_off,
                                              const int32_t* __restrict__ fit_ns,
                                              const int32_t* __restrict__ samples,
                                              const double* __restrict__ weights,
                                              const int64_t* __restrict__ weight_off,
                                              const double* const* __restrict__ wmodels,
                                              const double* __restrict__ pair_sq_thr, int flags,
                                              double* __restrict__ scratch, const int64_t* __restrict__ scratch_off,
                                              double* __restrict__ models_out, int8_t* __restrict__ ok_out,
                                              int32_t* __restrict__ lm_passes, int lds_rows) {
  __shared__ FitShared SH;
  const int f = blockIdx.x;
  const int lane = lane_id();
  const int pair = fit_pair[f];
  const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
  const int32_t* __restrict__ S = samples + fit_sample_off[f];
  const int ns = fit_ns[f];
  const bool normalize = flags & 1;
  const bool refine = !(flags & 2);
  const int wmode = (flags >> 8) & 3;
  const bool t0 = threadIdx.x == 0;
  if (ns < (refine ? 7 : 8)) {
    if (t0) {
      for (int k = 0; k < 3; ++k) ok_out[3 * f + k] = 0;
      if (lm_passes) {
        lm_passes[2 * f] = 0;
        lm_passes[2 * f + 1] = 0;
      }
    }
    return;
  }
  // normalised rows: the first lds_rows in LDS, the rest in scratch
  extern __shared__ double lds_pts[];
  double* const rows_g = scratch + scratch_off[f];
  // reduction staging of a 256-thread block after the rows
  const BlockRed B{lds_pts + 4 * (int64_t)lds_rows, lds_pts + 4 * (int64_t)lds_rows + kRedValues * 128};
  auto wrow = [&](int i) { return i < lds_rows ? lds_pts + 4 * i : rows_g + 4 * (int64_t)i; };
  double T1[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T2[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  // the point of sample entry `lane`, gathered once (entries 0 .. 63 of
  // every pass below are read from these registers)
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
  if (lane < ns) {
    const double* d = P + 4 * (int64_t)S[lane];
    g0 = d[0];
    g1 = d[1];
    g2 = d[2];
    g3 = d[3];
  }
  auto point = [&](int i, double* v) {
    if (i < 64) {
      v[0] = g0;
      v[1] = g1;
      v[2] = g2;
      v[3] = g3;
    } else {
      const double* d = P + 4 * (int64_t)S[i];
      v[0] = d[0];
      v[1] = d[1];
      v[2] = d[2];
      v[3] = d[3];
    }
  };
  if (normalize) {
    // Mass points and mean distances: sequential sums in sample order
    // (homography_estimator.h:236-270), evaluated by the whole wave.
    double msum[4];
    wave_seq_sum4(ns, [&](int i, double* v) { point(i, v); }, msum);
    const double ms0 = msum[0] / ns, ms1 = msum[1] / ns;
    const double md0 = msum[2] / ns, md1 = msum[3] / ns;
    double dsum[4];
    wave_seq_sum4(ns, [&](int i, double* v) {
      double d[4];
      point(i, d);
      const double dx1 = ms0 - d[0], dy1 = ms1 - d[1], dx2 = md0 - d[2], dy2 = md1 - d[3];
      v[0] = sqrt(dx1 * dx1 + dy1 * dy1);
      v[1] = sqrt(dx2 * dx2 + dy2 * dy2);
      v[2] = 0.0;
      v[3] = 0.0;
    }, dsum);
    const double as = dsum[0] / ns, ad = dsum[1] / ns;
    const double rs = 1.4142135623730951 / as, rd = 1.4142135623730951 / ad;
    for (int i = threadIdx.x; i < ns; i += NT) {
      double d[4];
      point(i, d);
      double* o = wrow(i);
      o[0] = (d[0] - ms0) * rs;
      o[1] = (d[1] - ms1) * rs;
      o[2] = (d[2] - md0) * rd;
      o[3] = (d[3] - md1) * rd;
    }
    const double t1[9] = {rs, 0, -rs * ms0, 0, rs, -rs * ms1, 0, 0, 1};
    const double t2[9] = {rd, 0, -rd * md0, 0, rd, -rd * md1, 0, 0, 1};
    for (int i = 0; i < 9; ++i) {
      T1[i] = t1[i];
      T2[i] = t2[i];
    }
  } else {
    for (int i = threadIdx.x; i < ns; i += NT) {
      double d[4];
      point(i, d);
      double* o = wrow(i);
      for (int k = 0; k < 4; ++k) o[k] = d[k];
    }
  }
  __syncthreads();
  FitRows R;
  R.lds = lds_pts;
  R.glob = rows_g;
  R.cap = lds_rows;
  R.n = ns;
  R.weight_mode = wmode;
  R.W = (wmode == 1) ? weights + weight_off[pair] : nullptr;
  R.wmodel = (wmode == 2) ? wmodels[f] : nullptr;
  R.wthr = pair_sq_thr[pair];
  // models of the fit in SH.F, in slot order; mask = the slots in use
  int mask = 0;
  if (ns >= 8) {
    wave_f8<NT>(R, SH, SH.F[0], B);
    mask = 1;
  } else if constexpr (NT == 64) {
    __syncthreads();  // rows are written by lane (row mod 64)
    double x0 = 0.0, y0 = 0.0, x1 = 0.0, y1 = 0.0;
    if (lane < 63) {
      const double* d = R.row(lane / 9);
      x0 = d[0];
      y0 = d[1];
      x1 = d[2];
      y1 = d[3];
    }
    double Fs[3][9];
    const int m7 = wave_f7(x0, y0, x1, y1, Fs);
    if (m7 > 0) {
      mask = m7;
      if (t0) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if ((m7 >> k) & 1)
            for (int i = 0; i < 9; ++i) SH.F[k][i] = Fs[k][i];
      }
    }
    __syncthreads();
  }
  // output slot of model k: the number of models before it (root order)
  int slot = 0;
  int jac_passes = 0, cost_passes = 0;
  for (int k = 0; k < 3; ++k) {
    if (!((mask >> k) & 1)) continue;
    double* Fk = SH.F[k];
    if (refine) wave_lm_f<NT>(R, SH, Fk, jac_passes, cost_passes, B);
    if (t0) {
      double G[9];
      for (int i = 0; i < 9; ++i) G[i] = Fk[i];
      if (normalize) {
        double T2t[9], tmp[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) T2t[3 * j + i] = T2[3 * i + j];
        mat3_mul(T2t, G, tmp);
        mat3_mul(tmp, T1, G);
        double sq = 0.0;
        for (int i = 0; i < 9; ++i) sq += G[i] * G[i];
        sq = sqrt(sq);
        for (int i = 0; i < 9; ++i) G[i] /= sq;
        if (G[8] < 0)
          for (int i = 0; i < 9; ++i) G[i] = -G[i];
      }
      for (int i = 0; i < 9; ++i) models_out[(3 * (int64_t)f + slot) * 9 + i] = G[i];
    }
    ++slot;
  }
  if (t0) {
    for (int k = 0; k < 3; ++k) ok_out[3 * f + k] = k < slot ? 2 : 0;
    // point passes of the refinement (work accounting of the profiler)
    if (lm_passes) {
      lm_passes[2 * f] = jac_passes;
      lm_passes[2 * f + 1] = cost_passes;
    }
  }
}

void launch_fit_f(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, const double* const* wmodels, const double* pair_sq_thr, int flags,
                  double* scratch, const int64_t* scratch_off, double* models_out, int8_t* ok_out, int num_fits,
                  int32_t* lm_passes, int max_ns) {
  if (num_fits <= 0) return;
  // fits of more than 64 points run on four waves (256-lane sums), the
  // others on one.  The launch's largest sample goes to LDS when it fits
  // (next to the static FitShared and the reduction staging), else its first
  // rows (a multiple of 64) with the rest in scratch.
  const bool wide = max_ns > 64;
  const int red_bytes = wide ? (kRedValues * 128 + kRedValues) * (int)sizeof(double) : 0;
  const int avail = lds_limit_bytes() - (int)sizeof(FitShared) - red_bytes - 256;
  const int limit_rows = (avail > 0 ? avail / (4 * (int)sizeof(double)) : 0) & ~63;
  const int lds_rows = max_ns <= limit_rows ? max_ns : limit_rows;
  const size_t shmem = (size_t)(lds_rows > 0 ? lds_rows : 1) * 4 * sizeof(double) + (size_t)red_bytes;
  if (wide)
    hipLaunchKernelGGL(k_fit_f<256>, dim3(num_fits), dim3(256), shmem, st, pts, pair_pt_off, fit_pair,
                       fit_sample_off, fit_ns, samples, weights, weight_off, wmodels, pair_sq_thr, flags, scratch,
                       scratch_off, models_out, ok_out, lm_passes, lds_rows);
  else
    hipLaunchKernelGGL(k_fit_f<64>, dim3(num_fits), dim3(64), shmem, st, pts, pair_pt_off, fit_pair,
                       fit_sample_off, fit_ns, samples, weights, weight_off, wmodels, pair_sq_thr, flags, scratch,
                       scratch_off, models_out, ok_out, lm_passes, lds_rows);
}

// ---------------------------------------------------------------------------
// DEGENSAC (fundamental_estimator.h:341-572)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cross3v(const double* u, const double* v, double* r) {
  r[0] = u[1] * v[2] - u[2] * v[1];
  r[1] = u[2] * v[0] - u[0] * v[2];
  r[2] = u[0] * v[1] - u[1] * v[0];
}
__device__ __forceinline__ double sqnorm3(const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

// H-degenerate sample test of every kept seven-point model, one thread per
// model slot (slot / 3 = candidate).  The epipole e (e_3 = 1) is the largest
// cross product of two columns of F (the left null vector), A = [e]x F; each
// of the five sample triplets defines H = A - e (M^-1 b)^T (M: the triplet's
// first-image points as rows, b_k = ((x2 x A x1) . (x2 x e)) / |x2 x e|^2),
// and the sample is H-degenerate when that H maps at least five of the seven
// points within the homography threshold (squared transfer error < 4).
// degen[slot] = 1 and degen_h[slot] = the first such H.
__global__ void k_degen_test(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                             const int32_t* __restrict__ cand_pair, const int32_t* __restrict__ samples,
                             const double* __restrict__ models, const int8_t* __restrict__ status,
                             int64_t num_slots, int8_t* __restrict__ degen, double* __restrict__ degen_h) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= num_slots) return;
  degen[q] = 0;
  if (status[q] != 2) return;
  const int64_t c = q / 3;
  const double* P = pts + 4 * pair_pt_off[cand_pair[c]];
  const int32_t* smp = samples + 7 * c;
  const double* F = models + 9 * q;
  const double c0[3] = {F[0], F[3], F[6]}, c1[3] = {F[1], F[4], F[7]}, c2[3] = {F[2], F[5], F[8]};
  double e01[3], e02[3], e12[3];
  cross3v(c0, c1, e01);
  cross3v(c0, c2, e02);
  cross3v(c1, c2, e12);
  const double* eb = e01;
  if (sqnorm3(e02) > sqnorm3(eb)) eb = e02;
  if (sqnorm3(e12) > sqnorm3(eb)) eb = e12;
  const double e[3] = {eb[0] / eb[2], eb[1] / eb[2], 1.0};
  const double ex[9] = {0, -e[2], e[1], e[2], 0, -e[0], -e[1], e[0], 0};
  double A[9];
  mat3_mul(ex, F, A);
  constexpr int kTriplets[15] = {0, 1, 2, 3, 4, 5, 0, 1, 6, 3, 4, 6, 2, 5, 6};
  for (int t = 0; t < 5; ++t) {
    const int32_t i1 = smp[kTriplets[3 * t]], i2 = smp[kTriplets[3 * t + 1]], i3 = smp[kTriplets[3 * t + 2]];
    const int32_t ids[3] = {i1, i2, i3};
    double b[3], M[9];
    for (int k = 0; k < 3; ++k) {
      const double* pk = P + 4 * (int64_t)ids[k];
      const double x1[3] = {pk[0], pk[1], 1.0}, x2[3] = {pk[2], pk[3], 1.0};
      double ce[3], Ax[3], c2x[3];
      cross3v(x2, e, ce);
      for (int r = 0; r < 3; ++r) Ax[r] = A[3 * r] * x1[0] + A[3 * r + 1] * x1[1] + A[3 * r + 2] * x1[2];
      cross3v(x2, Ax, c2x);
      b[k] = (c2x[0] * ce[0] + c2x[1] * ce[1] + c2x[2] * ce[2]) / sqnorm3(ce);
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
      for (int cc = 0; cc < 3; ++cc) Hm[3 * r + cc] = A[3 * r + cc] - e[r] * v[cc];
    int cnt = 3;
    for (int k = 0; k < 7; ++k) {
      const int32_t idx = smp[k];
      if (idx == i1 || idx == i2 || idx == i3) continue;
      if (h_sq_residual(P + 4 * (int64_t)idx, Hm) < 4.0) ++cnt;
    }
    if (cnt >= 5) {
      degen[q] = 1;
      for (int i = 0; i < 9; ++i) degen_h[9 * q + i] = Hm[i];
      return;
    }
  }
}

// Plane-and-parallax minimal solve (solver_fundamental_matrix_plane_and_
// parallax.h) of DEGENSAC's nested run, one thread per candidate: the
// epipole is the intersection of the lines (H x1) x x2 of the two
// correspondences, F = [e]x H, then the estimator's oriented-epipolar filter
// on the two points (fundamental_estimator.h:161-184).  pair_h[pair]: the
// pair's homography (row-major).  status 2 = model kept, 1 = degenerate
// epipole, 0 = rejected.
__global__ void k_pp_solve(const double* __restrict__ pts, const int64_t* __restrict__ pair_pt_off,
                           const int32_t* __restrict__ cand_pair, const int32_t* __restrict__ samples,
                           const double* const* __restrict__ pair_h, int64_t num_cand, double* __restrict__ models,
                           int8_t* __restrict__ mstatus, int8_t* __restrict__ nmodels) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= num_cand) return;
  mstatus[c] = 0;
  nmodels[c] = 0;
  const int pair = cand_pair[c];
  const int32_t* s = samples + 2 * c;
  const double* H = pair_h[pair];
  if (s[0] < 0 || s[1] < 0 || H == nullptr) return;
  const double* P = pts + 4 * pair_pt_off[pair];
  const double* p1 = P + 4 * (int64_t)s[0];
  const double* p2 = P + 4 * (int64_t)s[1];
  double a1[3], a2[3];
  for (int i = 0; i < 3; ++i) {
    a1[i] = H[3 * i] * p1[0] + H[3 * i + 1] * p1[1] + H[3 * i + 2] * 1.0;
    a2[i] = H[3 * i] * p2[0] + H[3 * i + 1] * p2[1] + H[3 * i + 2] * 1.0;
  }
  const double d1[3] = {p1[2], p1[3], 1.0}, d2[3] = {p2[2], p2[3], 1.0};
  double l1[3], l2[3], e[3];
  cross3v(a1, d1, l1);
  cross3v(a2, d2, l2);
  cross3v(l1, l2, e);
  if (fabs(e[2]) < 2.220446049250313e-16) {
    mstatus[c] = 1;
    return;
  }
  const double ex[9] = {0, -e[2], e[1], e[2], 0, -e[0], -e[1], e[0], 0};
  double F[9];
  mat3_mul(ex, H, F);
  if (!f_orientation_valid_idx(F, P, s, 2)) return;
  for (int i = 0; i < 9; ++i) models[9 * c + i] = F[i];
  mstatus[c] = 2;
  nmodels[c] = 1;
}

void launch_degen_test(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                       const int32_t* samples, const double* models, const int8_t* status, int64_t num_slots,
                       int8_t* degen, double* degen_h) {
  if (num_slots <= 0) return;
  hipLaunchKernelGGL(k_degen_test, dim3((unsigned)((num_slots + 255) / 256)), dim3(256), 0, st, pts, pair_pt_off,
                     cand_pair, samples, models, status, num_slots, degen, degen_h);
}

void launch_pp_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, const double* const* pair_h, int64_t num_cand, double* models,
                     int8_t* mstatus, int8_t* nmodels) {
  if (num_cand <= 0) return;
  hipLaunchKernelGGL(k_pp_solve, dim3((unsigned)((num_cand + 255) / 256)), dim3(256), 0, st, pts, pair_pt_off,
                     cand_pair, samples, pair_h, num_cand, models, mstatus, nmodels);
}

}  // namespace gcr
