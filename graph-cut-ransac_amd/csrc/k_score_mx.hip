// Candidate-model scoring on the matrix cores, with every decision and every
// summed residual taken in FP64.
//
// MSAC scoring (scoring_function.h:137-233) of a model over a pair's points is
// a dense contraction: for the Sampson distance (fundamental_estimator.h:
// 195-222) the numerator is C = x2^T F x1 = vec(F) . z with
// z = x2 (x) x1 = (x2x1, x2y1, x2, y2x1, y2y1, y2, x1, y1, 1), and the four
// denominator terms (F x1)_0,1 and (F^T x2)_0,1 are linear in z as well.  For
// the homography transfer error (homography_estimator.h:181-199) the
// quantities P = x2 w - u, Q = y2 w - v and W = w (u, v, w = H x1) are linear
// in z, with r^2 = (P^2 + Q^2) / W^2.  So one wave scores 32 models x 32
// points per step as five (Sampson) or three (transfer) 32x32 MFMA tiles
// whose reduction dimension holds the z terms.
//
// Precision.  Points are scaled by a per-pair power of two 1/s (exact) so
// that |x'| <= 1, models to F'' = S F S / m (S = diag(s, s, 1), m = max |.|,
// so |F''| <= 1; the inlier test C^2 < tau^2 den becomes C''^2 <
// (tau / s)^2 den'' for the Sampson distance, and likewise for the transfer
// error).  Each operand is split into FP16 hi + lo parts and the MFMA sums
// hi*hi + hi*lo + lo*hi in FP32 (27 terms, padded to 32: two 32x32x16 MFMAs
// per quantity).  The error of a quantity q = c . z against its exact value is
// below 2^-18.5 sum_t |c_t||z_t| + 2^-25 (9 + sum_t |z_t|) (operand splits
// 2^-22 relative each, the dropped lo*lo term 2^-22, FP32 rounding of z
// 2^-24, FP32 accumulation of 27 exact products 27 * 2^-24, FP16 subnormal
// spacing 2^-24); with |c_t| <= 1 the kernel uses delta = 2^-17 S + 2^-20,
// S the sum of |z_t| over the terms the quantity reads (C: (1 + |x1| + |y1|)
// (1 + |x2| + |y2|); (F x1): 1 + |x1| + |y1|; (F^T x2): 1 + |x2| + |y2|).
//
// A point is dropped for a model only when the bounds prove it an outlier:
// Sampson (|C| - dC)^2 > tau''^2 (1 + 2^-16) sum (|q| + dq)^2, transfer
// (|P| - dP)+^2 + (|Q| - dQ)+^2 > tau''^2 (1 + 2^-16) (|W| + dW)^2 (the
// factor 1 + 2^-16 covers the FP32 rounding of the test itself, about
// 10 * 2^-24).  Every other point -- all inliers and every point near the
// threshold -- is evaluated by the FP64 routines of the FP64 kernel
// (f_sampson_parts with its guard band, f_sq_sampson, f_sq_symmetric,
// h_sq_residual) and the MSAC value is accumulated in ascending point order
// on the model's lane.  Counts, values and symmetric counts are therefore the
// FP64 kernel's bits.  NaN / inf models or points never pass the rejection
// test, so they reach the FP64 path as before.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr float kDeltaRel = 0x1p-17f;
constexpr float kDeltaAbs = 0x1p-20f;
constexpr double kTauMargin = 1.0 + 0x1p-16;
constexpr int kCompactPerBlock = 2048;  // slots per block of the ordered compaction
constexpr int kMxBlocks = 2048;         // persistent grids: at most 8,192 waves
// list entries per rejection + exact pass: as many as keep the segment's
// candidate masks within kMxMaskBudget bytes, between kMxSegMin and
// kMxSegMax (each segment costs three launches; the masks bound the memory)
constexpr int64_t kMxSegMin = 65536, kMxSegMax = 1048576;
constexpr int64_t kMxMaskBudget = 512ll << 20;
constexpr int kMxSlice = 128;           // points per rejection work item
constexpr int kNumXcd = 8;

__device__ __forceinline__ f32x16 mfma16(const h8& a, const h8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int j = 0; j < 16; ++j) z[j] = 0.0f;
  return z;
}

// Reduction-dimension layout of one quantity (k = 0..31): k < 9 hi*hi of
// term k, 9..17 hi(coef)*lo(z) of term k-9, 18..26 lo(coef)*hi(z) of term
// k-18, 27..31 zero.  Lane half hf of MFMA kh holds k = 16 kh + 8 hf + e.
__device__ __forceinline__ h8 coef_slice(const _Float16* hi, const _Float16* lo, int kh, int hf) {
  h8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k0 = 16 * kh + e, k1 = 16 * kh + 8 + e;
    const _Float16 a0 = k0 < 9 ? hi[k0] : (k0 < 18 ? hi[k0 - 9] : (k0 < 27 ? lo[k0 - 18] : (_Float16)0.0f));
    const _Float16 a1 = k1 < 9 ? hi[k1] : (k1 < 18 ? hi[k1 - 9] : (k1 < 27 ? lo[k1 - 18] : (_Float16)0.0f));
    v[e] = hf ? a1 : a0;
  }
  return v;
}

// One model's coefficient vectors over z (row-major 3x3 M): Sampson
// quantities C, (F x1)_0, (F x1)_1, (F^T x2)_0, (F^T x2)_1; transfer
// quantities P, Q, W.
template <int KIND>
__device__ __forceinline__ void quantity_coefs(const double* F, double c[][9]) {
  constexpr int NQ = KIND == 1 ? 5 : 3;
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int t = 0; t < 9; ++t) c[q][t] = 0.0;
  if constexpr (KIND == 1) {
#pragma unroll
    for (int t = 0; t < 9; ++t) c[0][t] = F[t];
    c[1][6] = F[0]; c[1][7] = F[1]; c[1][8] = F[2];
    c[2][6] = F[3]; c[2][7] = F[4]; c[2][8] = F[5];
    c[3][2] = F[0]; c[3][5] = F[3]; c[3][8] = F[6];
    c[4][2] = F[1]; c[4][5] = F[4]; c[4][8] = F[7];
  } else {
    c[0][0] = F[6]; c[0][1] = F[7]; c[0][2] = F[8]; c[0][6] = -F[0]; c[0][7] = -F[1]; c[0][8] = -F[2];
    c[1][3] = F[6]; c[1][4] = F[7]; c[1][5] = F[8]; c[1][6] = -F[3]; c[1][7] = -F[4]; c[1][8] = -F[5];
    c[2][6] = F[6]; c[2][7] = F[7]; c[2][8] = F[8];
  }
}

}  // namespace

// ---- per-pair scale and per-point operands ------------------------------------

// 1 / s per pair: s = the smallest power of two >= max |coordinate| of the
// pair's points (1 when that is 0 or not finite).
__global__ void __launch_bounds__(256) k_mx_pair_scale(const double* __restrict__ pts,
                                                       const int64_t* __restrict__ pair_pt_off,
                                                       const int32_t* __restrict__ pair_n, int num_pairs,
                                                       double* __restrict__ inv_scale) {
  __shared__ double red[4];
  const int p = blockIdx.x;
  if (p >= num_pairs) return;
  const double* P = pts + 4 * pair_pt_off[p];
  const int64_t n4 = 4 * (int64_t)pair_n[p];
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n4; i += blockDim.x) {
    const double v = fabs(P[i]);
    m = (v > m || v != v) ? v : m;
  }
  for (int s = 32; s >= 1; s >>= 1) {
    const double o = __shfl_down(m, s, 64);
    m = (o > m || o != o) ? o : m;
  }
  if (lane_id() == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) mm = (red[w] > mm || red[w] != red[w]) ? red[w] : mm;
    double inv = 1.0;
    if (mm > 0.0 && mm <= 1e300) {
      int e = 0;
      const double f = frexp(mm, &e);  // mm = f 2^e, f in [0.5, 1)
      (void)f;
      inv = ldexp(1.0, -e);             // 2^e >= mm
    }
    inv_scale[p] = inv;
  }
}

// Per point: the B-operand slices of both lane halves (k order, FP16) and
// the three error bounds.
__global__ void __launch_bounds__(256) k_mx_points(const double* __restrict__ pts,
                                                   const int32_t* __restrict__ pt_pair,
                                                   const double* __restrict__ inv_scale, int64_t total,
                                                   MxPoint* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double is = inv_scale[pt_pair[i]];
  const double x1 = pts[4 * i] * is, y1 = pts[4 * i + 1] * is, x2 = pts[4 * i + 2] * is, y2 = pts[4 * i + 3] * is;
  const double zd[9] = {x2 * x1, x2 * y1, x2, y2 * x1, y2 * y1, y2, x1, y1, 1.0};
  _Float16 hi[9], lo[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const float z = (float)zd[t];
    hi[t] = (_Float16)z;
    lo[t] = (_Float16)(z - (float)hi[t]);
  }
  MxPoint o;
#pragma unroll
  for (int kh = 0; kh < 2; ++kh)
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = 16 * kh + 8 * hf + e;
        const _Float16 v = k < 9 ? hi[k] : (k < 18 ? lo[k - 9] : (k < 27 ? hi[k - 18] : (_Float16)0.0f));
        o.b[kh][hf][e] = __builtin_bit_cast(uint16_t, v);
      }
    }
  const float sa = 1.0f + fabsf((float)x1) + fabsf((float)y1);
  const float sb = 1.0f + fabsf((float)x2) + fabsf((float)y2);
  o.d[0] = kDeltaRel * (sa * sb) + kDeltaAbs;
  o.d[1] = kDeltaRel * sa + kDeltaAbs;
  o.d[2] = kDeltaRel * sb + kDeltaAbs;
  o.d[3] = 0.0f;
  out[i] = o;
}

// ---- ordered compaction of the active model slots ------------------------------

// Pass 1: inactive slots get their empty result; per block, the number of
// active slots.
__global__ void __launch_bounds__(256) k_mx_count(const int8_t* __restrict__ active, int64_t num_models,
                                                  int32_t* __restrict__ counts, double* __restrict__ values,
                                                  int32_t* __restrict__ sym_counts, int32_t* __restrict__ blk) {
  __shared__ int32_t part[4];
  const int64_t s0 = (int64_t)blockIdx.x * kCompactPerBlock;
  int c = 0;
  for (int e = 0; e < kCompactPerBlock / 256; ++e) {
    const int64_t s = s0 + (int64_t)e * 256 + threadIdx.x;
    if (s < num_models) {
      if (active[s] == 2) {
        ++c;
      } else {
        counts[s] = -1;
        values[s] = 0.0;
        if (sym_counts) sym_counts[s] = 0;
      }
    }
  }
  for (int s = 32; s >= 1; s >>= 1) c += __shfl_down(c, s, 64);
  if (lane_id() == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) blk[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Pass 2 (one block): exclusive scan of the block counts; the total goes to
// *list_count.
__global__ void __launch_bounds__(1024) k_mx_scan(int32_t* __restrict__ blk, int nblk,
                                                  uint32_t* __restrict__ list_count) {
  __shared__ int32_t buf[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblk; base += 1024) {
    const int i = base + (int)threadIdx.x;
    const int v = i < nblk ? blk[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblk) blk[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *list_count = (uint32_t)carry;
}

// Pass 3: each block writes its active slots, in slot order, from its offset.
__global__ void __launch_bounds__(256) k_mx_write(const int8_t* __restrict__ active, int64_t num_models,
                                                  const int32_t* __restrict__ blk, int32_t* __restrict__ list) {
  __shared__ int32_t wbase[4];
  __shared__ int32_t run;
  if (threadIdx.x == 0) run = blk[blockIdx.x];
  __syncthreads();
  const int64_t s0 = (int64_t)blockIdx.x * kCompactPerBlock;
  const int w = threadIdx.x >> 6;
  for (int e = 0; e < kCompactPerBlock / 256; ++e) {
    const int64_t s = s0 + (int64_t)e * 256 + threadIdx.x;
    const bool act = s < num_models && active[s] == 2;
    const uint64_t b = ballot64(act);
    if (lane_id() == 0) wbase[w] = __popcll(b);
    __syncthreads();
    int before = run;
    for (int k = 0; k < w; ++k) before += wbase[k];
    if (act)
      list[before + __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] =
          (int32_t)s;
    __syncthreads();
    if (threadIdx.x == 0) run += wbase[0] + wbase[1] + wbase[2] + wbase[3];
    __syncthreads();
  }
}

// ---- per-model A operands -----------------------------------------------------------

// One thread per (list entry of the segment, lane half): the entry's scaled,
// normalised coefficient vectors split into FP16 hi / lo and laid out as the
// MFMA A-operand slices of that lane half (NQ quantities x 2 MFMAs x 8).  A
// zero, inf or NaN model gets zero operands (nothing is rejected for it).
template <int KIND>
__global__ void __launch_bounds__(256) k_mx_coefs(const double* __restrict__ pair_inv_scale,
                                                  const double* __restrict__ models,
                                                  const int32_t* __restrict__ model_pair,
                                                  const int32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ list_count, int64_t seg_begin,
                                                  int64_t seg_len, h8* __restrict__ coefs, int slots_per) {
  constexpr int NQ = KIND == 1 ? 5 : 3;
  const int64_t total = (int64_t)*list_count;
  if (seg_begin >= total) return;
  const int64_t seg_n = total - seg_begin < seg_len ? total - seg_begin : seg_len;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 2 * seg_n) return;
  const int64_t e = g >> 1;
  const int hf = (int)(g & 1);
  const int64_t slot = list[seg_begin + e];
  const int p = model_pair[slots_per == 1 ? slot : slot / slots_per];
  double M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = models[9 * slot + k];
  const double is = pair_inv_scale[p];
  const double s = 1.0 / is;  // a power of two: exact
  const double sc[3] = {s, s, 1.0};
  double Fs[9], mx = 0.0;
  if constexpr (KIND == 1) {
    // F' = S F S: C, (F x1) and (F^T x2) in the scaled coordinates
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) Fs[3 * i + k] = M[3 * i + k] * sc[i] * sc[k];
  } else {
    // H' = S^-1 H S maps scaled x1 to scaled x2
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) Fs[3 * i + k] = M[3 * i + k] / sc[i] * sc[k];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) mx = fabs(Fs[k]) > mx ? fabs(Fs[k]) : mx;
  const bool lp_ok = mx > 0.0 && mx <= 1e300;
  const double inv_mx = lp_ok ? 1.0 / mx : 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) Fs[k] = Fs[k] * inv_mx;
  double cq[NQ][9];
  quantity_coefs<KIND>(Fs, cq);
  h8* out = coefs + (e * 2 + hf) * (2 * NQ);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    _Float16 hi[9], lo[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      hi[k] = (_Float16)cq[q][k];
      lo[k] = (_Float16)(cq[q][k] - (double)hi[k]);
    }
    out[2 * q] = coef_slice(hi, lo, 0, hf);
    out[2 * q + 1] = coef_slice(hi, lo, 1, hf);
  }
}

// ---- phase 1: rejection on the matrix cores ---------------------------------------

// Work item (tile t, slice j): the 32 list entries 32 t .. 32 t + 31 of the
// segment (slot order, so the entries of one pair are adjacent; a tile that
// spans pairs is handled pair by pair) against points 128 j .. 128 j + 127 of
// their pair.  Lane r < 32 owns entry r and writes, per 32-point chunk, the
// 32-bit word of its candidate mask (bit b: point base + b must be evaluated
// in FP64).  Words of points at or beyond the pair's n are not written.
template <int KIND>
__global__ void __launch_bounds__(256) k_mx_reject(const int64_t* __restrict__ pair_pt_off,
                                                   const int32_t* __restrict__ pair_n,
                                                   const double* __restrict__ pair_sq_thr,
                                                   const double* __restrict__ pair_inv_scale,
                                                   const MxPoint* __restrict__ mxp,
                                                   const h8* __restrict__ coefs,
                                                   const int32_t* __restrict__ model_pair,
                                                   const int32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ list_count, int64_t seg_begin,
                                                   int64_t seg_len, int slices, int words, uint32_t* __restrict__ mask,
                                                   int slots_per) {
  constexpr int NQ = KIND == 1 ? 5 : 3;
  const int lane = lane_id();
  const int hf = lane >> 5, r = lane & 31;
  const int64_t total = (int64_t)*list_count;
  if (seg_begin >= total) return;
  const int64_t seg_n = total - seg_begin < seg_len ? total - seg_begin : seg_len;
  const int64_t items = (seg_n + 31) / 32 * slices;
  // gridDim.x is a multiple of kNumXcd: workgroup b runs on XCD b % kNumXcd,
  // which takes a contiguous range of items (adjacent items share a pair)
  const int xcd = (int)(blockIdx.x % kNumXcd);
  const int64_t waves_per_xcd = (int64_t)(gridDim.x / kNumXcd) * 4;
  const int64_t w = (int64_t)(blockIdx.x / kNumXcd) * 4 + (threadIdx.x >> 6);
  const int64_t lo_u = items * xcd / kNumXcd, hi_u = items * (xcd + 1) / kNumXcd;
  for (int64_t u = lo_u + w; u < hi_u; u += waves_per_xcd) {
    const int64_t t = u / slices;
    const int j = (int)(u - t * slices);
    const int64_t e0 = 32 * t;  // within the segment
    const int cnt_all = (int)(seg_n - e0 < 32 ? seg_n - e0 : 32);
    const int64_t slot = r < cnt_all ? (int64_t)list[seg_begin + e0 + r] : -1;
    const int my_pair = slot >= 0 ? model_pair[slots_per == 1 ? slot : slot / slots_per] : -1;
    uint32_t* __restrict__ my_mask = mask + (e0 + r) * (int64_t)words;
    int done = 0;
    while (done < cnt_all) {
      const int p = __builtin_amdgcn_readlane(my_pair, done);
      const bool mine = r >= done && r < cnt_all && my_pair == p;
      const int sub_end = done + __popcll(ballot64(mine) & 0xffffffffull);
      const int n = pair_n[p];
      const int first = kMxSlice * j;
      if (first < n) {
        // the sub-tile's A operands (lane: row r, half hf); other rows zero
        h8 A[NQ][2];
        const h8* __restrict__ my_coefs = coefs + ((e0 + r) * 2 + hf) * (2 * NQ);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int kh = 0; kh < 2; ++kh) {
            if (mine) {
              A[q][kh] = my_coefs[2 * q + kh];
            } else {
#pragma unroll
              for (int e = 0; e < 8; ++e) A[q][kh][e] = (_Float16)0.0f;
            }
          }
        const double is = pair_inv_scale[p];
        const float tau2m = (float)(pair_sq_thr[p] * is * is * kTauMargin);
        const MxPoint* __restrict__ X = mxp + pair_pt_off[p];
        const int last = first + kMxSlice < n ? first + kMxSlice : n;
        for (int base = first; base < last; base += 32) {
          const int i = base + r;
          h8 b0, b1;
          float dq0 = 0.0f, dq1 = 0.0f, dq2 = 0.0f;
          const bool pin = i < n;
          if (pin) {
            const MxPoint& x = X[i];
            b0 = *reinterpret_cast<const h8*>(x.b[0][hf]);
            b1 = *reinterpret_cast<const h8*>(x.b[1][hf]);
            dq0 = x.d[0];
            dq1 = x.d[1];
            dq2 = x.d[2];
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              b0[e] = (_Float16)0.0f;
              b1[e] = (_Float16)0.0f;
            }
          }
          // accumulator slot j holds row 8 (j / 4) + 4 hf + (j % 4) of column r
          uint64_t keep[16];
          if constexpr (KIND == 1) {
            f32x16 den = zero16();
#pragma unroll
            for (int q = 1; q < 5; ++q) {
              f32x16 a = mfma16(A[q][0], b0, zero16());
              a = mfma16(A[q][1], b1, a);
              const float dlt = q < 3 ? dq1 : dq2;
#pragma unroll
              for (int jj = 0; jj < 16; ++jj) {
                const float uu = fabsf(a[jj]) + dlt;
                den[jj] = __builtin_fmaf(uu, uu, den[jj]);
              }
            }
            f32x16 c = mfma16(A[0][0], b0, zero16());
            c = mfma16(A[0][1], b1, c);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
              const float cl = fabsf(c[jj]) - dq0;
              const bool rej = cl > 0.0f && cl * cl > den[jj] * tau2m;
              keep[jj] = ballot64(pin && !rej);
            }
          } else {
            f32x16 pp = mfma16(A[0][0], b0, zero16());
            pp = mfma16(A[0][1], b1, pp);
            f32x16 qq = mfma16(A[1][0], b0, zero16());
            qq = mfma16(A[1][1], b1, qq);
            f32x16 ww = mfma16(A[2][0], b0, zero16());
            ww = mfma16(A[2][1], b1, ww);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
              const float pl = fmaxf(fabsf(pp[jj]) - dq0, 0.0f);
              const float ql = fmaxf(fabsf(qq[jj]) - dq0, 0.0f);
              const float wu = fabsf(ww[jj]) + dq1;
              const bool rej = __builtin_fmaf(pl, pl, ql * ql) > wu * wu * tau2m;
              keep[jj] = ballot64(pin && !rej);
            }
          }
          // lane r (< 32) stores row r's word: slot j = 4 (r / 8) + r % 4,
          // bits 32 ((r / 4) % 2) .. +31 of that slot's ballot
          const int jr = 4 * (r >> 3) + (r & 3), hr = (r >> 2) & 1;
          uint64_t sel = 0;
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) sel = jr == jj ? keep[jj] : sel;
          if (mine && hf == 0) my_mask[base >> 5] = (uint32_t)(sel >> (32 * hr));
        }
      }
      done = sub_end;
    }
  }
}

// ---- phase 2: the FP64 decisions and the in-order MSAC sums -----------------------

// One wave per list entry of the segment.  The candidate points (mask bits)
// are queued in point order and evaluated 64 at a time by the whole wave with
// the FP64 kernel's operations; inliers are accumulated in point order
// through the scalar unit (value -= r2), as the FP64 kernel does.
template <int KIND>
__global__ void __launch_bounds__(256) k_mx_exact(const double* __restrict__ pts,
                                                  const int64_t* __restrict__ pair_pt_off,
                                                  const int32_t* __restrict__ pair_n,
                                                  const double* __restrict__ pair_sq_thr,
                                                  const double* __restrict__ models,
                                                  const int32_t* __restrict__ model_pair,
                                                  const int32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ list_count, int64_t seg_begin,
                                                  int64_t seg_len, int words, const uint32_t* __restrict__ mask,
                                                  int32_t* __restrict__ counts, double* __restrict__ values,
                                                  int32_t* __restrict__ sym_counts, int slots_per) {
  __shared__ int32_t queue[4][128];
  const int lane = lane_id();
  int32_t* q = queue[threadIdx.x >> 6];
  const int64_t total = (int64_t)*list_count;
  if (seg_begin >= total) return;
  const int64_t seg_n = total - seg_begin < seg_len ? total - seg_begin : seg_len;
  const int xcd = (int)(blockIdx.x % kNumXcd);
  const int64_t waves_per_xcd = (int64_t)(gridDim.x / kNumXcd) * 4;
  const int64_t w = (int64_t)(blockIdx.x / kNumXcd) * 4 + (threadIdx.x >> 6);
  const int64_t lo_e = seg_n * xcd / kNumXcd, hi_e = seg_n * (xcd + 1) / kNumXcd;
  for (int64_t e = lo_e + w; e < hi_e; e += waves_per_xcd) {
    const int64_t slot = list[seg_begin + e];
    const int pair = model_pair[slots_per == 1 ? slot : slot / slots_per];
    const double* __restrict__ P = pts + 4 * pair_pt_off[pair];
    const int n = pair_n[pair];
    const double thr = pair_sq_thr[pair];
    double M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = models[9 * slot + k];
    const uint32_t* __restrict__ mk = mask + e * (int64_t)words;
    double acc = 0.0;
    int64_t cnt = 0, scnt = 0;
    int qn = 0;
    // evaluates the first cnt_flush queued candidates and accumulates the
    // inliers among them in queue (= point) order
    auto flush = [&](int cnt_flush) {
      bool in = false, sym_in = false;
      double r2 = 0.0;
      if (lane < cnt_flush) {
        const double* s = P + 4 * (int64_t)q[lane];
        if constexpr (KIND == 1) {
          double num, den;
          f_sampson_parts(s, M, num, den);
          const double t = thr * den;
          if (num < t * (1.0 - 1e-12)) in = true;
          else if (!(num > t * (1.0 + 1e-12))) in = (num / den) < thr;
          if (in) {
            r2 = f_sq_sampson(s, M);
            if (sym_counts) sym_in = f_sq_symmetric(s, M) < thr;
          }
        } else {
          r2 = h_sq_residual(s, M);
          in = r2 < thr;
        }
      }
      uint64_t b = ballot64(in);
      cnt += __popcll(b);
      if (KIND == 1 && sym_counts) scnt += __popcll(ballot64(sym_in));
      while (b) {
        const int l = __ffsll((unsigned long long)b) - 1;
        acc = acc - readlane_f64(r2, l);
        b &= b - 1;
      }
    };
    const int nw = (n + 31) >> 5;
    // the mask row is read 256 words (8,192 points) at a time, 4 words per
    // lane, coalesced; the walk then takes the words from the lanes that hold
    // them, so no memory access sits inside the point loop
    for (int w0 = 0; w0 < nw; w0 += 256) {
      uint32_t held[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int wi = w0 + 64 * k + lane;
        held[k] = wi < nw ? mk[wi] : 0u;
      }
      const int wend = nw - w0 < 256 ? nw - w0 : 256;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        for (int t = 0; t < 64 && 64 * k + t < wend; t += 2) {
          const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)held[k], t);
          const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)held[k], t + 1);
          const uint64_t bits = lo | (hi << 32);
          if (bits == 0) continue;
          const int wi = w0 + 64 * k + t;
          const bool c = (bits >> lane) & 1ull;
          if (c) q[qn + __popcll(bits & ((1ull << lane) - 1ull))] = 32 * wi + lane;
          qn += __popcll(bits);
          if (qn >= 64) {
            flush(64);
            const int rest = qn - 64;
            const int32_t moved = lane < rest ? q[64 + lane] : 0;
            if (lane < rest) q[lane] = moved;
            qn = rest;
          }
        }
      }
    }
    flush(qn);
    if (lane == 0) {
      counts[slot] = (int32_t)cnt;
      values[slot] = cnt == 0 ? 0.0 : (acc + (double)cnt * thr) / thr;
      if (sym_counts) sym_counts[slot] = (int32_t)scnt;
    }
  }
}

// ---- launchers --------------------------------------------------------------------

void launch_mx_prepare(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                       const int32_t* pt_pair, int num_pairs, int64_t total, double* inv_scale, MxPoint* out) {
  if (num_pairs <= 0 || total <= 0) return;
  hipLaunchKernelGGL(k_mx_pair_scale, dim3((unsigned)num_pairs), dim3(256), 0, st, pts, pair_pt_off, pair_n,
                     num_pairs, inv_scale);
  hipLaunchKernelGGL(k_mx_points, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pts, pt_pair, inv_scale,
                     total, out);
}

size_t mx_list_scratch_ints(int64_t num_models) {
  return (size_t)((num_models + kCompactPerBlock - 1) / kCompactPerBlock);
}

static int64_t mx_segment(int max_n) {
  const int64_t words = (max_n + 31) / 32 + 1;
  const int64_t fit = kMxMaskBudget / (4 * words);
  return fit < kMxSegMin ? kMxSegMin : (fit > kMxSegMax ? kMxSegMax : fit);
}

size_t mx_mask_words(int max_n) {
  // the candidate masks, then the segment's A operands (5 x 2 x 2 x 16 B per entry)
  const int64_t seg = mx_segment(max_n);
  return (size_t)seg * (size_t)((max_n + 31) / 32 + 1) + (size_t)seg * 80;
}

void launch_score_mx(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                     const double* pair_sq_thr, const double* pair_inv_scale, const MxPoint* mxp,
                     const double* models, const int32_t* model_pair, const int8_t* active, int64_t num_models,
                     int32_t* counts, double* values, int32_t* sym_counts, int slots_per, int32_t* work_list,
                     uint32_t* work_count, int32_t* blk_scratch, int max_n, uint32_t* mask) {
  if (num_models <= 0) return;
  const int nblk = (int)((num_models + kCompactPerBlock - 1) / kCompactPerBlock);
  hipLaunchKernelGGL(k_mx_count, dim3((unsigned)nblk), dim3(256), 0, st, active, num_models, counts, values,
                     sym_counts, blk_scratch);
  hipLaunchKernelGGL(k_mx_scan, dim3(1), dim3(1024), 0, st, blk_scratch, nblk, work_count);
  hipLaunchKernelGGL(k_mx_write, dim3((unsigned)nblk), dim3(256), 0, st, active, num_models, blk_scratch, work_list);
  const int slices = (max_n + kMxSlice - 1) / kMxSlice;
  const int words = (max_n + 31) / 32 + 1;
  const int64_t segment = mx_segment(max_n);
  h8* const coefs = reinterpret_cast<h8*>(mask + (size_t)segment * (size_t)words);
  // the active count is on the device; segments up to the slot count, and
  // each kernel returns at once past the end of the list
  for (int64_t seg = 0; seg < num_models; seg += segment) {
    const int64_t seg_len = num_models - seg < segment ? num_models - seg : segment;
    int64_t b1 = ((seg_len + 31) / 32 * slices + 3) / 4;
    if (b1 > kMxBlocks) b1 = kMxBlocks;
    b1 = (b1 + kNumXcd - 1) / kNumXcd * kNumXcd;
    int64_t b2 = (seg_len + 3) / 4;
    if (b2 > kMxBlocks) b2 = kMxBlocks;
    b2 = (b2 + kNumXcd - 1) / kNumXcd * kNumXcd;
    const unsigned bc = (unsigned)((2 * seg_len + 255) / 256);
    if (kind == 0) {
      hipLaunchKernelGGL(k_mx_coefs<0>, dim3(bc), dim3(256), 0, st, pair_inv_scale, models, model_pair, work_list,
                         work_count, seg, seg_len, coefs, slots_per);
      hipLaunchKernelGGL(k_mx_reject<0>, dim3((unsigned)b1), dim3(256), 0, st, pair_pt_off, pair_n, pair_sq_thr,
                         pair_inv_scale, mxp, coefs, model_pair, work_list, work_count, seg, seg_len, slices, words,
                         mask, slots_per);
      hipLaunchKernelGGL(k_mx_exact<0>, dim3((unsigned)b2), dim3(256), 0, st, pts, pair_pt_off, pair_n, pair_sq_thr,
                         models, model_pair, work_list, work_count, seg, seg_len, words, mask, counts, values,
                         sym_counts, slots_per);
    } else {
      hipLaunchKernelGGL(k_mx_coefs<1>, dim3(bc), dim3(256), 0, st, pair_inv_scale, models, model_pair, work_list,
                         work_count, seg, seg_len, coefs, slots_per);
      hipLaunchKernelGGL(k_mx_reject<1>, dim3((unsigned)b1), dim3(256), 0, st, pair_pt_off, pair_n, pair_sq_thr,
                         pair_inv_scale, mxp, coefs, model_pair, work_list, work_count, seg, seg_len, slices, words,
                         mask, slots_per);
      hipLaunchKernelGGL(k_mx_exact<1>, dim3((unsigned)b2), dim3(256), 0, st, pts, pair_pt_off, pair_n, pair_sq_thr,
                         models, model_pair, work_list, work_count, seg, seg_len, words, mask, counts, values,
                         sym_counts, slots_per);
    }
  }
}

}  // namespace gcr
