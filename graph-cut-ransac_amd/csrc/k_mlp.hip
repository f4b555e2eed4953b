// Device kernels of the learned inlier-probability network.
//
// There is no reference counterpart: the reference's hook for a learned
// quality score is the PROSAC sampler, which draws from prefixes of the
// points sorted by quality (inc/samplers/prosac_sampler.h:51-237).  These
// kernels compute that score per correspondence:
//   k_knn_bin, k_knn_scatter, k_knn_query
//                    8-nearest-neighbour features (per-pair grid of cells,
//                    one thread per query walking rings of cells)
//   k_mlp_predict    forward pass on the matrix cores
//   k_mlp_train      forward + backward-data pass of one mini-batch
//   k_mlp_wgrad      weight gradients (batch as the reduction dimension)
//   k_mlp_reduce     deterministic sum of the per-chunk partials + Adam
//
// Matrix-core formulation.  Activations are kept transposed: a wave owns 32
// points (the N = 32 columns of a 32x32x8 BF16 MFMA) and computes
// Z^T = W H^T, with W in the A operand (32 output neurons x 8 inputs per
// instruction) and H^T in B.  The 16 FP32 results a lane holds for one
// 32-neuron tile are the neurons 8*(r/4) + 4*(lane/32) + r%4 of point
// lane%32 -- exactly the B-operand slots of the next layer's k-blocks, so a
// layer's output (bias, GELU, BF16 rounding applied in registers) feeds the
// next layer with no data movement.  The backward-data pass does the same
// with W^T, and the weight gradients G = D H^T run the batch through the
// reduction dimension from neuron-major BF16 copies written by the forward
// pass.  Weights live in LDS with a 132-element row stride (no bank
// conflicts on the 8-byte A-operand reads).
#include <hip/hip_runtime.h>

#include "mlp_kernels.hpp"

namespace gcr {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int RS = 132;  // LDS row stride (elements) of the 128-wide matrices
constexpr int ZL = 64 * 64;  // floats of one layer's pre-activations per wave

struct FwdLds {
  uint16_t w1[HID * IN];
  uint16_t w2[HID * RS];
  uint16_t w3[HID * RS];
  float b1[HID], b2[HID], b3[HID], w4[HID];
  float b4, pad_[3];
};
struct TrainLds {
  FwdLds f;
  uint16_t w2t[HID * RS];
  uint16_t w3t[HID * RS];
};
static_assert(sizeof(FwdLds) % 16 == 0, "LDS layout");

__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// GELU (tanh form) and its derivative, operations in the oracle's order
__device__ inline float gelu_f(float z) {
  const float c = 0.7978845608028654f;
  const float u = c * (z + 0.044715f * (z * z * z));
  return 0.5f * z * (1.0f + tanhf(u));
}
__device__ inline float gelu_grad_f(float z) {
  const float c = 0.7978845608028654f;
  const float z2 = z * z;
  const float t = tanhf(c * (z + 0.044715f * (z2 * z)));
  return 0.5f * (1.0f + t) + 0.5f * z * ((1.0f - t * t) * (c * (1.0f + 0.134145f * z2)));
}
__device__ inline float sigmoid_f(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ inline float softplus_f(float z) {
  const float a = z < 0.0f ? -z : z;
  return (z > 0.0f ? z : 0.0f) + log1pf(expf(-a));
}

__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ inline f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = 0.0f;
  return z;
}

__device__ inline f32x16 mfma(const bf16x4& a, const bf16x4& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

// 128 x 128 BF16 matrix (row-major) -> LDS rows of RS elements
__device__ inline void stage_matrix(uint16_t* dst, const uint16_t* src) {
  for (int e = threadIdx.x; e < HID * (HID / 4); e += blockDim.x) {
    const int r = e >> 5, c4 = e & 31;
    *reinterpret_cast<uint2*>(dst + r * RS + c4 * 4) = *reinterpret_cast<const uint2*>(src + r * HID + c4 * 4);
  }
}

__device__ inline void stage_forward(FwdLds& s, const uint16_t* wpk, const float* bias) {
  for (int e = threadIdx.x; e < HID * IN / 4; e += blockDim.x)
    *reinterpret_cast<uint2*>(s.w1 + 4 * e) = *reinterpret_cast<const uint2*>(wpk + PK_W1 + 4 * e);
  stage_matrix(s.w2, wpk + PK_W2);
  stage_matrix(s.w3, wpk + PK_W3);
  for (int i = threadIdx.x; i < HID; i += blockDim.x) {
    s.b1[i] = bias[i];
    s.b2[i] = bias[HID + i];
    s.b3[i] = bias[2 * HID + i];
    s.w4[i] = bf2f(wpk[PK_W4 + i]);
  }
  if (threadIdx.x == 0) s.b4 = bias[3 * HID];
}

// bias + GELU + BF16 of one 32-neuron tile; the rounded activations go to
// the next layer's B operand (k-blocks 4*MT .. 4*MT+3) and, when training,
// the pre-activations to the wave's z buffer and the activations to the
// neuron-major copy used by the weight gradients
template <int MT, bool TRAIN>
__device__ __attribute__((always_inline)) inline void epilogue(const f32x16& acc, const float* bias, int half, int lane, bf16x4* hout, float* zl,
                                uint16_t* arow, int64_t bpad, int64_t col) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 32 * MT + 8 * (r >> 2) + 4 * half + (r & 3);
    const float z = acc[r] + bias[i];
    const uint16_t h = f2bf(gelu_f(z));
    hout[4 * MT + (r >> 2)][r & 3] = (short)h;
    if (TRAIN) {
      zl[(16 * MT + r) * 64 + lane] = z;
      arow[(int64_t)i * bpad + col] = h;
    }
  }
}

template <int MT, bool TRAIN>
__device__ __attribute__((always_inline)) inline void hidden_tile(const uint16_t* w, const float* bias, const bf16x4* hin, int j, int half, int lane,
                                   bf16x4* hout, float* zl, uint16_t* arow, int64_t bpad, int64_t col) {
  f32x16 acc = zero16();
  const uint16_t* wrow = w + (32 * MT + j) * RS + 4 * half;
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(wrow + 8 * kb);
    acc = mfma(a, hin[kb], acc);
  }
  epilogue<MT, TRAIN>(acc, bias, half, lane, hout, zl, arow, bpad, col);
}

// Forward pass of the wave's 32 points (lane % 32 = point, lane / 32 = half
// of the k-block); returns the logit (both halves hold it).  fptr: this
// lane's feature row, nullptr for a padding column.
// wj: lane % 32 as the weight-row index; the persistent inference loop
// passes it through an empty asm so the weight reads stay inside the loop
// (hoisted, the 128 weight registers per layer would spill).
template <bool TRAIN>
__device__ __attribute__((always_inline)) inline float forward_wave(const FwdLds& s, const float* fptr, int lane,
                                                                    int wj, float* zw, uint16_t* act, int64_t bpad,
                                                                    int64_t col) {
  const int j = wj, half = lane >> 5;
  float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (fptr) f = *reinterpret_cast<const float4*>(fptr + 4 * half);
  bf16x4 xb;
  xb[0] = (short)f2bf(f.x);
  xb[1] = (short)f2bf(f.y);
  xb[2] = (short)f2bf(f.z);
  xb[3] = (short)f2bf(f.w);
  if (TRAIN) {
#pragma unroll
    for (int e = 0; e < 4; ++e) act[(int64_t)(ROW_X + 4 * half + e) * bpad + col] = (uint16_t)xb[e];
  }
  bf16x4 ha[16], hb[16];
  // layer 1: one k-block (8 inputs)
  {
    const uint16_t* wrow = s.w1 + j * IN + 4 * half;
    uint16_t* a1 = TRAIN ? act + (int64_t)ROW_H1 * bpad : nullptr;
    float* z1 = TRAIN ? zw : nullptr;
    f32x16 acc0 = mfma(*reinterpret_cast<const bf16x4*>(wrow), xb, zero16());
    epilogue<0, TRAIN>(acc0, s.b1, half, lane, ha, z1, a1, bpad, col);
    f32x16 acc1 = mfma(*reinterpret_cast<const bf16x4*>(wrow + 32 * IN), xb, zero16());
    epilogue<1, TRAIN>(acc1, s.b1, half, lane, ha, z1, a1, bpad, col);
    f32x16 acc2 = mfma(*reinterpret_cast<const bf16x4*>(wrow + 64 * IN), xb, zero16());
    epilogue<2, TRAIN>(acc2, s.b1, half, lane, ha, z1, a1, bpad, col);
    f32x16 acc3 = mfma(*reinterpret_cast<const bf16x4*>(wrow + 96 * IN), xb, zero16());
    epilogue<3, TRAIN>(acc3, s.b1, half, lane, ha, z1, a1, bpad, col);
  }
  {
    uint16_t* a2 = TRAIN ? act + (int64_t)ROW_H2 * bpad : nullptr;
    float* z2 = TRAIN ? zw + ZL : nullptr;
    hidden_tile<0, TRAIN>(s.w2, s.b2, ha, j, half, lane, hb, z2, a2, bpad, col);
    hidden_tile<1, TRAIN>(s.w2, s.b2, ha, j, half, lane, hb, z2, a2, bpad, col);
    hidden_tile<2, TRAIN>(s.w2, s.b2, ha, j, half, lane, hb, z2, a2, bpad, col);
    hidden_tile<3, TRAIN>(s.w2, s.b2, ha, j, half, lane, hb, z2, a2, bpad, col);
  }
  {
    uint16_t* a3 = TRAIN ? act + (int64_t)ROW_H3 * bpad : nullptr;
    float* z3 = TRAIN ? zw + 2 * ZL : nullptr;
    hidden_tile<0, TRAIN>(s.w3, s.b3, hb, j, half, lane, ha, z3, a3, bpad, col);
    hidden_tile<1, TRAIN>(s.w3, s.b3, hb, j, half, lane, ha, z3, a3, bpad, col);
    hidden_tile<2, TRAIN>(s.w3, s.b3, hb, j, half, lane, ha, z3, a3, bpad, col);
    hidden_tile<3, TRAIN>(s.w3, s.b3, hb, j, half, lane, ha, z3, a3, bpad, col);
  }
  // output unit: this half's 64 neurons in ascending order, then the halves
  float sum = 0.0f;
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) sum = sum + s.w4[8 * kb + 4 * half + e] * bf2f((uint16_t)ha[kb][e]);
  }
  const float other = __shfl_xor(sum, 32);
  return (half == 0 ? sum + other : other + sum) + s.b4;
}

// backward-data tile: D_l = (W^T D_{l+1}) * gelu'(Z_l) for 32 neurons
template <int MT, bool OUT>
__device__ __attribute__((always_inline)) inline void backward_tile(const uint16_t* wt, const bf16x4* din, int j, int half, int lane, const float* zl,
                                     bf16x4* dout, uint16_t* drow, int64_t bpad, int64_t col) {
  f32x16 acc = zero16();
  const uint16_t* wrow = wt + (32 * MT + j) * RS + 4 * half;
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) acc = mfma(*reinterpret_cast<const bf16x4*>(wrow + 8 * kb), din[kb], acc);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 32 * MT + 8 * (r >> 2) + 4 * half + (r & 3);
    const float d = acc[r] * gelu_grad_f(zl[(16 * MT + r) * 64 + lane]);
    const uint16_t b = f2bf(d);
    if (OUT) dout[4 * MT + (r >> 2)][r & 3] = (short)b;
    drow[(int64_t)i * bpad + col] = b;
  }
}

}  // namespace

int lds_forward_bytes() { return (int)sizeof(FwdLds); }
int lds_train_bytes() { return (int)sizeof(TrainLds); }

// ---- features --------------------------------------------------------------
//
// Per correspondence: normalised coordinates, the descriptor ratio and the
// mean displacement (x2 - x1, y2 - y1) of its 8 nearest first-image
// neighbours minus its own, as the oracle's index-order scan defines them.
// The neighbour search is restricted by a grid: a pair's
// first-image points are binned into square cells (about four points per
// cell, indices clamped to the grid), sorted by cell, and each query walks
// rings of cells around its own cell until its eighth-best distance is below
// a lower bound of the distance of every point outside the rings.  Candidates
// are ranked by (distance, index): the index-order scan keeps, among
// equal distances, the earlier index, so both build the same ordered list of
// the 8 smallest (distance, index) pairs and the same sums.
//
// Lower bound: a point in a cell of column index below c has
// fl(x * inv_cs) < c, so x < c / inv_cs up to one rounding of the product;
// one in a column above c has x >= (c + 1) / inv_cs likewise (clamping moves
// a point only into the edge cell on its own side, which keeps both
// inequalities).  The bound is reduced by a margin far above those roundings
// and the float rounding of the distance, and the search stops only when the
// eighth distance is strictly below it, so no point outside the rings can
// enter the list.

__device__ __forceinline__ int knn_axis_cell(float v, float inv_cs, int g) {
  const float f = floorf(v * inv_cs);
  if (!(f >= 0.0f)) return 0;  // negative or NaN
  if (f >= (float)g) return g - 1;
  return (int)f;
}

__global__ __launch_bounds__(256) void k_knn_bin(const double* __restrict__ corr, const int64_t* __restrict__ pair_off,
                                                 const KnnGrid* __restrict__ grids,
                                                 const int32_t* __restrict__ tile_pair,
                                                 const int32_t* __restrict__ tile_start, int32_t* __restrict__ cell_count,
                                                 int32_t* __restrict__ cell_of, int32_t* __restrict__ rank) {
  const int p = tile_pair[blockIdx.x];
  const int64_t base = pair_off[p];
  const int64_t n = pair_off[p + 1] - base;
  const int64_t q = (int64_t)tile_start[blockIdx.x] + threadIdx.x;
  if (q >= n) return;
  const KnnGrid g = grids[p];
  const double* r = corr + 4 * (base + q);
  const float x = (float)r[0], y = (float)r[1];
  const int c = knn_axis_cell(y, g.inv_cs, g.gy) * g.gx + knn_axis_cell(x, g.inv_cs, g.gx);
  cell_of[base + q] = c;
  rank[base + q] = atomicAdd(cell_count + g.cell_base + c, 1);
}

__global__ __launch_bounds__(256) void k_knn_scatter(const double* __restrict__ corr,
                                                     const int64_t* __restrict__ pair_off,
                                                     const KnnGrid* __restrict__ grids,
                                                     const int32_t* __restrict__ tile_pair,
                                                     const int32_t* __restrict__ tile_start,
                                                     const int32_t* __restrict__ cell_off,
                                                     const int32_t* __restrict__ cell_of,
                                                     const int32_t* __restrict__ rank, float4* __restrict__ spts,
                                                     int32_t* __restrict__ sidx) {
  const int p = tile_pair[blockIdx.x];
  const int64_t base = pair_off[p];
  const int64_t n = pair_off[p + 1] - base;
  const int64_t q = (int64_t)tile_start[blockIdx.x] + threadIdx.x;
  if (q >= n) return;
  const KnnGrid g = grids[p];
  const double* r = corr + 4 * (base + q);
  const float x = (float)r[0], y = (float)r[1];
  const int64_t pos = (int64_t)cell_off[g.cell_base + cell_of[base + q]] + rank[base + q];
  spts[pos] = make_float4(x, y, (float)r[2] - x, (float)r[3] - y);
  sidx[pos] = (int32_t)q;
}

// one thread per sorted position (neighbouring threads query neighbouring
// cells); writes the features of the point at that position
__global__ __launch_bounds__(256) void k_knn_query(const double* __restrict__ corr,
                                                   const int64_t* __restrict__ pair_off,
                                                   const int32_t* __restrict__ sizes, const double* __restrict__ ratio,
                                                   const KnnGrid* __restrict__ grids,
                                                   const int32_t* __restrict__ tile_pair,
                                                   const int32_t* __restrict__ tile_start,
                                                   const int32_t* __restrict__ cell_off,
                                                   const float4* __restrict__ spts, const int32_t* __restrict__ sidx,
                                                   float* __restrict__ feat) {
  const int p = tile_pair[blockIdx.x];
  const int64_t base = pair_off[p];
  const int64_t n = pair_off[p + 1] - base;
  const int64_t sp = (int64_t)tile_start[blockIdx.x] + threadIdx.x;
  if (sp >= n) return;
  const KnnGrid g = grids[p];
  const float4 me = spts[base + sp];
  const int32_t q = sidx[base + sp];
  const float xq = me.x, yq = me.y, uxq = me.z, uyq = me.w;
  const int cx = knn_axis_cell(xq, g.inv_cs, g.gx), cy = knn_axis_cell(yq, g.inv_cs, g.gy);
  const double cs = 1.0 / (double)g.inv_cs;
  const double margin = 1e-5 * (fabs((double)xq) + fabs((double)yq) + cs * (double)(g.gx + g.gy + 2)) + 1e-3;
  const int32_t* coff = cell_off + g.cell_base;
  float td[8], tx[8], ty[8];
  int32_t ti[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    td[k] = INFINITY;
    ti[k] = -1;
    tx[k] = 0.0f;
    ty[k] = 0.0f;
  }
  auto visit = [&](int cell) {
    const int32_t b0 = coff[cell], b1 = coff[cell + 1];
    for (int32_t s = b0; s < b1; ++s) {
      const float4 cj = spts[s];
      const int32_t j = sidx[s];
      const float dx = xq - cj.x, dy = yq - cj.y;
      const float d = dx * dx + dy * dy;
      if (j == q || !(d < td[7] || (d == td[7] && j < ti[7]))) continue;
#pragma unroll
      for (int k = 7; k > 0; --k) {
        if (d < td[k - 1] || (d == td[k - 1] && j < ti[k - 1])) {
          td[k] = td[k - 1];
          ti[k] = ti[k - 1];
          tx[k] = tx[k - 1];
          ty[k] = ty[k - 1];
        } else if (d < td[k] || (d == td[k] && j < ti[k])) {
          td[k] = d;
          ti[k] = j;
          tx[k] = cj.z;
          ty[k] = cj.w;
        }
      }
      if (d < td[0] || (d == td[0] && j < ti[0])) {
        td[0] = d;
        ti[0] = j;
        tx[0] = cj.z;
        ty[0] = cj.w;
      }
    }
  };
  const int rmax = g.gx > g.gy ? g.gx : g.gy;
  for (int R = 0; R <= rmax; ++R) {
    const int i0 = cx - R, i1 = cx + R, j0 = cy - R, j1 = cy + R;
    for (int j = j0 < 0 ? 0 : j0; j <= (j1 < g.gy ? j1 : g.gy - 1); ++j) {
      if (j == j0 || j == j1) {
        for (int i = i0 < 0 ? 0 : i0; i <= (i1 < g.gx ? i1 : g.gx - 1); ++i) visit(j * g.gx + i);
      } else {
        if (i0 >= 0) visit(j * g.gx + i0);
        if (i1 < g.gx && i1 != i0) visit(j * g.gx + i1);
      }
    }
    // lower bound of the distance of every point outside rings 0..R
    const bool lo_x = i0 > 0, hi_x = i1 < g.gx - 1, lo_y = j0 > 0, hi_y = j1 < g.gy - 1;
    if (!(lo_x || hi_x || lo_y || hi_y)) break;  // every cell visited
    double b = INFINITY;
    if (lo_x) b = fmin(b, (double)xq - (double)i0 * cs);
    if (hi_x) b = fmin(b, (double)(i1 + 1) * cs - (double)xq);
    if (lo_y) b = fmin(b, (double)yq - (double)j0 * cs);
    if (hi_y) b = fmin(b, (double)(j1 + 1) * cs - (double)yq);
    b -= margin;
    if (b > 0.0 && (double)td[7] < b * b * (1.0 - 1e-5)) break;
  }
  float sx = 0.0f, sy = 0.0f;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (ti[k] >= 0) {
      sx = sx + tx[k];
      sy = sy + ty[k];
      ++c;
    }
  }
  const int h1 = sizes[4 * p + 0], w1 = sizes[4 * p + 1], h2 = sizes[4 * p + 2], w2 = sizes[4 * p + 3];
  const float s1x = 2.0f / (float)w1, s1y = 2.0f / (float)h1, s2x = 2.0f / (float)w2, s2y = 2.0f / (float)h2;
  const float sd = 2.0f / (float)(w1 > h1 ? w1 : h1);
  float f5 = 0.0f, f6 = 0.0f;
  if (c > 0) {
    const float fc = (float)c;
    f5 = (sx / fc - uxq) * sd;
    f6 = (sy / fc - uyq) * sd;
  }
  const double* r = corr + 4 * (base + q);
  const float x2q = (float)r[2], y2q = (float)r[3];
  float4* out = reinterpret_cast<float4*>(feat + IN * (base + q));
  out[0] = make_float4(xq * s1x - 1.0f, yq * s1y - 1.0f, x2q * s2x - 1.0f, y2q * s2y - 1.0f);
  out[1] = make_float4((float)ratio[base + q], f5, f6, 0.0f);
}

void launch_knn_bin(hipStream_t st, const double* corr, const int64_t* pair_off, const KnnGrid* grids,
                    const int32_t* tile_pair, const int32_t* tile_start, int tiles, int32_t* cell_count,
                    int32_t* cell_of, int32_t* rank) {
  if (tiles <= 0) return;
  hipLaunchKernelGGL(k_knn_bin, dim3(tiles), dim3(256), 0, st, corr, pair_off, grids, tile_pair, tile_start,
                     cell_count, cell_of, rank);
}

void launch_knn_scatter(hipStream_t st, const double* corr, const int64_t* pair_off, const KnnGrid* grids,
                        const int32_t* tile_pair, const int32_t* tile_start, int tiles, const int32_t* cell_off,
                        const int32_t* cell_of, const int32_t* rank, float4* spts, int32_t* sidx) {
  if (tiles <= 0) return;
  hipLaunchKernelGGL(k_knn_scatter, dim3(tiles), dim3(256), 0, st, corr, pair_off, grids, tile_pair, tile_start,
                     cell_off, cell_of, rank, spts, sidx);
}

void launch_knn_query(hipStream_t st, const double* corr, const int64_t* pair_off, const int32_t* sizes,
                      const double* ratio, const KnnGrid* grids, const int32_t* tile_pair, const int32_t* tile_start,
                      int tiles, const int32_t* cell_off, const float4* spts, const int32_t* sidx, float* feat) {
  if (tiles <= 0) return;
  hipLaunchKernelGGL(k_knn_query, dim3(tiles), dim3(256), 0, st, corr, pair_off, sizes, ratio, grids, tile_pair,
                     tile_start, cell_off, spts, sidx, feat);
}


// ---- packing, forward, training -------------------------------------------

__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ wpk,
                                                  float* __restrict__ bias) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < HID * HID) {
    const int r = t / HID, c = t % HID;
    const uint16_t a = f2bf(p[OFF_W2 + t]), b = f2bf(p[OFF_W3 + t]);
    wpk[PK_W2 + t] = a;
    wpk[PK_W2T + c * HID + r] = a;
    wpk[PK_W3 + t] = b;
    wpk[PK_W3T + c * HID + r] = b;
  }
  if (t < HID * IN) wpk[PK_W1 + t] = f2bf(p[OFF_W1 + t]);
  if (t < HID) {
    wpk[PK_W4 + t] = f2bf(p[OFF_W4 + t]);
    bias[t] = p[OFF_B1 + t];
    bias[HID + t] = p[OFF_B2 + t];
    bias[2 * HID + t] = p[OFF_B3 + t];
  }
  if (t == 0) bias[3 * HID] = p[OFF_B4];
}

__global__ __launch_bounds__(256) void k_mlp_predict(const uint16_t* __restrict__ wpk, const float* __restrict__ bias,
                                                     const float* __restrict__ feat, int64_t n,
                                                     float* __restrict__ logit, float* __restrict__ prob) {
  extern __shared__ __align__(16) unsigned char mlp_lds_f[];
  FwdLds& s = *reinterpret_cast<FwdLds*>(mlp_lds_f);
  stage_forward(s, wpk, bias);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t b0 = (int64_t)blockIdx.x * POINTS_PER_BLOCK; b0 < n; b0 += (int64_t)gridDim.x * POINTS_PER_BLOCK) {
    const int64_t col = b0 + 32 * w + (lane & 31);
    const bool valid = col < n;
    int wj = lane & 31;
    asm volatile("" : "+v"(wj));
    const float z = forward_wave<false>(s, valid ? feat + IN * col : nullptr, lane, wj, nullptr, nullptr, 0, 0);
    if ((lane >> 5) == 0 && valid) {
      logit[col] = z;
      if (prob) prob[col] = sigmoid_f(z);
    }
  }
}

__global__ __launch_bounds__(256) void k_mlp_train(const uint16_t* __restrict__ wpk, const float* __restrict__ bias,
                                                   const float* __restrict__ feat, const float* __restrict__ labels,
                                                   int64_t batch, int64_t bpad, int64_t sample_n, uint64_t seed,
                                                   int64_t step, float inv_batch, float* __restrict__ zbuf,
                                                   uint16_t* __restrict__ act, float* __restrict__ d4out,
                                                   float* __restrict__ lossv) {
  extern __shared__ __align__(16) unsigned char mlp_lds_t[];
  TrainLds& s = *reinterpret_cast<TrainLds*>(mlp_lds_t);
  stage_forward(s.f, wpk, bias);
  stage_matrix(s.w2t, wpk + PK_W2T);
  stage_matrix(s.w3t, wpk + PK_W3T);
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, j = lane & 31, half = lane >> 5;
  const int64_t wave_g = (int64_t)blockIdx.x * 4 + w;
  const int64_t col = wave_g * 32 + j;
  const bool valid = col < batch;
  int64_t idx = col;
  if (valid && sample_n > 0)
    idx = (int64_t)(mix64(seed ^ mix64((uint64_t)step * 0x100000001B3ull + (uint64_t)col)) % (uint64_t)sample_n);
  float* zw = zbuf + wave_g * (3 * ZL);
  const float z4 = forward_wave<true>(s.f, valid ? feat + IN * idx : nullptr, lane, j, zw, act, bpad, col);
  const float y = valid ? labels[idx] : 0.0f;
  const float d4 = valid ? (sigmoid_f(z4) - y) * inv_batch : 0.0f;
  if (half == 0) {
    d4out[col] = d4;
    lossv[col] = valid ? (softplus_f(z4) - y * z4) * inv_batch : 0.0f;
  }
  // D3 = w4 d4 * gelu'(Z3)
  bf16x4 da[16], db[16];
  uint16_t* d3row = act + (int64_t)ROW_D3 * bpad;
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = 8 * kb + 4 * half + e;
      const float d = (s.f.w4[i] * d4) * gelu_grad_f(zw[2 * ZL + (4 * kb + e) * 64 + lane]);
      const uint16_t b = f2bf(d);
      da[kb][e] = (short)b;
      d3row[(int64_t)i * bpad + col] = b;
    }
  }
  uint16_t* d2row = act + (int64_t)ROW_D2 * bpad;
  backward_tile<0, true>(s.w3t, da, j, half, lane, zw + ZL, db, d2row, bpad, col);
  backward_tile<1, true>(s.w3t, da, j, half, lane, zw + ZL, db, d2row, bpad, col);
  backward_tile<2, true>(s.w3t, da, j, half, lane, zw + ZL, db, d2row, bpad, col);
  backward_tile<3, true>(s.w3t, da, j, half, lane, zw + ZL, db, d2row, bpad, col);
  uint16_t* d1row = act + (int64_t)ROW_D1 * bpad;
  backward_tile<0, false>(s.w2t, db, j, half, lane, zw, nullptr, d1row, bpad, col);
  backward_tile<1, false>(s.w2t, db, j, half, lane, zw, nullptr, d1row, bpad, col);
  backward_tile<2, false>(s.w2t, db, j, half, lane, zw, nullptr, d1row, bpad, col);
  backward_tile<3, false>(s.w2t, db, j, half, lane, zw, nullptr, d1row, bpad, col);
}

// Weight gradients of one chunk of WGRAD_CHUNK batch columns.  Wave w owns
// output rows 32w .. 32w+31 of each 128 x 128 gradient (four 32x32 tiles);
// biases, w4, b4 and the loss are sequential column-order sums.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t* __restrict__ act, const float* __restrict__ d4,
                                                   const float* __restrict__ lossv, int64_t bpad,
                                                   float* __restrict__ part) {
  const int c = blockIdx.x;
  const int64_t p0 = (int64_t)c * WGRAD_CHUNK;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, j = lane & 31, half = lane >> 5;
  float* out = part + (int64_t)c * PART_STRIDE;
#pragma unroll
  for (int job = 0; job < 2; ++job) {
    const uint16_t* D = act + (int64_t)(job == 0 ? ROW_D3 : ROW_D2) * bpad;
    const uint16_t* H = act + (int64_t)(job == 0 ? ROW_H2 : ROW_H1) * bpad;
    const int off = job == 0 ? OFF_W3 : OFF_W2;
    f32x16 acc[4] = {zero16(), zero16(), zero16(), zero16()};
    const uint16_t* arow = D + (int64_t)(32 * w + j) * bpad + p0 + 4 * half;
    const uint16_t* brow = H + (int64_t)j * bpad + p0 + 4 * half;
    for (int kb = 0; kb < WGRAD_CHUNK / 8; ++kb) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(arow + 8 * kb);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        acc[nt] = mfma(a, *reinterpret_cast<const bf16x4*>(brow + (int64_t)32 * nt * bpad + 8 * kb), acc[nt]);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = 32 * w + 8 * (r >> 2) + 4 * half + (r & 3);
        out[off + k * HID + 32 * nt + j] = acc[nt][r];
      }
    }
  }
  {
    // W1: D1 x X^T, input columns j < 8
    f32x16 acc = zero16();
    const uint16_t* arow = act + (int64_t)(ROW_D1 + 32 * w + j) * bpad + p0 + 4 * half;
    const uint16_t* brow = act + (int64_t)(ROW_X + (j < IN ? j : 0)) * bpad + p0 + 4 * half;
    bf16x4 zb;
    zb[0] = zb[1] = zb[2] = zb[3] = 0;
    for (int kb = 0; kb < WGRAD_CHUNK / 8; ++kb) {
      const bf16x4 a = *reinterpret_cast<const bf16x4*>(arow + 8 * kb);
      const bf16x4 b = j < IN ? *reinterpret_cast<const bf16x4*>(brow + 8 * kb) : zb;
      acc = mfma(a, b, acc);
    }
    if (j < IN) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int k = 32 * w + 8 * (r >> 2) + 4 * half + (r & 3);
        out[OFF_W1 + k * IN + j] = acc[r];
      }
    }
  }
  for (int row = threadIdx.x; row < 4 * HID + 2; row += blockDim.x) {
    float sum = 0.0f;
    if (row < 3 * HID) {
      const int layer = row / HID, i = row % HID;
      const uint16_t* src = act + (int64_t)(ROW_D1 + row) * bpad + p0;
      for (int t = 0; t < WGRAD_CHUNK; ++t) sum = sum + bf2f(src[t]);
      out[(layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : OFF_B3) + i] = sum;
    } else if (row < 4 * HID) {
      const int i = row - 3 * HID;
      const uint16_t* src = act + (int64_t)(ROW_H3 + i) * bpad + p0;
      for (int t = 0; t < WGRAD_CHUNK; ++t) sum = sum + d4[p0 + t] * bf2f(src[t]);
      out[OFF_W4 + i] = sum;
    } else if (row == 4 * HID) {
      for (int t = 0; t < WGRAD_CHUNK; ++t) sum = sum + d4[p0 + t];
      out[OFF_B4] = sum;
    } else {
      for (int t = 0; t < WGRAD_CHUNK; ++t) sum = sum + lossv[p0 + t];
      out[NPARAM] = sum;
    }
  }
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int chunks,
                                                    float* __restrict__ grad, float* __restrict__ loss_out, int adam,
                                                    float* __restrict__ params, float* __restrict__ m,
                                                    float* __restrict__ v, float lr, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > NPARAM) return;
  float g = 0.0f;
  for (int c = 0; c < chunks; ++c) g = g + part[(int64_t)c * PART_STRIDE + q];
  if (q == NPARAM) {
    if (loss_out) *loss_out = g;
    return;
  }
  if (grad) grad[q] = g;
  if (adam) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[q] + (1.0f - b1) * g;
    const float vv = b2 * v[q] + (1.0f - b2) * (g * g);
    m[q] = mm;
    v[q] = vv;
    const float mh = mm / c1, vh = vv / c2;
    params[q] = params[q] - lr * mh / (sqrtf(vh) + eps);
  }
}

// ---- guided-batch helpers ----------------------------------------------------

__device__ inline int find_pair(const int64_t* pair_off, int num_pairs, int64_t k) {
  int lo = 0, hi = num_pairs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pair_off[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void k_local_iota(const int64_t* __restrict__ pair_off, int num_pairs, int64_t total,
                             int32_t* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int p = find_pair(pair_off, num_pairs, pair_off[0] + k);
  out[k] = (int32_t)(pair_off[0] + k - pair_off[p]);
}

__global__ void k_gather_rows(const double* __restrict__ corr, const int32_t* __restrict__ perm,
                              const int64_t* __restrict__ pair_off, int num_pairs, int64_t total,
                              double* __restrict__ sorted) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const int p = find_pair(pair_off, num_pairs, pair_off[0] + k);
  const int64_t src = pair_off[p] - pair_off[0] + perm[k];
#pragma unroll
  for (int c = 0; c < 4; ++c) sorted[4 * k + c] = corr[4 * src + c];
}

// ---- launchers ------------------------------------------------------------------

void launch_pack(hipStream_t st, const float* params, uint16_t* wpk, float* bias) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(HID * HID / 256), dim3(256), 0, st, params, wpk, bias);
}

void launch_predict(hipStream_t st, const uint16_t* wpk, const float* bias, const float* feat, int64_t n,
                    float* logit, float* prob) {
  if (n <= 0) return;
  int64_t blocks = (n + POINTS_PER_BLOCK - 1) / POINTS_PER_BLOCK;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)blocks), dim3(256), sizeof(FwdLds), st, wpk, bias, feat, n,
                     logit, prob);
}

void launch_train_fwd_bwd(hipStream_t st, const uint16_t* wpk, const float* bias, const float* feat,
                          const float* labels, int64_t batch, int64_t bpad, int64_t sample_n, uint64_t seed,
                          int64_t step, float inv_batch, float* zbuf, uint16_t* act, float* d4, float* lossv) {
  hipLaunchKernelGGL(k_mlp_train, dim3((unsigned)(bpad / POINTS_PER_BLOCK)), dim3(256), sizeof(TrainLds), st, wpk,
                     bias, feat, labels, batch, bpad, sample_n, seed, step, inv_batch, zbuf, act, d4, lossv);
}

void launch_wgrad(hipStream_t st, const uint16_t* act, const float* d4, const float* lossv, int64_t bpad,
                  float* part) {
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((unsigned)(bpad / WGRAD_CHUNK)), dim3(256), 0, st, act, d4, lossv, bpad,
                     part);
}

void launch_reduce_adam(hipStream_t st, const float* part, int chunks, float* grad, float* loss_out, int adam,
                        float* params, float* m, float* v, float lr, float c1, float c2) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, part, chunks, grad, loss_out,
                     adam, params, m, v, lr, c1, c2);
}

void launch_local_iota(hipStream_t st, const int64_t* pair_off, int num_pairs, int64_t total, int32_t* out) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_local_iota, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pair_off, num_pairs,
                     total, out);
}

void launch_gather_rows(hipStream_t st, const double* corr, const int32_t* perm, const int64_t* pair_off,
                        int num_pairs, int64_t total, double* sorted) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, corr, perm, pair_off,
                     num_pairs, total, sorted);
}

}  // namespace mlp
}  // namespace gcr
