// Learned inlier-probability network: layout constants and kernel launchers.
//
// Network 8-128-128-128-1 (7 features + a zero pad), GELU (tanh form) on the
// hidden layers, sigmoid output.  Parameters are one FP32 array of
// NPARAM entries (row-major W[out][in]):
//   W1 [128][8] | b1 [128] | W2 [128][128] | b2 | W3 [128][128] | b3 | w4 [128] | b4
// The kernels run on BF16 copies of the weights (packed by launch_mlp_pack)
// with FP32 accumulation on the matrix cores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gcr {
namespace mlp {

constexpr int IN = 8;
constexpr int HID = 128;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + HID * IN;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID * HID;
constexpr int OFF_W4 = OFF_B3 + HID;
constexpr int OFF_B4 = OFF_W4 + HID;
constexpr int NPARAM = OFF_B4 + 1;
// one row of the per-chunk gradient partials: NPARAM gradients + the loss
constexpr int PART_STRIDE = NPARAM + 3;  // 34,308: 16-byte rows

// packed BF16 weights (uint16 offsets): W1, W2, W3, W2^T, W3^T, w4
constexpr int PK_W1 = 0;
constexpr int PK_W2 = PK_W1 + HID * IN;
constexpr int PK_W3 = PK_W2 + HID * HID;
constexpr int PK_W2T = PK_W3 + HID * HID;
constexpr int PK_W3T = PK_W2T + HID * HID;
constexpr int PK_W4 = PK_W3T + HID * HID;
constexpr int PK_TOTAL = PK_W4 + HID;
// FP32 biases: b1 | b2 | b3 | b4
constexpr int PB_TOTAL = 3 * HID + 1;

// training activations, BF16, neuron-major rows of `bpad` columns:
// X (8 rows) | H1 | H2 | H3 | D1 | D2 | D3 (128 rows each)
constexpr int ROW_X = 0, ROW_H1 = 8, ROW_H2 = ROW_H1 + HID, ROW_H3 = ROW_H2 + HID, ROW_D1 = ROW_H3 + HID,
              ROW_D2 = ROW_D1 + HID, ROW_D3 = ROW_D2 + HID, ACT_ROWS = ROW_D3 + HID;

constexpr int POINTS_PER_BLOCK = 128;  // 4 waves x 32 points
constexpr int WGRAD_CHUNK = 512;       // batch columns per weight-gradient workgroup

// dynamic LDS of the kernels (bytes)
int lds_forward_bytes();
int lds_train_bytes();

// Features (7 + a zero pad per correspondence) with a per-pair grid
// (k_mlp.hip): pair p's points are
// binned into grids[p].gx x grids[p].gy cells of side 1 / inv_cs, whose
// counters start at cell_base in the global cell arrays.  bin: cell_count
// (zeroed by the caller) counts, cell_of / rank per point; the caller forms
// cell_off, the exclusive prefix sum of cell_count (one entry more than the
// cells); scatter: points sorted by cell into spts (x1, y1, x2 - x1,
// y2 - y1) and sidx (index in the pair); query: the features.
struct KnnGrid {
  float inv_cs;
  int32_t gx, gy;
  int32_t pad_;
  int64_t cell_base;
};
void launch_knn_bin(hipStream_t st, const double* corr, const int64_t* pair_off, const KnnGrid* grids,
                    const int32_t* tile_pair, const int32_t* tile_start, int tiles, int32_t* cell_count,
                    int32_t* cell_of, int32_t* rank);
void launch_knn_scatter(hipStream_t st, const double* corr, const int64_t* pair_off, const KnnGrid* grids,
                        const int32_t* tile_pair, const int32_t* tile_start, int tiles, const int32_t* cell_off,
                        const int32_t* cell_of, const int32_t* rank, float4* spts, int32_t* sidx);
void launch_knn_query(hipStream_t st, const double* corr, const int64_t* pair_off, const int32_t* sizes,
                      const double* ratio, const KnnGrid* grids, const int32_t* tile_pair, const int32_t* tile_start,
                      int tiles, const int32_t* cell_off, const float4* spts, const int32_t* sidx, float* feat);
// params (FP32) -> packed BF16 weights + FP32 biases
void launch_pack(hipStream_t st, const float* params, uint16_t* wpk, float* bias);
// logits of n points; prob (nullable) = sigmoid(logit)
void launch_predict(hipStream_t st, const uint16_t* wpk, const float* bias, const float* feat, int64_t n,
                    float* logit, float* prob);
// forward + backward-data of `batch` columns padded to `bpad` (a multiple of
// WGRAD_CHUNK).  Column c < batch reads point idx(c): c itself when
// sample_n == 0, else a hash of (seed, step, c) modulo sample_n.
void launch_train_fwd_bwd(hipStream_t st, const uint16_t* wpk, const float* bias, const float* feat,
                          const float* labels, int64_t batch, int64_t bpad, int64_t sample_n, uint64_t seed,
                          int64_t step, float inv_batch, float* zbuf, uint16_t* act, float* d4, float* lossv);
// per-chunk gradient partials: part[c * PART_STRIDE + q]
void launch_wgrad(hipStream_t st, const uint16_t* act, const float* d4, const float* lossv, int64_t bpad,
                  float* part);
// sums the partials of `chunks` chunks in chunk order; grad (nullable),
// loss_out (nullable) = the summed loss; adam != 0 applies one Adam step
// with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t
void launch_reduce_adam(hipStream_t st, const float* part, int chunks, float* grad, float* loss_out, int adam,
                        float* params, float* m, float* v, float lr, float c1, float c2);
// out[k] = k - (first row of k's pair), rows relative to pair_off[0]
void launch_local_iota(hipStream_t st, const int64_t* pair_off, int num_pairs, int64_t total, int32_t* out);
// sorted[k] = corr[first row of k's pair + perm[k]] (perm holds local indices)
void launch_gather_rows(hipStream_t st, const double* corr, const int32_t* perm, const int64_t* pair_off,
                        int num_pairs, int64_t total, double* sorted);

}  // namespace mlp
}  // namespace gcr
