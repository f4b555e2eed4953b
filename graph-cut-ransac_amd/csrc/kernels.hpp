// Kernel launchers of the engine (one stream argument each; all pointers are
// device pointers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gcr {

// minimal solvers
void launch_h4_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* status, int8_t* valid);

void launch_f7_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* mstatus, int8_t* nmodels);
void launch_fit_f(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, const double* const* wmodels, const double* pair_sq_thr, int flags,
                  double* scratch, const int64_t* scratch_off, double* models_out, int8_t* ok_out, int num_fits,
                  int32_t* lm_passes, int max_ns);

// scoring (kind 0 homography transfer error, 1 Sampson).  With active
// statuses, no masks and work_list (num_models int32) + work_count (one
// uint32) given, inactive slots get their empty result from a gather kernel
// and only the active slots are scored, from a list.
void launch_score(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                  const double* pair_sq_thr, const double* models, const int32_t* model_pair, const int8_t* active,
                  int64_t num_models, int32_t* counts, double* values, uint8_t* masks, const int64_t* mask_off,
                  int32_t* sym_counts = nullptr, bool candidates = false, int slots_per = 1,
                  int32_t* work_list = nullptr, uint32_t* work_count = nullptr);

// Candidate scoring on the matrix cores (FP16 hi/lo MFMA rejection of
// proven outliers, FP64 in point order for every other point; same results as
// launch_score).  MxPoint: one point's B-operand slices (FP16 bits, k order,
// [MFMA][lane half][8]) and error bounds; launch_mx_prepare fills them and
// the per-pair power-of-two inverse scale once per engine run.
struct alignas(16) MxPoint {
  uint16_t b[2][2][8];
  float d[4];
};
void launch_mx_prepare(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                       const int32_t* pt_pair, int num_pairs, int64_t total, double* inv_scale, MxPoint* out);
// ints of blk_scratch that launch_score_mx needs for num_models slots, and
// words of its candidate mask for pairs of at most max_n points
size_t mx_list_scratch_ints(int64_t num_models);
size_t mx_mask_words(int max_n);
void launch_score_mx(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                     const double* pair_sq_thr, const double* pair_inv_scale, const MxPoint* mxp,
                     const double* models, const int32_t* model_pair, const int8_t* active, int64_t num_models,
                     int32_t* counts, double* values, int32_t* sym_counts, int slots_per, int32_t* work_list,
                     uint32_t* work_count, int32_t* blk_scratch, int max_n, uint32_t* mask);

// non-minimal fits
void launch_fit_h(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, double* scratch, const int64_t* scratch_off, double* models_out,
                  int8_t* ok_out, int num_fits, int max_ns);

// DEGENSAC: H-degenerate test of the seven-point models (slot / 3 =
// candidate, samples 7 per candidate) and the plane-and-parallax minimal
// solver of the nested run (samples 2 per candidate, pair_h per pair)
void launch_degen_test(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                       const int32_t* samples, const double* models, const int8_t* status, int64_t num_slots,
                       int8_t* degen, double* degen_h);
void launch_pp_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, const double* const* pair_h, int64_t num_cand, double* models,
                     int8_t* mstatus, int8_t* nmodels);

// dynamic LDS bytes one workgroup may allocate on the current device
int lds_limit_bytes();

// neighbourhood, energies, min-cut, compaction
void launch_grid_keys(hipStream_t st, const double* pts, const int32_t* pt_pair, const int64_t* pair_pt_off,
                      const double* pair_cell, const int64_t* pair_div, int64_t total, uint64_t* keys,
                      int32_t* vals);
void launch_grid_runs(hipStream_t st, const uint64_t* skeys, const int32_t* svals, const int32_t* pt_pair,
                      const int64_t* pair_pt_off, const int32_t* pair_n, int64_t total, int32_t* run_start,
                      int32_t* run_len, int32_t* pos_of, int64_t* out_deg);
void launch_grid_cells(hipStream_t st, const int32_t* run_start, const int32_t* run_len, const int32_t* pt_pair,
                       const int64_t* pair_pt_off, int64_t total, int32_t* cell_pair, int32_t* cell_start,
                       int32_t* cell_len, int32_t* num_cells);
void launch_lo_dist(hipStream_t st, int kind, const double* pts, const int32_t* pt_pair, const int8_t* pair_active,
                    const double* pair_lo_thr, const double* lo_models, int64_t total, double* dist, int8_t* near);
// pair_radius[p] != 0: pair p's edges are its radius graph's (r_off / r_nbr
// out-neighbours, adj_off incident-arc offsets); else its grid cells'
void launch_lo_caps(hipStream_t st, const int32_t* pt_pair, const int8_t* pair_active, const int64_t* pair_pt_off,
                    const double* pair_lambda, const int32_t* pair_qbits, const double* dist, const int8_t* near,
                    const int32_t* svals, const int32_t* pos_of, const int32_t* run_len, const int64_t* edge_off,
                    const int8_t* pair_radius, const int64_t* r_off, const int32_t* r_nbr, const int64_t* adj_off,
                    int64_t total, int64_t* rf, int64_t* rb, int64_t* node_src, int64_t* node_sink,
                    uint8_t* sink_out);
// grid cells of up to this many points run with their state in LDS; larger
// ones use gh / ghn (one int32 per point of the batch) and node_src /
// node_sink in place.  pair_fail[p] = 1 when a cell of pair p did not
// converge.
constexpr int kMaxLdsCell = 2048;
void launch_mincut_cells(hipStream_t st, const int32_t* cell_pair, const int32_t* cell_start, const int32_t* cell_len,
                         int num_cells, const int8_t* pair_active, const double* pair_lambda,
                         const int64_t* pair_pt_off, const int32_t* svals, const int64_t* edge_off, int64_t* rf,
                         int64_t* rb, int64_t* node_src, int64_t* node_sink, int32_t* gh, int32_t* ghn,
                         uint8_t* sink_out, int32_t* pair_fail, int max_cell);
// radius neighbourhood (k_radius.hip): tiles of 256 query points of one pair
void launch_radius_search(hipStream_t st, bool fill, const double* pts, const int32_t* tile_pair,
                          const int32_t* tile_first, int num_tiles, const int64_t* pair_pt_off, const int32_t* pair_n,
                          const float* pair_r2, int64_t* out_cnt, int64_t* in_cnt, const int64_t* r_off,
                          int32_t* r_nbr, uint32_t* r_key);
void launch_radius_adjacency(hipStream_t st, const int32_t* pt_pair, const int8_t* pair_radius,
                             const int64_t* pair_pt_off, const int64_t* r_off, const int32_t* r_nbr,
                             const int64_t* edge_off, const int64_t* adj_off, const int64_t* out_cnt,
                             int32_t* in_cursor, int64_t total, int64_t* adj_arc, int32_t* adj_other);
// a[i] += b[i]
void launch_add_i64(hipStream_t st, int64_t* a, const int64_t* b, int64_t n);
// one workgroup per listed pair; node state ex / rt (node_src / node_sink,
// in place), h / hn and two frontiers (queue, 2 per point) in global memory
void launch_mincut_radius(hipStream_t st, const int32_t* rpairs, int num_rpairs, const int8_t* pair_active,
                          const double* pair_lambda, const int64_t* pair_pt_off, const int32_t* pair_n,
                          const int64_t* adj_off, const int64_t* adj_arc, const int32_t* adj_other, int64_t* rf,
                          int64_t* rb, int64_t* ex, int64_t* rt, int32_t* h, int32_t* hn, int32_t* queue,
                          uint8_t* sink_out, int32_t* pair_fail);
void launch_compact(hipStream_t st, const uint8_t* mask, const int64_t* list_off, const int32_t* list_n, int num_lists,
                    int32_t* out, int32_t* out_count);

// Space partitioning (dev_sp.hpp): a pair's filled grid cells in the grid
// container's order; start / len index the pair's cell-order permutation.
struct SpCell {
  uint64_t idx[4];  // the cell's axis indices (first point's floor(x / cell size), as size_t)
  int32_t start, len;
};
struct SpPair {
  int64_t cell_off;  // first cell of the pair in the batch's cell array
  int32_t ncells, pad;
  double cs[4];      // grid cell sizes
  double P;          // divisions per axis
  double tau;        // truncated threshold (3/2 threshold)
};
// MSAC over the kept cells' points in cell order (scoring_function.h:186-214)
// of every homography slot of a space-partitioning pair (pair_sp[pair] != 0,
// status 2): counts / values overwritten, kept point counts to sel.
void launch_score_sp(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                     const double* pair_sq_thr, const int8_t* pair_sp, const SpPair* sp_pairs, const SpCell* sp_cells,
                     const int32_t* sp_perm, const int32_t* sp_cell_of, const double* models, const int32_t* cpair,
                     const int8_t* status, int64_t num_slots, int32_t* counts, double* values, int32_t* sel,
                     int max_cells);
// the inliers of item i's model (models + 9 i, pair item_pair[i]) among its
// kept cells, in cell order, into lists + list_off[i]; count to counts[i];
// items with item_sp[i] == 0 are left alone
void launch_sp_list(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* pair_n,
                    const double* thr, const SpPair* sp_pairs, const SpCell* sp_cells, const int32_t* sp_perm,
                    const int32_t* sp_cell_of, const double* models, const int32_t* item_pair, const int8_t* item_sp,
                    const int64_t* list_off, int num_items, int32_t* lists, int32_t* counts, int max_cells);

// SPRT walk (k_sprt.hip): one request per pair, filled by the host and
// updated in place by the device
struct SprtItem {
  const double* models;    // the chunk's candidate models (9 per slot)
  const int32_t* slots;    // model slots to test, in the loop's order
  const double* a_table;   // A for delta = k / N at the current epsilon (N + 1 entries)
  int32_t num_slots, pair, kind, delta_changes;
  int64_t pool_idx;        // running pool index (in / out)
  double eps, delta, A;    // the test's current history (delta, A in / out)
  double best_value;       // so-far-the-best score value
  int32_t stop, last;      // slot index that stopped the walk (accepted, above best) / last tested, or -1
  int64_t last_start, last_tested, last_count;
  double last_value;
  int32_t last_accepted;
  int32_t sp;              // the models go through the space partitioning selector first
  int64_t best_count;      // so-far-the-best inlier number (the selector's skip rule)
};
// sp_list: per point of the batch, the tested inliers of the last model an
// SP walk tested (in test order, from the pair's base)
void launch_sprt_walk(hipStream_t st, SprtItem* items, int num_items, const double* pts, const int64_t* pair_pt_off,
                      const int32_t* pair_n, const double* pair_sq_thr, const int32_t* pools, const SpPair* sp_pairs,
                      const SpCell* sp_cells, const int32_t* sp_perm, const int32_t* sp_cell_of, int32_t* sp_list,
                      int max_cells);

// Candidate results filtered on the device (k_util.hip): per item (one
// pair's candidates base .. base + k - 1, S slots each), the model slots
// whose value beats the item's best value, in slot order, and each
// candidate's last model's count.  For homographies (h != 0) a slot holds a
// model when its status is 2, and nmodels is written here.
// The uniform sampler's generator state of one pair on the device
// (std::mt19937's 624 words and the position of the next word, 624: the
// block is used up).
struct DevMT {
  uint32_t x[624];
  int32_t pos;
  int32_t pad_[3];
};
// One pair's draws of a candidate chunk: count samples of m indices into
// samples[(base + s) * m + k], in nseg segments from segs[seg_off].
struct DevSampleItem {
  int32_t pair;
  int32_t m;
  int64_t base;
  int64_t count;
  int32_t seg_off;
  int32_t nseg;
};
// count consecutive samples whose first `draws` indices are unique draws in
// [0, range) and whose last index is `fixed` when fixed >= 0; range 0: the
// sampler refuses (no draws, every index -1)
struct DevSampleSeg {
  int64_t start;  // first sample of the segment (from the item's first)
  int64_t count;
  uint32_t range;
  int32_t draws;
  int32_t fixed;
  int32_t pad_;
};
void launch_mt_seed(hipStream_t st, const uint32_t* seeds, const int32_t* pairs, int num, DevMT* states);
void launch_sample_streams(hipStream_t st, const DevSampleItem* items, const DevSampleSeg* segs, int num_items,
                           DevMT* states, int32_t* samples);

struct SlotRec {
  int32_t slot, count;  // slot relative to the item's first candidate
  double value;
  int32_t sym;
  int8_t valid, degen, pad[2];
  int32_t sample0;  // first index of the candidate's sample
  int32_t pad2;
};
void launch_chunk_filter(hipStream_t st, int h, int S, int num_items, const int64_t* item_base,
                         const int32_t* item_ncand, const double* item_best, const int8_t* status, int8_t* nmodels,
                         const int32_t* counts, const double* values, const int32_t* sym, const int8_t* valid,
                         const int8_t* degen, const int32_t* samples, int m, int32_t* last_count, SlotRec* recs,
                         int32_t* rec_count, SlotRec* packed);

// small utilities
void launch_fill_cand_pair(hipStream_t st, int num_items, const int64_t* base, const int32_t* count,
                           const int32_t* pair, int32_t* cand_pair);
void launch_copy_models(hipStream_t st, const double* const* src, double* const* dst, int count);
// a[off[i] .. off[i] + len[i]) = 0 for i < count
void launch_zero_ranges(hipStream_t st, double* a, const int64_t* off, const int32_t* len, int count, int max_len);
void launch_irls_weights(hipStream_t st, int kind, const double* pts, const int64_t* pair_pt_off,
                         const double* pair_sq_thr, const int32_t* item_pair, const double* const* item_model,
                         const int32_t* inliers, const int64_t* inl_off, const int32_t* inl_n, double* weights,
                         int num_items, int zero, int max_n);

// essential matrices (k_ess.hip)
void launch_e5_solve(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* cand_pair,
                     const int32_t* samples, int64_t num_cand, double* models, int8_t* mstatus, int8_t* nmodels);
void launch_fit_e(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                  const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples, const double* weights,
                  const int64_t* weight_off, int weight_mode, double* scratch, const int64_t* scratch_off,
                  double* models_out, int8_t* ok_out, int num_fits);
void launch_ba_e(hipStream_t st, const double* pts, const int64_t* pair_pt_off, const int32_t* fit_pair,
                 const int64_t* fit_sample_off, const int32_t* fit_ns, const int32_t* samples,
                 const double* const* wmodels, const double* post_sq, double* scratch, const int64_t* scratch_off,
                 double* models_out, int8_t* ok_out, int num_fits);
void launch_prepare_points(hipStream_t st, const double* pix, const int32_t* pt_pair, const int8_t* norm_pair,
                           const double* kinv, int64_t total, double* out);

}  // namespace gcr
