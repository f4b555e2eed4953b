// Graph-cut local optimisation kernels.
//
// Neighbourhood (grid_neighborhood_graph.h:229-301): every point gets the
// cell key sum_d floor(x_d / cell_d) * n^d; points sharing a key form a
// clique.  After a stable segmented sort by key, each cell is a run of
// ascending point indices, so the out-edges of point p (the same-cell points
// q > p, visited by the reference in ascending q) are the run entries after
// p's sorted position, and edge (p, q) lives at edge_off[p] + (pos(q) -
// pos(p) - 1).
//
// Energy (GCRANSAC.h:913-1021 through energy.h:204-254 and graph.h:388-426):
// per point the squared residual r2, d = clamp(r2 / tau_lo, 0, 1) with
// tau_lo = t*t*9/4, the unary term, then one add_tweights(p, 0, lambda*e00)
// per out-edge with exactly the reference's operation sequence, giving the
// node's terminal residual tr_cap; the arc pair gets capacities lambda - A
// (p->q) and lambda (q->p).
//
// Min-cut: the grid graph is a disjoint union of per-cell cliques joined only
// through the terminals, so the max-flow decomposes by cell.  One wavefront
// runs push-relabel on one cell with heights and excesses in LDS and
// capacities quantised to 2^-S fixed point (S from the graph size; any
// positive capacity stays >= 1 unit).  When no node with excess can reach the
// sink, the final reverse BFS from the sink (exact distance labels) gives the
// set of nodes that reach the sink through residual arcs, which is the SINK
// set Boykov-Kolmogorov's max-flow leaves (graph.h:478-488): it is the same
// for every maximum (pre)flow.  Single-point cells are decided by the sign of
// tr_cap in the energy kernel.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

__device__ __forceinline__ int64_t quantise(double c, int S) {
  if (!(c > 0.0)) return 0;
  const int64_t q = (int64_t)rint(ldexp(c, S));
  return q < 1 ? 1 : q;
}

// ---- neighbourhood -------------------------------------------------------
__global__ void k_grid_keys(const double* __restrict__ pts, const int32_t* __restrict__ pt_pair,
                            const int64_t* __restrict__ pair_pt_off, const double* __restrict__ pair_cell,
                            const int64_t* __restrict__ pair_div, int64_t total, uint64_t* __restrict__ keys,
                            int32_t* __restrict__ vals) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  const double* p = pts + 4 * g;
  const uint64_t div = (uint64_t)pair_div[pair];
  if (div == 0) {
    // no neighbourhood graph for this pair: every point is its own cell
    keys[g] = (uint64_t)(g - pair_pt_off[pair]);
    vals[g] = (int32_t)(g - pair_pt_off[pair]);
    return;
  }
  uint64_t key = 0, off = 1;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const uint64_t idx = (uint64_t)(int64_t)floor(p[d] / pair_cell[4 * pair + d]);
    key += off * idx;
    off *= div;
  }
  keys[g] = key;
  vals[g] = (int32_t)(g - pair_pt_off[pair]);
}

// Run boundaries of the sorted keys; pos_of maps a point to its sorted
// position; out_deg = number of later points in the same cell.
__global__ void k_grid_runs(const uint64_t* __restrict__ skeys, const int32_t* __restrict__ svals,
                            const int32_t* __restrict__ pt_pair, const int64_t* __restrict__ pair_pt_off,
                            const int32_t* __restrict__ pair_n, int64_t total, int32_t* __restrict__ run_start,
                            int32_t* __restrict__ run_len, int32_t* __restrict__ pos_of,
                            int64_t* __restrict__ out_deg) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  const int64_t base = pair_pt_off[pair];
  const int n = pair_n[pair];
  const int s = (int)(g - base);
  const uint64_t k = skeys[g];
  int a = s;
  while (a > 0 && skeys[base + a - 1] == k) --a;
  int b = s + 1;
  while (b < n && skeys[base + b] == k) ++b;
  run_start[g] = a;
  run_len[g] = b - a;
  const int p = svals[g];
  pos_of[base + p] = s;
  out_deg[base + p] = (int64_t)(b - s - 1);
}

// Cells with >= 2 points, appended in any order.
__global__ void k_grid_cells(const int32_t* __restrict__ run_start, const int32_t* __restrict__ run_len,
                             const int32_t* __restrict__ pt_pair, const int64_t* __restrict__ pair_pt_off,
                             int64_t total, int32_t* __restrict__ cell_pair, int32_t* __restrict__ cell_start,
                             int32_t* __restrict__ cell_len, int32_t* __restrict__ num_cells) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  const int s = (int)(g - pair_pt_off[pair]);
  if (run_start[g] == s && run_len[g] >= 2) {
    const int slot = atomicAdd(num_cells, 1);
    cell_pair[slot] = pair;
    cell_start[slot] = s;
    cell_len[slot] = run_len[g];
  }
}

// ---- energies --------------------------------------------------------------
// Per point of the active pairs: d = clamp(r2 / tau_lo, 0, 1) and the unary
// side (r2 <= tau_lo).  lo_pair_of_pt < 0 marks points of inactive pairs.
template <int KIND>
__global__ void k_lo_dist(const double* __restrict__ pts, const int32_t* __restrict__ pt_pair,
                          const int8_t* __restrict__ pair_active, const double* __restrict__ pair_lo_thr,
                          const double* __restrict__ lo_models, int64_t total, double* __restrict__ dist,
                          int8_t* __restrict__ near) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  if (!pair_active[pair]) return;
  double M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = lo_models[9 * pair + k];
  const double r2 = sq_residual<KIND>(pts + 4 * g, M);
  const double thr = pair_lo_thr[pair];
  dist[g] = clamp01(r2 / thr);
  near[g] = r2 <= thr ? 1 : 0;
}

__device__ __forceinline__ void add_tweights(double& tr, double cap_source, double cap_sink) {
  const double delta = tr;
  if (delta > 0) cap_source += delta;
  else cap_sink -= delta;
  tr = cap_source - cap_sink;
}

// Terminal residuals and arc capacities, quantised.  Single-point cells are
// labelled here.
__global__ void k_lo_caps(const int32_t* __restrict__ pt_pair, const int8_t* __restrict__ pair_active,
                          const int64_t* __restrict__ pair_pt_off, const double* __restrict__ pair_lambda,
                          const int32_t* __restrict__ pair_qbits, const double* __restrict__ dist,
                          const int8_t* __restrict__ near, const int32_t* __restrict__ svals,
                          const int32_t* __restrict__ pos_of, const int32_t* __restrict__ run_len,
                          const int64_t* __restrict__ edge_off, const int8_t* __restrict__ pair_radius,
                          const int64_t* __restrict__ r_off, const int32_t* __restrict__ r_nbr,
                          const int64_t* __restrict__ adj_off, int64_t total, int64_t* __restrict__ rf,
                          int64_t* __restrict__ rb, int64_t* __restrict__ node_src, int64_t* __restrict__ node_sink,
                          uint8_t* __restrict__ sink_out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  if (!pair_active[pair]) return;
  const int64_t base = pair_pt_off[pair];
  const double lambda = pair_lambda[pair];
  const int S = pair_qbits[pair];
  const double one_minus_lambda = 1.0 - lambda;
  const double dp = dist[g];
  const double e = 1 - dp;
  double tr = 0.0;
  if (near[g]) add_tweights(tr, 0, one_minus_lambda * e);
  else add_tweights(tr, one_minus_lambda * (1 - e), 0);
  const int s = pos_of[g];
  const bool radius = pair_radius[pair] != 0;
  if (lambda > 0) {
    // out-edges: the later points of the grid cell, in ascending index, or
    // the radius graph's later neighbours in (distance, index) order
    const int64_t eo = edge_off[g];
    const int64_t ne = edge_off[g + 1] - eo;
    const int64_t ro = radius ? r_off[g] : 0;
    for (int64_t j = 0; j < ne; ++j) {
      const int q = radius ? r_nbr[ro + j] : svals[base + s + 1 + j];
      const double e00 = 0.5 * (dp + dist[base + q]);
      const double A = e00 * lambda;
      add_tweights(tr, 0.0, A);
      const double B = lambda - A;
      const double C = lambda - 0.0;
      rf[eo + j] = quantise(B, S);
      rb[eo + j] = quantise(C, S);
    }
  }
  node_src[g] = tr > 0 ? quantise(tr, S) : 0;
  node_sink[g] = tr < 0 ? quantise(-tr, S) : 0;
  // a point with no edge (alone in its cell, no neighbour within the radius,
  // or lambda == 0) is SINK iff it has a sink arc
  const bool isolated =
      (lambda > 0) ? (radius ? adj_off[g + 1] == adj_off[g] : run_len[base + s] < 2) : true;
  if (isolated) sink_out[g] = tr < 0 ? 1 : 0;
}

// ---- per-cell push-relabel ---------------------------------------------------
// Cells of up to kMaxLdsCell points keep their node state in LDS (one wave);
// larger cells keep it in global memory, indexed by point, and run on four
// waves.  A cell that has not converged after 256 + 64 * nc push / relabel
// rounds marks its pair failed.
struct CellView {
  const int64_t* edge_off;
  int64_t* rf;
  int64_t* rb;
  // arc a -> b between the cell's points of batch-global indices ga, gb at
  // cell positions a, b
  __device__ __forceinline__ int64_t* arc(int a, int b, int64_t ga, int64_t gb) const {
    return a < b ? rf + edge_off[ga] + (b - a - 1) : rb + edge_off[gb] + (a - b - 1);
  }
};

template <int NT>
__device__ __forceinline__ bool cell_any(bool v) {
  if constexpr (NT == 64) return __any(v);
  else return __syncthreads_or(v ? 1 : 0) != 0;
}

template <int NT, bool GLOBAL>
__global__ void __launch_bounds__(NT) k_mincut_cells(const int32_t* __restrict__ cell_pair,
                                                     const int32_t* __restrict__ cell_start,
                                                     const int32_t* __restrict__ cell_len,
                                                     const int8_t* __restrict__ pair_active,
                                                     const double* __restrict__ pair_lambda,
                                                     const int64_t* __restrict__ pair_pt_off,
                                                     const int32_t* __restrict__ svals,
                                                     const int64_t* __restrict__ edge_off, int64_t* rf, int64_t* rb,
                                                     int64_t* node_src, int64_t* node_sink, int32_t* gh,
                                                     int32_t* ghn, uint8_t* __restrict__ sink_out,
                                                     int32_t* __restrict__ pair_fail, int lds_cell) {
  const int c = blockIdx.x;
  const int pair = cell_pair[c];
  // lambda == 0: the graph has no edges and k_lo_caps labelled every point
  if (!pair_active[pair] || !(pair_lambda[pair] > 0)) return;
  const int nc = cell_len[c];
  // each cell runs in exactly one of the two launches
  if (GLOBAL ? nc <= lds_cell : nc > lds_cell) return;
  const int lane = threadIdx.x;
  const int64_t base = pair_pt_off[pair];
  const int start = cell_start[c];
  extern __shared__ int64_t smem64[];
  // node state: LDS arrays indexed by cell position, or the batch's arrays
  // indexed by point (node_src / node_sink are this labelling's scratch)
  int64_t *ex, *rt;
  int32_t *h, *hn, *pid;
  if constexpr (GLOBAL) {
    ex = node_src;
    rt = node_sink;
    h = gh;
    hn = ghn;
    pid = nullptr;
  } else {
    ex = smem64;
    rt = ex + lds_cell;
    h = (int32_t*)(rt + lds_cell);
    hn = h + lds_cell;
    pid = hn + lds_cell;
    for (int a = lane; a < nc; a += NT) {
      const int64_t gi = base + svals[base + start + a];
      pid[a] = (int32_t)gi;  // batch-global point index (fits: total points < 2^31)
      ex[a] = node_src[gi];
      rt[a] = node_sink[gi];
    }
    __syncthreads();
  }
  // state slot and batch-global index of the point at cell position a
  auto gid = [&](int a) -> int64_t {
    if constexpr (GLOBAL) return base + svals[base + start + a];
    else return pid[a];
  };
  auto at = [&](int a) -> int64_t {
    if constexpr (GLOBAL) return base + svals[base + start + a];
    else return a;
  };
  const CellView cv{edge_off, rf, rb};
  const int INF = nc + 1;
  const int64_t max_rounds = 256 + 64 * (int64_t)nc;
  int64_t rounds = 0;
  bool failed = false;
  while (true) {
    // Exact distance-to-sink labels by reverse BFS over residual arcs.
    for (int a = lane; a < nc; a += NT) h[at(a)] = rt[at(a)] > 0 ? 1 : INF;
    __syncthreads();
    for (int level = 1; level < INF; ++level) {
      bool found = false;
      for (int a = lane; a < nc; a += NT) {
        int v = h[at(a)];
        if (v == INF) {
          const int64_t ga = gid(a);
          for (int b = 0; b < nc; ++b)
            if (h[at(b)] == level && *cv.arc(a, b, ga, gid(b)) > 0) {
              v = level + 1;
              found = true;
              break;
            }
        }
        hn[at(a)] = v;
      }
      __syncthreads();
      for (int a = lane; a < nc; a += NT) h[at(a)] = hn[at(a)];
      __syncthreads();
      if (!cell_any<NT>(found)) break;
    }
    bool act = false;
    for (int a = lane; a < nc; a += NT) act |= (ex[at(a)] > 0 && h[at(a)] < INF);
    if (!cell_any<NT>(act)) break;
    if (rounds > max_rounds) {
      failed = true;
      break;
    }
    for (int it = 0; it < 4; ++it, ++rounds) {
      // push phase: heights fixed, an arc pair is touched only by its
      // admissible end
      for (int a = lane; a < nc; a += NT) {
        const int64_t sa = at(a);
        const int ha = h[sa];
        if (ha >= INF) continue;
        int64_t e = ex[sa];
        if (e <= 0) continue;
        const int64_t ga = gid(a);
        int64_t pushed = 0;
        if (ha == 1 && rt[sa] > 0) {
          const int64_t d = e < rt[sa] ? e : rt[sa];
          rt[sa] -= d;
          e -= d;
          pushed += d;
        }
        for (int b = 0; b < nc && e > 0; ++b) {
          if (b == a || h[at(b)] != ha - 1) continue;
          const int64_t gb = gid(b);
          int64_t* r = cv.arc(a, b, ga, gb);
          const int64_t rc = *r;
          if (rc <= 0) continue;
          const int64_t d = e < rc ? e : rc;
          *r = rc - d;
          int64_t* rr = cv.arc(b, a, gb, ga);
          *rr = *rr + d;
          e -= d;
          pushed += d;
          atomicAdd((unsigned long long*)&ex[at(b)], (unsigned long long)d);
        }
        if (pushed) atomicAdd((unsigned long long*)&ex[sa], (unsigned long long)(-pushed));
      }
      __syncthreads();
      // relabel phase (all new heights from the old ones)
      for (int a = lane; a < nc; a += NT) {
        const int64_t sa = at(a);
        int v = h[sa];
        if (v < INF && ex[sa] > 0) {
          const int64_t ga = gid(a);
          int mn = INF;
          if (rt[sa] > 0) mn = 0;
          for (int b = 0; b < nc; ++b) {
            if (b == a) continue;
            if (*cv.arc(a, b, ga, gid(b)) > 0 && h[at(b)] < mn) mn = h[at(b)];
          }
          if (mn != v - 1) v = (mn + 1 < INF) ? mn + 1 : INF;
        }
        hn[sa] = v;
      }
      __syncthreads();
      for (int a = lane; a < nc; a += NT) h[at(a)] = hn[at(a)];
      __syncthreads();
    }
  }
  if (failed) {
    if (lane == 0) pair_fail[pair] = 1;
    return;
  }
  for (int a = lane; a < nc; a += NT) sink_out[gid(a)] = h[at(a)] < INF ? 1 : 0;
}

// ---- ordered compaction of a mask into an index list, one wave per list ----
// Each lane takes 16 consecutive mask bytes per pass (1,024 per wave), so
// the list order is the lane order; a wave prefix sum of the lanes' counts
// places each lane's indices.
constexpr int kCompactPerLane = 16;
__global__ void __launch_bounds__(64) k_compact(const uint8_t* __restrict__ mask, const int64_t* __restrict__ list_off,
                                               const int32_t* __restrict__ list_n, int32_t* __restrict__ out,
                                               int32_t* __restrict__ out_count) {
  const int l = blockIdx.x;
  const int lane = threadIdx.x;
  const uint8_t* m = mask + list_off[l];
  int32_t* o = out + list_off[l];
  const int n = list_n[l];
  int cnt = 0;
  for (int base = 0; base < n; base += 64 * kCompactPerLane) {
    const int first = base + lane * kCompactPerLane;
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < kCompactPerLane; ++j) {
      const int i = first + j;
      if (i < n && m[i]) bits |= 1u << j;
    }
    const int c = __popc(bits);
    // inclusive prefix over the wave
    int incl = c;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      const int v = __shfl_up(incl, s, 64);
      if (lane >= s) incl += v;
    }
    int w = cnt + incl - c;
    while (bits) {
      const int j = __ffs(bits) - 1;
      o[w++] = first + j;
      bits &= bits - 1u;
    }
    cnt += __shfl(incl, 63, 64);
  }
  if (lane == 0) out_count[l] = cnt;
}

// ---- launchers ---------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_grid_keys(hipStream_t st, const double* pts, const int32_t* pt_pair, const int64_t* pair_pt_off,
                      const double* pair_cell, const int64_t* pair_div, int64_t total, uint64_t* keys,
                      int32_t* vals) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_grid_keys, dim3(blocks_for(total, 256)), dim3(256), 0, st, pts, pt_pair, pair_pt_off, pair_cell,
                     pair_div, total, keys, vals);
}

void launch_grid_runs(hipStream_t st, const uint64_t* skeys, const int32_t* svals, const int32_t* pt_pair,
                      const int64_t* pair_pt_off, const int32_t* pair_n, int64_t total, int32_t* run_start,
                      int32_t* run_len, int32_t* pos_of, int64_t* out_deg) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_grid_runs, dim3(blocks_for(total, 256)), dim3(256), 0, st, skeys, svals, pt_pair, pair_pt_off,
                     pair_n, total, run_start, run_len, pos_of, out_deg);
}

void launch_grid_cells(hipStream_t st, const int32_t* run_start, const int32_t* run_len, const int32_t* pt_pair,
                       const int64_t* pair_pt_off, int64_t total, int32_t* cell_pair, int32_t* cell_start,
                       int32_t* cell_len, int32_t* num_cells) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_grid_cells, dim3(blocks_for(total, 256)), dim3(256), 0, st, run_start, run_len, pt_pair,
                     pair_pt_off, total, cell_pair, cell_start, cell_len, num_cells);
}

void launch_lo_dist(hipStream_t st, int kind, const double* pts, const int32_t* pt_pair, const int8_t* pair_active,
                    const double* pair_lo_thr, const double* lo_models, int64_t total, double* dist, int8_t* near) {
  if (total <= 0) return;
  if (kind == 0)
    hipLaunchKernelGGL(k_lo_dist<0>, dim3(blocks_for(total, 256)), dim3(256), 0, st, pts, pt_pair, pair_active,
                       pair_lo_thr, lo_models, total, dist, near);
  else
    hipLaunchKernelGGL(k_lo_dist<1>, dim3(blocks_for(total, 256)), dim3(256), 0, st, pts, pt_pair, pair_active,
                       pair_lo_thr, lo_models, total, dist, near);
}

void launch_lo_caps(hipStream_t st, const int32_t* pt_pair, const int8_t* pair_active, const int64_t* pair_pt_off,
                    const double* pair_lambda, const int32_t* pair_qbits, const double* dist, const int8_t* near,
                    const int32_t* svals, const int32_t* pos_of, const int32_t* run_len, const int64_t* edge_off,
                    const int8_t* pair_radius, const int64_t* r_off, const int32_t* r_nbr, const int64_t* adj_off,
                    int64_t total, int64_t* rf, int64_t* rb, int64_t* node_src, int64_t* node_sink,
                    uint8_t* sink_out) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_lo_caps, dim3(blocks_for(total, 256)), dim3(256), 0, st, pt_pair, pair_active, pair_pt_off,
                     pair_lambda, pair_qbits, dist, near, svals, pos_of, run_len, edge_off, pair_radius, r_off, r_nbr,
                     adj_off, total, rf, rb, node_src, node_sink, sink_out);
}

void launch_mincut_cells(hipStream_t st, const int32_t* cell_pair, const int32_t* cell_start, const int32_t* cell_len,
                         int num_cells, const int8_t* pair_active, const double* pair_lambda,
                         const int64_t* pair_pt_off, const int32_t* svals, const int64_t* edge_off, int64_t* rf,
                         int64_t* rb, int64_t* node_src, int64_t* node_sink, int32_t* gh, int32_t* ghn,
                         uint8_t* sink_out, int32_t* pair_fail, int max_cell) {
  if (num_cells <= 0) return;
  const int lds_cell = max_cell < kMaxLdsCell ? max_cell : kMaxLdsCell;
  const size_t shmem = (size_t)lds_cell * (8 + 8 + 4 + 4 + 4);
  hipLaunchKernelGGL((k_mincut_cells<64, false>), dim3(num_cells), dim3(64), shmem, st, cell_pair, cell_start,
                     cell_len, pair_active, pair_lambda, pair_pt_off, svals, edge_off, rf, rb, node_src, node_sink,
                     gh, ghn, sink_out, pair_fail, lds_cell);
  if (max_cell > kMaxLdsCell)
    hipLaunchKernelGGL((k_mincut_cells<256, true>), dim3(num_cells), dim3(256), 0, st, cell_pair, cell_start,
                       cell_len, pair_active, pair_lambda, pair_pt_off, svals, edge_off, rf, rb, node_src, node_sink,
                       gh, ghn, sink_out, pair_fail, lds_cell);
}

void launch_compact(hipStream_t st, const uint8_t* mask, const int64_t* list_off, const int32_t* list_n, int num_lists,
                    int32_t* out, int32_t* out_count) {
  if (num_lists <= 0) return;
  hipLaunchKernelGGL(k_compact, dim3(num_lists), dim3(64), 0, st, mask, list_off, list_n, out, out_count);
}

}  // namespace gcr
