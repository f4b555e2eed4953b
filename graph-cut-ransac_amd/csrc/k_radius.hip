// Radius neighbourhood and the min-cut on its graph.
//
// Neighbourhood (flann_neighborhood_graph.h:100-139, searched exhaustively):
// the reference converts the (x1, y1, x2, y2) rows to float and keeps, for
// every point, the points whose float squared distance
// ((dx0^2 + dx1^2) + dx2^2) + dx3^2 is at most fl(r)^2, in (distance, index)
// order, minus the first entry.  The relation is symmetric and a list never
// drops a later point, so the labelling's edges are exactly the pairs p < q
// within the radius, added at p in p's list order.  Per point p this file
// builds
//   r_off / r_nbr  the out-neighbours q > p (pair-relative), sorted by
//                  (distance, q) with a stable segmented sort on the distance
//                  bits after the fill kernel writes them in ascending q;
//   adj_off / adj  every incident arc: (edge << 1 | dir, other end), dir 0
//                  for an out-edge (p is the edge's first end), 1 otherwise.
// The capacity arrays share the engine's edge numbering: the out-edge j of p
// is edge edge_off[p] + j, with residuals rf (p -> q) and rb (q -> p).
//
// Min-cut: one workgroup per pair runs push-relabel over the pair's graph
// nodes (points with at least one edge; the others are decided by the sign
// of their terminal residual when the capacities are built) with node state
// in global memory: exact distance-to-sink labels from a frontier BFS over
// residual arcs, then four synchronous push / relabel rounds, repeated until
// no node with excess reaches the sink.  The nodes that still reach the sink
// through residual arcs are the SINK set (graph.h:478-488); that set is the
// same for every maximum preflow, so it does not depend on the order in which
// pushes happen.  Capacities are the engine's 2^-S fixed point, so the
// arithmetic is exact.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace gcr {

// FLANN's L2<float>: float differences, squared and summed left to right,
// with no fused multiply-add.
__device__ __forceinline__ float radius_sq_distance(const float4 a, const float4 b) {
#pragma clang fp contract(off)
  const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
  float s = d0 * d0 + d1 * d1;
  s = s + d2 * d2;
  s = s + d3 * d3;
  return 0.0f + s;
}

__device__ __forceinline__ float4 to_float4(const double* p) {
  return make_float4((float)p[0], (float)p[1], (float)p[2], (float)p[3]);
}

constexpr int kSearchThreads = 256;

// One workgroup per (pair, 256 query points); the pair's points stream
// through LDS 256 at a time.  COUNT: per point the numbers of neighbours of
// higher and lower index.  FILL: the out-neighbours in ascending index with
// their distance bits (the sort key).
template <bool FILL>
__global__ void __launch_bounds__(kSearchThreads)
    k_radius_search(const double* __restrict__ pts, const int32_t* __restrict__ tile_pair,
                    const int32_t* __restrict__ tile_first, const int64_t* __restrict__ pair_pt_off,
                    const int32_t* __restrict__ pair_n, const float* __restrict__ pair_r2,
                    int64_t* __restrict__ out_cnt, int64_t* __restrict__ in_cnt, const int64_t* __restrict__ r_off,
                    int32_t* __restrict__ r_nbr, uint32_t* __restrict__ r_key) {
  __shared__ float4 tile[kSearchThreads];
  const int pair = tile_pair[blockIdx.x];
  const int first = tile_first[blockIdx.x];
  const int64_t base = pair_pt_off[pair];
  const int n = pair_n[pair];
  const float r2 = pair_r2[pair];
  const int p = first + (int)threadIdx.x;
  const bool live = p < n;
  const float4 me = live ? to_float4(pts + 4 * (base + p)) : make_float4(0.f, 0.f, 0.f, 0.f);
  int n_out = 0, n_in = 0;
  int64_t w = 0;
  if (FILL && live) w = r_off[base + p];
  for (int t0 = 0; t0 < n; t0 += kSearchThreads) {
    const int q = t0 + (int)threadIdx.x;
    if (q < n) tile[threadIdx.x] = to_float4(pts + 4 * (base + q));
    __syncthreads();
    const int m = n - t0 < kSearchThreads ? n - t0 : kSearchThreads;
    if (live) {
      for (int k = 0; k < m; ++k) {
        const int qq = t0 + k;
        if (qq == p) continue;
        const float d2 = radius_sq_distance(me, tile[k]);
        if (!(d2 <= r2)) continue;
        if (qq > p) {
          if (FILL) {
            r_nbr[w] = qq;
            r_key[w] = __float_as_uint(d2);
            ++w;
          } else {
            ++n_out;
          }
        } else if (!FILL) {
          ++n_in;
        }
      }
    }
    __syncthreads();
  }
  if (!FILL && live) {
    out_cnt[base + p] = n_out;
    in_cnt[base + p] = n_in;
  }
}

// Incident arcs: each radius point writes its out-arcs at the front of its
// own adjacency range (in list order) and appends the matching in-arc to
// the other end's range after that end's out-arcs.
__global__ void k_radius_adjacency(const int32_t* __restrict__ pt_pair, const int8_t* __restrict__ pair_radius,
                                   const int64_t* __restrict__ pair_pt_off, const int64_t* __restrict__ r_off,
                                   const int32_t* __restrict__ r_nbr, const int64_t* __restrict__ edge_off,
                                   const int64_t* __restrict__ adj_off, const int64_t* __restrict__ out_cnt,
                                   int32_t* __restrict__ in_cursor, int64_t total, int64_t* __restrict__ adj_arc,
                                   int32_t* __restrict__ adj_other) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int pair = pt_pair[g];
  if (!pair_radius[pair]) return;
  const int64_t base = pair_pt_off[pair];
  const int64_t ro = r_off[g];
  const int64_t ne = r_off[g + 1] - ro;
  const int64_t eo = edge_off[g];
  const int64_t ao = adj_off[g];
  for (int64_t j = 0; j < ne; ++j) {
    const int64_t qg = base + r_nbr[ro + j];
    const int64_t e = eo + j;
    adj_arc[ao + j] = e << 1;
    adj_other[ao + j] = (int32_t)qg;
    const int slot = atomicAdd(in_cursor + qg, 1);
    const int64_t at = adj_off[qg] + out_cnt[qg] + slot;
    adj_arc[at] = (e << 1) | 1;
    adj_other[at] = (int32_t)g;
  }
}

// ---- push-relabel on one pair's radius graph ------------------------------
constexpr int kCsrThreads = 256;

__device__ __forceinline__ bool block_any(bool v) { return __syncthreads_or(v ? 1 : 0) != 0; }

__global__ void __launch_bounds__(kCsrThreads)
    k_mincut_radius(const int32_t* __restrict__ rpairs, const int8_t* __restrict__ pair_active,
                    const double* __restrict__ pair_lambda, const int64_t* __restrict__ pair_pt_off,
                    const int32_t* __restrict__ pair_n, const int64_t* __restrict__ adj_off,
                    const int64_t* __restrict__ adj_arc, const int32_t* __restrict__ adj_other, int64_t* rf,
                    int64_t* rb, int64_t* ex, int64_t* rt, int32_t* h, int32_t* hn, int32_t* queue,
                    uint8_t* __restrict__ sink_out, int32_t* __restrict__ pair_fail) {
  const int pair = rpairs[blockIdx.x];
  if (!pair_active[pair] || !(pair_lambda[pair] > 0)) return;
  const int64_t base = pair_pt_off[pair];
  const int n = pair_n[pair];
  const int INF = n + 1;
  // give-up bound: outer iterations (one exact relabel, four push / relabel
  // rounds); the pair is reported failed beyond it
  const int64_t max_outer = 64 + 16 * (int64_t)n;
  const int tid = threadIdx.x;
  int32_t* qa = queue + 2 * base;  // two frontiers of n entries
  int32_t* qb = qa + n;
  __shared__ int s_len[2];
  // residual of the arc leaving the adjacency entry's owner, and of its sister
  auto res_out = [&](int64_t arc) -> int64_t* { return (arc & 1) ? rb + (arc >> 1) : rf + (arc >> 1); };
  auto res_in = [&](int64_t arc) -> int64_t* { return (arc & 1) ? rf + (arc >> 1) : rb + (arc >> 1); };
  int64_t outer = 0;
  bool failed = false;
  while (true) {
    // exact distance-to-sink labels: frontier BFS over residual arcs u -> v
    if (tid == 0) s_len[0] = 0;
    __syncthreads();
    for (int a = tid; a < n; a += kCsrThreads) {
      const int64_t g = base + a;
      if (adj_off[g + 1] == adj_off[g]) continue;
      const bool t = rt[g] > 0;
      h[g] = t ? 1 : INF;
      if (t) qa[atomicAdd(&s_len[0], 1)] = a;
    }
    __syncthreads();
    int cur = 0, level = 1;
    while (true) {
      const int len = s_len[cur];
      if (len == 0) break;
      int32_t* fr = cur ? qb : qa;
      int32_t* nx = cur ? qa : qb;
      __syncthreads();
      if (tid == 0) s_len[cur ^ 1] = 0;
      __syncthreads();
      for (int i = tid; i < len; i += kCsrThreads) {
        const int64_t v = base + fr[i];
        for (int64_t k = adj_off[v]; k < adj_off[v + 1]; ++k) {
          const int64_t u = adj_other[k];
          if (*res_in(adj_arc[k]) <= 0) continue;  // residual u -> v
          if (atomicCAS(h + u, INF, level + 1) == INF) nx[atomicAdd(&s_len[cur ^ 1], 1)] = (int32_t)(u - base);
        }
      }
      __syncthreads();
      cur ^= 1;
      ++level;
    }
    bool act = false;
    for (int a = tid; a < n; a += kCsrThreads) {
      const int64_t g = base + a;
      if (adj_off[g + 1] == adj_off[g]) continue;
      act |= ex[g] > 0 && h[g] < INF;
    }
    if (!block_any(act)) break;
    if (++outer > max_outer) {
      failed = true;
      break;
    }
    for (int it = 0; it < 4; ++it) {
      // push: heights fixed, an arc pair is touched only by its higher end
      for (int a = tid; a < n; a += kCsrThreads) {
        const int64_t g = base + a;
        const int ha = h[g];
        if (ha >= INF || adj_off[g + 1] == adj_off[g]) continue;
        int64_t e = ex[g];
        if (e <= 0) continue;
        int64_t pushed = 0;
        if (ha == 1 && rt[g] > 0) {
          const int64_t d = e < rt[g] ? e : rt[g];
          rt[g] -= d;
          e -= d;
          pushed += d;
        }
        for (int64_t k = adj_off[g]; k < adj_off[g + 1] && e > 0; ++k) {
          const int64_t b = adj_other[k];
          if (h[b] != ha - 1) continue;
          int64_t* r = res_out(adj_arc[k]);
          const int64_t rc = *r;
          if (rc <= 0) continue;
          const int64_t d = e < rc ? e : rc;
          *r = rc - d;
          int64_t* rr = res_in(adj_arc[k]);
          *rr = *rr + d;
          e -= d;
          pushed += d;
          atomicAdd((unsigned long long*)(ex + b), (unsigned long long)d);
        }
        if (pushed) atomicAdd((unsigned long long*)(ex + g), (unsigned long long)(-pushed));
      }
      __syncthreads();
      // relabel from the old heights
      for (int a = tid; a < n; a += kCsrThreads) {
        const int64_t g = base + a;
        if (adj_off[g + 1] == adj_off[g]) continue;
        int v = h[g];
        if (v < INF && ex[g] > 0) {
          int mn = INF;
          if (rt[g] > 0) mn = 0;
          for (int64_t k = adj_off[g]; k < adj_off[g + 1]; ++k) {
            if (*res_out(adj_arc[k]) <= 0) continue;
            const int hb = h[adj_other[k]];
            mn = hb < mn ? hb : mn;
          }
          if (mn != v - 1) v = (mn + 1 < INF) ? mn + 1 : INF;
        }
        hn[g] = v;
      }
      __syncthreads();
      for (int a = tid; a < n; a += kCsrThreads) {
        const int64_t g = base + a;
        if (adj_off[g + 1] != adj_off[g]) h[g] = hn[g];
      }
      __syncthreads();
    }
  }
  if (failed) {
    if (tid == 0) pair_fail[pair] = 1;
    return;
  }
  for (int a = tid; a < n; a += kCsrThreads) {
    const int64_t g = base + a;
    if (adj_off[g + 1] != adj_off[g]) sink_out[g] = h[g] < INF ? 1 : 0;
  }
}

__global__ void k_add_i64(int64_t* __restrict__ a, const int64_t* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

// ---- launchers ---------------------------------------------------------------
void launch_add_i64(hipStream_t st, int64_t* a, const int64_t* b, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n);
}

static inline unsigned blocks_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_radius_search(hipStream_t st, bool fill, const double* pts, const int32_t* tile_pair,
                          const int32_t* tile_first, int num_tiles, const int64_t* pair_pt_off, const int32_t* pair_n,
                          const float* pair_r2, int64_t* out_cnt, int64_t* in_cnt, const int64_t* r_off,
                          int32_t* r_nbr, uint32_t* r_key) {
  if (num_tiles <= 0) return;
  if (fill)
    hipLaunchKernelGGL(k_radius_search<true>, dim3(num_tiles), dim3(kSearchThreads), 0, st, pts, tile_pair,
                       tile_first, pair_pt_off, pair_n, pair_r2, out_cnt, in_cnt, r_off, r_nbr, r_key);
  else
    hipLaunchKernelGGL(k_radius_search<false>, dim3(num_tiles), dim3(kSearchThreads), 0, st, pts, tile_pair,
                       tile_first, pair_pt_off, pair_n, pair_r2, out_cnt, in_cnt, r_off, r_nbr, r_key);
}

void launch_radius_adjacency(hipStream_t st, const int32_t* pt_pair, const int8_t* pair_radius,
                             const int64_t* pair_pt_off, const int64_t* r_off, const int32_t* r_nbr,
                             const int64_t* edge_off, const int64_t* adj_off, const int64_t* out_cnt,
                             int32_t* in_cursor, int64_t total, int64_t* adj_arc, int32_t* adj_other) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_radius_adjacency, dim3(blocks_of(total, 256)), dim3(256), 0, st, pt_pair, pair_radius,
                     pair_pt_off, r_off, r_nbr, edge_off, adj_off, out_cnt, in_cursor, total, adj_arc, adj_other);
}

void launch_mincut_radius(hipStream_t st, const int32_t* rpairs, int num_rpairs, const int8_t* pair_active,
                          const double* pair_lambda, const int64_t* pair_pt_off, const int32_t* pair_n,
                          const int64_t* adj_off, const int64_t* adj_arc, const int32_t* adj_other, int64_t* rf,
                          int64_t* rb, int64_t* ex, int64_t* rt, int32_t* h, int32_t* hn, int32_t* queue,
                          uint8_t* sink_out, int32_t* pair_fail) {
  if (num_rpairs <= 0) return;
  hipLaunchKernelGGL(k_mincut_radius, dim3(num_rpairs), dim3(kCsrThreads), 0, st, rpairs, pair_active, pair_lambda,
                     pair_pt_off, pair_n, adj_off, adj_arc, adj_other, rf, rb, ex, rt, h, hn, queue, sink_out,
                     pair_fail);
}

}  // namespace gcr
