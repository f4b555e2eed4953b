// Host side of the learned inlier-probability network and the guided batch:
// C ABI (include/gcr.h, "learned guidance"), device workspace, training loop.
//
// The guided batch computes every pair's features and logits on the device,
// orders each pair's correspondences by descending logit (a stable
// segmented radix sort, ties in index order), gathers the rows in that
// order and runs the batched engine with the PROSAC sampler on them -- the
// input contract of the reference's PROSAC mode, which expects points
// ordered by quality (inc/samplers/prosac_sampler.h:51-237; the drivers
// pass the caller's order through, py/gcransac_python.cpp:4344-4359).
// Masks are mapped back to the caller's order.  Every model is still judged
// by full-set MSAC scoring inside the engine.
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "kernels.hpp"
#include "mlp_kernels.hpp"

namespace {

using namespace gcr::mlp;

int32_t fail(int32_t code, const std::string& msg) {
  gcr::set_last_error(msg);
  return code;
}

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw gcr::EngineError(e == hipErrorOutOfMemory ? GCR_ERR_OUT_OF_MEMORY : GCR_ERR_DEVICE,
                           std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct Buf {
  T* p = nullptr;
  size_t cap = 0;
  void ensure(size_t n) {
    if (n <= cap) return;
    if (p) hip_ok(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    hip_ok(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
    cap = n;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

// One workspace per host thread (device buffers grow, never shrink).
struct Workspace {
  int device = -1;
  hipStream_t st = nullptr;
  Buf<double> corr, ratio, sorted;
  Buf<int64_t> off;
  Buf<int32_t> sizes, tile_pair, tile_start, vals, perm;
  Buf<float> feat, labels, params, m, v, logit, skeys, zbuf, d4, lossv, part, grad, loss;
  Buf<uint16_t> wpk, act;
  Buf<float> bias;
  Buf<unsigned char> tmp;
  Buf<KnnGrid> grids;
  Buf<int32_t> cell_count, cell_off, cell_of, rank, sidx;
  Buf<float4> spts;
  std::vector<int32_t> tile_pair_h, tile_start_h;  // host sources of in-flight uploads
  std::vector<KnnGrid> grids_h;
  // the sort's permutation read back while the batch runs (pinned, so the
  // copy runs beside the batch), and the event that marks the sorted rows
  int32_t* perm_h = nullptr;
  size_t perm_h_cap = 0;
  hipEvent_t sorted_ev = nullptr;
  int32_t* perm_host(size_t n) {
    if (n > perm_h_cap) {
      if (perm_h) hip_ok(hipHostFree(perm_h), "hipHostFree");
      perm_h = nullptr;
      perm_h_cap = 0;
      hip_ok(hipHostMalloc((void**)&perm_h, std::max<size_t>(n, 1) * sizeof(int32_t), hipHostMallocDefault),
             "hipHostMalloc");
      perm_h_cap = n;
    }
    return perm_h;
  }
  void init() {
    int dev = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    if (st && dev == device) return;
    if (sorted_ev) (void)hipEventDestroy(sorted_ev);
    if (st) (void)hipStreamDestroy(st);
    hip_ok(hipStreamCreate(&st), "hipStreamCreate");
    hip_ok(hipEventCreateWithFlags(&sorted_ev, hipEventDisableTiming), "hipEventCreate");
    device = dev;
  }
  ~Workspace() {
    if (perm_h) (void)hipHostFree(perm_h);
    if (sorted_ev) (void)hipEventDestroy(sorted_ev);
    if (st) (void)hipStreamDestroy(st);
  }
};

Workspace& ws() {
  static thread_local Workspace w;
  w.init();
  return w;
}

template <class F>
int32_t guarded(F&& f) {
  try {
    gcr::set_last_error("");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(GCR_ERR_DEVICE, "no HIP device available");
    return f();
  } catch (const gcr::EngineError& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(GCR_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(GCR_ERR_INTERNAL, e.what());
  }
}

template <class T>
void h2d(Workspace& w, T* dst, const T* src, size_t n) {
  if (n) hip_ok(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, w.st), "copy to device");
}
template <class T>
void d2h(Workspace& w, T* dst, const T* src, size_t n) {
  if (n) hip_ok(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, w.st), "copy to host");
}
void sync(Workspace& w) { hip_ok(hipStreamSynchronize(w.st), "stream synchronize"); }

void check_lds(int bytes) {
  if (gcr::lds_limit_bytes() < bytes)
    throw gcr::EngineError(GCR_ERR_DEVICE, "device LDS per workgroup below " + std::to_string(bytes) + " bytes");
}

// Features of num_pairs pairs (rel: zero-based offsets) from device rows.
void features(Workspace& w, int num_pairs, const std::vector<int64_t>& rel, const double* d_corr,
              const double* d_ratio, const int32_t* sizes, float* d_feat) {
  std::vector<int32_t>& tp = w.tile_pair_h;
  std::vector<int32_t>& ts = w.tile_start_h;
  tp.clear();
  ts.clear();
  for (int p = 0; p < num_pairs; ++p)
    for (int64_t s = 0; s < rel[(size_t)p + 1] - rel[(size_t)p]; s += 256) {
      tp.push_back(p);
      ts.push_back((int32_t)s);
    }
  w.off.ensure(rel.size());
  w.sizes.ensure(4 * (size_t)num_pairs);
  w.tile_pair.ensure(tp.size());
  w.tile_start.ensure(ts.size());
  h2d(w, w.off.p, rel.data(), rel.size());
  h2d(w, w.sizes.p, sizes, 4 * (size_t)num_pairs);
  h2d(w, w.tile_pair.p, tp.data(), tp.size());
  h2d(w, w.tile_start.p, ts.data(), ts.size());
  // per-pair grids of about four points per cell over the first image
  std::vector<KnnGrid>& g = w.grids_h;
  g.resize((size_t)num_pairs);
  int64_t cells = 0;
  for (int p = 0; p < num_pairs; ++p) {
    const int64_t n = rel[(size_t)p + 1] - rel[(size_t)p];
    const double h1 = sizes[4 * p + 0], w1 = sizes[4 * p + 1];
    KnnGrid& k = g[(size_t)p];
    double cs = 1.0;
    k.gx = k.gy = 1;
    if (w1 > 0 && h1 > 0) {
      cs = std::max(std::sqrt(4.0 * w1 * h1 / (double)n), std::max(w1, h1) / 512.0);
      k.gx = (int32_t)std::min(512.0, std::max(1.0, std::ceil(w1 / cs)));
      k.gy = (int32_t)std::min(512.0, std::max(1.0, std::ceil(h1 / cs)));
    }
    k.inv_cs = (float)(1.0 / cs);
    k.pad_ = 0;
    k.cell_base = cells;
    cells += (int64_t)k.gx * k.gy;
  }
  const size_t total = (size_t)rel.back();
  w.grids.ensure(g.size());
  w.cell_count.ensure((size_t)cells + 1);
  w.cell_off.ensure((size_t)cells + 1);
  w.cell_of.ensure(total);
  w.rank.ensure(total);
  w.sidx.ensure(total);
  w.spts.ensure(total);
  h2d(w, w.grids.p, g.data(), g.size());
  hip_ok(hipMemsetAsync(w.cell_count.p, 0, sizeof(int32_t) * ((size_t)cells + 1), w.st), "zero cell counts");
  const int tiles = (int)tp.size();
  launch_knn_bin(w.st, d_corr, w.off.p, w.grids.p, w.tile_pair.p, w.tile_start.p, tiles, w.cell_count.p,
                 w.cell_of.p, w.rank.p);
  size_t tmp_bytes = 0;
  hip_ok(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, w.cell_count.p, w.cell_off.p, (int)(cells + 1), w.st),
         "cell scan (size)");
  w.tmp.ensure(tmp_bytes);
  hip_ok(hipcub::DeviceScan::ExclusiveSum(w.tmp.p, tmp_bytes, w.cell_count.p, w.cell_off.p, (int)(cells + 1), w.st),
         "cell scan");
  launch_knn_scatter(w.st, d_corr, w.off.p, w.grids.p, w.tile_pair.p, w.tile_start.p, tiles, w.cell_off.p,
                     w.cell_of.p, w.rank.p, w.spts.p, w.sidx.p);
  launch_knn_query(w.st, d_corr, w.off.p, w.sizes.p, d_ratio, w.grids.p, w.tile_pair.p, w.tile_start.p, tiles,
                   w.cell_off.p, w.spts.p, w.sidx.p, d_feat);
  hip_ok(hipGetLastError(), "feature kernel launch");
}

void pack(Workspace& w, const float* params) {
  w.params.ensure(NPARAM);
  w.wpk.ensure(PK_TOTAL);
  w.bias.ensure(PB_TOTAL);
  h2d(w, w.params.p, params, NPARAM);
  launch_pack(w.st, w.params.p, w.wpk.p, w.bias.p);
}

int64_t padded(int64_t batch) { return (batch + WGRAD_CHUNK - 1) / WGRAD_CHUNK * WGRAD_CHUNK; }

void train_buffers(Workspace& w, int64_t bpad) {
  w.zbuf.ensure((size_t)(bpad / 32) * 3 * 64 * 64);
  w.act.ensure((size_t)ACT_ROWS * (size_t)bpad);
  w.d4.ensure((size_t)bpad);
  w.lossv.ensure((size_t)bpad);
  w.part.ensure((size_t)(bpad / WGRAD_CHUNK) * PART_STRIDE);
}

int check_pairs(int32_t num_pairs, const int64_t* offsets, std::vector<int64_t>& rel) {
  if (num_pairs <= 0 || !offsets) return fail(GCR_ERR_INVALID_ARGUMENT, "num_pairs must be positive");
  rel.resize((size_t)num_pairs + 1);
  for (int i = 0; i <= num_pairs; ++i) {
    rel[(size_t)i] = offsets[i] - offsets[0];
    if (i > 0 && rel[(size_t)i] <= rel[(size_t)i - 1])
      return fail(GCR_ERR_INVALID_ARGUMENT, "every pair needs at least one row (offsets must increase)");
  }
  if (rel.back() > (int64_t)std::numeric_limits<int32_t>::max())
    return fail(GCR_ERR_INVALID_ARGUMENT, "too many rows in one batch");
  return GCR_OK;
}

}  // namespace

extern "C" {

int32_t gcr_mlp_num_params(void) { return NPARAM; }

int32_t gcr_mlp_features(int32_t num_pairs, const int64_t* offsets, const double* corr, const int32_t* sizes,
                         const double* ratios, float* feat_out) {
  std::vector<int64_t> rel;
  if (int32_t rc = check_pairs(num_pairs, offsets, rel)) return rc;
  if (!corr || !sizes || !ratios || !feat_out) return fail(GCR_ERR_INVALID_ARGUMENT, "null pointer argument");
  return guarded([&]() -> int32_t {
    Workspace& w = ws();
    const size_t total = (size_t)rel.back();
    w.corr.ensure(4 * total);
    w.ratio.ensure(total);
    w.feat.ensure(IN * total);
    h2d(w, w.corr.p, corr + 4 * offsets[0], 4 * total);
    h2d(w, w.ratio.p, ratios + offsets[0], total);
    features(w, num_pairs, rel, w.corr.p, w.ratio.p, sizes, w.feat.p);
    d2h(w, feat_out, w.feat.p, IN * total);
    sync(w);
    return GCR_OK;
  });
}

int32_t gcr_mlp_predict(const float* params, const float* feat, int64_t n, float* logit_out, float* prob_out) {
  if (!params || !feat || !logit_out || n <= 0) return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
  return guarded([&]() -> int32_t {
    check_lds(lds_forward_bytes());
    Workspace& w = ws();
    pack(w, params);
    w.feat.ensure(IN * (size_t)n);
    w.logit.ensure((size_t)n);
    w.skeys.ensure((size_t)n);
    h2d(w, w.feat.p, feat, IN * (size_t)n);
    launch_predict(w.st, w.wpk.p, w.bias.p, w.feat.p, n, w.logit.p, prob_out ? w.skeys.p : nullptr);
    hip_ok(hipGetLastError(), "predict launch");
    d2h(w, logit_out, w.logit.p, (size_t)n);
    if (prob_out) d2h(w, prob_out, w.skeys.p, (size_t)n);
    sync(w);
    return GCR_OK;
  });
}

int32_t gcr_mlp_gradient(const float* params, const float* feat, const float* labels, int64_t n, float* grad_out,
                         float* loss_out) {
  if (!params || !feat || !labels || !grad_out || n <= 0) return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
  return guarded([&]() -> int32_t {
    check_lds(lds_train_bytes());
    Workspace& w = ws();
    pack(w, params);
    const int64_t bpad = padded(n);
    train_buffers(w, bpad);
    w.feat.ensure(IN * (size_t)n);
    w.labels.ensure((size_t)n);
    w.grad.ensure(NPARAM);
    w.loss.ensure(1);
    h2d(w, w.feat.p, feat, IN * (size_t)n);
    h2d(w, w.labels.p, labels, (size_t)n);
    launch_train_fwd_bwd(w.st, w.wpk.p, w.bias.p, w.feat.p, w.labels.p, n, bpad, 0, 0, 0, 1.0f / (float)n, w.zbuf.p,
                         w.act.p, w.d4.p, w.lossv.p);
    launch_wgrad(w.st, w.act.p, w.d4.p, w.lossv.p, bpad, w.part.p);
    launch_reduce_adam(w.st, w.part.p, (int)(bpad / WGRAD_CHUNK), w.grad.p, w.loss.p, 0, nullptr, nullptr, nullptr,
                       0.0f, 1.0f, 1.0f);
    hip_ok(hipGetLastError(), "gradient launches");
    d2h(w, grad_out, w.grad.p, NPARAM);
    float loss = 0.0f;
    d2h(w, &loss, w.loss.p, 1);
    sync(w);
    if (loss_out) *loss_out = loss;
    return GCR_OK;
  });
}

int32_t gcr_mlp_train(float* params, float* adam_m, float* adam_v, int64_t adam_step, const float* feat,
                      const float* labels, int64_t n, int32_t batch, int32_t steps, float lr, uint64_t seed,
                      float* loss_out) {
  if (!params || !adam_m || !adam_v || !feat || !labels || n <= 0 || batch <= 0 || steps < 0 || adam_step < 0)
    return fail(GCR_ERR_INVALID_ARGUMENT, "bad arguments");
  return guarded([&]() -> int32_t {
    check_lds(lds_train_bytes());
    Workspace& w = ws();
    const int64_t bpad = padded(batch);
    train_buffers(w, bpad);
    w.params.ensure(NPARAM);
    w.m.ensure(NPARAM);
    w.v.ensure(NPARAM);
    w.wpk.ensure(PK_TOTAL);
    w.bias.ensure(PB_TOTAL);
    w.feat.ensure(IN * (size_t)n);
    w.labels.ensure((size_t)n);
    w.loss.ensure((size_t)std::max(steps, 1));
    h2d(w, w.params.p, params, NPARAM);
    h2d(w, w.m.p, adam_m, NPARAM);
    h2d(w, w.v.p, adam_v, NPARAM);
    h2d(w, w.feat.p, feat, IN * (size_t)n);
    h2d(w, w.labels.p, labels, (size_t)n);
    const float inv = 1.0f / (float)batch;
    for (int32_t s = 0; s < steps; ++s) {
      const int64_t t = adam_step + s + 1;
      const float c1 = 1.0f - std::pow(0.9f, (float)t), c2 = 1.0f - std::pow(0.999f, (float)t);
      launch_pack(w.st, w.params.p, w.wpk.p, w.bias.p);
      launch_train_fwd_bwd(w.st, w.wpk.p, w.bias.p, w.feat.p, w.labels.p, batch, bpad, n, seed, t, inv, w.zbuf.p,
                           w.act.p, w.d4.p, w.lossv.p);
      launch_wgrad(w.st, w.act.p, w.d4.p, w.lossv.p, bpad, w.part.p);
      launch_reduce_adam(w.st, w.part.p, (int)(bpad / WGRAD_CHUNK), nullptr, w.loss.p + s, 1, w.params.p, w.m.p,
                         w.v.p, lr, c1, c2);
    }
    hip_ok(hipGetLastError(), "training launches");
    d2h(w, params, w.params.p, NPARAM);
    d2h(w, adam_m, w.m.p, NPARAM);
    d2h(w, adam_v, w.v.p, NPARAM);
    if (loss_out && steps > 0) d2h(w, loss_out, w.loss.p, (size_t)steps);
    sync(w);
    return GCR_OK;
  });
}

int32_t gcr_find_batch_guided(int32_t num_pairs, const int32_t* kinds, const int64_t* offsets, const double* corr,
                              int32_t inputs_on_device, const int32_t* sizes, const double* K, const double* ratios,
                              const float* mlp_params, const gcr_params_t* params, double* models_out,
                              uint8_t* masks_out, int32_t* num_inliers_out, gcr_stats_t* stats, float* logits_out) {
  std::vector<int64_t> rel;
  if (int32_t rc = check_pairs(num_pairs, offsets, rel)) return rc;
  if (!kinds || !corr || !sizes || !ratios || !mlp_params || !params || !models_out || !masks_out ||
      !num_inliers_out)
    return fail(GCR_ERR_INVALID_ARGUMENT, "null pointer argument");
  for (int i = 0; i < num_pairs; ++i) {
    gcr_params_t p = params[i];
    p.sampler = 1;  // overridden below: PROSAC over the network's order
    const double* K1 = K ? K + 18 * i : nullptr;
    if (int32_t rc = gcr::check_pair(i, kinds[i], rel[(size_t)i + 1] - rel[(size_t)i], p, K1, K1 ? K1 + 9 : nullptr))
      return rc;
  }
  return guarded([&]() -> int32_t {
    check_lds(lds_forward_bytes());
    const auto t0 = std::chrono::steady_clock::now();
    Workspace& w = ws();
    const size_t total = (size_t)rel.back();
    std::vector<gcr::PairSpec> specs((size_t)num_pairs);
    for (int i = 0; i < num_pairs; ++i) {
      gcr::PairSpec& s = specs[(size_t)i];
      s.kind = kinds[i];
      s.offset = rel[(size_t)i];
      s.n = rel[(size_t)i + 1] - rel[(size_t)i];
      s.h1 = sizes[4 * i + 0];
      s.w1 = sizes[4 * i + 1];
      s.h2 = sizes[4 * i + 2];
      s.w2 = sizes[4 * i + 3];
      if (K) {
        std::memcpy(s.K1, K + 18 * i, sizeof(s.K1));
        std::memcpy(s.K2, K + 18 * i + 9, sizeof(s.K2));
      }
      s.params = params[i];
      s.params.sampler = 1;  // PROSAC over the network's order
    }
    const double* d_corr = corr + 4 * offsets[0];
    const double* d_ratio = ratios + offsets[0];
    if (!inputs_on_device) {
      w.corr.ensure(4 * total);
      w.ratio.ensure(total);
      h2d(w, w.corr.p, d_corr, 4 * total);
      h2d(w, w.ratio.p, d_ratio, total);
      d_corr = w.corr.p;
      d_ratio = w.ratio.p;
    }
    w.feat.ensure(IN * total);
    w.logit.ensure(total);
    w.skeys.ensure(total);
    w.vals.ensure(total);
    w.perm.ensure(total);
    w.sorted.ensure(4 * total);
    features(w, num_pairs, rel, d_corr, d_ratio, sizes, w.feat.p);
    pack(w, mlp_params);
    launch_predict(w.st, w.wpk.p, w.bias.p, w.feat.p, (int64_t)total, w.logit.p, nullptr);
    gcr::mlp::launch_local_iota(w.st, w.off.p, num_pairs, (int64_t)total, w.vals.p);
    size_t tmp_bytes = 0;
    hip_ok(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(nullptr, tmp_bytes, w.logit.p, w.skeys.p, w.vals.p,
                                                                 w.perm.p, (int)total, num_pairs, w.off.p,
                                                                 w.off.p + 1, 0, 32, w.st),
           "segmented sort (size)");
    w.tmp.ensure(tmp_bytes);
    hip_ok(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(w.tmp.p, tmp_bytes, w.logit.p, w.skeys.p, w.vals.p,
                                                                 w.perm.p, (int)total, num_pairs, w.off.p,
                                                                 w.off.p + 1, 0, 32, w.st),
           "segmented sort");
    gcr::mlp::launch_gather_rows(w.st, d_corr, w.perm.p, w.off.p, num_pairs, (int64_t)total, w.sorted.p);
    hip_ok(hipGetLastError(), "guidance launches");
    // the batch starts once the sorted rows exist; the permutation (needed
    // only to put the masks back in the caller's order) comes back meanwhile
    hip_ok(hipEventRecord(w.sorted_ev, w.st), "hipEventRecord");
    int32_t* const perm = w.perm_host(total);
    d2h(w, perm, w.perm.p, total);
    hip_ok(hipEventSynchronize(w.sorted_ev), "hipEventSynchronize");
    std::vector<gcr::PairResult> res;
    gcr::run_batch(specs, w.sorted.p, true, res);
    if (logits_out) d2h(w, logits_out, w.logit.p, total);
    sync(w);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < num_pairs; ++i) {
      const gcr::PairResult& r = res[(size_t)i];
      std::memcpy(models_out + 9 * i, r.model, sizeof(double) * 9);
      const int64_t o = rel[(size_t)i];
      uint8_t* mk = masks_out + o;
      std::memset(mk, 0, (size_t)specs[(size_t)i].n);
      for (int32_t k : r.inliers) mk[perm[(size_t)(o + k)]] = 1;
      num_inliers_out[i] = r.status;
      if (stats) {
        stats[i] = r.stats;
        stats[i].processing_time = secs;
      }
    }
    return GCR_OK;
  });
}

}  // extern "C"
