// gcr_score_models_noalloc: MSAC scoring of k models over n points with no
// device allocation.  Inputs and outputs are the caller's device buffers; the
// scoring kernel's index tables and 32-bit counts live in one file-scope
// device array of fixed size (`scratch`), which each GPU holds until the
// process ends.  Every call on a GPU reuses that GPU's array, so calls that
// overlap on one GPU race; calls on one stream are ordered.  The tables are
// written by a kernel at the start of every call (no memset), so the entry
// point can also be recorded into a HIP graph.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"
#include "kernels.hpp"

namespace {

constexpr int64_t kMaxModels = GCR_SCORE_NOALLOC_MAX_MODELS;

// Scratch layout (bytes):
//   [0, 8)      pair point offset (int64, 0)
//   [8, 12)     pair point count n (int32)
//   [16, 24)    squared truncated threshold (double)
//   [256, ...)  model -> pair table (int32 x kMaxModels, all 0)
//   then        mask offsets (int64 x kMaxModels, j * n)
//   then        counts (int32 x kMaxModels), widened into the caller's int64
constexpr size_t kOffPair = 256;
constexpr size_t kOffMask = kOffPair + 4 * kMaxModels;
constexpr size_t kOffCount = kOffMask + 8 * kMaxModels;
constexpr size_t kScratchBytes = kOffCount + 4 * kMaxModels;
static_assert(kOffMask % 8 == 0 && kOffCount % 8 == 0, "scratch alignment");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {nullptr};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int32_t scratch_for_current_device(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    gcr::set_last_error("no current HIP device");
    return GCR_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      gcr::set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return GCR_ERR_DEVICE;
    }
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return GCR_OK;
}

__global__ void k_noalloc_tables(unsigned char* s, int32_t n, double sq_thr, int64_t k) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) {
    *reinterpret_cast<int64_t*>(s) = 0;
    *reinterpret_cast<int32_t*>(s + 8) = n;
    *reinterpret_cast<double*>(s + 16) = sq_thr;
  }
  if (j < k) {
    reinterpret_cast<int32_t*>(s + kOffPair)[j] = 0;
    reinterpret_cast<int64_t*>(s + kOffMask)[j] = j * (int64_t)n;
  }
}

__global__ void k_noalloc_widen(const int32_t* c32, int64_t* c64, int64_t k) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) c64[j] = c32[j];
}

}  // namespace

extern "C" int32_t gcr_score_models_noalloc(int32_t kind, const double* pts, int64_t n, const double* models,
                                            int64_t k, double sq_thr, int64_t* counts, double* values,
                                            uint8_t* masks, void* stream) {
  if (!pts || !models || !counts || !values) {
    gcr::set_last_error("null pointer argument");
    return GCR_ERR_INVALID_ARGUMENT;
  }
  if (kind != 0 && kind != 1) {
    gcr::set_last_error("kind must be 0 (homography) or 1 (Sampson)");
    return GCR_ERR_INVALID_ARGUMENT;
  }
  if (n <= 0 || n > (int64_t)INT32_MAX) {
    gcr::set_last_error("n must be in [1, 2^31 - 1]");
    return GCR_ERR_INVALID_ARGUMENT;
  }
  if (k <= 0 || k > kMaxModels) {
    gcr::set_last_error("k must be in [1, " + std::to_string(kMaxModels) +
                        "]: the no-allocation scratch holds that many models");
    return GCR_ERR_INVALID_ARGUMENT;
  }
  if (masks && k > INT64_MAX / n) {
    gcr::set_last_error("k * n overflows");
    return GCR_ERR_INVALID_ARGUMENT;
  }
  unsigned char* s = nullptr;
  if (int32_t rc = scratch_for_current_device(&s)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned blocks = (unsigned)((k + 255) / 256);
  hipLaunchKernelGGL(k_noalloc_tables, dim3(blocks), dim3(256), 0, st, s, (int32_t)n, sq_thr, k);
  int32_t* c32 = reinterpret_cast<int32_t*>(s + kOffCount);
  gcr::launch_score(st, kind, pts, reinterpret_cast<const int64_t*>(s), reinterpret_cast<const int32_t*>(s + 8),
                    reinterpret_cast<const double*>(s + 16), models, reinterpret_cast<const int32_t*>(s + kOffPair),
                    nullptr, k, c32, values, masks, reinterpret_cast<const int64_t*>(s + kOffMask));
  hipLaunchKernelGGL(k_noalloc_widen, dim3(blocks), dim3(256), 0, st, c32, counts, k);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gcr::set_last_error(std::string("scoring launch: ") + hipGetErrorString(e));
    return GCR_ERR_DEVICE;
  }
  return GCR_OK;
}
