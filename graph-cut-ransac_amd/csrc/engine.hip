// Batched Graph-Cut RANSAC engine: coroutine control flow per pair, batched
// device requests across pairs.  See engine.hpp.
#include "engine.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <coroutine>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "host_pool.hpp"
#include "host_samplers.hpp"
#include "kernels.hpp"

namespace gcr {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    if (e == hipErrorOutOfMemory) throw EngineError(GCR_ERR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
    throw EngineError(GCR_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  }
}

// ---------------------------------------------------------------------------
// Kernel profiler: HIP events recorded on the engine stream around the
// launches of each kernel family, read after the step's synchronisation.
// ---------------------------------------------------------------------------
struct ProfEntry {
  double ms = 0.0;
  int64_t launches = 0;
  double work = 0.0;  // family-specific unit (hypothesis-points for scoring)
};
static bool g_prof_on = false;

// Candidate chunks are scored by the matrix-core kernel (launch_score_mx)
// unless GCR_SCORE_MFMA=0 selects the FP64 kernel; both give the same
// counts, values and symmetric counts.  Read once per engine run.
static bool use_mx_scoring() {
  const char* e = std::getenv("GCR_SCORE_MFMA");
  return !(e && e[0] == '0');
}
// The uniform sampler's draws stay on the host, drawn ahead while the
// device works, unless GCR_DEVICE_SAMPLING=1 moves them to the device
// (k_sample.hip); both give the same samples.  Measured on MI355X at 1,024
// fundamental pairs: 3,996 / 4,063 pairs/s with host draws drawn ahead
// against 3,808 / 3,787 with device draws, whose kernel (one wavefront per
// pair walking its stream) sits on each chunk's critical path.
// PROSAC's draws stay on the host unless GCR_DEVICE_PROSAC=1: its subsets
// grow by about one point per sample for the first ~1,700 samples of a
// 4,000-point pair, so the device batches cut short at repeated indices and
// span many segments; measured on MI355X (learned-guidance batch, 1,024
// pairs) the device draws took 275 us per chunk launch and 4,539 pairs/s
// against 5,202 with host draws.
static bool use_device_sampling() {
  const char* e = std::getenv("GCR_DEVICE_SAMPLING");
  return e && e[0] == '1';
}
static bool use_device_prosac() {
  const char* e = std::getenv("GCR_DEVICE_PROSAC");
  return e && e[0] == '1';
}
static std::map<std::string, ProfEntry> g_prof;
static std::mutex g_prof_mu;  // engines of concurrent batch groups add to g_prof

void profile_enable(bool on) { g_prof_on = on; }
void profile_reset() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof.clear();
}
bool profile_read(const std::string& name, double* ms, int64_t* launches, double* work) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto it = g_prof.find(name);
  if (it == g_prof.end()) return false;
  *ms = it->second.ms;
  *launches = it->second.launches;
  *work = it->second.work;
  return true;
}

namespace {

struct ProfSpan {
  std::string name;
  hipEvent_t a, b;
  double work;
  std::string sub;  // optional second entry that gets the same time
  double sub_work;
};

struct Profiler {
  std::vector<ProfSpan> open;
  std::vector<hipEvent_t> spare;  // events of collected spans, reused
  hipStream_t st = nullptr;
  void begin(hipStream_t s) { st = s; }
  ~Profiler() {
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    for (auto& sp : open) {
      (void)hipEventDestroy(sp.a);
      (void)hipEventDestroy(sp.b);
    }
  }
  hipEvent_t event() {
    if (!spare.empty()) {
      hipEvent_t e = spare.back();
      spare.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    check_hip(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  // records the start event; returns an index to pass to end()
  int start(const char* name, double work, const char* sub = nullptr, double sub_work = 0.0) {
    if (!g_prof_on) return -1;
    ProfSpan sp{name, event(), event(), work, sub ? sub : "", sub_work};
    check_hip(hipEventRecord(sp.a, st), "hipEventRecord");
    open.push_back(sp);
    return (int)open.size() - 1;
  }
  void stop(int i) {
    if (i < 0) return;
    check_hip(hipEventRecord(open[(size_t)i].b, st), "hipEventRecord");
  }
  // after a stream synchronisation
  void collect() {
    if (open.empty()) return;
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& sp : open) {
      float ms = 0.f;
      check_hip(hipEventElapsedTime(&ms, sp.a, sp.b), "hipEventElapsedTime");
      ProfEntry& e = g_prof[sp.name];
      e.ms += ms;
      e.launches += 1;
      e.work += sp.work;
      if (!sp.sub.empty()) {
        ProfEntry& u = g_prof[sp.sub];
        u.ms += ms;
        u.launches += 1;
        u.work += sp.sub_work;
      }
      spare.push_back(sp.a);
      spare.push_back(sp.b);
    }
    open.clear();
  }
};

// Host wall time of a step outside the engine loop's phases ("host_*").
void prof_host(const char* name, std::chrono::steady_clock::time_point since) {
  if (!g_prof_on) return;
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - since).count();
  std::lock_guard<std::mutex> g(g_prof_mu);
  ProfEntry& e = g_prof[name];
  e.ms += ms;
  e.launches += 1;
}

// Work (e.g. hypothesis-points) known only after a span's results are back.
void prof_add_work(const char* name, double work) {
  if (!g_prof_on) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof[name].work += work;
}

// Algorithmic FP64 flops of one non-minimal fit of ns points (add, sub, mul,
// div and sqrt count one each), from the operations of the fit as written:
//  homography: Hartley normalisation 26/pt, two DLT rows 12/pt, column-pivoted
//    Householder QR of the 2ns x 9 system 2 m n^2 - 2 n^3 / 3 (m = 2 ns, n = 9),
//    back substitution and denormalisation ~200;
//  fundamental: normalisation 26/pt (flag bit 0); eight-point A^T A 98/pt +
//    9 x 9 QR ~1100, or the seven-point solve ~3000 for ns = 7; each LM
//    Jacobian pass 306/pt (residual, 9 x 7 chain rule, 7 + 28 accumulations),
//    each cost pass 35/pt, ~600 per step (7 x 7 Cholesky, solves, rotation
//    update);
//  essential (Stewenius): weighted A^T A 100/pt + ~16,000 (QR, 10 x 20
//    constraints, LU, 10 x 10 eigenvalues, null vectors); the driver's pose
//    bundle adjustment is not counted.
double fit_flops(int kind, int flags, int ns, int jac_passes, int cost_passes) {
  const double n = (double)ns;
  if (kind == GCR_PROBLEM_HOMOGRAPHY) return (26.0 + 12.0) * n + (2.0 * (2.0 * n) * 81.0 - 486.0) + 200.0;
  if (kind == GCR_PROBLEM_FUNDAMENTAL) {
    double f = (flags & 1) ? 26.0 * n : 0.0;
    f += ns >= 8 ? 98.0 * n + 1100.0 : 3000.0;
    f += (306.0 * n) * jac_passes + (35.0 * n + 600.0) * cost_passes;
    return f;
  }
  if (((flags >> 8) & 3) == 3) return 0.0;
  return 100.0 * n + 16000.0;
}

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  T* ensure(size_t n) {
    if (n == 0) n = 1;
    if (n > cap) {
      if (p) check_hip(hipFree(p), "hipFree");
      p = nullptr;
      check_hip(hipMalloc((void**)&p, n * sizeof(T)), "hipMalloc");
      cap = n;
    }
    return p;
  }
};

template <class T>
struct HostBuf {
  T* p = nullptr;
  size_t cap = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  HostBuf(HostBuf&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  HostBuf& operator=(HostBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipHostFree(p);
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  T* ensure(size_t n) {
    if (n == 0) n = 1;
    if (n > cap) {
      if (p) check_hip(hipHostFree(p), "hipHostFree");
      p = nullptr;
      check_hip(hipHostMalloc((void**)&p, n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
      cap = n;
    }
    return p;
  }
};

// Device model arrays handed to the pairs of a batch.  A pair keeps the
// candidate models of its last chunk (and the models of its last fit) while
// it is suspended on other requests, and later slot copies read from them, so
// each chunk / fit step takes its own array from the pool; an array returns
// to the pool when the last request holding it is gone.
struct ModelPool {
  std::vector<std::unique_ptr<DevBuf<double>>> idle;
  std::mutex mu;  // arrays are released from the host pool's threads
  ModelPool() = default;
  ModelPool(ModelPool&& o) noexcept : idle(std::move(o.idle)) {}
  ModelPool& operator=(ModelPool&& o) noexcept {
    idle = std::move(o.idle);
    return *this;
  }
  std::shared_ptr<DevBuf<double>> get(size_t n) {
    std::unique_ptr<DevBuf<double>> b;
    {
      std::lock_guard<std::mutex> g(mu);
      size_t pick = idle.size();
      for (size_t i = 0; i < idle.size(); ++i)
        if (idle[i]->cap >= n && (pick == idle.size() || idle[i]->cap < idle[pick]->cap)) pick = i;
      if (pick < idle.size()) {
        b = std::move(idle[pick]);
        idle.erase(idle.begin() + (std::ptrdiff_t)pick);
      }
    }
    if (!b) {
      b = std::make_unique<DevBuf<double>>();
      b->ensure(n);
    }
    return std::shared_ptr<DevBuf<double>>(b.release(), [this](DevBuf<double>* q) {
      std::lock_guard<std::mutex> g(mu);
      idle.emplace_back(q);
    });
  }
};

// ---------------------------------------------------------------------------
// Scores and model slots
// ---------------------------------------------------------------------------
struct Score {
  int64_t count = 0;
  double value = 0.0;
};
inline bool operator<(const Score& a, const Score& b) { return a.value < b.value; }

enum Slot { S_BEST = 0, S_CONT0 = 1, S_CONT1 = 2, S_LOBEST = 3, S_IRLS = 4, S_TMP = 5, NSLOT = 6 };
// Per pair: the run's NSLOT slots, the nested run's NSLOT slots (DEGENSAC's
// plane-and-parallax GC-RANSAC, slot base NSLOT) and its homography.
constexpr int NSLOT_ALL = 2 * NSLOT + 1;
constexpr int S_PPH_ABS = 2 * NSLOT;

// Estimators a pair's loop can run: the problem kinds and the
// plane-and-parallax fundamental-matrix estimator of DEGENSAC's nested run
// (two-point solver with a fixed homography, eight-point non-minimal fits).
constexpr int EST_PP = 3;
int est_sample_size(int est) {
  return est == GCR_PROBLEM_HOMOGRAPHY ? 4 : (est == GCR_PROBLEM_FUNDAMENTAL ? 7 : (est == EST_PP ? 2 : 5));
}
int est_solutions(int est) {
  return est == GCR_PROBLEM_HOMOGRAPHY ? 1 : (est == GCR_PROBLEM_FUNDAMENTAL ? 3 : (est == EST_PP ? 1 : 10));
}
// residual / fit kind of an estimator: plane-and-parallax models are
// fundamental matrices
int est_kind(int est) { return est == EST_PP ? GCR_PROBLEM_FUNDAMENTAL : est; }

// ---------------------------------------------------------------------------
// Coroutine plumbing
// ---------------------------------------------------------------------------
struct RootTask {
  struct promise_type {
    std::exception_ptr error;
    RootTask get_return_object() { return RootTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    std::suspend_always final_suspend() noexcept { return {}; }
    void return_void() {}
    void unhandled_exception() { error = std::current_exception(); }
  };
  std::coroutine_handle<promise_type> h;
};

template <class T>
struct SubTask {
  struct promise_type {
    T value{};
    std::exception_ptr error;
    std::coroutine_handle<> cont;
    SubTask get_return_object() { return SubTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    struct Final {
      bool await_ready() noexcept { return false; }
      std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept { return h.promise().cont; }
      void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void return_value(T v) { value = v; }
    void unhandled_exception() { error = std::current_exception(); }
  };
  std::coroutine_handle<promise_type> h;
  bool await_ready() { return false; }
  std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) {
    h.promise().cont = c;
    return h;
  }
  T await_resume() {
    std::exception_ptr e = h.promise().error;
    T v = h.promise().value;
    h.destroy();
    if (e) std::rethrow_exception(e);
    return v;
  }
};

// One device phase of an engine step (the candidate chunks of one
// estimator, the labellings, the fits, ...).  The phase enqueues its copies
// and kernels on the step's stream, then awaits the step's synchronisation
// (co_await Engine::step_sync()) before it reads its results back.  The
// engine starts every phase of a step before it synchronises once, so the
// host packs one phase while the device runs the one before.  A phase may
// await another phase task (a nested stage); an exception is kept in the
// promise and rethrown by the awaiting phase or by the engine.
struct PhaseTask {
  struct promise_type {
    std::exception_ptr error;
    std::coroutine_handle<> cont;
    PhaseTask get_return_object() { return PhaseTask{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    struct Final {
      bool await_ready() noexcept { return false; }
      std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
        if (h.promise().cont) return h.promise().cont;
        return std::noop_coroutine();
      }
      void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void return_void() {}
    void unhandled_exception() { error = std::current_exception(); }
  };
  std::coroutine_handle<promise_type> h;
  explicit PhaseTask(std::coroutine_handle<promise_type> hh) : h(hh) {}
  PhaseTask(PhaseTask&& o) noexcept : h(o.h) { o.h = nullptr; }
  PhaseTask& operator=(PhaseTask&&) = delete;
  PhaseTask(const PhaseTask&) = delete;
  ~PhaseTask() {
    if (h) h.destroy();
  }
  bool await_ready() { return false; }
  std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) {
    h.promise().cont = c;
    return h;
  }
  void await_resume() {
    if (h.promise().error) std::rethrow_exception(h.promise().error);
  }
};

// The point where a phase waits for the step's synchronisation: the phase's
// innermost coroutine parks itself on the engine's list.
struct StepSync {
  std::vector<std::coroutine_handle<>>* parked;
  bool await_ready() noexcept { return false; }
  void await_suspend(std::coroutine_handle<> h) { parked->push_back(h); }
  void await_resume() noexcept {}
};

enum ReqType { REQ_NONE = 0, REQ_CHUNK, REQ_LABEL, REQ_FIT, REQ_MASKLIST, REQ_SPRT, REQ_NTYPES };

struct Request {
  ReqType type = REQ_NONE;
  // REQ_CHUNK: candidate minimal samples (num_cand x m), or drawn on the
  // device from the pair's uniform sampler stream (dev_sample: empty)
  std::vector<int32_t> samples;
  int64_t num_cand = 0;
  bool dev_sample = false;
  std::vector<SampleSeg> sample_segs;  // dev_sample: the draws, as segments
  // the host draws the samples from the pair's main sampler straight into
  // the step's upload block (samples empty)
  bool host_draw = false;
  // REQ_FIT: fits (sample lists), optional IRLS weights from (weight_model,
  // samples of fit 0)
  std::vector<std::vector<int32_t>> fits;
  bool weighted = false;
  const double* weight_model = nullptr;
  bool want_list = false;
  // REQ_MASKLIST
  const double* model = nullptr;
  bool post = false;
  int fit_flags = 0;  // bit 0: Hartley normalisation; bits 8-9: weight mode (F fits)
  // results (per model slot; slots_per slots per candidate / fit)
  int slots_per = 1;
  std::vector<int8_t> status, valid;
  std::vector<int8_t> nmodels;  // chunk: models per candidate
  std::vector<int32_t> counts, sym;
  std::vector<double> values;
  const double* dev_models = nullptr;  // chunk / fit model array of this request (9 per slot)
  std::shared_ptr<DevBuf<double>> models_hold;  // keeps dev_models allocated
  std::vector<int32_t> list;           // label / want_list / masklist result
  double model_host[9] = {0};
  // REQ_CHUNK of a DEGENSAC pair: per slot, H-degenerate flag and homography
  std::vector<int8_t> degen;
  const double* dev_degen_h = nullptr;
  // REQ_FIT / REQ_MASKLIST: residual / fit kind and squared threshold other
  // than the pair's current estimator's (-1 / < 0: none)
  int kind_override = -1;
  double sq_override = -1.0;
  // REQ_SPRT: the walk over these model slots of dev_models (in / out)
  std::vector<int32_t> sprt_slots;
  SprtItem sprt{};
  // REQ_CHUNK of a space-partitioning pair: per slot, the points in the
  // cells the selector kept
  std::vector<int32_t> sel;
  // REQ_CHUNK filtered on the device (loops scoring every model in full):
  // the input best value; the slots whose value beats it, in slot order,
  // and each candidate's last model's count.  A slot whose value does not
  // beat the best at the chunk's start cannot beat the best later (it only
  // grows), so the loop needs nothing else of it but the inlier container's
  // count, which a candidate's last model sets.
  double filter_best = 0.0;
  bool filtered = false;
  std::vector<SlotRec> recs;
  std::vector<int32_t> last_count;
  // REQ_MASKLIST: the inliers in the selector's cells (cell order), or the
  // list the last SPRT walk wrote (walked >= 0 entries)
  bool sp_list = false;
  int64_t walked = -1;
};

struct PairCtx;

struct DeviceAwait {
  PairCtx* p;
  bool await_ready() noexcept { return false; }
  void await_suspend(std::coroutine_handle<> h) noexcept;
  void await_resume() noexcept {}
};

// One of the loop's two inlier containers (inc/GCRANSAC.h:232-236): the
// model slot its list is regenerated from and the list's length.  The list's
// order: pool_start >= 0, an SPRT test's list in pool order from pool_start
// over its first pool_tested positions; kListAscending, point order;
// kListCellOrder, the model's inliers in its space-partitioning cells, in
// cell order; kListWalked, the list the last SPRT walk over cells wrote on
// the device.
constexpr int64_t kListAscending = -1, kListCellOrder = -2, kListWalked = -3;
struct Container {
  int slot;
  int64_t count = 0;
  bool has = false;
  int64_t pool_start = -1;
  int64_t pool_tested = 0;
};

// SPRT state of a pair's main loop (preemption_sprt.h:141-190): the random
// pool, the running pool index and the current history; a_table holds A for
// delta = k / N at the current epsilon, uploaded to a_dev.
struct SprtHost {
  std::vector<int32_t> pool;
  int64_t pool_idx = 0;
  double eps = 0, delta = 0, A = 0;
  double t_M = 0, m_S = 1;
  std::vector<double> a_table;
  double* a_dev = nullptr;
  bool table_dirty = true;
  // preemption_sprt.h:299-330
  double threshold_A(double epsilon, double d) const {
    const double C = (1 - d) * std::log((1 - d) / (1 - epsilon)) + d * (std::log(d / epsilon));
    const double K = (t_M * C) / m_S + 1;
    double An_1 = K, An = K;
    for (unsigned int i = 0; i < 10; ++i) {
      An = K + std::log(An_1);
      if (std::fabs(An - An_1) < 1.5e-8) break;
      An_1 = An;
    }
    return An;
  }
  void fill_table(size_t n) {
    a_table.assign(n + 1, 0.0);
    for (size_t k = 1; k <= n; ++k) a_table[k] = threshold_A(eps, (double)k / (double)n);
    table_dirty = true;
  }
};

struct PairCtx {
  int index = 0;
  PairSpec spec;
  PairResult* out = nullptr;
  // derived settings
  int m = 4;
  int maxsol = 1;  // models per minimal sample
  int64_t n = 0;
  double threshold = 1, conf = 0.99, lambda = 0, tau = 0, sq_thr = 0, lo_thr = 0, log_prob = 0;
  double post_sq = 0;      // squared threshold of the driver's post-hoc bundle adjustment
  double min_ratio = 0.5;  // validity: symmetric-epipolar share of the Sampson inliers
  size_t min_iters = 0, max_iters = 0, lo_number = 50;
  int qbits = 40;
  std::unique_ptr<Sampler> main_sampler;
  std::unique_ptr<UniformSampler> lo_sampler;
  // the loop's current estimator and samplers (the run's, or DEGENSAC's
  // nested run's: one sampler for both roles, as the reference passes it)
  int est = 0;
  bool degensac = false;  // fundamental: DEGENSAC on every accepted model
  Sampler* main_s = nullptr;
  UniformSampler* lo_s = nullptr;
  // samples of main_sampler drawn ahead while the device works (ahead_m
  // indices each, in stream order from ahead_pos); a host-drawn chunk takes
  // them first.  ahead_hint: how many the next chunk is expected to take.
  std::vector<int32_t> ahead;
  size_t ahead_pos = 0, ahead_m = 0, ahead_hint = 0;
  bool dev_sampling = false;  // main_sampler (uniform or PROSAC) draws on the device
  double time_limit = 0.0;    // seconds from t_start (GCRANSAC::setFPS), 0: none
  std::chrono::steady_clock::time_point t_start;
  uint32_t main_seed_v = 0;
  std::unique_ptr<UniformSampler> nested_sampler;
  int slot_base = 0;
  bool params_dirty = false;  // sq_thr / lo_thr / lambda changed: upload before the next step
  bool radius = false;        // labelled on the radius neighbourhood
  bool sprt_on = false;       // the main loop verifies models with SPRT (off in DEGENSAC's nested run)
  bool sp_on = false;         // homographies: space partitioning selector before scoring
  SprtHost sprt;
  bool no_graph = false;      // radius neighbourhood without an edge: no model, no run
  bool label_failed = false;  // the last labelling's min-cut gave up
  // coroutine
  Request req;
  std::coroutine_handle<> handle;
  RootTask root;
  bool finished = false;
  // model slots (device), pending slot copies
  double* slots = nullptr;  // NSLOT_ALL x 9
  std::array<const double*, NSLOT_ALL> pending{};  // per slot: source of its pending copy (nullptr: none)
  // statistics
  size_t iteration = 0, graph_cut_number = 0, lo_count = 0;
  int64_t candidates_solved = 0, models_scored = 0;

  double* slot(int s) const { return slots + 9 * (slot_base + s); }
  double* abs_slot(int k) const { return slots + 9 * k; }
  // Slot s (of the current slot base) will hold the model at src (executed
  // before the next device step).
  void set_slot(int s, const double* src) { set_abs_slot(slot_base + s, src); }
  void set_abs_slot(int k_dst, const double* src) {
    for (int k = 0; k < NSLOT_ALL; ++k)
      if (src == abs_slot(k)) {
        if (pending[(size_t)k]) src = pending[(size_t)k];
        break;
      }
    if (src == abs_slot(k_dst)) {
      pending[(size_t)k_dst] = nullptr;  // the slot already holds it
      return;
    }
    pending[(size_t)k_dst] = src;
  }
};

void DeviceAwait::await_suspend(std::coroutine_handle<> h) noexcept { p->handle = h; }

// inc/GCRANSAC.h:157-172
size_t iteration_number(size_t inliers, size_t n, size_t m, double log_prob) {
  const double q = std::pow(static_cast<double>(inliers) / n, m);
  const double log2 = std::log(1 - q);
  if (std::abs(log2) < std::numeric_limits<double>::epsilon()) return std::numeric_limits<size_t>::max();
  const double iter = log_prob / log2;
  return static_cast<size_t>(iter) + 1;
}

// Model slots of one fit: homography 1, fundamental 3 (seven-point fits give
// up to three), essential 10 (Stewenius), essential bundle adjustment 40
// (up to four poses of each of up to ten initial models).
int fit_slots(int kind, int flags) {
  if (kind == GCR_PROBLEM_HOMOGRAPHY) return 1;
  if (kind == GCR_PROBLEM_FUNDAMENTAL) return 3;
  return ((flags >> 8) & 3) == 3 ? 40 : 10;
}

int sample_size_of(int kind) { return kind == GCR_PROBLEM_HOMOGRAPHY ? 4 : (kind == GCR_PROBLEM_FUNDAMENTAL ? 7 : 5); }

// 3x3 inverse by the adjugate (row-major), expanded along the first row.
void mat3_inverse_host(const double* a, double* inv) {
  const double c00 = a[4] * a[8] - a[5] * a[7];
  const double c01 = a[5] * a[6] - a[3] * a[8];
  const double c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  const double id = 1.0 / det;
  inv[0] = c00 * id;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  inv[3] = c01 * id;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  inv[6] = c02 * id;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

int mincut_bits(size_t nodes, size_t edges) {
  const double bound = (double)nodes + 3.0 * (double)edges + 1.0;
  int lg = 0;
  while (std::ldexp(1.0, lg) < bound) ++lg;
  return 61 - lg;
}

// ---------------------------------------------------------------------------
// Per-pair control flow (inc/GCRANSAC.h), H path
// ---------------------------------------------------------------------------
struct Engine;

struct PairRunner {
  PairCtx& P;

  // The points of an ascending inlier list met in the SPRT pool's order from
  // start over its first `tested` positions: the list an SPRT test wrote
  // (preemption_sprt.h:236-258).
  std::vector<int32_t> pool_order(const std::vector<int32_t>& asc, int64_t start, int64_t tested) const {
    const int64_t n = P.n;
    std::vector<uint8_t> in((size_t)n, 0);
    for (int32_t i : asc) in[(size_t)i] = 1;
    std::vector<int32_t> out;
    for (int64_t t = 0; t < tested; ++t) {
      int64_t pos = start + t;
      if (pos >= n) pos -= n;
      const int32_t idx = P.sprt.pool[(size_t)pos];
      if (in[(size_t)idx]) out.push_back(idx);
    }
    return out;
  }

  // --- requests ---
  DeviceAwait chunk(std::vector<int32_t>&& samples, int64_t num, double best_value) {
    P.req = Request();
    P.req.type = REQ_CHUNK;
    P.req.samples = std::move(samples);
    P.req.num_cand = num;
    P.req.filter_best = best_value;
    return DeviceAwait{&P};
  }
  // the next num samples of the pair's main sampler, drawn on the host by
  // the step that uploads them
  DeviceAwait chunk_host(int64_t num, double best_value) {
    P.req = Request();
    P.req.type = REQ_CHUNK;
    P.req.host_draw = true;
    P.req.num_cand = num;
    P.req.filter_best = best_value;
    return DeviceAwait{&P};
  }
  // the next num samples of the pair's main sampler, whose random draws the
  // device makes (segs: the sampler's plan of them)
  DeviceAwait chunk_dev(int64_t num, double best_value, std::vector<SampleSeg>&& segs) {
    P.req = Request();
    P.req.type = REQ_CHUNK;
    P.req.dev_sample = true;
    P.req.sample_segs = std::move(segs);
    P.req.num_cand = num;
    P.req.filter_best = best_value;
    return DeviceAwait{&P};
  }
  DeviceAwait label() {
    P.req = Request();
    P.req.type = REQ_LABEL;
    return DeviceAwait{&P};
  }
  // Non-minimal fits.  weighted: IRLS weights from (wmodel, samples of fit
  // 0) stored by point index.  flags: see Request::fit_flags.
  DeviceAwait fit(std::vector<std::vector<int32_t>>&& fits, bool weighted, const double* wmodel, bool want_list,
                  int flags) {
    P.req = Request();
    P.req.type = REQ_FIT;
    P.req.fits = std::move(fits);
    P.req.weighted = weighted;
    P.req.weight_model = wmodel;
    P.req.want_list = want_list;
    P.req.fit_flags = flags;
    return DeviceAwait{&P};
  }
  // post: score at the post-hoc adjustment's threshold instead of tau^2
  // SPRT walk over model slots of one candidate chunk (dev_models)
  DeviceAwait sprt_walk(const double* dev_models, std::vector<int32_t>&& slots, double best_value,
                        int64_t best_count) {
    P.req = Request();
    P.req.type = REQ_SPRT;
    P.req.sprt_slots = std::move(slots);
    SprtItem& it = P.req.sprt;
    it.models = dev_models;
    it.num_slots = (int32_t)P.req.sprt_slots.size();
    it.pair = P.index;
    it.kind = est_kind(P.est) == GCR_PROBLEM_HOMOGRAPHY ? 0 : 1;
    it.pool_idx = P.sprt.pool_idx;
    it.eps = P.sprt.eps;
    it.delta = P.sprt.delta;
    it.A = P.sprt.A;
    it.best_value = best_value;
    it.sp = P.sp_on ? 1 : 0;
    it.best_count = best_count;
    return DeviceAwait{&P};
  }
  // the inliers of a model in its space-partitioning cells, in cell order
  DeviceAwait masklist_cells(const double* model) {
    DeviceAwait a = masklist(model);
    P.req.sp_list = true;
    return a;
  }
  // the list the last SPRT walk over cells wrote (count entries)
  DeviceAwait masklist_walked(const double* model, int64_t count) {
    DeviceAwait a = masklist(model);
    P.req.walked = count;
    return a;
  }
  DeviceAwait masklist(const double* model, bool post = false) {
    P.req = Request();
    P.req.type = REQ_MASKLIST;
    P.req.model = model;
    P.req.post = post;
    return DeviceAwait{&P};
  }
  // fit with another estimator's non-minimal solver (kind) on this pair
  DeviceAwait fit_as(std::vector<std::vector<int32_t>>&& fits, int kind, int flags) {
    DeviceAwait a = fit(std::move(fits), false, nullptr, false, flags);
    P.req.kind_override = kind;
    return a;
  }
  // inlier list of a model of another residual kind at squared threshold sq
  DeviceAwait masklist_as(const double* model, int kind, double sq) {
    P.req = Request();
    P.req.type = REQ_MASKLIST;
    P.req.model = model;
    P.req.kind_override = kind;
    P.req.sq_override = sq;
    return DeviceAwait{&P};
  }

  // flags of the estimator's own non-minimal fit
  int nonminimal_flags(bool weighted) const {
    if (P.est == GCR_PROBLEM_HOMOGRAPHY) return 0;
    if (P.est == EST_PP) return 1 | 2 | ((weighted ? 1 : 0) << 8);
    return 1 | ((weighted ? 1 : 0) << 8);
  }

  // isValidModel of the estimator for a scored chunk model slot (rec: the
  // slot's record of a filtered chunk).
  bool valid_model(const Request& r, size_t slot, const SlotRec* rec = nullptr) const {
    if (P.est == GCR_PROBLEM_HOMOGRAPHY) return (rec ? rec->valid : r.valid[slot]) != 0;
    // fundamental_estimator.h:287-338 / essential_estimator.h:279-305:
    // symmetric epipolar agreement of at least max(m, ratio * inliers) of the
    // Sampson inliers (ratio 0.5 for F, 0.1 for E)
    const double need = std::max((double)P.m, (double)(rec ? rec->count : r.counts[slot]) * P.min_ratio);
    return (size_t)(rec ? rec->sym : r.sym[slot]) >= (size_t)need;
  }

  // inc/GCRANSAC.h:780-910.  Returns true when the best model improved.
  SubTask<bool> local_optimisation(Score& best, Container& cont) {
    const size_t inlier_limit = 7 * (size_t)P.m;
    Score max_score = best;
    P.set_slot(S_LOBEST, P.slot(S_BEST));
    int64_t best_inl_count = 0;
    bool have_lo_best = false;
    ++P.lo_count;  // counted twice per local optimisation (:485, :805)
    std::vector<int32_t> cur(inlier_limit);
    while (++P.graph_cut_number < 10) {
      bool updated = false;
      co_await label();
      if (P.label_failed)
        throw PairError(GCR_ERR_INTERNAL, "pair " + std::to_string(P.index) +
                                              ": the device min-cut did not converge within its round bound");
      std::vector<int32_t> inliers = std::move(P.req.list);
      const size_t sample_size = std::min(inlier_limit, inliers.size());
      std::vector<std::vector<int32_t>> fits;
      if (sample_size < inliers.size()) {
        for (size_t t = 0; t < P.lo_number; ++t) {
          P.lo_s->sample(inliers.data(), inliers.size(), cur.data(), sample_size);
          fits.emplace_back(cur.begin(), cur.begin() + sample_size);
        }
      } else if ((size_t)P.m < inliers.size() && P.lo_number > 0) {
        // every trial fits all inliers: identical fits, only the first can
        // improve the score
        fits.emplace_back(inliers);
      }
      if (!fits.empty()) {
        co_await fit(std::move(fits), false, nullptr, false, nonminimal_flags(false));
        Request& r = P.req;
        for (size_t k = 0; k < r.status.size(); ++k) {
          if (r.status[k] != 2) continue;
          const Score s{r.counts[k], r.values[k]};
          if (max_score < s) {
            updated = true;
            max_score = s;
            P.set_slot(S_LOBEST, r.dev_models + 9 * k);
            best_inl_count = s.count;
            have_lo_best = true;
          }
        }
      }
      if (!updated) break;
    }
    if (best < max_score && have_lo_best) {
      best = max_score;
      P.set_slot(S_BEST, P.slot(S_LOBEST));
      P.set_slot(cont.slot, P.slot(S_LOBEST));
      cont.count = best_inl_count;
      cont.pool_start = kListAscending;
      cont.has = true;
      co_return true;
    }
    co_return false;
  }

  // inc/GCRANSAC.h:630-758.  inliers is the container list (modified in
  // place as the reference does); slot S_IRLS holds the input/output model.
  SubTask<bool> irls(std::vector<int32_t>& inliers) {
    const size_t m = (size_t)P.m;
    if (inliers.size() <= m) co_return false;
    size_t iterations = 0;
    Score best_score;
    while (++iterations < 10) {
      std::vector<std::vector<int32_t>> fits;
      fits.push_back(inliers);
      co_await fit(std::move(fits), true, P.slot(S_IRLS), true, nonminimal_flags(true));
      Request& r = P.req;
      int nm = 0;
      for (size_t k = 0; k < r.status.size(); ++k) nm += r.status[k] == 2;
      if (nm == 0) break;
      if (nm == 1) {
        std::vector<int32_t> tmp = std::move(r.list);
        if (tmp.size() < m) break;
        if ((size_t)r.counts[0] <= inliers.size()) break;
        P.set_slot(S_IRLS, r.dev_models);
        inliers.swap(tmp);
      } else {
        // several models (essential matrices, :716-753): a model is taken
        // when it has more inliers than the current set; the reference swaps
        // the taken model's inliers in at once, so a later model must beat
        // the taken one strictly (a tie keeps the earlier model)
        bool updated = false;
        size_t cur_size = inliers.size();
        for (size_t k = 0; k < r.status.size(); ++k) {
          if (r.status[k] != 2) continue;
          if ((size_t)r.counts[k] < m) continue;
          if ((size_t)r.counts[k] <= cur_size) continue;
          if (r.counts[k] >= best_score.count) {
            updated = true;
            best_score = Score{r.counts[k], r.values[k]};
            cur_size = (size_t)r.counts[k];
            P.set_slot(S_IRLS, r.dev_models + 9 * k);
          }
        }
        if (!updated) break;
        co_await masklist(P.slot(S_IRLS));
        inliers = std::move(P.req.list);
      }
    }
    co_return iterations > 1;
  }

  // Result of one GCRANSAC::run (inc/GCRANSAC.h:202-627) of the pair's
  // current estimator; the model is in slot S_BEST.
  struct CoreResult {
    bool found = false;
    Score best;
    std::vector<int32_t> inliers;  // statistics.inliers
  };

  // inc/GCRANSAC.h:202-627 with the pair's current estimator, samplers and
  // slot base.
  SubTask<CoreResult> core() {
    const size_t m = (size_t)P.m;
    const size_t N = (size_t)P.n;
    const size_t S = (size_t)P.maxsol;
    size_t max_iteration = iteration_number(1, N, m, P.log_prob);
    Score best;
    bool have_best = false;
    Container cont[2] = {Container{S_CONT0}, Container{S_CONT1}};
    int offset = 0;
    // candidate stream (one request's results)
    Request cr;
    size_t pos = 0, avail = 0, consumed = 0;
    // DEGENSAC's nested run passes one sampler as the main and the local
    // optimisation sampler.  The main loop draws a chunk ahead, so before the
    // local optimisation draws, the sampler is put back to its state after
    // the candidates the loop has actually consumed (replayed from a copy
    // taken at the chunk's start).
    const bool shared = static_cast<Sampler*>(P.lo_s) == P.main_s;
    // the top-level run's uniform or PROSAC sampler draws on the device (its
    // random stream lives there; the host object only plans the draws)
    const bool dev_draws = P.dev_sampling && !shared && P.main_s == P.main_sampler.get();
    const bool timed = P.time_limit > 0.0 && P.main_s == P.main_sampler.get();
    UniformSampler chunk_start(0, 1);
    size_t chunk_used = 0;
    // samplers with state-changing updates: the state at the chunk's start,
    // to replay the chunk when the loop updates a sample a second time
    std::unique_ptr<Sampler> main_start;
    auto resync_shared = [&]() {
      if (!shared || pos >= avail) return;
      *P.lo_s = chunk_start;
      std::vector<int32_t> tmp(m);
      for (size_t c = 0; c < chunk_used; ++c) P.lo_s->sample(nullptr, N, tmp.data(), m);
      pos = avail;
    };
    // SPRT (preemption_sprt.h): the device walk has tested the chunk's model
    // slots before walk_end; it stopped at stop_slot (an accepted model above
    // the best; -1: none), whose count, value and pool start it returned.
    const bool sprt = P.sprt_on;
    int64_t walk_end = 0, stop_slot = -1, stop_count = 0, stop_start = 0;
    size_t rec_pos = 0;  // filtered chunks: the next record of interest
    double stop_value = 0.0;
    // exclusive end of the candidates the loop goes through from candidate
    // c0 on with the current iteration limit (each candidate is one
    // iteration; 100 failed draws end an iteration; for samplers with
    // updates such an iteration drops the rest of the chunk)
    auto loop_end = [&](size_t c0, size_t max_iteration_now) -> size_t {
      const size_t limit_it = std::min(max_iteration_now, P.max_iters);
      size_t it = P.iteration;
      size_t c = c0 + 1;
      while (P.min_iters > it || it < limit_it) {
        ++it;
        int u = -1;
        bool g = false;
        while (++u < 100) {
          if (c >= avail) return avail;
          if (cr.nmodels[c++] > 0) {
            g = true;
            break;
          }
        }
        it += (size_t)u;
        if (!g && P.main_s->has_updates()) return c;
      }
      return c;
    };

    while (P.min_iters > P.iteration || P.iteration < std::min(max_iteration, P.max_iters)) {
      bool do_lo = false;
      ++P.iteration;
      int unsuccessful = -1;
      bool got = false;
      size_t cand = 0;
      while (++unsuccessful < 100) {
        if (pos >= avail) {
          // next chunk: a multiple of what the loop has used so far, bounded
          // by the iterations it can still run
          const size_t limit = std::max(P.min_iters, std::min(max_iteration, P.max_iters));
          const size_t remaining = limit > P.iteration ? limit - P.iteration + 1 : 1;
          size_t want = std::max<size_t>(256, consumed);
          want = std::min(want, remaining + remaining / 8 + 32);
          want = std::min<size_t>(want, 16384);
          chunk_used = 0;
          std::vector<SampleSeg> segs;
          if (dev_draws && P.main_s->plan(want, N, m, segs)) {
            co_await chunk_dev((int64_t)want, best.value, std::move(segs));
          } else if (!shared && !P.main_s->has_updates()) {
            // the next chunk's likely size (the rule above, had this chunk
            // been used up), for the draws ahead
            const size_t rem_after = remaining > want ? remaining - want : 0;
            P.ahead_hint = rem_after == 0 ? 0
                                          : std::min<size_t>(std::min<size_t>(std::max<size_t>(256, consumed + want),
                                                                              rem_after + rem_after / 8 + 32),
                                                             16384);
            co_await chunk_host((int64_t)want, best.value);
          } else {
            std::vector<int32_t> samples(want * m);
            if (shared) chunk_start = *P.lo_s;
            if (P.main_s->has_updates()) main_start = P.main_s->clone();
            P.main_s->sample_chunk(N, samples.data(), want, m);
            co_await chunk(std::move(samples), (int64_t)want, best.value);
          }
          cr = std::move(P.req);
          pos = 0;
          avail = want;
          walk_end = 0;
          stop_slot = -1;
          rec_pos = 0;
        }
        cand = pos++;
        ++consumed;
        ++chunk_used;
        if (cr.nmodels[cand] > 0) {
          got = true;
          break;
        }
      }
      P.iteration += (size_t)unsuccessful;
      if (!got && P.main_s->has_updates() && main_start) {
        // all 100 draws failed: the reference updates the last sample once
        // more (inc/GCRANSAC.h:474-478), which changes every later draw.
        // Restore the chunk's start, replay the consumed draws, update the
        // last one again and drop the rest of the chunk.
        Sampler* live = P.main_s;
        std::unique_ptr<Sampler> replay = main_start->clone();
        std::vector<int32_t> tmp(m);
        for (size_t c = 0; c < chunk_used; ++c) replay->sample(nullptr, N, tmp.data(), m);
        replay->update_again(tmp.data(), m);
        if (live == P.main_sampler.get()) {
          P.main_sampler = std::move(replay);
          P.main_s = P.main_sampler.get();
        }
        pos = avail;
      }
      if (got) {
        for (int j = 0; j < cr.nmodels[cand]; ++j) {
          const size_t slot = cand * S + (size_t)j;
          Score s;
          int64_t sprt_start = -1;
          if (sprt) {
            if ((int64_t)slot >= walk_end) {
              // walk the device test over the models the loop meets from
              // here: to the end of this candidate when a local
              // optimisation follows it, else to the end of the run or chunk
              const size_t cend = do_lo ? cand + 1 : loop_end(cand, max_iteration);
              std::vector<int32_t> slots;
              for (size_t c = cand; c < cend; ++c)
                for (int jj = (c == cand ? j : 0); jj < cr.nmodels[c]; ++jj) slots.push_back((int32_t)(c * S + jj));
              const std::vector<int32_t> walked = slots;
              co_await sprt_walk(cr.dev_models, std::move(slots), best.value, best.count);
              const SprtItem it = P.req.sprt;
              P.sprt.pool_idx = it.pool_idx;
              P.sprt.delta = it.delta;
              P.sprt.A = it.A;
              if (it.last >= 0) {
                // the list the last tested model wrote (preemption_sprt.h:220)
                P.set_slot(cont[offset].slot, cr.dev_models + 9 * (size_t)walked[(size_t)it.last]);
                cont[offset].count = it.last_count;
                cont[offset].has = true;
                cont[offset].pool_start = P.sp_on ? kListWalked : it.last_start;
                cont[offset].pool_tested = it.last_tested;
              }
              if (it.stop >= 0) {
                stop_slot = walked[(size_t)it.stop];
                stop_count = it.last_count;
                stop_value = it.last_value;
                stop_start = it.last_start;
                walk_end = stop_slot + 1;
              } else {
                stop_slot = -1;
                walk_end = (int64_t)(cend * S);
              }
            }
            if ((int64_t)slot != stop_slot) continue;  // rejected, or accepted at most the best
            stop_slot = -1;
            // accepted above the best: the test's new history
            // (preemption_sprt.h:259-270)
            P.sprt.eps = (double)stop_count / (double)N;
            P.sprt.A = P.sprt.threshold_A(P.sprt.eps, P.sprt.delta);
            P.sprt.fill_table(N);
            s = Score{stop_count, stop_value};
            sprt_start = P.sp_on ? -1 : stop_start;
            // an accepted model was tested on all its cells' points: its list
            // can be rebuilt, unlike a walked list that later walks overwrite
            if (P.sp_on) cont[offset].pool_start = kListCellOrder;
          } else if (cr.filtered) {
            while (rec_pos < cr.recs.size() && cr.recs[rec_pos].slot < (int32_t)slot) ++rec_pos;
            const bool hit = rec_pos < cr.recs.size() && cr.recs[rec_pos].slot == (int32_t)slot;
            P.set_slot(cont[offset].slot, cr.dev_models + 9 * slot);
            cont[offset].count = hit ? cr.recs[rec_pos].count : (j + 1 == cr.nmodels[cand] ? cr.last_count[cand] : 0);
            cont[offset].has = true;
            cont[offset].pool_start = kListAscending;
            if (!hit) continue;  // at most the best at the chunk's start
            s = Score{cr.recs[rec_pos].count, cr.recs[rec_pos].value};
          } else {
            // inc/GCRANSAC.h:343-367: the selector's cells hold fewer points
            // than the best model's inliers
            if (P.sp_on && (int64_t)cr.sel[slot] < best.count) continue;
            s = Score{cr.counts[slot], cr.values[slot]};
            P.set_slot(cont[offset].slot, cr.dev_models + 9 * slot);
            cont[offset].count = s.count;
            cont[offset].has = true;
            cont[offset].pool_start = P.sp_on ? kListCellOrder : kListAscending;
          }
          if (best < s && valid_model(cr, slot, cr.filtered ? &cr.recs[rec_pos] : nullptr)) {
            Score acc = s;
            const double* acc_model = cr.dev_models + 9 * slot;
            const bool degen_slot = cr.filtered ? (P.degensac && cr.recs[rec_pos].degen != 0)
                                                : (P.degensac && !cr.degen.empty() && cr.degen[slot]);
            if (degen_slot) {
              // fundamental_estimator.h:466-572 for an H-degenerate sample
              const DegenResult dr = co_await degensac(acc_model, s.count, cr.dev_degen_h + 9 * slot,
                                                       cr.filtered ? cr.recs[rec_pos].sample0 : cr.samples[cand * m],
                                                       sprt_start);
              if (dr.rc == 0) continue;  // isValidModel false
              if (dr.rc == 2) {
                // the model was replaced and rescored into this container
                acc = dr.score;
                acc_model = P.slot(S_TMP);
                P.set_slot(cont[offset].slot, acc_model);
                cont[offset].count = acc.count;
                cont[offset].pool_start = kListAscending;
              }
            }
            offset = (offset + 1) % 2;
            P.set_slot(S_BEST, acc_model);
            best = acc;
            have_best = true;
            do_lo = P.iteration > 20 && (size_t)best.count > m;
            max_iteration = iteration_number((size_t)best.count, N, m, P.log_prob);
          }
        }
      }
      if (do_lo) {
        ++P.lo_count;
        resync_shared();
        co_await local_optimisation(best, cont[offset]);
        max_iteration = iteration_number((size_t)best.count, N, m, P.log_prob);
      }
      // GCRANSAC.h:503-516: a desired FPS ends the loop after the iteration
      // that exceeds the time limit (the top-level run's; the nested
      // DEGENSAC run has none, as the reference's nested object)
      if (timed && std::chrono::duration<double>(std::chrono::steady_clock::now() - P.t_start).count() >
                       P.time_limit) {
        P.min_iters = 0;
        max_iteration = 0;
        break;
      }
    }

    CoreResult res;
    res.best = best;
    if ((size_t)best.count <= m || !have_best) co_return res;  // model not found (:521-527)
    resync_shared();
    if (P.lo_count == 0) {
      ++P.lo_count;
      co_await local_optimisation(best, cont[offset]);
    }
    if (cont[offset].count != best.count) offset = (offset + 1) % 2;
    if (cont[offset].count != best.count) {
      P.set_slot(cont[offset].slot, P.slot(S_BEST));
      cont[offset].count = best.count;
      cont[offset].pool_start = kListAscending;
    }
    const int64_t order = cont[offset].pool_start;
    if (order == kListWalked) co_await masklist_walked(P.slot(cont[offset].slot), cont[offset].count);
    else if (order == kListCellOrder) co_await masklist_cells(P.slot(cont[offset].slot));
    else co_await masklist(P.slot(cont[offset].slot));
    std::vector<int32_t> cont_list = std::move(P.req.list);
    if (order >= 0) cont_list = pool_order(cont_list, order, cont[offset].pool_tested);

    bool applied = false;
    P.set_slot(S_IRLS, P.slot(S_BEST));
    const bool success = co_await irls(cont_list);
    if (success) {
      P.set_slot(cont[offset].slot, P.slot(S_IRLS));
      cont[offset].count = (int64_t)cont_list.size();
      co_await masklist(P.slot(S_IRLS));
      const int idx = (offset + 1) % 2;
      const Score cur{(int64_t)P.req.list.size(), P.req.values.empty() ? 0.0 : P.req.values[0]};
      P.set_slot(cont[idx].slot, P.slot(S_IRLS));
      cont[idx].count = cur.count;
      if (best < cur) {
        applied = true;
        P.set_slot(S_BEST, P.slot(S_IRLS));
        offset = idx;
        cont_list = std::move(P.req.list);
      }
    }
    if (!applied) {
      // one least-squares fit on the first best.count container entries
      std::vector<int32_t> sample(cont_list.begin(),
                                  cont_list.begin() + std::min<size_t>((size_t)best.count, cont_list.size()));
      std::vector<std::vector<int32_t>> fits;
      fits.push_back(std::move(sample));
      co_await fit(std::move(fits), false, nullptr, false, nonminimal_flags(false));
      Request r = std::move(P.req);
      for (size_t k = 0; k < r.status.size(); ++k) {
        if (r.status[k] != 2) continue;
        const Score cur{r.counts[k], r.values[k]};
        if (best < cur) {
          P.set_slot(S_BEST, r.dev_models + 9 * k);
          co_await masklist(P.slot(S_BEST));
          cont_list = std::move(P.req.list);
        }
      }
    }
    res.found = true;
    res.best = best;
    res.inliers = std::move(cont_list);
    co_return res;
  }

  // State of the loop that DEGENSAC's nested run replaces and restores.
  struct LoopState {
    int est, m, maxsol, slot_base;
    double threshold, tau, sq_thr, lo_thr, lambda, conf, log_prob, min_ratio;
    size_t min_iters, max_iters, lo_number, iteration, graph_cut_number, lo_count;
    bool degensac, sprt_on;
    Sampler* main_s;
    UniformSampler* lo_s;
  };
  LoopState save_loop() const {
    return LoopState{P.est,       P.m,         P.maxsol,   P.slot_base,         P.threshold, P.tau,
                     P.sq_thr,    P.lo_thr,    P.lambda,   P.conf,              P.log_prob,  P.min_ratio,
                     P.min_iters, P.max_iters, P.lo_number, P.iteration,        P.graph_cut_number,
                     P.lo_count,  P.degensac,  P.sprt_on,  P.main_s,            P.lo_s};
  }
  void restore_loop(const LoopState& L) {
    P.est = L.est;
    P.m = L.m;
    P.maxsol = L.maxsol;
    P.slot_base = L.slot_base;
    P.threshold = L.threshold;
    P.tau = L.tau;
    P.sq_thr = L.sq_thr;
    P.lo_thr = L.lo_thr;
    P.lambda = L.lambda;
    P.conf = L.conf;
    P.log_prob = L.log_prob;
    P.min_ratio = L.min_ratio;
    P.min_iters = L.min_iters;
    P.max_iters = L.max_iters;
    P.lo_number = L.lo_number;
    P.iteration = L.iteration;
    P.graph_cut_number = L.graph_cut_number;
    P.lo_count = L.lo_count;
    P.degensac = L.degensac;
    P.sprt_on = L.sprt_on;
    P.main_s = L.main_s;
    P.lo_s = L.lo_s;
    P.params_dirty = true;
  }

  struct DegenResult {
    int rc = 0;  // 0: model rejected, 1: kept, 2: replaced (slot S_TMP, score)
    Score score;
  };

  // fundamental_estimator.h:466-572: the model's MSAC inliers that the
  // sample's homography also explains (squared transfer error < 4), that
  // homography refitted to them by least squares, and a nested GC-RANSAC
  // with the plane-and-parallax estimator (threshold = this loop's truncated
  // threshold, spatial weight 0, confidence 0.99, the other settings at their
  // defaults, one uniform sampler for both roles seeded from the run's seed
  // and the sample's first index); its model replaces the input when it has
  // more inliers.
  // pool_start >= 0: the model's inlier list is an SPRT list (pool order
  // from pool_start, all points tested), so the homography's inliers are
  // taken in that order (fundamental_estimator.h:510-518 walks the list).
  SubTask<DegenResult> degensac(const double* model, int64_t model_inliers, const double* degen_h,
                                int32_t sample0, int64_t pool_start = -1) {
    DegenResult out;
    co_await masklist(model);
    std::vector<int32_t> f_in = std::move(P.req.list);
    if (pool_start >= 0) f_in = pool_order(f_in, pool_start, P.n);
    co_await masklist_as(degen_h, GCR_PROBLEM_HOMOGRAPHY, 4.0);
    const std::vector<int32_t> h_in = std::move(P.req.list);
    std::vector<int32_t> hin;
    {
      std::vector<uint8_t> in_h((size_t)P.n, 0);
      for (int32_t idx : h_in) in_h[(size_t)idx] = 1;
      for (int32_t idx : f_in)
        if (in_h[(size_t)idx]) hin.push_back(idx);
    }
    static const bool trace = std::getenv("GCR_TRACE_DEGENSAC") != nullptr;
    if (trace)
      std::fprintf(stderr, "degensac: pair %d sample0 %d model inliers %lld, H inliers %zu of %zu\n", P.index,
                   (int)sample0, (long long)model_inliers, hin.size(), f_in.size());
    if (hin.size() < 4) co_return out;
    {
      std::vector<std::vector<int32_t>> fits;
      fits.push_back(std::move(hin));
      co_await fit_as(std::move(fits), GCR_PROBLEM_HOMOGRAPHY, 0);
    }
    {
      Request r = std::move(P.req);
      if (r.status.empty() || r.status[0] != 2) co_return out;
      P.set_abs_slot(S_PPH_ABS, r.dev_models);
    }
    const LoopState outer = save_loop();
    P.est = EST_PP;
    P.m = est_sample_size(EST_PP);
    P.maxsol = est_solutions(EST_PP);
    P.slot_base = NSLOT;
    P.threshold = outer.tau;
    P.tau = 3.0 / 2.0 * P.threshold;
    P.sq_thr = P.tau * P.tau;
    P.lo_thr = P.threshold * P.threshold * 9 / 4;
    P.lambda = 0.0;
    P.conf = 0.99;
    P.log_prob = std::log(1.0 - P.conf);
    P.min_ratio = 0.0;
    P.min_iters = 20;
    P.max_iters = std::numeric_limits<size_t>::max();
    P.lo_number = 10;
    P.iteration = 0;
    P.graph_cut_number = 0;
    P.lo_count = 0;
    P.degensac = false;
    P.sprt_on = false;  // the nested run verifies every model in full
    const uint32_t nseed = (uint32_t)((P.spec.params.seed * 2654435761ull) ^ (uint64_t)sample0);
    P.nested_sampler.reset(new UniformSampler(nseed, (size_t)P.n));
    P.main_s = P.nested_sampler.get();
    P.lo_s = P.nested_sampler.get();
    P.params_dirty = true;
    CoreResult nr = co_await core();
    const double* nested_best = P.slot(S_BEST);
    if (trace)
      std::fprintf(stderr, "degensac: nested run %s, %zu inliers, %zu iterations\n", nr.found ? "found" : "none",
                   nr.inliers.size(), P.iteration);
    restore_loop(outer);
    out.rc = 1;
    if (nr.found && (int64_t)nr.inliers.size() > model_inliers) {
      P.set_slot(S_TMP, nested_best);
      co_await masklist(P.slot(S_TMP));
      out.rc = 2;
      out.score = Score{(int64_t)P.req.list.size(), P.req.values.empty() ? 0.0 : P.req.values[0]};
    }
    co_return out;
  }

  // inc/GCRANSAC.h:202-627 followed by the output stage of the driver.
  RootTask run() {
    CoreResult cres = co_await core();
    PairResult& R = *P.out;
    Score best = cres.best;
    if (!cres.found) {
      // model not found (:521-527)
      R.inliers.clear();
      if (P.spec.kind != GCR_PROBLEM_HOMOGRAPHY && P.spec.driver_stage) set_identity(R.model);
      R.status = 0;
      finish_stats(best);
      co_return;
    }
    std::vector<int32_t> cont_list = std::move(cres.inliers);
    double final_score = best.value;
    if (P.spec.kind == GCR_PROBLEM_FUNDAMENTAL && P.spec.driver_stage) {
      // py/gcransac_python.cpp:1832-1887: weighted bundle adjustment on the
      // inliers in pixel coordinates, kept when its MSAC value is higher
      if (cont_list.size() > 7) {
        std::vector<std::vector<int32_t>> fits;
        fits.push_back(cont_list);
        co_await fit(std::move(fits), false, P.slot(S_BEST), false, 2 << 8);
        Request r = std::move(P.req);
        for (size_t k = 0; k < r.status.size(); ++k) {
          if (r.status[k] != 2) continue;
          if (final_score < r.values[k]) {
            final_score = r.values[k];
            P.set_slot(S_BEST, r.dev_models + 9 * k);
            co_await masklist(P.slot(S_BEST));
            cont_list = std::move(P.req.list);
          }
        }
      }
    }
    bool identity = P.spec.driver_stage && P.spec.kind == GCR_PROBLEM_FUNDAMENTAL && cont_list.size() <= 7;
    if (P.spec.kind == GCR_PROBLEM_ESSENTIAL && P.spec.driver_stage) {
      // py/gcransac_python.cpp:2766-2823: pose bundle adjustment of the
      // inliers (weights from the model by inlier position), every refined
      // pose scored at the adjustment's truncated threshold; kept when its
      // MSAC value is higher
      if (cont_list.size() > 5) {
        std::vector<std::vector<int32_t>> fits;
        fits.push_back(cont_list);
        co_await fit(std::move(fits), false, P.slot(S_BEST), false, 3 << 8);
        Request r = std::move(P.req);
        for (size_t k = 0; k < r.status.size(); ++k) {
          if (r.status[k] != 2) continue;
          if (final_score < r.values[k]) {
            final_score = r.values[k];
            P.set_slot(S_BEST, r.dev_models + 9 * k);
            co_await masklist(P.slot(S_BEST), true);
            cont_list = std::move(P.req.list);
          }
        }
      } else {
        identity = true;
      }
    }
    co_await masklist(P.slot(S_BEST));
    std::memcpy(R.model, P.req.model_host, sizeof(R.model));
    if (identity) set_identity(R.model);
    R.inliers = std::move(cont_list);
    R.status = (int)R.inliers.size();
    finish_stats(best);
    R.stats.score = final_score;
  }

  static void set_identity(double* m) {
    for (int i = 0; i < 9; ++i) m[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }

  void finish_stats(const Score& best) {
    PairResult& R = *P.out;
    R.stats.iteration_number = (int64_t)P.iteration;
    R.stats.local_optimization_number = (int64_t)P.lo_count;
    R.stats.graph_cut_number = (int64_t)P.graph_cut_number;
    R.stats.num_inliers = (int64_t)R.inliers.size();
    R.stats.score = best.value;
    R.stats.candidates_solved = P.candidates_solved;
    R.stats.models_scored = P.models_scored;
  }
};

// ---------------------------------------------------------------------------
// Driver: device state and batched execution of requests
// ---------------------------------------------------------------------------
// Every device and pinned host array of an engine run.  Arrays only grow, so
// a workspace left by one call is handed to the next call on the same device
// (no allocation in steady state); see workspace_take / workspace_put.
// A HIP stream kept with a cached workspace, so that calls reuse it instead
// of creating and destroying a stream each time.
struct OwnedStream {
  hipStream_t s = nullptr;
  OwnedStream() = default;
  OwnedStream(const OwnedStream&) = delete;
  OwnedStream& operator=(const OwnedStream&) = delete;
  OwnedStream(OwnedStream&& o) noexcept : s(o.s) { o.s = nullptr; }
  OwnedStream& operator=(OwnedStream&& o) noexcept {
    if (this != &o) {
      if (s) (void)hipStreamDestroy(s);
      s = o.s;
      o.s = nullptr;
    }
    return *this;
  }
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct Workspace {
  OwnedStream ws_stream;  // the engine's stream (created on first use)
  ModelPool model_pool;
  // device arrays
  DevBuf<double> d_pts_own;
  const double* d_pts = nullptr;   // correspondences as given (pixels): the grid is built on these
  const double* d_rpts = nullptr;  // the points RANSAC runs on (K-normalised for essential pairs)
  DevBuf<double> d_rpts_own;
  DevBuf<int32_t> d_pt_pair;
  DevBuf<int64_t> d_pair_pt_off;
  DevBuf<int32_t> d_pair_n;
  DevBuf<double> d_pair_sq_thr, d_pair_lo_thr, d_pair_lambda, d_pair_cell, d_pair_post_sq, d_kinv;
  DevBuf<int8_t> d_norm_pair;
  DevBuf<int64_t> d_pair_div;
  DevBuf<int32_t> d_pair_qbits;
  DevBuf<double> d_slots;
  // neighbourhood
  DevBuf<uint64_t> d_keys, d_skeys;
  DevBuf<int32_t> d_vals, d_svals, d_run_start, d_run_len, d_pos_of;
  DevBuf<int64_t> d_out_deg, d_edge_off;
  DevBuf<int32_t> d_cell_pair, d_cell_start, d_cell_len, d_num_cells;
  DevBuf<char> d_tmp;
  DevBuf<int64_t> d_rf, d_rb, d_node_src, d_node_sink;
  // radius neighbourhood (pairs with neighborhood 1): per-pair flag and
  // squared float radius; per point the out / in neighbour counts, the
  // sorted out-neighbour lists and the incident-arc lists; min-cut state
  DevBuf<int8_t> d_pair_radius;
  DevBuf<float> d_pair_r2;
  DevBuf<int64_t> d_out_cnt, d_in_cnt, d_r_off, d_adj_off, d_adj_arc;
  DevBuf<int32_t> d_r_nbr, d_r_nbr_s, d_adj_other, d_in_cursor, d_tiles, d_rpairs;
  DevBuf<uint32_t> d_r_key, d_r_key_s;
  DevBuf<int32_t> d_h, d_hn, d_queue;  // node heights of the global-memory min-cuts, BFS frontiers
  // labelling
  DevBuf<double> d_dist;
  DevBuf<int8_t> d_near;
  DevBuf<uint8_t> d_sink;
  // chunk
  // per estimator (a step's chunks of several estimators are in flight
  // together): a chunk's samples and pairs; its per-slot results; the
  // scoring kernel's active-slot list, its counter and block offsets
  DevBuf<char> d_cand_in[4], d_cand_res[4], d_cand_work[4];
  DevBuf<double> d_pair_inv_scale;     // matrix-core scoring: per-pair power-of-two inverse scale
  DevBuf<MxPoint> d_mxp;               // matrix-core scoring: per-point operands and error bounds
  bool mx_on = false;                  // this run scores candidate chunks on the matrix cores
  int mx_max_n = 0;                    // largest pair of the run (candidate-mask rows)
  DevBuf<uint32_t> d_mx_mask;          // matrix-core scoring: candidate masks of one list segment
  bool dev_sampling_on = false;        // this run draws uniform samples on the device
  DevBuf<DevMT> d_mt;                  // per pair: the uniform sampler's generator state
  DevBuf<uint32_t> d_mt_seed;
  DevBuf<int32_t> d_mt_pair;
  // fits
  DevBuf<double> d_fit_scratch, d_weights, d_stage;
  // pinned staging of the candidate chunks, per estimator (the chunks of
  // the estimators of one step are in flight together)
  HostBuf<char> h_cand_in[4], h_cand_res[4];
  // pinned inlier lists read back by the labelling, fit and mask-list phases
  HostBuf<int32_t> h_lists_lab, h_lists_fit, h_lists_ml;
  DevBuf<char> d_lab_up;  // a labelling step's inputs and zeroed outputs (one upload)
  HostBuf<int32_t> h_lab_out;
  DevBuf<const double*> d_pair_h;
  std::vector<const double*> pair_h_host;  // staging of d_pair_h (alive until the step's sync)
  DevBuf<uint8_t> d_masks;
  DevBuf<int32_t> d_lists;
  // a fit step's index arrays (one upload) and results (one download;
  // per F fit it includes the LM point passes, Jacobian and cost, when
  // profiling)
  DevBuf<char> d_fit_up, d_fit_res;
  HostBuf<char> h_res;
  // pinned upload staging of the labelling, fit and mask-list phases
  HostBuf<char> h_pack_lab, h_pack_fit, h_pack_ml;
  // masklist: one upload, one result block; d_ml_models stages the
  // labelling's models
  DevBuf<double> d_ml_models;
  DevBuf<char> d_ml_up, d_ml_res;
  HostBuf<char> h_ml_res;
  // copies
  DevBuf<char> d_rc_up;  // the pending slot copies' source / destination tables
  HostBuf<char> h_rc_stage;
  // host staging
  HostBuf<char> h_stage;
  // SPRT: per point the pool of its pair (pair-relative indices), per pair
  // an A table of n + 1 entries; the walk requests' upload / download
  DevBuf<int32_t> d_sprt_pool;
  DevBuf<double> d_sprt_tables;
  // filtered candidate results (device scratch, pinned records), per
  // estimator: the records are read back after the step's other phases ran
  DevBuf<char> d_filt[4];
  HostBuf<SlotRec> h_filt_recs[4];
  // space partitioning: per pair its cells (container order), per point of
  // a pair its cell-order permutation and cell rank, the walked lists
  DevBuf<int8_t> d_pair_sp;
  DevBuf<SpPair> d_sp_pairs;
  DevBuf<SpCell> d_sp_cells;
  DevBuf<int32_t> d_sp_perm, d_sp_cell_of, d_sp_list;
  int sp_max_cells = 0;
  DevBuf<int8_t> d_ml_sp;
  std::vector<int8_t> ml_sp_flags;
  std::vector<int8_t> sp_stage_flags;
  std::vector<SpPair> sp_stage_pairs;
  std::vector<SpCell> sp_stage_cells;
  std::vector<int32_t> sp_stage_perm, sp_stage_cell_of;
  DevBuf<char> d_sprt_up;
  HostBuf<char> h_sprt_stage;

};

std::mutex g_ws_mu;
std::map<int, std::vector<std::unique_ptr<Workspace>>> g_ws_cache;

std::unique_ptr<Workspace> workspace_take(int dev) {
  std::lock_guard<std::mutex> g(g_ws_mu);
  auto& v = g_ws_cache[dev];
  if (v.empty()) return nullptr;
  std::unique_ptr<Workspace> w = std::move(v.back());
  v.pop_back();
  return w;
}

void workspace_put(int dev, std::unique_ptr<Workspace> w) {
  std::lock_guard<std::mutex> g(g_ws_mu);
  g_ws_cache[dev].push_back(std::move(w));
}

struct Engine : Workspace {
  int device = 0;
  hipStream_t stream = nullptr;
  // the step's phases parked at their synchronisation point
  std::vector<std::coroutine_handle<>> parked;
  StepSync step_sync() { return StepSync{&parked}; }
  HostPool* pool = &HostPool::get();  // host threads for the pairs' control flow
  Profiler prof;
  std::vector<PairCtx> pairs;
  int64_t total = 0;
  int max_cell = 1;
  int num_cells = 0;
  std::vector<int32_t> radius_pairs;  // pairs labelled on their radius graph
  int64_t radius_edges = 0;           // edges of all radius graphs

  // Radius neighbour lists (flann_neighborhood_graph.h:100-139): counts per
  // point, out-lists filled in ascending index and stably sorted by distance,
  // the arc offsets; the radius out-degrees join the grid's (zero for radius
  // pairs) so that edge_off numbers every pair's edges.
  void build_radius_lists(const std::vector<int32_t>& nn) {
    d_out_cnt.ensure(total);
    d_in_cnt.ensure(total);
    check_hip(hipMemsetAsync(d_out_cnt.p, 0, (size_t)total * sizeof(int64_t), stream), "memset counts");
    check_hip(hipMemsetAsync(d_in_cnt.p, 0, (size_t)total * sizeof(int64_t), stream), "memset counts");
    std::vector<int32_t> tiles;  // (pair, first query) per workgroup
    for (int32_t i : radius_pairs)
      for (int32_t f = 0; f < nn[(size_t)i]; f += 256) {
        tiles.push_back(i);
        tiles.push_back(f);
      }
    const int nt = (int)(tiles.size() / 2);
    std::vector<int32_t> tp((size_t)nt), tf((size_t)nt);
    for (int t = 0; t < nt; ++t) {
      tp[(size_t)t] = tiles[2 * (size_t)t];
      tf[(size_t)t] = tiles[2 * (size_t)t + 1];
    }
    d_tiles.ensure(2 * (size_t)std::max(1, nt));
    h2d(d_tiles.p, tp.data(), (size_t)nt);
    h2d(d_tiles.p + nt, tf.data(), (size_t)nt);
    launch_radius_search(stream, false, d_pts, d_tiles.p, d_tiles.p + nt, nt, d_pair_pt_off.p, d_pair_n.p,
                         d_pair_r2.p, d_out_cnt.p, d_in_cnt.p, nullptr, nullptr, nullptr);
    // out-degrees into the engine's edge numbering; in_cnt becomes the degree
    launch_add_i64(stream, d_out_deg.p, d_out_cnt.p, total);
    launch_add_i64(stream, d_in_cnt.p, d_out_cnt.p, total);
    d_r_off.ensure(total + 1);
    d_adj_off.ensure(total + 1);
    check_hip(hipMemsetAsync(d_r_off.p, 0, sizeof(int64_t), stream), "memset r_off");
    check_hip(hipMemsetAsync(d_adj_off.p, 0, sizeof(int64_t), stream), "memset adj_off");
    size_t b1 = 0, b2 = 0;
    check_hip(hipcub::DeviceScan::InclusiveSum(nullptr, b1, d_out_cnt.p, d_r_off.p + 1, (int)total, stream),
              "scan (size)");
    check_hip(hipcub::DeviceScan::InclusiveSum(nullptr, b2, d_in_cnt.p, d_adj_off.p + 1, (int)total, stream),
              "scan (size)");
    d_tmp.ensure(std::max(b1, b2));
    check_hip(hipcub::DeviceScan::InclusiveSum(d_tmp.p, b1, d_out_cnt.p, d_r_off.p + 1, (int)total, stream), "scan");
    check_hip(hipcub::DeviceScan::InclusiveSum(d_tmp.p, b2, d_in_cnt.p, d_adj_off.p + 1, (int)total, stream), "scan");
    d2h(&radius_edges, d_r_off.p + total, 1);
    sync();
    if (radius_edges > (int64_t)std::numeric_limits<int32_t>::max() / 2)
      throw EngineError(GCR_ERR_OUT_OF_MEMORY, "radius neighbourhood of " + std::to_string(radius_edges) +
                                                   " edges exceeds the list sort's 2^30 entries");
    const size_t ne = (size_t)std::max<int64_t>(1, radius_edges);
    d_r_nbr.ensure(ne);
    d_r_nbr_s.ensure(ne);
    d_r_key.ensure(ne);
    d_r_key_s.ensure(ne);
    launch_radius_search(stream, true, d_pts, d_tiles.p, d_tiles.p + nt, nt, d_pair_pt_off.p, d_pair_n.p,
                         d_pair_r2.p, nullptr, nullptr, d_r_off.p, d_r_nbr.p, d_r_key.p);
    if (radius_edges > 0) {
      size_t sb = 0;
      check_hip(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, sb, d_r_key.p, d_r_key_s.p, d_r_nbr.p,
                                                            d_r_nbr_s.p, (int)radius_edges, (int)total, d_r_off.p,
                                                            d_r_off.p + 1, 0, 32, stream),
                "segmented sort (size)");
      d_tmp.ensure(sb);
      check_hip(hipcub::DeviceSegmentedRadixSort::SortPairs(d_tmp.p, sb, d_r_key.p, d_r_key_s.p, d_r_nbr.p,
                                                            d_r_nbr_s.p, (int)radius_edges, (int)total, d_r_off.p,
                                                            d_r_off.p + 1, 0, 32, stream),
                "segmented sort (radius lists)");
    }
  }


  Engine() {
    check_hip(hipGetDevice(&device), "hipGetDevice");
    if (std::unique_ptr<Workspace> w = workspace_take(device)) static_cast<Workspace&>(*this) = std::move(*w);
  }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  ~Engine() {
    // requests holding model-pool arrays go first, then the device work
    // drains before the arrays are handed to the next call
    for (auto& P : pairs)
      if (P.root.h) {
        P.root.h.destroy();
        P.root.h = nullptr;
      }
    pairs.clear();
    if (stream) (void)hipStreamSynchronize(stream);  // the stream stays with the workspace
    stream = nullptr;
    d_pts = nullptr;
    d_rpts = nullptr;
    workspace_put(device, std::make_unique<Workspace>(std::move(static_cast<Workspace&>(*this))));
  }

  template <class T>
  void h2d(T* dst, const T* src, size_t n) {
    if (n == 0) return;
    check_hip(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
  }
  template <class T>
  void d2h(T* dst, const T* src, size_t n) {
    if (n == 0) return;
    check_hip(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
  }
  void sync() {
    if (g_prof_on) {
      // blocked time of the step's synchronisations ("host_syncwait"): the
      // part of the host phases spent waiting for the device
      const auto a = std::chrono::steady_clock::now();
      check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
      std::lock_guard<std::mutex> g(g_prof_mu);
      ProfEntry& e = g_prof["host_syncwait"];
      e.ms += ms;
      e.launches += 1;
    } else {
      check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    lab_busy = fit_busy = ml_busy = rc_busy = false;
    prof.collect();
  }

  // Host arrays of one step uploaded with a single copy: add() lays each
  // array out (256-byte aligned) in the pinned staging block h_pack, and
  // upload() copies the block to `dev` and sets every array's device
  // address.  The staging block is refilled only after the copy that read it
  // has completed (a stream synchronisation in between).
  struct Packed {
    struct Part {
      const void* src;
      size_t bytes, off;
      std::function<void(char*)> set;
    };
    std::vector<Part> parts;
    size_t bytes = 0;
    template <class T>
    void add(T*& out, const T* src, size_t n) {
      const size_t b = n * sizeof(T);
      parts.push_back(Part{src, b, bytes, [&out](char* at) { out = reinterpret_cast<T*>(at); }});
      bytes += (b + 255) & ~(size_t)255;
    }
  };
  bool lab_busy = false, fit_busy = false, ml_busy = false, rc_busy = false;
  void upload(Packed& pk, DevBuf<char>& dev, HostBuf<char>& stage, bool& busy) {
    if (busy) check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const size_t total = std::max<size_t>(pk.bytes, 256);
    char* const d = dev.ensure(total);
    char* const h = stage.ensure(total);
    for (auto& part : pk.parts) {
      if (part.bytes) std::memcpy(h + part.off, part.src, part.bytes);
      part.set(d + part.off);
    }
    if (pk.bytes) {
      check_hip(hipMemcpyAsync(d, h, pk.bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
      busy = true;
    }
  }

  // One transfer of the spans [off[i], off[i] + len[i]) of d_lists into the
  // pinned h_lists (valid after the next sync); returns the first copied
  // offset, to subtract from off[i] when reading h_lists.
  int64_t d2h_list_spans(HostBuf<int32_t>& h_lists, const std::vector<int64_t>& off,
                         const std::vector<int32_t>& len) {
    int64_t lo = std::numeric_limits<int64_t>::max(), hi = 0;
    for (size_t i = 0; i < off.size(); ++i)
      if (len[i] > 0) {
        lo = std::min(lo, off[i]);
        hi = std::max(hi, off[i] + (int64_t)len[i]);
      }
    if (hi <= lo) return 0;
    h_lists.ensure((size_t)(hi - lo));
    d2h(h_lists.p, d_lists.p + lo, (size_t)(hi - lo));
    return lo;
  }

  // ---- set-up: pair tables and the grid neighbourhood -------------------
  void setup(const std::vector<PairSpec>& specs, const double* corr, bool on_device) {
    if (!ws_stream.s) check_hip(hipStreamCreateWithFlags(&ws_stream.s, hipStreamNonBlocking), "hipStreamCreate");
    stream = ws_stream.s;
    prof.begin(stream);
    dev_sampling_on = use_device_sampling();
    const int np = (int)specs.size();
    pairs.resize(np);
    total = 0;
    for (int i = 0; i < np; ++i) total += specs[i].n;
    if (on_device) {
      d_pts = corr;  // pairs are contiguous (checked by the caller)
    } else {
      d_pts_own.ensure((size_t)total * 4);
      d_pts = d_pts_own.p;
    }
    std::vector<int64_t> off(np);
    std::vector<int32_t> nn(np), qb(np);
    std::vector<double> sq(np), lo(np), lam(np), cell(4 * (size_t)np);
    std::vector<int64_t> dv(np);
    std::vector<int32_t> pt_pair((size_t)total);
    std::vector<double> kinv(18 * (size_t)np, 0.0), post(np, 0.0);
    std::vector<int8_t> normp(np, 0), rad(np, 0);
    std::vector<float> r2(np, 0.0f);
    bool any_norm = false;
    int64_t acc = 0;
    for (int i = 0; i < np; ++i) {
      const PairSpec& s = specs[i];
      PairCtx& P = pairs[i];
      P.index = i;
      P.spec = s;
      P.m = sample_size_of(s.kind);
      P.maxsol = s.kind == GCR_PROBLEM_HOMOGRAPHY ? 1 : (s.kind == GCR_PROBLEM_FUNDAMENTAL ? 3 : 10);
      P.n = s.n;
      P.threshold = s.params.threshold;
      P.min_ratio = 0.5;
      if (s.kind == GCR_PROBLEM_ESSENTIAL && !s.driver_stage) {
        P.min_ratio = 0.1;  // essential_estimator.h:82; points and threshold as given
      } else if (s.kind == GCR_PROBLEM_ESSENTIAL) {
        // py/gcransac_python.cpp:2626-2628,2723: RANSAC on K-normalised
        // points with threshold / (mean focal length); the post-hoc
        // adjustment's truncated threshold is 3 / 2 * threshold / normaliser
        const double normaliser = 0.25 * (s.K1[0] + s.K1[4] + s.K2[0] + s.K2[4]);
        P.threshold = s.params.threshold / normaliser;
        const double trunc = 3.0 / 2.0 * s.params.threshold / normaliser;
        P.post_sq = trunc * trunc;
        P.min_ratio = 0.1;  // essential_estimator.h:82
        mat3_inverse_host(s.K1, kinv.data() + 18 * (size_t)i);
        mat3_inverse_host(s.K2, kinv.data() + 18 * (size_t)i + 9);
        normp[(size_t)i] = 1;
        any_norm = true;
      }
      P.conf = s.params.confidence;
      P.lambda = s.params.spatial_coherence_weight;
      P.tau = 3.0 / 2.0 * P.threshold;
      P.sq_thr = P.tau * P.tau;
      if (s.kind == GCR_PROBLEM_FUNDAMENTAL) P.post_sq = P.sq_thr;
      P.lo_thr = P.threshold * P.threshold * 9 / 4;
      P.log_prob = std::log(1.0 - P.conf);
      P.min_iters = (size_t)std::max<int64_t>(0, s.params.min_iters);
      P.max_iters = (size_t)std::max<int64_t>(0, s.params.max_iters);
      P.lo_number = (size_t)std::max<int32_t>(0, s.params.lo_number);
      P.time_limit = s.params.time_limit > 0.0 ? s.params.time_limit : 0.0;
      if (s.params.sample_fn) {
        // the caller's host sampler: its draws stay on the host
        P.main_sampler.reset(new CallbackSampler(s.params.sample_fn, s.params.sample_ctx));
      } else if (s.params.sampler == 0) {
        P.main_sampler.reset(new UniformSampler(main_seed(s.params.seed), (size_t)s.n));
        P.dev_sampling = dev_sampling_on;
        P.main_seed_v = main_seed(s.params.seed);
      } else if (s.params.sampler == 1) {
        P.main_sampler.reset(new ProsacSampler(main_seed(s.params.seed), (size_t)s.n, (size_t)P.m));
        P.dev_sampling = dev_sampling_on && use_device_prosac();
        P.main_seed_v = main_seed(s.params.seed);
      } else if (s.params.sampler == 3) {
        P.main_sampler.reset(new ImportanceSampler(main_seed(s.params.seed), s.probs, (size_t)s.n));
      } else {
        // py/gcransac_python.cpp:4349-4358: probabilities divided by their maximum
        std::vector<double> pr(s.probs, s.probs + s.n);
        double mx = 0;
        for (double v : pr) mx = mx < v ? v : mx;
        for (double& v : pr) v /= mx;
        P.main_sampler.reset(new AdaptiveReorderingSampler(main_seed(s.params.seed), pr.data(), (size_t)s.n,
                                                           s.params.sampler_variance));
      }
      P.lo_sampler.reset(new UniformSampler(lo_seed(s.params.seed), (size_t)s.n));
      P.main_s = P.main_sampler.get();
      P.lo_s = P.lo_sampler.get();
      // findHomography_'s space partitioning selector (py/gcransac_python.cpp:
      // 4282-4286, 4396-4423) on the grid, which the driver always builds
      // (the template API's selector slot: the same selector on the caller's grid)
      P.sp_on = s.kind == GCR_PROBLEM_HOMOGRAPHY && s.params.use_space_partitioning != 0;
      P.sprt_on = s.params.use_sprt != 0;
      if (P.sprt_on) {
        // preemption_sprt.h:141-190.  The pool shuffle draws from the
        // rand() stream seeded with the main sampler's seed (the adaptive
        // reordering sampler's own stream, before its first draw).  t_M is
        // the reference's (x / m) / x, i.e. 1 / m; findHomography_ builds
        // the test with the class default ratio 0.01
        // (py/gcransac_python.cpp:4390-4393); the template API passes the
        // ratio its SPRT object was built with.
        GlibcRandom local(main_seed(s.params.seed));
        GlibcRandom* g = &local;
        if (!s.params.sample_fn && s.params.sampler == 4)
          g = &static_cast<AdaptiveReorderingSampler*>(P.main_sampler.get())->stream();
        const size_t n = (size_t)s.n;
        P.sprt.pool.resize(n);
        for (size_t k = 0; k < n; ++k) P.sprt.pool[k] = (int32_t)k;
        int max = (int)n;
        for (size_t k = 0; k < n; ++k) {
          const size_t r = (size_t)(g->next() % max);
          const int32_t t = P.sprt.pool[r];
          --max;
          P.sprt.pool[r] = P.sprt.pool[(size_t)max];
          P.sprt.pool[(size_t)max] = t;
        }
        const double ratio = s.kind == GCR_PROBLEM_HOMOGRAPHY && !s.cell_override
                                 ? 0.01
                                 : s.params.min_inlier_ratio_for_sprt;
        P.sprt.t_M = 1.0 / (double)P.m;
        P.sprt.m_S = (double)P.maxsol;
        P.sprt.pool_idx = 0;
        P.sprt.delta = ratio / 10.0;
        P.sprt.eps = ratio;
        P.sprt.A = P.sprt.threshold_A(P.sprt.eps, P.sprt.delta);
        P.sprt.fill_table(n);
      }
      P.est = s.kind;
      // the fundamental-matrix estimator of the drivers and of the template
      // API applies DEGENSAC (fundamental_estimator.h:91-93, use_degensac)
      P.degensac = s.kind == GCR_PROBLEM_FUNDAMENTAL && s.degensac;
      off[i] = acc;
      nn[i] = (int32_t)s.n;
      sq[i] = P.sq_thr;
      post[i] = P.post_sq;
      lo[i] = P.lo_thr;
      lam[i] = P.lambda;
      const double divs = (double)(size_t)s.params.neighborhood_size;
      const bool grid = P.lambda > std::numeric_limits<double>::epsilon();
      // the radius neighbourhood: built by the F / E drivers when the spatial
      // weight is above eps (py/gcransac_python.cpp:1658-1681, 2583-2604), by
      // findHomography_ always unless space partitioning switches it to the
      // grid (:4282-4301)
      P.radius = s.params.neighborhood == 1 && !s.cell_override &&
                 (s.kind == GCR_PROBLEM_HOMOGRAPHY ? !s.params.use_space_partitioning : grid);
      rad[(size_t)i] = P.radius ? 1 : 0;
      {
        const float r = (float)s.params.neighborhood_size;
        r2[(size_t)i] = r * r;
      }
      if (s.cell_override) {
        for (int d = 0; d < 4; ++d) cell[4 * i + d] = s.cell[d];
      } else if (s.kind == GCR_PROBLEM_HOMOGRAPHY) {
        const double eps = std::numeric_limits<double>::epsilon();
        cell[4 * i + 0] = (s.w1 + eps) / divs;
        cell[4 * i + 1] = (s.h1 + eps) / divs;
        cell[4 * i + 2] = (s.w2 + eps) / divs;
        cell[4 * i + 3] = (s.h2 + eps) / divs;
      } else {
        cell[4 * i + 0] = s.w1 / divs;
        cell[4 * i + 1] = s.h1 / divs;
        cell[4 * i + 2] = s.w2 / divs;
        cell[4 * i + 3] = s.h2 / divs;
      }
      dv[i] = grid && !P.radius ? (int64_t)(size_t)s.params.neighborhood_size : 0;
      for (int64_t k = 0; k < s.n; ++k) pt_pair[(size_t)(acc + k)] = i;
      acc += s.n;
    }
    bool any_grid_src = false;
    for (const auto& sp : specs) any_grid_src |= sp.grid_corr != nullptr;
    if (any_grid_src && on_device) throw EngineError(GCR_ERR_INVALID_ARGUMENT, "separate grid points need host input");
    if (!on_device) {
      // the grid source (d_pts): the pair's grid points when given
      int64_t o = 0;
      for (int i = 0; i < np; ++i) {
        const double* src = specs[i].grid_corr ? specs[i].grid_corr : corr + 4 * specs[i].offset;
        h2d(d_pts_own.p + 4 * o, src, (size_t)specs[i].n * 4);
        o += specs[i].n;
      }
    }
    d_pt_pair.ensure(total);
    d_pair_pt_off.ensure(np);
    d_pair_n.ensure(np);
    d_pair_sq_thr.ensure(np);
    d_pair_lo_thr.ensure(np);
    d_pair_lambda.ensure(np);
    d_pair_cell.ensure(4 * (size_t)np);
    d_pair_div.ensure(np);
    d_pair_qbits.ensure(np);
    h2d(d_pt_pair.p, pt_pair.data(), (size_t)total);
    h2d(d_pair_pt_off.p, off.data(), np);
    h2d(d_pair_n.p, nn.data(), np);
    h2d(d_pair_sq_thr.p, sq.data(), np);
    h2d(d_pair_lo_thr.p, lo.data(), np);
    h2d(d_pair_lambda.p, lam.data(), np);
    h2d(d_pair_cell.p, cell.data(), 4 * (size_t)np);
    h2d(d_pair_div.p, dv.data(), np);
    d_pair_radius.ensure(np);
    d_pair_r2.ensure(np);
    h2d(d_pair_radius.p, rad.data(), np);
    h2d(d_pair_r2.p, r2.data(), np);
    radius_pairs.clear();
    for (int i = 0; i < np; ++i)
      if (rad[(size_t)i]) radius_pairs.push_back(i);
    d_pair_post_sq.ensure(np);
    h2d(d_pair_post_sq.p, post.data(), np);
    // space partitioning: the grid cells in the order the reference's grid
    // container (an unordered_map from cell key, filled in point order)
    // iterates them, each cell's points and axis indices
    {
      std::vector<int8_t> spf((size_t)np, 0);
      std::vector<SpPair> spp((size_t)np);
      std::vector<SpCell> cells;
      std::vector<int32_t> perm, cell_of;
      bool any = false;
      sp_max_cells = 0;
      for (int i = 0; i < np; ++i) {
        PairCtx& P = pairs[(size_t)i];
        if (!P.sp_on) continue;
        any = true;
        spf[(size_t)i] = 1;
        const PairSpec& sp = specs[(size_t)i];
        const int64_t n = sp.n;
        std::vector<double> host_pts;
        const double* src;
        if (on_device) {
          host_pts.resize(4 * (size_t)n);
          d2h(host_pts.data(), d_pts + 4 * off[(size_t)i], 4 * (size_t)n);
          sync();
          src = host_pts.data();
        } else {
          src = sp.grid_corr ? sp.grid_corr : corr + 4 * sp.offset;
        }
        const double* cs = cell.data() + 4 * (size_t)i;
        const uint64_t div = (uint64_t)(size_t)sp.params.neighborhood_size;
        std::unordered_map<size_t, std::pair<std::vector<int32_t>, std::array<uint64_t, 4>>> grid;
        for (int64_t r = 0; r < n; ++r) {
          std::array<uint64_t, 4> idx;
          uint64_t key = 0, mul = 1;
          for (int d = 0; d < 4; ++d) {
            idx[(size_t)d] = (uint64_t)(int64_t)std::floor(src[4 * r + d] / cs[d]);
            key += mul * idx[(size_t)d];
            mul *= div;
          }
          auto& c = grid[(size_t)key];
          if (c.first.empty()) c.second = idx;
          c.first.push_back((int32_t)r);
        }
        SpPair& pp = spp[(size_t)i];
        pp.cell_off = (int64_t)cells.size();
        pp.ncells = (int32_t)grid.size();
        for (int d = 0; d < 4; ++d) pp.cs[d] = cs[d];
        pp.P = (double)div;
        pp.tau = P.tau;
        perm.resize((size_t)(off[(size_t)i] + n));
        cell_of.resize((size_t)(off[(size_t)i] + n));
        int32_t at = 0, rank = 0;
        for (const auto& kv : grid) {
          SpCell c;
          for (int d = 0; d < 4; ++d) c.idx[d] = kv.second.second[(size_t)d];
          c.start = at;
          c.len = (int32_t)kv.second.first.size();
          for (int32_t q : kv.second.first) {
            perm[(size_t)(off[(size_t)i] + at)] = q;
            cell_of[(size_t)(off[(size_t)i] + at)] = rank;
            ++at;
          }
          cells.push_back(c);
          ++rank;
        }
        sp_max_cells = std::max(sp_max_cells, (int)grid.size());
      }
      // the host arrays stay alive until the setup's copies complete
      sp_stage_flags = std::move(spf);
      sp_stage_pairs = std::move(spp);
      sp_stage_cells = std::move(cells);
      sp_stage_perm = std::move(perm);
      sp_stage_cell_of = std::move(cell_of);
      d_pair_sp.ensure((size_t)np);
      h2d(d_pair_sp.p, sp_stage_flags.data(), (size_t)np);
      if (any) {
        d_sp_pairs.ensure((size_t)np);
        d_sp_cells.ensure(sp_stage_cells.size());
        d_sp_perm.ensure(total);
        d_sp_cell_of.ensure(total);
        d_sp_list.ensure(total);
        h2d(d_sp_pairs.p, sp_stage_pairs.data(), (size_t)np);
        h2d(d_sp_cells.p, sp_stage_cells.data(), sp_stage_cells.size());
        h2d(d_sp_perm.p, sp_stage_perm.data(), sp_stage_perm.size());
        h2d(d_sp_cell_of.p, sp_stage_cell_of.data(), sp_stage_cell_of.size());
      }
    }
    // SPRT: the pairs' pools and A tables
    {
      int64_t ntab = 0;
      for (auto& P : pairs)
        if (P.sprt_on) ntab += P.n + 1;
      if (ntab > 0) {
        d_sprt_pool.ensure(total);
        d_sprt_tables.ensure((size_t)ntab);
        int64_t t = 0;
        for (int i = 0; i < np; ++i) {
          PairCtx& P = pairs[(size_t)i];
          if (!P.sprt_on) continue;
          h2d(d_sprt_pool.p + off[(size_t)i], P.sprt.pool.data(), P.sprt.pool.size());
          P.sprt.a_dev = d_sprt_tables.p + t;
          P.sprt.table_dirty = true;
          t += P.n + 1;
        }
      }
    }
    // RANSAC points: K^-1-normalised rows for essential pairs (utils.h:65-110)
    if (any_grid_src) {
      // RANSAC runs on the correspondences themselves
      d_rpts_own.ensure((size_t)total * 4);
      int64_t o = 0;
      for (int i = 0; i < np; ++i) {
        h2d(d_rpts_own.p + 4 * o, corr + 4 * specs[i].offset, (size_t)specs[i].n * 4);
        o += specs[i].n;
      }
      d_rpts = d_rpts_own.p;
    } else if (any_norm) {
      d_rpts_own.ensure((size_t)total * 4);
      d_kinv.ensure(18 * (size_t)np);
      d_norm_pair.ensure(np);
      h2d(d_kinv.p, kinv.data(), 18 * (size_t)np);
      h2d(d_norm_pair.p, normp.data(), np);
      launch_prepare_points(stream, d_pts, d_pt_pair.p, d_norm_pair.p, d_kinv.p, total, d_rpts_own.p);
      d_rpts = d_rpts_own.p;
    } else {
      d_rpts = d_pts;
    }
    // matrix-core scoring operands of the RANSAC points
    mx_on = use_mx_scoring();
    if (mx_on) {
      mx_max_n = 0;
      for (int i = 0; i < np; ++i) mx_max_n = std::max(mx_max_n, (int)nn[(size_t)i]);
      d_mx_mask.ensure(mx_mask_words(mx_max_n));
      d_pair_inv_scale.ensure(np);
      d_mxp.ensure((size_t)total);
      launch_mx_prepare(stream, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pt_pair.p, np, total, d_pair_inv_scale.p,
                        d_mxp.p);
    }
    // generator states of the pairs whose uniform sampler draws on the device
    {
      std::vector<uint32_t> seeds;
      std::vector<int32_t> which;
      for (int i = 0; i < np; ++i)
        if (pairs[i].dev_sampling) {
          seeds.push_back(pairs[i].main_seed_v);
          which.push_back(i);
        }
      if (!which.empty()) {
        d_mt.ensure((size_t)np);
        d_mt_seed.ensure(seeds.size());
        d_mt_pair.ensure(which.size());
        h2d(d_mt_seed.p, seeds.data(), seeds.size());
        h2d(d_mt_pair.p, which.data(), which.size());
        launch_mt_seed(stream, d_mt_seed.p, d_mt_pair.p, (int)which.size(), d_mt.p);
        sync();  // the host vectors above are the copies' sources
      }
    }
    // the masks and inlier lists of the labelling, fit and mask-list phases
    // (at most one list per pair each): sized once, so that no phase
    // reallocates them while an earlier phase of its step is in flight
    d_masks.ensure((size_t)total);
    d_lists.ensure((size_t)total);
    d_slots.ensure((size_t)np * NSLOT_ALL * 9);
    check_hip(hipMemsetAsync(d_slots.p, 0, (size_t)np * NSLOT_ALL * 9 * sizeof(double), stream), "memset slots");
    for (int i = 0; i < np; ++i) pairs[i].slots = d_slots.p + (size_t)i * NSLOT_ALL * 9;

    // grid keys, per-pair stable sort, runs, edge offsets, multi-point cells
    d_keys.ensure(total);
    d_vals.ensure(total);
    d_skeys.ensure(total);
    d_svals.ensure(total);
    launch_grid_keys(stream, d_pts, d_pt_pair.p, d_pair_pt_off.p, d_pair_cell.p, d_pair_div.p, total, d_keys.p,
                     d_vals.p);
    {
      std::vector<int64_t> seg(np + 1);
      for (int i = 0; i < np; ++i) seg[i] = off[i];
      seg[np] = total;
      DevBuf<int64_t> d_seg;
      d_seg.ensure(np + 1);
      h2d(d_seg.p, seg.data(), np + 1);
      size_t tmp_bytes = 0;
      check_hip(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp_bytes, d_keys.p, d_skeys.p, d_vals.p,
                                                            d_svals.p, (int)total, np, d_seg.p, d_seg.p + 1, 0, 64,
                                                            stream),
                "segmented sort (size)");
      d_tmp.ensure(tmp_bytes);
      check_hip(hipcub::DeviceSegmentedRadixSort::SortPairs(d_tmp.p, tmp_bytes, d_keys.p, d_skeys.p, d_vals.p,
                                                            d_svals.p, (int)total, np, d_seg.p, d_seg.p + 1, 0, 64,
                                                            stream),
                "segmented sort");
      d_run_start.ensure(total);
      d_run_len.ensure(total);
      d_pos_of.ensure(total);
      d_out_deg.ensure(total);
      d_edge_off.ensure(total + 1);
      launch_grid_runs(stream, d_skeys.p, d_svals.p, d_pt_pair.p, d_pair_pt_off.p, d_pair_n.p, total, d_run_start.p,
                       d_run_len.p, d_pos_of.p, d_out_deg.p);
      if (!radius_pairs.empty()) build_radius_lists(nn);
      check_hip(hipMemsetAsync(d_edge_off.p, 0, sizeof(int64_t), stream), "memset edge_off");
      size_t scan_bytes = 0;
      check_hip(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, d_out_deg.p, d_edge_off.p + 1, (int)total, stream),
                "scan (size)");
      DevBuf<char> scan_tmp;
      scan_tmp.ensure(scan_bytes);
      check_hip(hipcub::DeviceScan::InclusiveSum(scan_tmp.p, scan_bytes, d_out_deg.p, d_edge_off.p + 1, (int)total,
                                                 stream),
                "scan");
      d_cell_pair.ensure(total);
      d_cell_start.ensure(total);
      d_cell_len.ensure(total);
      d_num_cells.ensure(1);
      check_hip(hipMemsetAsync(d_num_cells.p, 0, sizeof(int32_t), stream), "memset cells");
      launch_grid_cells(stream, d_run_start.p, d_run_len.p, d_pt_pair.p, d_pair_pt_off.p, total, d_cell_pair.p,
                        d_cell_start.p, d_cell_len.p, d_num_cells.p);
      std::vector<int64_t> eoff((size_t)total + 1);
      int32_t ncell = 0;
      d2h(eoff.data(), d_edge_off.p, (size_t)total + 1);
      d2h(&ncell, d_num_cells.p, 1);
      sync();
      num_cells = ncell;
      std::vector<int32_t> clen(std::max(1, ncell));
      d2h(clen.data(), d_cell_len.p, (size_t)ncell);
      sync();
      max_cell = 2;
      for (int c = 0; c < ncell; ++c) max_cell = std::max(max_cell, clen[c]);
      for (int i = 0; i < np; ++i) {
        const int64_t e = eoff[(size_t)(off[i] + nn[i])] - eoff[(size_t)off[i]];
        qb[i] = mincut_bits((size_t)nn[i], (size_t)e);
        pairs[i].qbits = qb[i];
        // flann_neighborhood_graph.h:131-138: no edge, graph not initialised;
        // the drivers return 0 inliers (py/gcransac_python.cpp:1683-1691,
        // 2608-2619, 4309-4317)
        if (pairs[i].radius && e == 0) pairs[i].no_graph = true;
      }
      if (!radius_pairs.empty()) {
        d_adj_arc.ensure((size_t)std::max<int64_t>(1, 2 * radius_edges));
        d_adj_other.ensure((size_t)std::max<int64_t>(1, 2 * radius_edges));
        d_in_cursor.ensure(total);
        check_hip(hipMemsetAsync(d_in_cursor.p, 0, (size_t)total * sizeof(int32_t), stream), "memset cursor");
        launch_radius_adjacency(stream, d_pt_pair.p, d_pair_radius.p, d_pair_pt_off.p, d_r_off.p, d_r_nbr_s.p,
                                d_edge_off.p, d_adj_off.p, d_out_cnt.p, d_in_cursor.p, total, d_adj_arc.p,
                                d_adj_other.p);
        d_rpairs.ensure(radius_pairs.size());
        h2d(d_rpairs.p, radius_pairs.data(), radius_pairs.size());
        d_queue.ensure(2 * (size_t)total);
      }
      if (!radius_pairs.empty() || max_cell > kMaxLdsCell) {
        d_h.ensure(total);
        d_hn.ensure(total);
      }
      h2d(d_pair_qbits.p, qb.data(), np);
      const int64_t etot = eoff[(size_t)total];
      d_rf.ensure((size_t)etot);
      d_rb.ensure((size_t)etot);
      d_node_src.ensure(total);
      d_node_sink.ensure(total);
      d_dist.ensure(total);
      d_near.ensure(total);
      d_sink.ensure(total);
      d_weights.ensure(total);
      check_hip(hipMemsetAsync(d_weights.p, 0, (size_t)total * sizeof(double), stream), "memset weights");
      sync();
    }
  }

  // ---- request execution -------------------------------------------------
  // source / destination tables of the pending slot copies (reused)
  std::vector<const double*> copy_src_h;
  std::vector<double*> copy_dst_h;

  void run_copies() {
    std::vector<const double*> src;
    std::vector<double*> dst;
    for (auto& P : pairs)
      for (int k = 0; k < NSLOT_ALL; ++k)
        if (P.pending[(size_t)k]) {
          src.push_back(P.pending[(size_t)k]);
          dst.push_back(P.abs_slot(k));
        }
    for (auto& P : pairs) P.pending.fill(nullptr);
    // thresholds / spatial weight of a pair whose loop switched estimator
    // (DEGENSAC's nested run); the sources are the pairs' own fields, which
    // stay unchanged until the step has synchronised
    for (auto& P : pairs)
      if (P.params_dirty) {
        P.params_dirty = false;
        h2d(d_pair_sq_thr.p + P.index, &P.sq_thr, 1);
        h2d(d_pair_lo_thr.p + P.index, &P.lo_thr, 1);
        h2d(d_pair_lambda.p + P.index, &P.lambda, 1);
      }
    if (src.empty()) return;
    // two-phase: stage sources first so a slot that is both a source and a
    // destination in this batch is read before it is written
    const int c = (int)src.size();
    DevBuf<double>& stage = d_stage;
    stage.ensure((size_t)c * 9);
    copy_src_h.resize(2 * (size_t)c);
    copy_dst_h.resize(2 * (size_t)c);
    for (int i = 0; i < c; ++i) {
      copy_src_h[(size_t)i] = src[(size_t)i];
      copy_dst_h[(size_t)i] = stage.p + 9 * (size_t)i;
      copy_src_h[(size_t)(c + i)] = stage.p + 9 * (size_t)i;
      copy_dst_h[(size_t)(c + i)] = dst[(size_t)i];
    }
    Packed up;
    const double** u_src = nullptr;
    double** u_dst = nullptr;
    up.add(u_src, copy_src_h.data(), 2 * (size_t)c);
    up.add(u_dst, copy_dst_h.data(), 2 * (size_t)c);
    upload(up, d_rc_up, h_rc_stage, rc_busy);  // own staging: no wait for the step's later uploads
    launch_copy_models(stream, u_src, u_dst, c);
    launch_copy_models(stream, u_src + c, u_dst + c, c);
  }

  // Minimal solves and scores of the candidate samples of every pair in the
  // step; one launch pair per problem kind.
  // The filtered read-back of a chunk launch (see Request::filtered):
  // models per candidate, each candidate's last model's count and the
  // records of the slots above each item's best, packed in item order.
  PhaseTask exec_chunk_filtered(std::vector<PairCtx*>& items, int kind, int m, int S, int64_t tot,
                                const std::vector<int64_t>& base, const int8_t* d_status, int8_t* d_nmodels,
                                const int32_t* d_counts, const double* d_values, const int32_t* d_sym,
                                const int8_t* d_valid, const int8_t* d_degen, const int32_t* d_samples,
                                const int64_t* d_ibase, const int32_t* d_incand, const double* d_ibest,
                                const std::shared_ptr<DevBuf<double>>& cand_models, double* d_cand_models,
                                double* d_cand_degen_h) {
    const size_t ni = items.size();
    const size_t ns = (size_t)tot * S;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // device: last counts | record counts | packed records | per-item records
    const size_t f_rc = al(4 * (size_t)tot), f_pk = f_rc + al(4 * ni), f_rs = f_pk + al(sizeof(SlotRec) * ns),
                 f_end = f_rs + sizeof(SlotRec) * ns;
    char* const df = d_filt[kind].ensure(f_end);
    int32_t* const d_last = reinterpret_cast<int32_t*>(df);
    int32_t* const d_rc = reinterpret_cast<int32_t*>(df + f_rc);
    SlotRec* const d_packed = reinterpret_cast<SlotRec*>(df + f_pk);
    SlotRec* const d_recs = reinterpret_cast<SlotRec*>(df + f_rs);
    launch_chunk_filter(stream, kind == GCR_PROBLEM_HOMOGRAPHY ? 1 : 0, S, (int)ni, d_ibase, d_incand, d_ibest,
                        d_status, d_nmodels, d_counts, d_values, d_sym, d_valid, d_degen, d_samples, m, d_last, d_recs,
                        d_rc, d_packed);
    // host: models per candidate | the block's head (last counts, record
    // counts and the first `spec` packed records, one copy).  Few slots beat
    // a pair's best, so the records usually come back with the counts; more
    // than `spec` of them take a second copy after the counts are known.
    const size_t spec = std::min(ns, 16 * ni + 256);
    const size_t h_blk = al((size_t)tot), head = f_pk + sizeof(SlotRec) * spec;
    char* const hs = h_cand_res[kind].ensure(h_blk + head);
    d2h(reinterpret_cast<int8_t*>(hs), d_nmodels, (size_t)tot);
    d2h(hs + h_blk, df, head);
    co_await step_sync();
    const int8_t* nm = reinterpret_cast<const int8_t*>(hs);
    const int32_t* lc = reinterpret_cast<const int32_t*>(hs + h_blk);
    const int32_t* rc = reinterpret_cast<const int32_t*>(hs + h_blk + f_rc);
    std::vector<int64_t> roff(ni + 1, 0);
    for (size_t i = 0; i < ni; ++i) roff[i + 1] = roff[i] + rc[i];
    const SlotRec* hr = reinterpret_cast<const SlotRec*>(hs + h_blk + f_pk);
    if ((size_t)roff[ni] > spec) {
      SlotRec* const all = h_filt_recs[kind].ensure((size_t)roff[ni]);
      d2h(all, d_packed, (size_t)roff[ni]);
      co_await step_sync();
      hr = all;
    }
    pool->run(ni, [&](size_t i) {
      Request& r = items[i]->req;
      const size_t b = (size_t)base[i], k = (size_t)r.num_cand;
      r.slots_per = S;
      r.filtered = true;
      r.nmodels.assign(nm + b, nm + b + k);
      r.last_count.assign(lc + b, lc + b + k);
      r.recs.assign(hr + roff[i], hr + roff[i + 1]);
      r.dev_models = d_cand_models + 9 * b * S;
      r.models_hold = cand_models;
      if (items[i]->degensac) r.dev_degen_h = d_cand_degen_h + 9 * b * S;
      items[i]->candidates_solved += (int64_t)k;
      int64_t kept = 0;
      for (size_t c = 0; c < k; ++c) kept += nm[b + c];
      items[i]->models_scored += kept;
    });
    if (g_prof_on) {
      double hp = 0.0;
      for (size_t i = 0; i < ni; ++i) {
        int64_t kept = 0;
        for (size_t c = 0; c < (size_t)items[i]->req.num_cand; ++c) kept += nm[(size_t)base[i] + c];
        hp += (double)kept * (double)items[i]->n;
      }
      prof_add_work(kind == GCR_PROBLEM_HOMOGRAPHY ? "score_h" : (kind == GCR_PROBLEM_ESSENTIAL ? "score_e" : "score_f"),
                    hp);
    }
  }

  // The candidate chunks of the pairs whose current estimator is `kind`.
  PhaseTask exec_chunk(std::vector<PairCtx*> items, int kind) {
    const int m = est_sample_size(kind);
    const int S = est_solutions(kind);
    bool degen_any = false;
    for (auto* p : items) degen_any |= p->degensac;
    int64_t tot = 0;
    std::vector<int64_t> base(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
      base[i] = tot;
      tot += items[i]->req.num_cand;
    }
    // samples and their pairs go up in one copy from one pinned block
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    // loops that score every model in full get their results filtered on
    // the device (all pairs of the launch, or none)
    bool filt = true;
    for (auto* p : items) filt &= !p->sprt_on && !p->sp_on;
    const size_t ni = items.size();
    // items whose samples the device draws (the rest bring host samples)
    std::vector<DevSampleItem> dsi;
    std::vector<DevSampleSeg> dss;
    for (size_t i = 0; i < ni; ++i)
      if (items[i]->req.dev_sample) {
        DevSampleItem d{};
        d.pair = items[i]->index;
        d.m = m;
        d.base = base[i];
        d.count = items[i]->req.num_cand;
        d.seg_off = (int32_t)dss.size();
        int64_t st0 = 0;
        for (const SampleSeg& g : items[i]->req.sample_segs) {
          dss.push_back(DevSampleSeg{st0, g.count, g.range, g.draws, g.fixed, 0});
          st0 += g.count;
        }
        d.nseg = (int32_t)(dss.size() - (size_t)d.seg_off);
        dsi.push_back(d);
      }
    const size_t in_pair = al(4 * (size_t)tot * m), in_ib = in_pair + al(4 * (size_t)tot),
                 in_in = in_ib + al(8 * ni), in_bv = in_in + al(4 * ni), in_ip = in_bv + al(8 * ni),
                 in_ds = in_ip + al(4 * ni),
                 in_sg = in_ds + al(sizeof(DevSampleItem) * dsi.size()),
                 in_end = in_sg + sizeof(DevSampleSeg) * dss.size();
    char* const hin = h_cand_in[kind].ensure(in_end);
    int32_t* samples = reinterpret_cast<int32_t*>(hin);
    int64_t* ibase = reinterpret_cast<int64_t*>(hin + in_ib);
    int32_t* incand = reinterpret_cast<int32_t*>(hin + in_in);
    double* ibest = reinterpret_cast<double*>(hin + in_bv);
    int32_t* ipair = reinterpret_cast<int32_t*>(hin + in_ip);
    if (!dsi.empty()) {
      std::memcpy(hin + in_ds, dsi.data(), sizeof(DevSampleItem) * dsi.size());
      std::memcpy(hin + in_sg, dss.data(), sizeof(DevSampleSeg) * dss.size());
    }
    for (size_t i = 0; i < ni; ++i) {
      ibase[i] = base[i];
      incand[i] = (int32_t)items[i]->req.num_cand;
      ibest[i] = items[i]->req.filter_best;
      ipair[i] = items[i]->index;
    }
    pool->run(items.size(), [&](size_t i) {
      Request& r = items[i]->req;
      PairCtx& Q = *items[i];
      if (r.host_draw) {
        int32_t* dst = samples + m * base[i];
        size_t need = (size_t)r.num_cand;
        if (Q.main_s == Q.main_sampler.get() && Q.ahead_m == (size_t)m && !Q.ahead.empty()) {
          // the samples drawn ahead come first (the same stream)
          const size_t have = Q.ahead.size() / (size_t)m - Q.ahead_pos, take = std::min(have, need);
          std::memcpy(dst, Q.ahead.data() + Q.ahead_pos * (size_t)m, sizeof(int32_t) * (size_t)m * take);
          Q.ahead_pos += take;
          dst += (size_t)m * take;
          need -= take;
          if (Q.ahead_pos * (size_t)m == Q.ahead.size()) {
            Q.ahead.clear();
            Q.ahead_pos = 0;
          }
        }
        if (need) Q.main_s->sample_chunk((size_t)Q.n, dst, need, (size_t)m);
      } else if (!r.dev_sample)
        std::memcpy(samples + m * base[i], r.samples.data(), sizeof(int32_t) * m * r.num_cand);
    });
    const size_t ns = (size_t)tot * S;
    char* const din = d_cand_in[kind].ensure(in_end);
    int32_t* const d_samples = reinterpret_cast<int32_t*>(din);
    int32_t* const d_cpair = reinterpret_cast<int32_t*>(din + in_pair);
    const int64_t* const d_ibase = reinterpret_cast<const int64_t*>(din + in_ib);
    const int32_t* const d_incand = reinterpret_cast<const int32_t*>(din + in_in);
    const double* const d_ibest = reinterpret_cast<const double*>(din + in_bv);
    // model slots, then (DEGENSAC pairs) the slots' degenerate homographies
    std::shared_ptr<DevBuf<double>> cand_models = model_pool.get(ns * 9 * (degen_any ? 2 : 1));
    double* const d_cand_models = cand_models->p;
    double* const d_cand_degen_h = d_cand_models + 9 * ns;
    // per model slot results in one block, read back with one copy:
    // status | valid | models per candidate | degenerate | count | sym | value
    const size_t r_vd = al(ns), r_nm = r_vd + al(ns), r_dg = r_nm + al((size_t)tot), r_cn = r_dg + al(ns),
                 r_sy = r_cn + al(4 * ns), r_vl = r_sy + al(4 * ns), r_sl = r_vl + al(8 * ns),
                 r_end = r_sl + 4 * ns;
    char* const dres = d_cand_res[kind].ensure(r_end);
    int8_t* const d_status = reinterpret_cast<int8_t*>(dres);
    int8_t* const d_valid = reinterpret_cast<int8_t*>(dres + r_vd);
    int8_t* const d_nmodels = reinterpret_cast<int8_t*>(dres + r_nm);
    int8_t* const d_degen = reinterpret_cast<int8_t*>(dres + r_dg);
    int32_t* const d_counts = reinterpret_cast<int32_t*>(dres + r_cn);
    int32_t* const d_sym = reinterpret_cast<int32_t*>(dres + r_sy);
    double* const d_values = reinterpret_cast<double*>(dres + r_vl);
    int32_t* const d_sel = reinterpret_cast<int32_t*>(dres + r_sl);
    bool sp_any = false;
    for (auto* p : items) sp_any |= p->sp_on;
    const bool mx = mx_on;
    char* const dwork = d_cand_work[kind].ensure(al(4 * ns) + 256 + al(4 * mx_list_scratch_ints((int64_t)ns)));
    int32_t* const d_work_list = reinterpret_cast<int32_t*>(dwork);
    uint32_t* const d_work_count = reinterpret_cast<uint32_t*>(dwork + al(4 * ns));
    int32_t* const d_blk = reinterpret_cast<int32_t*>(dwork + al(4 * ns) + 256);
    check_hip(hipMemcpyAsync(din, hin, in_end, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    // each candidate's pair, from the items' first candidates and counts
    launch_fill_cand_pair(stream, (int)ni, d_ibase, d_incand, reinterpret_cast<const int32_t*>(din + in_ip),
                          d_cpair);
    if (!dsi.empty()) {
      const int ts = prof.start("sample", (double)tot);
      launch_sample_streams(stream, reinterpret_cast<const DevSampleItem*>(din + in_ds),
                            reinterpret_cast<const DevSampleSeg*>(din + in_sg), (int)dsi.size(), d_mt.p, d_samples);
      prof.stop(ts);
    }
    if (kind == GCR_PROBLEM_HOMOGRAPHY) {
      const int t0 = prof.start("solve_h4", (double)tot);
      launch_h4_solve(stream, d_rpts, d_pair_pt_off.p, d_cpair, d_samples, tot, d_cand_models,
                      d_status, d_valid);
      prof.stop(t0);
      const int t1 = prof.start("score_h", 0.0);
      if (mx)
        launch_score_mx(stream, 0, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pair_sq_thr.p, d_pair_inv_scale.p,
                        d_mxp.p, d_cand_models, d_cpair, d_status, tot, d_counts, d_values, nullptr, 1,
                        d_work_list, d_work_count, d_blk, mx_max_n, d_mx_mask.p);
      else
        launch_score(stream, 0, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pair_sq_thr.p, d_cand_models, d_cpair,
                     d_status, tot, d_counts, d_values, nullptr, nullptr, nullptr, true, 1, d_work_list,
                     d_work_count);
      prof.stop(t1);
      if (sp_any) {
        // space-partitioning pairs: the score over the selector's cells
        const int t2 = prof.start("score_sp", 0.0);
        launch_score_sp(stream, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pair_sq_thr.p, d_pair_sp.p, d_sp_pairs.p,
                        d_sp_cells.p, d_sp_perm.p, d_sp_cell_of.p, d_cand_models, d_cpair, d_status, tot, d_counts,
                        d_values, d_sel, sp_max_cells);
        prof.stop(t2);
      }
    } else {
      if (kind == GCR_PROBLEM_FUNDAMENTAL) {
        const int t0 = prof.start("solve_f7", (double)tot);
        launch_f7_solve(stream, d_rpts, d_pair_pt_off.p, d_cpair, d_samples, tot, d_cand_models,
                        d_status, d_nmodels);
        prof.stop(t0);
      } else if (kind == EST_PP) {
        // each pair's homography for the plane-and-parallax solver
        pair_h_host.assign(pairs.size(), nullptr);
        for (auto* p : items) pair_h_host[(size_t)p->index] = p->abs_slot(S_PPH_ABS);
        d_pair_h.ensure(pairs.size());
        h2d(d_pair_h.p, pair_h_host.data(), pairs.size());
        const int t0 = prof.start("solve_pp", (double)tot);
        launch_pp_solve(stream, d_rpts, d_pair_pt_off.p, d_cpair, d_samples, d_pair_h.p, tot,
                        d_cand_models, d_status, d_nmodels);
        prof.stop(t0);
      } else {
        const int t0 = prof.start("solve_e5", (double)tot);
        launch_e5_solve(stream, d_rpts, d_pair_pt_off.p, d_cpair, d_samples, tot, d_cand_models,
                        d_status, d_nmodels);
        prof.stop(t0);
      }
      const int t1 = prof.start(kind == GCR_PROBLEM_ESSENTIAL ? "score_e" : "score_f", 0.0);
      if (mx)
        launch_score_mx(stream, 1, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pair_sq_thr.p, d_pair_inv_scale.p,
                        d_mxp.p, d_cand_models, d_cpair, d_status, (int64_t)ns, d_counts, d_values, d_sym, S,
                        d_work_list, d_work_count, d_blk, mx_max_n, d_mx_mask.p);
      else
        launch_score(stream, 1, d_rpts, d_pair_pt_off.p, d_pair_n.p, d_pair_sq_thr.p, d_cand_models,
                     d_cpair, d_status, (int64_t)ns, d_counts, d_values, nullptr,
                     nullptr, d_sym, true, S, d_work_list, d_work_count);
      prof.stop(t1);
      if (degen_any) {
        launch_degen_test(stream, d_rpts, d_pair_pt_off.p, d_cpair, d_samples, d_cand_models,
                          d_status, (int64_t)ns, d_degen, d_cand_degen_h);
      }
    }
    if (filt) {
      co_await exec_chunk_filtered(items, kind, m, S, tot, base, d_status, d_nmodels, d_counts, d_values, d_sym,
                                   d_valid, degen_any ? d_degen : nullptr, d_samples, d_ibase, d_incand, d_ibest,
                                   cand_models, d_cand_models, d_cand_degen_h);
      co_return;
    }
    char* const hres = h_cand_res[kind].ensure(r_end);
    d2h(hres, dres, r_end);
    // samples drawn on the device or into the upload block go to the host
    // loop (DEGENSAC seeds its nested run with a sample's first index)
    if (!dsi.empty()) d2h(samples, d_samples, (size_t)tot * m);
    co_await step_sync();
    for (size_t i = 0; i < ni; ++i) {
      Request& r = items[i]->req;
      if (r.dev_sample || r.host_draw)
        r.samples.assign(samples + m * base[i], samples + m * (base[i] + r.num_cand));
    }
    int8_t* st = reinterpret_cast<int8_t*>(hres);
    int8_t* vd = reinterpret_cast<int8_t*>(hres + r_vd);
    int8_t* nm = reinterpret_cast<int8_t*>(hres + r_nm);
    int8_t* dg = degen_any ? reinterpret_cast<int8_t*>(hres + r_dg) : nullptr;
    int32_t* cn = reinterpret_cast<int32_t*>(hres + r_cn);
    int32_t* sy = reinterpret_cast<int32_t*>(hres + r_sy);
    double* vl = reinterpret_cast<double*>(hres + r_vl);
    int32_t* sl = reinterpret_cast<int32_t*>(hres + r_sl);
    if (kind == GCR_PROBLEM_HOMOGRAPHY)
      for (size_t c = 0; c < (size_t)tot; ++c) nm[c] = st[c] == 2 ? 1 : 0;
    // hypothesis-points actually scored (profiling): model slots with
    // status 2 (the others exit after one test), counted per item below
    std::vector<double> hp_item(items.size(), 0.0);
    pool->run(items.size(), [&](size_t i) {
      Request& r = items[i]->req;
      const size_t b = (size_t)base[i], k = (size_t)r.num_cand;
      r.slots_per = S;
      r.nmodels.assign(nm + b, nm + b + k);
      r.status.assign(st + b * S, st + (b + k) * S);
      r.counts.assign(cn + b * S, cn + (b + k) * S);
      r.values.assign(vl + b * S, vl + (b + k) * S);
      if (kind == GCR_PROBLEM_HOMOGRAPHY) r.valid.assign(vd + b, vd + b + k);
      else r.sym.assign(sy + b * S, sy + (b + k) * S);
      if (items[i]->sp_on) r.sel.assign(sl + b * S, sl + (b + k) * S);
      r.dev_models = d_cand_models + 9 * b * S;
      r.models_hold = cand_models;
      if (items[i]->degensac) {
        r.degen.assign(dg + b * S, dg + (b + k) * S);
        r.dev_degen_h = d_cand_degen_h + 9 * b * S;
      }
      items[i]->candidates_solved += (int64_t)k;
      int64_t kept = 0;
      for (size_t c = 0; c < k; ++c) kept += nm[b + c];
      items[i]->models_scored += kept;
      if (g_prof_on) {
        int64_t act = 0;
        for (size_t q = b * S; q < (b + k) * S; ++q) act += st[q] == 2;
        hp_item[i] = (double)act * (double)items[i]->n;
      }
    });
    if (g_prof_on) {
      double hp = 0.0;
      for (double v : hp_item) hp += v;
      prof_add_work(kind == GCR_PROBLEM_HOMOGRAPHY ? "score_h" : (kind == GCR_PROBLEM_ESSENTIAL ? "score_e" : "score_f"),
                    hp);
    }
  }

  PhaseTask exec_label(std::vector<PairCtx*> items) {
    const int np = (int)pairs.size();
    const int k = (int)items.size();
    std::vector<int8_t> act(np, 0);
    for (auto* p : items) act[p->index] = 1;
    // lo models are the pairs' S_LOBEST slots: gather into a contiguous array
    DevBuf<double>& lm = d_ml_models;
    lm.ensure((size_t)np * 9);
    std::vector<const double*> src;
    std::vector<double*> dst;
    for (auto* p : items) {
      src.push_back(p->slot(S_LOBEST));
      dst.push_back(lm.p + 9 * (size_t)p->index);
    }
    // distances per residual kind (active flags per kind in one array)
    bool kinds[3] = {false, false, false};
    for (auto* p : items) kinds[est_kind(p->est)] = true;
    std::vector<int8_t> ak(3 * (size_t)np, 0);
    for (auto* p : items) ak[(size_t)est_kind(p->est) * np + p->index] = 1;
    // ordered lists of SINK points per active pair
    std::vector<int64_t> loff, pbase(np);
    std::vector<int32_t> ln;
    int64_t acc = 0;
    for (int i = 0; i < np; ++i) {
      pbase[i] = acc;
      acc += pairs[i].n;
    }
    for (auto* p : items) {
      loff.push_back(pbase[p->index]);
      ln.push_back((int32_t)p->n);
    }
    // one upload: the inputs, then the outputs' block (list lengths and the
    // min-cut failure flag, uploaded as zeros) that comes back in one copy
    std::vector<int32_t> out_init((size_t)k + (size_t)np, 0);
    Packed up;
    int8_t *u_act = nullptr, *u_ak = nullptr;
    const double** u_src = nullptr;
    double** u_dst = nullptr;
    int64_t* u_loff = nullptr;
    int32_t *u_ln = nullptr, *u_out = nullptr;
    up.add(u_act, act.data(), (size_t)np);
    up.add(u_src, src.data(), src.size());
    up.add(u_dst, dst.data(), dst.size());
    up.add(u_ak, ak.data(), ak.size());
    up.add(u_loff, loff.data(), loff.size());
    up.add(u_ln, ln.data(), ln.size());
    up.add(u_out, out_init.data(), out_init.size());
    upload(up, d_lab_up, h_pack_lab, lab_busy);
    int32_t* const d_counts = u_out;
    int32_t* const d_failw = u_out + k;  // per pair: its min-cut gave up
    launch_copy_models(stream, u_src, u_dst, k);
    for (int kk = 0; kk < 3; ++kk) {
      if (!kinds[kk]) continue;
      launch_lo_dist(stream, kk == 0 ? 0 : 1, d_rpts, d_pt_pair.p, u_ak + (size_t)kk * np, d_pair_lo_thr.p, lm.p, total,
                     d_dist.p, d_near.p);
    }
    const int t0 = prof.start("energy", 0.0);
    launch_lo_caps(stream, d_pt_pair.p, u_act, d_pair_pt_off.p, d_pair_lambda.p, d_pair_qbits.p, d_dist.p, d_near.p,
                   d_svals.p, d_pos_of.p, d_run_len.p, d_edge_off.p, d_pair_radius.p, d_r_off.p, d_r_nbr_s.p,
                   d_adj_off.p, total, d_rf.p, d_rb.p, d_node_src.p, d_node_sink.p, d_sink.p);
    prof.stop(t0);
    const int t1 = prof.start("mincut", 0.0);
    launch_mincut_cells(stream, d_cell_pair.p, d_cell_start.p, d_cell_len.p, num_cells, u_act, d_pair_lambda.p,
                        d_pair_pt_off.p, d_svals.p, d_edge_off.p, d_rf.p, d_rb.p, d_node_src.p, d_node_sink.p, d_h.p,
                        d_hn.p, d_sink.p, d_failw, max_cell);
    if (!radius_pairs.empty())
      launch_mincut_radius(stream, d_rpairs.p, (int)radius_pairs.size(), u_act, d_pair_lambda.p, d_pair_pt_off.p,
                           d_pair_n.p, d_adj_off.p, d_adj_arc.p, d_adj_other.p, d_rf.p, d_rb.p, d_node_src.p,
                           d_node_sink.p, d_h.p, d_hn.p, d_queue.p, d_sink.p, d_failw);
    prof.stop(t1);
    d_lists.ensure((size_t)total);
    launch_compact(stream, d_sink.p, u_loff, u_ln, k, d_lists.p, d_counts);
    int32_t* const hout = h_lab_out.ensure((size_t)k + (size_t)np);
    d2h(hout, d_counts, (size_t)k + (size_t)np);
    const int64_t lbase = d2h_list_spans(h_lists_lab, loff, ln);
    co_await step_sync();
    for (int i = 0; i < k; ++i) {
      PairCtx* P = items[(size_t)i];
      P->label_failed = hout[(size_t)k + (size_t)P->index] != 0;
      const int32_t* b = h_lists_lab.p + (loff[(size_t)i] - lbase);
      if (P->label_failed) P->req.list.clear();
      else P->req.list.assign(b, b + hout[i]);
    }
  }

  // Non-minimal fits and their scores.  Fits are grouped into launches of
  // equal (kind, flags, weighting); each fit owns slots_per model slots.
  PhaseTask exec_fit(std::vector<PairCtx*> items) {
    const int np = (int)pairs.size();
    std::vector<int64_t> pbase(np);
    {
      int64_t acc = 0;
      for (int i = 0; i < np; ++i) {
        pbase[i] = acc;
        acc += pairs[i].n;
      }
    }
    struct FitRec {
      size_t item;
      int f;
      int kind;
      int flags;
      bool weighted;
      bool big;  // more than 64 points: own launch, so small fits keep a small LDS footprint
    };
    std::vector<FitRec> recs;
    for (size_t i = 0; i < items.size(); ++i) {
      Request& r = items[i]->req;
      bool big = false;
      for (const auto& fit : r.fits) big |= fit.size() > 64;
      for (size_t f = 0; f < r.fits.size(); ++f)
        recs.push_back(FitRec{i, (int)f, r.kind_override >= 0 ? r.kind_override : est_kind(items[i]->est),
                              r.fit_flags, r.weighted, big});
    }
    // group by launch key, keeping the per-item order within a group (every
    // fit of one request has the request's key)
    std::stable_sort(recs.begin(), recs.end(), [](const FitRec& a, const FitRec& b) {
      if (a.kind != b.kind) return a.kind < b.kind;
      if (a.weighted != b.weighted) return a.weighted < b.weighted;
      if (a.flags != b.flags) return a.flags < b.flags;
      return a.big < b.big;
    });
    const int nf = (int)recs.size();
    std::vector<int32_t> fpair(nf), fns(nf), fsamples;
    std::vector<int64_t> fsoff(nf), fscr(nf), fslot(nf), fmask;
    std::vector<int32_t> spair;
    std::vector<const double*> fwm(nf, nullptr);
    std::vector<int> fS(nf, 1);
    int64_t scr = 0, slot_total = 0, mask_total = 0;
    std::vector<std::vector<size_t>> item_fits(items.size());
    for (int j = 0; j < nf; ++j) {
      const FitRec& fr = recs[j];
      Request& r = items[fr.item]->req;
      const std::vector<int32_t>& smp = r.fits[fr.f];
      fpair[j] = items[fr.item]->index;
      fns[j] = (int32_t)smp.size();
      fsoff[j] = (int64_t)fsamples.size();
      fsamples.insert(fsamples.end(), smp.begin(), smp.end());
      fscr[j] = scr;
      const bool ba = fr.kind == GCR_PROBLEM_ESSENTIAL && ((fr.flags >> 8) & 3) == 3;
      if (fr.kind == GCR_PROBLEM_ESSENTIAL)
        scr += (ba ? 10 : 9) * (int64_t)std::max<size_t>(smp.size(), 8);
      else
        scr += 18 * (int64_t)std::max<size_t>(smp.size(), 8);
      const int S = fit_slots(fr.kind, fr.flags);
      fS[j] = S;
      fslot[j] = slot_total;
      for (int k = 0; k < S; ++k) {
        spair.push_back(fpair[j]);
        fmask.push_back(-1);
      }
      if (r.want_list && fr.f == 0) {
        fmask[(size_t)slot_total] = mask_total;  // list of the fit's first model
        mask_total += items[fr.item]->n;
      }
      slot_total += S;
      fwm[j] = r.weight_model;
      item_fits[fr.item].push_back((size_t)j);
    }
    d_fit_scratch.ensure((size_t)scr);
    std::shared_ptr<DevBuf<double>> fit_models = model_pool.get((size_t)slot_total * 9);
    double* const d_fit_models = fit_models->p;
    d_masks.ensure((size_t)std::max<int64_t>(mask_total, 1));
    // IRLS weights (GCRANSAC.h:655-679): zero the pair's array, then Tukey
    // weights of the request's list (fit 0) stored by point index
    std::vector<int32_t> wpair, wn, zn;
    std::vector<int64_t> woff, zoff;
    int maxz = 0;
    std::vector<const double*> wmodel;
    std::vector<int> wkind;
    int maxn = 0;
    for (size_t i = 0; i < items.size(); ++i) {
      Request& r = items[i]->req;
      if (!r.weighted) continue;
      zoff.push_back(pbase[items[i]->index]);
      zn.push_back((int32_t)items[i]->n);
      maxz = std::max(maxz, (int)items[i]->n);
      wpair.push_back(items[i]->index);
      woff.push_back(fsoff[item_fits[i][0]]);
      wn.push_back((int32_t)r.fits[0].size());
      wmodel.push_back(r.weight_model);
      wkind.push_back(est_kind(items[i]->est));
      maxn = std::max(maxn, (int)r.fits[0].size());
    }
    // ordered inlier lists of the want_list fits
    std::vector<int64_t> loff;
    std::vector<int32_t> ln;
    std::vector<size_t> lidx;
    for (size_t i = 0; i < items.size(); ++i)
      if (items[i]->req.want_list) {
        loff.push_back(fmask[(size_t)fslot[item_fits[i][0]]]);
        ln.push_back((int32_t)items[i]->n);
        lidx.push_back(i);
      }
    // the step's index arrays go up in one copy
    const bool count_work = g_prof_on;
    Packed up;
    int32_t *u_pair = nullptr, *u_ns = nullptr, *u_samples = nullptr, *u_slot_pair = nullptr, *u_wpair = nullptr,
            *u_wn = nullptr, *u_ln = nullptr;
    int64_t *u_soff = nullptr, *u_scr = nullptr, *u_mask_off = nullptr, *u_woff = nullptr, *u_wo = nullptr,
            *u_loff = nullptr;
    const double **u_wm = nullptr, **u_wmodel = nullptr;
    up.add(u_pair, fpair.data(), (size_t)nf);
    up.add(u_ns, fns.data(), (size_t)nf);
    up.add(u_soff, fsoff.data(), (size_t)nf);
    up.add(u_scr, fscr.data(), (size_t)nf);
    up.add(u_samples, fsamples.data(), fsamples.size());
    up.add(u_slot_pair, spair.data(), (size_t)slot_total);
    up.add(u_mask_off, fmask.data(), (size_t)slot_total);
    up.add(u_wm, fwm.data(), (size_t)nf);
    up.add(u_wpair, wpair.data(), wpair.size());
    up.add(u_woff, woff.data(), woff.size());
    up.add(u_wn, wn.data(), wn.size());
    up.add(u_wmodel, wmodel.data(), wmodel.size());
    up.add(u_wo, pbase.data(), (size_t)np);
    up.add(u_loff, loff.data(), loff.size());
    up.add(u_ln, ln.data(), ln.size());
    int64_t* u_zoff = nullptr;
    int32_t* u_zn = nullptr;
    up.add(u_zoff, zoff.data(), zoff.size());
    up.add(u_zn, zn.data(), zn.size());
    upload(up, d_fit_up, h_pack_fit, fit_busy);
    // the results come back in one copy: status, count and value per model
    // slot, the inlier-list lengths and (profiling) the LM pass counts
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t r_cn = al((size_t)slot_total), r_vl = r_cn + al(4 * (size_t)slot_total),
                 r_lc = r_vl + al(8 * (size_t)slot_total), r_lm = r_lc + al(4 * loff.size()),
                 r_end = r_lm + (count_work ? 8 * (size_t)nf : 0);
    char* const res = d_fit_res.ensure(std::max<size_t>(r_end, 256));
    int8_t* const d_ok = reinterpret_cast<int8_t*>(res);
    int32_t* const d_cn = reinterpret_cast<int32_t*>(res + r_cn);
    double* const d_vl = reinterpret_cast<double*>(res + r_vl);
    int32_t* const d_lc = reinterpret_cast<int32_t*>(res + r_lc);
    int32_t* const d_lm = count_work ? reinterpret_cast<int32_t*>(res + r_lm) : nullptr;
    if (!wpair.empty()) {
      // the weighted pairs' arrays start at zero
      launch_zero_ranges(stream, d_weights.p, u_zoff, u_zn, (int)zoff.size(), maxz);
      // one launch per residual kind over contiguous runs
      size_t a = 0;
      while (a < wpair.size()) {
        size_t b = a;
        const int rk = wkind[a] == 0 ? 0 : 1;
        while (b < wpair.size() && (wkind[b] == 0 ? 0 : 1) == rk) ++b;
        int mx = 0;
        for (size_t t = a; t < b; ++t) mx = std::max(mx, (int)wn[t]);
        launch_irls_weights(stream, rk, d_rpts, d_pair_pt_off.p, d_pair_sq_thr.p, u_wpair + a, u_wmodel + a,
                            u_samples, u_woff + a, u_wn + a, d_weights.p, (int)(b - a), 0, mx);
        a = b;
      }
    }
    // fit launches over runs of equal key
    // the profiler's work for the fit family is the algorithmic flop count
    // (fit_flops); the refinement's pass counts come back from the device
    int start = 0;
    while (start < nf) {
      int end = start;
      while (end < nf && recs[end].kind == recs[start].kind && recs[end].weighted == recs[start].weighted &&
             recs[end].flags == recs[start].flags && recs[end].big == recs[start].big)
        ++end;
      const FitRec& fr = recs[start];
      // one span per kernel launch; the sub-family (LO / IRLS / driver
      // post-hoc bundle adjustment) gets the same time, work = points
      const int wm = (fr.flags >> 8) & 3;
      const char* sub = (wm >= 2) ? "fit_post" : ((fr.weighted || wm == 1) ? "fit_irls" : "fit_lo");
      double pts_sum = 0.0;
      for (int j = start; j < end; ++j) pts_sum += fns[(size_t)j];
      const int tfit = prof.start("fit", 0.0, sub, pts_sum);
      int max_ns = 0;
      for (int j = start; j < end; ++j) max_ns = std::max(max_ns, (int)fns[(size_t)j]);
      if (fr.kind == GCR_PROBLEM_HOMOGRAPHY) {
        const bool use_w = fr.weighted || ((fr.flags >> 8) & 3) == 1;
        launch_fit_h(stream, d_rpts, d_pair_pt_off.p, u_pair + start, u_soff + start, u_ns + start, u_samples,
                     use_w ? d_weights.p : nullptr, u_wo, d_fit_scratch.p, u_scr + start,
                     d_fit_models + 9 * (size_t)fslot[start], d_ok + fslot[start], end - start, max_ns);
      } else if (fr.kind == GCR_PROBLEM_FUNDAMENTAL) {
        launch_fit_f(stream, d_rpts, d_pair_pt_off.p, u_pair + start, u_soff + start, u_ns + start, u_samples,
                     d_weights.p, u_wo, u_wm + start, d_pair_sq_thr.p, fr.flags, d_fit_scratch.p, u_scr + start,
                     d_fit_models + 9 * (size_t)fslot[start], d_ok + fslot[start], end - start,
                     count_work ? d_lm + 2 * (size_t)start : nullptr, max_ns);
      } else if (((fr.flags >> 8) & 3) == 3) {
        launch_ba_e(stream, d_rpts, d_pair_pt_off.p, u_pair + start, u_soff + start, u_ns + start, u_samples,
                    u_wm + start, d_pair_post_sq.p, d_fit_scratch.p, u_scr + start,
                    d_fit_models + 9 * (size_t)fslot[start], d_ok + fslot[start], end - start);
      } else {
        const bool use_w = fr.weighted || ((fr.flags >> 8) & 3) == 1;
        launch_fit_e(stream, d_rpts, d_pair_pt_off.p, u_pair + start, u_soff + start, u_ns + start, u_samples,
                     d_weights.p, u_wo, use_w ? 1 : 0, d_fit_scratch.p, u_scr + start,
                     d_fit_models + 9 * (size_t)fslot[start], d_ok + fslot[start], end - start);
      }
      prof.stop(tfit);
      start = end;
    }
    // scores of every model slot, per residual kind
    {
      // the driver's post-hoc bundle adjustment of essential pairs scores at
      // its own truncated threshold (py/gcransac_python.cpp:2797-2807)
      auto post_of = [&](int j) { return recs[j].kind == GCR_PROBLEM_ESSENTIAL && ((recs[j].flags >> 8) & 3) == 3; };
      int a = 0;
      while (a < nf) {
        int b = a;
        const int rk = recs[a].kind == 0 ? 0 : 1;
        const bool post = post_of(a);
        while (b < nf && (recs[b].kind == 0 ? 0 : 1) == rk && post_of(b) == post) ++b;
        const int64_t s0 = fslot[a], s1 = b < nf ? fslot[b] : slot_total;
        launch_score(stream, rk, d_rpts, d_pair_pt_off.p, d_pair_n.p, post ? d_pair_post_sq.p : d_pair_sq_thr.p,
                     d_fit_models + 9 * s0, u_slot_pair + s0, d_ok + s0, s1 - s0, d_cn + s0, d_vl + s0,
                     mask_total > 0 ? d_masks.p : nullptr, u_mask_off + s0, nullptr);
        a = b;
      }
    }
    if (!loff.empty()) {
      d_lists.ensure((size_t)mask_total);
      launch_compact(stream, d_masks.p, u_loff, u_ln, (int)loff.size(), d_lists.p, d_lc);
    }
    char* const hres = h_res.ensure(std::max<size_t>(r_end, 256));
    d2h(hres, res, r_end);
    std::vector<std::vector<int32_t>> lists(loff.size());
    int64_t lbase = 0;
    if (!loff.empty()) lbase = d2h_list_spans(h_lists_fit, loff, ln);
    co_await step_sync();
    std::vector<int8_t> ok((size_t)slot_total);
    std::vector<int32_t> cn((size_t)slot_total), lcount(loff.size());
    std::vector<double> vl((size_t)slot_total);
    std::memcpy(ok.data(), hres, (size_t)slot_total);
    std::memcpy(cn.data(), hres + r_cn, 4 * (size_t)slot_total);
    std::memcpy(vl.data(), hres + r_vl, 8 * (size_t)slot_total);
    if (!loff.empty()) std::memcpy(lcount.data(), hres + r_lc, 4 * loff.size());
    std::vector<int32_t> lm;
    if (count_work) {
      lm.resize((size_t)2 * (size_t)nf);
      std::memcpy(lm.data(), hres + r_lm, 8 * (size_t)nf);
    }
    for (size_t j = 0; j < loff.size(); ++j) {
      const int32_t* b = h_lists_fit.p + (loff[j] - lbase);
      lists[j].assign(b, b + lcount[j]);
    }
    if (count_work) {
      double flops = 0.0;
      for (int j = 0; j < nf; ++j) {
        const bool f_kind = recs[(size_t)j].kind == GCR_PROBLEM_FUNDAMENTAL;
        flops += fit_flops(recs[(size_t)j].kind, recs[(size_t)j].flags, fns[(size_t)j],
                           f_kind ? lm[2 * (size_t)j] : 0, f_kind ? lm[2 * (size_t)j + 1] : 0);
      }
      prof_add_work("fit", flops);
    }
    for (size_t i = 0; i < items.size(); ++i) {
      Request& r = items[i]->req;
      // fits of one item are contiguous in the sorted order (same key)
      const size_t j0 = item_fits[i][0];
      const int S = fS[j0];
      r.slots_per = S;
      r.status.clear();
      r.counts.clear();
      r.values.clear();
      for (size_t f = 0; f < item_fits[i].size(); ++f) {
        const size_t j = item_fits[i][f];
        for (int k = 0; k < S; ++k) {
          const size_t sl = (size_t)fslot[j] + k;
          r.status.push_back(ok[sl]);
          r.counts.push_back(cn[sl]);
          r.values.push_back(vl[sl]);
        }
      }
      r.dev_models = d_fit_models + 9 * (size_t)fslot[j0];
      r.models_hold = fit_models;
      items[i]->models_scored += (int64_t)r.status.size();
    }
    for (size_t j = 0; j < lidx.size(); ++j) items[lidx[j]]->req.list = std::move(lists[j]);
  }

  // SPRT walks (k_sprt.hip), one wavefront per pair: the items and their
  // slot lists go up in one copy, the items come back in one copy.
  PhaseTask exec_sprt(std::vector<PairCtx*> items) {
    const int k = (int)items.size();
    for (auto* P : items)
      if (P->sprt.table_dirty) {
        h2d(P->sprt.a_dev, P->sprt.a_table.data(), P->sprt.a_table.size());
        P->sprt.table_dirty = false;
      }
    std::vector<SprtItem> it((size_t)k);
    std::vector<int32_t> slots;
    std::vector<size_t> soff((size_t)k);
    for (int i = 0; i < k; ++i) {
      soff[(size_t)i] = slots.size();
      const std::vector<int32_t>& sl = items[(size_t)i]->req.sprt_slots;
      slots.insert(slots.end(), sl.begin(), sl.end());
      it[(size_t)i] = items[(size_t)i]->req.sprt;
      it[(size_t)i].a_table = items[(size_t)i]->sprt.a_dev;
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_items = al(sizeof(SprtItem) * (size_t)k);
    const size_t total_b = b_items + 4 * std::max<size_t>(1, slots.size());
    char* const d = d_sprt_up.ensure(total_b);
    char* const h = h_sprt_stage.ensure(total_b);
    int32_t* const d_slots = reinterpret_cast<int32_t*>(d + b_items);
    for (int i = 0; i < k; ++i) it[(size_t)i].slots = d_slots + soff[(size_t)i];
    std::memcpy(h, it.data(), sizeof(SprtItem) * (size_t)k);
    if (!slots.empty()) std::memcpy(h + b_items, slots.data(), 4 * slots.size());
    check_hip(hipMemcpyAsync(d, h, total_b, hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    const int t0 = prof.start("sprt", 0.0);
    launch_sprt_walk(stream, reinterpret_cast<SprtItem*>(d), k, d_rpts, d_pair_pt_off.p, d_pair_n.p,
                     d_pair_sq_thr.p, d_sprt_pool.p, d_sp_pairs.p, d_sp_cells.p, d_sp_perm.p, d_sp_cell_of.p,
                     d_sp_list.p, sp_max_cells);
    prof.stop(t0);
    check_hip(hipMemcpyAsync(h, d, sizeof(SprtItem) * (size_t)k, hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync D2H");
    co_await step_sync();
    const SprtItem* back = reinterpret_cast<const SprtItem*>(h);
    for (int i = 0; i < k; ++i) items[(size_t)i]->req.sprt = back[i];
  }

  PhaseTask exec_masklist(std::vector<PairCtx*> items) {
    const int k = (int)items.size();
    // results of the step in one block, read back with one copy: the staged
    // models, their values, the list lengths (and the scores' counts)
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t r_vl = al(72 * (size_t)k), r_lc = r_vl + al(8 * (size_t)k), r_sc = r_lc + al(4 * (size_t)k),
                 r_end = r_sc + 4 * (size_t)k;
    char* const dres = d_ml_res.ensure(r_end);
    double* const d_models = reinterpret_cast<double*>(dres);
    double* const d_values = reinterpret_cast<double*>(dres + r_vl);
    int32_t* const d_lcount = reinterpret_cast<int32_t*>(dres + r_lc);
    int32_t* const d_scount = reinterpret_cast<int32_t*>(dres + r_sc);
    std::vector<const double*> src(k);
    std::vector<double*> dst(k);
    std::vector<int32_t> mp(k), ln(k);
    std::vector<int64_t> moff(k);
    int64_t acc = 0;
    for (int i = 0; i < k; ++i) {
      src[i] = items[i]->req.model;
      dst[i] = d_models + 9 * (size_t)i;
      mp[i] = items[i]->index;
      moff[i] = acc;
      ln[i] = (int32_t)items[i]->n;
      acc += items[i]->n;
    }
    const std::vector<int64_t>& loff = moff;  // each item's list is compacted from its own mask
    // squared threshold per pair of this launch (tau^2, or the post-hoc
    // adjustment's threshold for a post list)
    std::vector<double> thr(pairs.size(), 0.0);
    for (int i = 0; i < k; ++i) {
      const Request& r = items[i]->req;
      thr[(size_t)items[i]->index] = r.sq_override >= 0.0 ? r.sq_override : (r.post ? items[i]->post_sq : items[i]->sq_thr);
    }
    Packed up;
    const double** u_src = nullptr;
    double** u_dst = nullptr;
    int32_t *u_pair = nullptr, *u_ln = nullptr;
    int64_t* u_moff = nullptr;
    double* u_thr = nullptr;
    up.add(u_src, src.data(), (size_t)k);
    up.add(u_dst, dst.data(), (size_t)k);
    up.add(u_pair, mp.data(), (size_t)k);
    up.add(u_moff, moff.data(), (size_t)k);
    up.add(u_ln, ln.data(), (size_t)k);
    up.add(u_thr, thr.data(), thr.size());
    upload(up, d_ml_up, h_pack_ml, ml_busy);
    launch_copy_models(stream, u_src, u_dst, k);
    d_masks.ensure((size_t)acc);
    d_lists.ensure((size_t)acc);
    // residual kind per item: runs of equal kind in item order
    {
      int a = 0;
      while (a < k) {
        int b = a;
        auto rk_of = [&](int i) {
          const int kk = items[i]->req.kind_override >= 0 ? items[i]->req.kind_override : est_kind(items[i]->est);
          return kk == 0 ? 0 : 1;
        };
        const int rk = rk_of(a);
        while (b < k && rk_of(b) == rk) ++b;
        launch_score(stream, rk, d_rpts, d_pair_pt_off.p, d_pair_n.p, u_thr, d_models + 9 * (size_t)a, u_pair + a,
                     nullptr, b - a, d_scount + a, d_values + a, d_masks.p, u_moff + a, nullptr);
        a = b;
      }
    }
    launch_compact(stream, d_masks.p, u_moff, u_ln, k, d_lists.p, d_lcount);
    // lists in the selector's cell order, and lists an SPRT walk over cells
    // left on the device
    {
      std::vector<int8_t> isp((size_t)k, 0);
      bool any_sp = false;
      std::vector<int64_t> pbase(pairs.size(), 0);
      for (size_t q = 1; q < pairs.size(); ++q) pbase[q] = pbase[q - 1] + pairs[q - 1].n;
      for (int i = 0; i < k; ++i) {
        const Request& r = items[(size_t)i]->req;
        if (r.sp_list) {
          isp[(size_t)i] = 1;
          any_sp = true;
        }
        if (r.walked > 0)
          check_hip(hipMemcpyAsync(d_lists.p + moff[(size_t)i], d_sp_list.p + pbase[(size_t)items[(size_t)i]->index],
                                   sizeof(int32_t) * (size_t)r.walked, hipMemcpyDeviceToDevice, stream),
                    "hipMemcpyAsync D2D");
      }
      if (any_sp) {
        ml_sp_flags = std::move(isp);
        d_ml_sp.ensure((size_t)k);
        h2d(d_ml_sp.p, ml_sp_flags.data(), (size_t)k);
        launch_sp_list(stream, d_rpts, d_pair_pt_off.p, d_pair_n.p, u_thr, d_sp_pairs.p, d_sp_cells.p, d_sp_perm.p,
                       d_sp_cell_of.p, d_models, u_pair, d_ml_sp.p, u_moff, k, d_lists.p, d_lcount, sp_max_cells);
      }
    }
    char* const hres = h_ml_res.ensure(r_end);
    d2h(hres, dres, r_end);
    const int64_t lbase = d2h_list_spans(h_lists_ml, loff, ln);
    co_await step_sync();
    const double* mh = reinterpret_cast<const double*>(hres);
    const double* vl = reinterpret_cast<const double*>(hres + r_vl);
    const int32_t* cn = reinterpret_cast<const int32_t*>(hres + r_lc);
    for (int i = 0; i < k; ++i) {
      Request& r = items[i]->req;
      const int32_t* b = h_lists_ml.p + (loff[i] - lbase);
      r.list.assign(b, b + (r.walked >= 0 ? (int32_t)r.walked : cn[i]));
      r.counts.assign(1, cn[i]);
      r.values.assign(1, vl[i]);
      std::memcpy(r.model_host, mh + 9 * (size_t)i, sizeof(r.model_host));
    }
  }

  void run(std::vector<PairResult>& results) {
    std::vector<PairRunner> runners;
    runners.reserve(pairs.size());
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < pairs.size(); ++i) {
      pairs[i].out = &results[i];
      pairs[i].t_start = t0;
      runners.push_back(PairRunner{pairs[i]});
    }
    for (size_t i = 0; i < pairs.size(); ++i) {
      if (pairs[i].no_graph) {
        // no edge in the radius neighbourhood: 0 inliers, no run
        pairs[i].finished = true;
        continue;
      }
      pairs[i].root = runners[i].run();
      pairs[i].handle = pairs[i].root.h;
    }
    {
      const auto ts = std::chrono::steady_clock::now();
      std::vector<PairCtx*> all;
      for (auto& P : pairs)
        if (!P.finished) all.push_back(&P);
      resume_all(all);
      prof_host("host_start", ts);
    }
    while (true) {
      std::vector<PairCtx*> by[REQ_NTYPES];
      bool any = false;
      for (auto& P : pairs)
        if (!P.finished) {
          by[P.req.type].push_back(&P);
          any = true;
        }
      if (!any) break;
      // host wall time per phase ("host_*" profile families, launches =
      // engine steps)
      auto t0 = std::chrono::steady_clock::now();
      auto lap = [&](const char* name) {
        const auto t1 = std::chrono::steady_clock::now();
        if (g_prof_on) {
          std::lock_guard<std::mutex> g(g_prof_mu);
          ProfEntry& e = g_prof[name];
          e.ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
          e.launches += 1;
        }
        t0 = t1;
      };
      run_copies();
      lap("host_copies");
      // every phase enqueues its work (in this order on the one stream),
      // then the step synchronises once and the phases read their results;
      // a phase with a second stage (the filtered chunk records) parks again
      std::vector<PhaseTask> phases;
      phases.reserve(8);
      parked.clear();
      auto start = [&](PhaseTask t, const char* name) {
        phases.push_back(std::move(t));
        phases.back().h.resume();
        lap(name);
      };
      for (int kind = 0; kind < 4; ++kind) {  // the pairs' current estimators
        std::vector<PairCtx*> items;
        for (auto* p : by[REQ_CHUNK])
          if (p->est == kind) items.push_back(p);
        if (!items.empty()) start(exec_chunk(std::move(items), kind), "host_chunk");
      }
      if (!by[REQ_LABEL].empty()) start(exec_label(by[REQ_LABEL]), "host_label");
      if (!by[REQ_FIT].empty()) start(exec_fit(by[REQ_FIT]), "host_fit");
      if (!by[REQ_MASKLIST].empty()) start(exec_masklist(by[REQ_MASKLIST]), "host_masklist");
      if (!by[REQ_SPRT].empty()) start(exec_sprt(by[REQ_SPRT]), "host_sprt");
      draw_ahead();
      lap("host_ahead");
      finish_phases(phases);
      lap("host_finish");
      std::vector<PairCtx*> step;
      for (int t = REQ_CHUNK; t < REQ_NTYPES; ++t) step.insert(step.end(), by[t].begin(), by[t].end());
      resume_all(step);
      lap("host_resume");
    }
    for (auto& P : pairs)
      if (P.root.h) {
        P.root.h.destroy();
        P.root.h = nullptr;
      }
  }

  // While the step's kernels run: the next chunk's samples of every pair
  // whose host sampler may draw ahead (PROSAC, uniform on the host) and that
  // is in its main loop; a later chunk takes them in stream order, so the
  // samples and the run are the same as drawn at the request.
  void draw_ahead() {
    std::vector<PairCtx*> list;
    for (auto& P : pairs)
      if (!P.finished && P.req.type == REQ_CHUNK && P.req.host_draw && P.ahead_hint > 0 &&
          P.main_s == P.main_sampler.get() && P.main_sampler->draws_ahead())
        list.push_back(&P);
    pool->run(list.size(), [&](size_t i) {
      PairCtx& P = *list[i];
      const size_t m = (size_t)P.m;
      if (P.ahead_m != m) {
        P.ahead.clear();
        P.ahead_pos = 0;
        P.ahead_m = m;
      }
      const size_t have = P.ahead.size() / m - P.ahead_pos;
      if (have >= P.ahead_hint) return;
      if (P.ahead_pos > 0) {
        P.ahead.erase(P.ahead.begin(), P.ahead.begin() + (std::ptrdiff_t)(P.ahead_pos * m));
        P.ahead_pos = 0;
      }
      const size_t add = P.ahead_hint - have, old = P.ahead.size();
      P.ahead.resize(old + add * m);
      P.main_sampler->sample_chunk((size_t)P.n, P.ahead.data() + old, add, m);
    });
  }

  // Synchronises the step and resumes its parked phases until every phase
  // has read its results; rethrows the first phase's error.
  void finish_phases(std::vector<PhaseTask>& phases) {
    while (!parked.empty()) {
      sync();
      std::vector<std::coroutine_handle<>> now;
      now.swap(parked);
      for (auto h : now) h.resume();
    }
    for (auto& t : phases)
      if (t.h.promise().error) std::rethrow_exception(t.h.promise().error);
  }

  // One phase on its own (the single-request entry points).
  void run_phase(PhaseTask t) {
    std::vector<PhaseTask> phases;
    parked.clear();
    phases.push_back(std::move(t));
    phases.back().h.resume();
    finish_phases(phases);
  }

  // The pairs are independent: their coroutines run concurrently on the host
  // pool (each generates its own samples and touches only its own state).
  void resume_all(std::vector<PairCtx*>& list) {
    pool->run(list.size(), [&](size_t i) { resume(*list[i]); });
  }

  void resume(PairCtx& P) {
    P.req.type = REQ_NONE;
    P.handle.resume();
    if (P.root.h.done()) {
      P.finished = true;
      if (std::exception_ptr err = P.root.h.promise().error) {
        try {
          std::rethrow_exception(err);
        } catch (const PairError& e) {
          *P.out = PairResult();
          P.out->status = e.code;
          P.out->error = e.what();
        }
      }
    } else if (P.req.type == REQ_NONE) {
      throw EngineError(GCR_ERR_INTERNAL, "pair suspended without a device request");
    }
  }
};

}  // namespace

void score_models(int kind, const double* pts, int64_t n, const double* models, int64_t k, double sq_thr,
                  int64_t* counts, double* values, uint8_t* masks) {
  hipStream_t st;
  check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  DevBuf<double> d_p, d_m, d_thr, d_v;
  DevBuf<int64_t> d_off, d_moff;
  DevBuf<int32_t> d_n, d_mp, d_c;
  DevBuf<uint8_t> d_mask;
  d_p.ensure((size_t)n * 4);
  d_m.ensure((size_t)k * 9);
  d_thr.ensure(1);
  d_v.ensure((size_t)k);
  d_off.ensure(1);
  d_moff.ensure((size_t)k);
  d_n.ensure(1);
  d_mp.ensure((size_t)k);
  d_c.ensure((size_t)k);
  std::vector<int32_t> mp((size_t)k, 0);
  std::vector<int64_t> moff((size_t)k);
  for (int64_t j = 0; j < k; ++j) moff[(size_t)j] = j * n;
  const int64_t zero = 0;
  const int32_t nn = (int32_t)n;
  check_hip(hipMemcpyAsync(d_p.p, pts, sizeof(double) * 4 * n, hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_m.p, models, sizeof(double) * 9 * k, hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_thr.p, &sq_thr, sizeof(double), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_off.p, &zero, sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_n.p, &nn, sizeof(int32_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_mp.p, mp.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_moff.p, moff.data(), sizeof(int64_t) * k, hipMemcpyHostToDevice, st), "H2D");
  if (masks) d_mask.ensure((size_t)(k * n));
  launch_score(st, kind, d_p.p, d_off.p, d_n.p, d_thr.p, d_m.p, d_mp.p, nullptr, k, d_c.p, d_v.p,
               masks ? d_mask.p : nullptr, d_moff.p);
  std::vector<int32_t> c((size_t)k);
  check_hip(hipMemcpyAsync(c.data(), d_c.p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipMemcpyAsync(values, d_v.p, sizeof(double) * k, hipMemcpyDeviceToHost, st), "D2H");
  if (masks) check_hip(hipMemcpyAsync(masks, d_mask.p, (size_t)(k * n), hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipStreamSynchronize(st), "sync");
  for (int64_t j = 0; j < k; ++j) counts[j] = c[(size_t)j];
  (void)hipStreamDestroy(st);
}

void sampler_draws(int32_t sampler, uint32_t seed, int64_t n, int32_t m, const int64_t* chunks, int32_t num_chunks,
                   int32_t* out) {
  std::unique_ptr<Sampler> plan_s;
  if (sampler == 0) plan_s.reset(new UniformSampler(seed, (size_t)n));
  else plan_s.reset(new ProsacSampler(seed, (size_t)n, (size_t)m));
  hipStream_t st;
  check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  int64_t total = 0;
  for (int32_t i = 0; i < num_chunks; ++i) total += chunks[i];
  std::vector<DevSampleItem> items((size_t)num_chunks);
  std::vector<DevSampleSeg> segs;
  int64_t b = 0;
  for (int32_t i = 0; i < num_chunks; ++i) {
    std::vector<SampleSeg> g;
    plan_s->plan((size_t)chunks[i], (size_t)n, (size_t)m, g);
    items[(size_t)i] = DevSampleItem{0, m, b, chunks[i], (int32_t)segs.size(), (int32_t)g.size()};
    int64_t st0 = 0;
    for (const SampleSeg& x : g) {
      segs.push_back(DevSampleSeg{st0, x.count, x.range, x.draws, x.fixed, 0});
      st0 += x.count;
    }
    b += chunks[i];
  }
  DevBuf<DevMT> d_mt;
  DevBuf<uint32_t> d_seed;
  DevBuf<int32_t> d_pair, d_out;
  DevBuf<DevSampleItem> d_items;
  DevBuf<DevSampleSeg> d_segs;
  d_mt.ensure(1);
  d_seed.ensure(1);
  d_pair.ensure(1);
  d_out.ensure((size_t)std::max<int64_t>(1, total * m));
  d_items.ensure(std::max<size_t>(1, items.size()));
  d_segs.ensure(std::max<size_t>(1, segs.size()));
  const int32_t zero = 0;
  check_hip(hipMemcpyAsync(d_seed.p, &seed, sizeof(uint32_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_pair.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_items.p, items.data(), sizeof(DevSampleItem) * items.size(), hipMemcpyHostToDevice, st),
            "H2D");
  if (!segs.empty())
    check_hip(hipMemcpyAsync(d_segs.p, segs.data(), sizeof(DevSampleSeg) * segs.size(), hipMemcpyHostToDevice, st),
              "H2D");
  launch_mt_seed(st, d_seed.p, d_pair.p, 1, d_mt.p);
  // one launch per chunk: the stream continues across launches
  for (int32_t i = 0; i < num_chunks; ++i) launch_sample_streams(st, d_items.p + i, d_segs.p, 1, d_mt.p, d_out.p);
  check_hip(hipGetLastError(), "sampler launch");
  if (total > 0)
    check_hip(hipMemcpyAsync(out, d_out.p, sizeof(int32_t) * (size_t)(total * m), hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipStreamSynchronize(st), "sync");
  (void)hipStreamDestroy(st);
}

void solve_minimal(int kind, const double* pts, int64_t n, const int32_t* samples, int64_t c, double* models,
                   int32_t* status) {
  if (kind == GCR_PROBLEM_FUNDAMENTAL || kind == GCR_PROBLEM_ESSENTIAL) {
    const int m = kind == GCR_PROBLEM_FUNDAMENTAL ? 7 : 5;
    const int S = kind == GCR_PROBLEM_FUNDAMENTAL ? 3 : 10;
    hipStream_t st;
    check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    DevBuf<double> d_p, d_m;
    DevBuf<int64_t> d_off;
    DevBuf<int32_t> d_cp, d_s;
    DevBuf<int8_t> d_st, d_nm;
    d_p.ensure((size_t)n * 4);
    d_m.ensure((size_t)c * 9 * S);
    d_off.ensure(1);
    d_cp.ensure((size_t)c);
    d_s.ensure((size_t)c * m);
    d_st.ensure((size_t)c * S);
    d_nm.ensure((size_t)c);
    const int64_t zero = 0;
    check_hip(hipMemcpyAsync(d_p.p, pts, sizeof(double) * 4 * n, hipMemcpyHostToDevice, st), "H2D");
    check_hip(hipMemcpyAsync(d_off.p, &zero, sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
    check_hip(hipMemsetAsync(d_cp.p, 0, sizeof(int32_t) * c, st), "memset");
    check_hip(hipMemcpyAsync(d_s.p, samples, sizeof(int32_t) * m * c, hipMemcpyHostToDevice, st), "H2D");
    check_hip(hipMemsetAsync(d_m.p, 0, sizeof(double) * 9 * S * c, st), "memset");
    if (kind == GCR_PROBLEM_FUNDAMENTAL) launch_f7_solve(st, d_p.p, d_off.p, d_cp.p, d_s.p, c, d_m.p, d_st.p, d_nm.p);
    else launch_e5_solve(st, d_p.p, d_off.p, d_cp.p, d_s.p, c, d_m.p, d_st.p, d_nm.p);
    std::vector<int8_t> s8((size_t)c * S);
    check_hip(hipMemcpyAsync(models, d_m.p, sizeof(double) * 9 * S * c, hipMemcpyDeviceToHost, st), "D2H");
    check_hip(hipMemcpyAsync(s8.data(), d_st.p, (size_t)c * S, hipMemcpyDeviceToHost, st), "D2H");
    check_hip(hipStreamSynchronize(st), "sync");
    for (int64_t j = 0; j < c * S; ++j) status[j] = s8[(size_t)j];
    (void)hipStreamDestroy(st);
    return;
  }
  if (kind != GCR_PROBLEM_HOMOGRAPHY) throw EngineError(GCR_ERR_UNSUPPORTED, "minimal solver kind not built");
  hipStream_t st;
  check_hip(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  DevBuf<double> d_p, d_m;
  DevBuf<int64_t> d_off;
  DevBuf<int32_t> d_cp, d_s;
  DevBuf<int8_t> d_st, d_vd;
  d_p.ensure((size_t)n * 4);
  d_m.ensure((size_t)c * 9);
  d_off.ensure(1);
  d_cp.ensure((size_t)c);
  d_s.ensure((size_t)c * 4);
  d_st.ensure((size_t)c);
  d_vd.ensure((size_t)c);
  const int64_t zero = 0;
  check_hip(hipMemcpyAsync(d_p.p, pts, sizeof(double) * 4 * n, hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemcpyAsync(d_off.p, &zero, sizeof(int64_t), hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemsetAsync(d_cp.p, 0, sizeof(int32_t) * c, st), "memset");
  check_hip(hipMemcpyAsync(d_s.p, samples, sizeof(int32_t) * 4 * c, hipMemcpyHostToDevice, st), "H2D");
  check_hip(hipMemsetAsync(d_m.p, 0, sizeof(double) * 9 * c, st), "memset");
  launch_h4_solve(st, d_p.p, d_off.p, d_cp.p, d_s.p, c, d_m.p, d_st.p, d_vd.p);
  std::vector<int8_t> s8((size_t)c);
  check_hip(hipMemcpyAsync(models, d_m.p, sizeof(double) * 9 * c, hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipMemcpyAsync(s8.data(), d_st.p, (size_t)c, hipMemcpyDeviceToHost, st), "D2H");
  check_hip(hipStreamSynchronize(st), "sync");
  for (int64_t j = 0; j < c; ++j) status[j] = s8[(size_t)j];
  (void)hipStreamDestroy(st);
}

int fit_nonminimal(int kind, const double* pts, int64_t n, const int32_t* sample, int64_t ns, const double* weights,
                   int weight_mode, int normalize, double threshold, double* models_out) {
  // one pair holding the points; the fit request goes through exec_fit
  std::vector<PairSpec> specs(1);
  PairSpec& s = specs[0];
  s.kind = kind;
  s.n = n;
  s.params.threshold = threshold;
  s.params.confidence = 0.99;
  s.params.neighborhood_size = 1;
  if (kind == GCR_PROBLEM_ESSENTIAL) {
    // points are given normalised: identity intrinsics (normaliser 1)
    for (int i = 0; i < 9; ++i) s.K1[i] = s.K2[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  Engine e;
  e.setup(specs, pts, false);
  PairCtx& P = e.pairs[0];
  P.req = Request();
  P.req.type = REQ_FIT;
  P.req.fits.emplace_back(sample, sample + ns);
  if (kind == GCR_PROBLEM_HOMOGRAPHY) {
    if (weight_mode == 1) {
      // weights are read by sample position from the point-indexed array
      check_hip(hipMemcpyAsync(e.d_weights.p, weights, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, e.stream),
                "H2D");
      P.req.weighted = false;
      P.req.fit_flags = 1 << 8;
    }
  } else if (kind == GCR_PROBLEM_ESSENTIAL) {
    int flags = 0;
    if (weight_mode == 1) {
      // Stewenius reads the weights by the sample's point index
      check_hip(hipMemcpyAsync(e.d_weights.p, weights, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, e.stream),
                "H2D");
      flags = 1 << 8;
    } else if (weight_mode == 3) {
      // the driver's pose bundle adjustment, weights from the model in `weights`
      check_hip(hipMemcpyAsync(P.slot(S_TMP), weights, sizeof(double) * 9, hipMemcpyHostToDevice, e.stream), "H2D");
      P.req.weight_model = P.slot(S_TMP);
      flags = 3 << 8;
    }
    P.req.fit_flags = flags;
  } else {
    int flags = normalize ? 1 : 0;
    if (weight_mode == 1) {
      check_hip(hipMemcpyAsync(e.d_weights.p, weights, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, e.stream),
                "H2D");
      flags |= 1 << 8;
    } else if (weight_mode == 2) {
      // post-hoc weights from a model stored in slot S_TMP
      check_hip(hipMemcpyAsync(P.slot(S_TMP), weights, sizeof(double) * 9, hipMemcpyHostToDevice, e.stream), "H2D");
      P.req.weight_model = P.slot(S_TMP);
      flags |= 2 << 8;
    }
    P.req.fit_flags = flags;
  }
  std::vector<PairCtx*> items{&P};
  e.run_phase(e.exec_fit(items));
  int k = 0;
  std::vector<double> host((size_t)P.req.status.size() * 9);
  check_hip(hipMemcpy(host.data(), P.req.dev_models, sizeof(double) * host.size(), hipMemcpyDeviceToHost), "D2H");
  for (size_t j = 0; j < P.req.status.size(); ++j)
    if (P.req.status[j] == 2) {
      std::memcpy(models_out + 9 * k, host.data() + 9 * j, sizeof(double) * 9);
      ++k;
    }
  return k;
}

void labeling(int kind, const double* pts, int64_t n, const double* model, double lambda, double threshold,
              const double* cell, int64_t divisions, uint8_t* sink) {
  std::vector<PairSpec> specs(1);
  PairSpec& s = specs[0];
  s.kind = kind;
  s.offset = 0;
  s.n = n;
  s.params.threshold = threshold;
  s.params.confidence = 0.99;
  s.params.spatial_coherence_weight = lambda;
  s.params.neighborhood_size = (double)divisions;
  s.params.lo_number = 50;
  s.cell_override = true;
  for (int d = 0; d < 4; ++d) s.cell[d] = cell[d];
  Engine e;
  e.setup(specs, pts, false);
  check_hip(hipMemcpyAsync(e.pairs[0].slot(S_LOBEST), model, sizeof(double) * 9, hipMemcpyHostToDevice, e.stream),
            "H2D");
  std::vector<PairCtx*> items{&e.pairs[0]};
  e.run_phase(e.exec_label(items));
  std::memset(sink, 0, (size_t)n);
  for (int32_t i : e.pairs[0].req.list) sink[i] = 1;
}

// A batch of 16 pairs or more runs as two groups of consecutive pairs (three
// from 768 pairs on), each in its own engine (stream, workspace, host pool)
// on its own host thread, so that one group's host steps overlap the other
// groups' kernels.  A pair's result does not depend on the other pairs of its
// batch, so the grouping changes no result.  GCR_BATCH_GROUPS sets the number
// of groups.  Four groups and more lose: their host threads compete for the
// host cores.
static size_t batch_groups(size_t num_pairs) {
  size_t g = std::min<size_t>(num_pairs >= 768 ? 3 : 2, num_pairs / 8);
  return std::max<size_t>(1, std::min(g, num_pairs));
}

// Batch-shape autotuner.  Which group count gives the most pairs/s depends on
// the balance of host and device work in the batch.  Measured on MI355X at
// 1,024 pairs with 8 hardware queues (GPU_MAX_HW_QUEUES, so that every group
// stream has its own; pairs/s with 3 / 4 / 5 groups): learned-guidance PROSAC
// fundamental batch 6,227 / 6,587 / 5,574, uniform fundamental 2,775 / 2,997
// / 2,731; with the default 4 queues two group streams share one and 2 / 3
// groups gave 5,010 / 4,713 (guided) and 2,194 / 2,390 (fundamental).  So
// batches of 256 pairs or more are keyed by their shape (problem kinds,
// samplers, pair count and mean correspondence count, in powers of two).
// The first four calls of a shape run 3, 4, 3 and 4 groups; the first two
// grow the workspaces and are not counted, the rates of the last two decide,
// and every later call runs the faster count.  GCR_BATCH_GROUPS (a fixed
// count) or GCR_BATCH_TUNE=0 (the fixed rule above) turn it off.
class BatchTuner {
 public:
  using Shape = std::array<int64_t, 4>;
  static BatchTuner& get() {
    static BatchTuner t;
    return t;
  }
  static Shape shape_of(const std::vector<PairSpec>& pairs) {
    int64_t kinds = 0, samplers = 0, pts = 0;
    for (const auto& p : pairs) {
      kinds |= int64_t(1) << (p.kind & 15);
      samplers |= int64_t(1) << (p.params.sampler & 15);
      pts += p.n;
    }
    auto lg = [](int64_t v) {
      int b = 0;
      while (v > 1) {
        v >>= 1;
        ++b;
      }
      return (int64_t)b;
    };
    return Shape{kinds, samplers, lg((int64_t)pairs.size()), lg(pts / std::max<int64_t>(1, (int64_t)pairs.size()))};
  }
  // the group count of the next call of this shape and whether its rate counts
  size_t choose(const Shape& s, bool& measured) {
    std::lock_guard<std::mutex> lock(mu_);
    Entry& e = entries_[s];
    const int64_t k = e.calls++;
    if (k < 4) {
      measured = k >= 2;
      return kSchedule[k];
    }
    measured = false;
    return e.rate[3] >= e.rate[4] ? 3 : 4;
  }
  void record(const Shape& s, size_t g, double pairs_per_s) {
    std::lock_guard<std::mutex> lock(mu_);
    if (g < 8) entries_[s].rate[g] = pairs_per_s;
  }

 private:
  static constexpr size_t kSchedule[4] = {3, 4, 3, 4};
  struct Entry {
    int64_t calls = 0;
    std::array<double, 8> rate{};
  };
  std::mutex mu_;
  std::map<Shape, Entry> entries_;
};

static void run_groups(const std::vector<PairSpec>& pairs, const double* corr, bool corr_on_device,
                       std::vector<PairResult>& results, size_t groups);

static std::atomic<int> g_last_groups{0};
int last_batch_groups() { return g_last_groups.load(); }

void run_batch(const std::vector<PairSpec>& pairs, const double* corr, bool corr_on_device,
               std::vector<PairResult>& results) {
  results.assign(pairs.size(), PairResult());
  size_t groups = batch_groups(pairs.size());
  const char* fixed = std::getenv("GCR_BATCH_GROUPS");
  const char* tune_env = std::getenv("GCR_BATCH_TUNE");
  if (fixed) {
    groups = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, std::atoi(fixed)), pairs.size()));
  } else if (pairs.size() >= 256 && !(tune_env && std::atoi(tune_env) == 0)) {
    const BatchTuner::Shape shape = BatchTuner::shape_of(pairs);
    bool measured = false;
    groups = BatchTuner::get().choose(shape, measured);
    g_last_groups = (int)groups;
    const auto t0 = std::chrono::steady_clock::now();
    run_groups(pairs, corr, corr_on_device, results, groups);
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (measured && sec > 0) BatchTuner::get().record(shape, groups, (double)pairs.size() / sec);
    return;
  }
  g_last_groups = (int)groups;
  run_groups(pairs, corr, corr_on_device, results, groups);
}

static void run_groups(const std::vector<PairSpec>& pairs, const double* corr, bool corr_on_device,
                       std::vector<PairResult>& results, size_t groups) {
  if (groups <= 1) {
    Engine e;
    e.setup(pairs, corr, corr_on_device);
    e.run(results);
    return;
  }
  int device = 0;
  check_hip(hipGetDevice(&device), "hipGetDevice");
  const size_t np = pairs.size();
  std::vector<std::exception_ptr> errors(groups);
  std::vector<std::thread> threads;
  threads.reserve(groups);
  for (size_t g = 0; g < groups; ++g) {
    const size_t lo = g * np / groups, hi = (g + 1) * np / groups;
    threads.emplace_back([&, g, lo, hi] {
      try {
        check_hip(hipSetDevice(device), "hipSetDevice");
        const std::vector<PairSpec> sub(pairs.begin() + (std::ptrdiff_t)lo, pairs.begin() + (std::ptrdiff_t)hi);
        // device input: the group's rows follow the earlier groups' rows
        const double* base = corr;
        if (corr_on_device)
          for (size_t i = 0; i < lo; ++i) base += 4 * pairs[i].n;
        std::vector<PairResult> part(sub.size());
        {
          const auto ts = std::chrono::steady_clock::now();
          Engine e;
          e.pool = &HostPool::group(g, groups);
          e.setup(sub, base, corr_on_device);
          prof_host("host_setup", ts);
          e.run(part);
        }
        for (size_t i = 0; i < part.size(); ++i) results[lo + i] = std::move(part[i]);
      } catch (...) {
        errors[g] = std::current_exception();
      }
    });
  }
  for (auto& t : threads) t.join();
  for (auto& e : errors)
    if (e) std::rethrow_exception(e);
}

}  // namespace gcr
