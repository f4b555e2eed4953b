// A small persistent pool of host threads for the per-pair control flow of a
// batch: the pairs' coroutines are independent (each touches only its own
// state, samplers and pending slot copies), so the engine resumes the pairs
// of one step concurrently.  One job runs at a time; a caller that finds the
// pool busy (another engine in another thread) runs its job itself.  A batch
// split into groups that run concurrently (one engine per group) gives each
// group a pool of its own (HostPool::group).
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gcr {

class HostPool {
 public:
  static HostPool& get() {
    static HostPool pool;
    return pool;
  }

  // fn(i) for every i in [0, n); returns when all are done.  The first
  // exception thrown by fn is rethrown here after the job has drained.
  void run(size_t n, const std::function<void(size_t)>& fn) {
    std::unique_lock<std::mutex> job(job_mu_, std::try_to_lock);
    if (!job.owns_lock() || threads_.empty() || n < 2) {
      for (size_t i = 0; i < n; ++i) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      error_ = nullptr;
      active_ = threads_.size();
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> g(mu_);
    done_cv_.wait(g, [&] { return active_ == 0; });
    fn_ = nullptr;
    if (error_) std::rethrow_exception(error_);
  }

  size_t threads() const { return threads_.size() + 1; }

  // Host threads the pools may use together: hardware threads up to 16, or
  // GCR_HOST_THREADS.
  static size_t budget() {
    unsigned hw = std::thread::hardware_concurrency();
    size_t want = hw > 1 ? std::min<size_t>(hw, 16) : 1;
    if (const char* e = std::getenv("GCR_HOST_THREADS")) want = (size_t)std::max(1, std::atoi(e));
    return want;
  }

  // The pool of group g when a batch runs as `count` concurrent groups: each
  // gets an equal share of the thread budget (its caller counts as one).
  static HostPool& group(size_t g, size_t count) {
    static std::mutex mu;
    static std::map<std::pair<size_t, size_t>, std::unique_ptr<HostPool>> pools;
    std::lock_guard<std::mutex> lock(mu);
    std::unique_ptr<HostPool>& p = pools[{count, g}];
    if (!p) p.reset(new HostPool(std::max<size_t>(1, budget() / std::max<size_t>(1, count))));
    return *p;
  }

 private:
  HostPool() : HostPool(budget()) {}
  explicit HostPool(size_t want) {
    for (size_t t = 1; t < want; ++t) threads_.emplace_back([this] { worker(); });
  }
 public:
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

 private:
  void drain() {
    for (size_t i; (i = next_.fetch_add(1)) < n_;) {
      try {
        (*fn_)(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(err_mu_);
        if (!error_) error_ = std::current_exception();
      }
    }
  }

  void worker() {
    uint64_t seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      drain();
      {
        std::lock_guard<std::mutex> g(mu_);
        if (--active_ == 0) done_cv_.notify_one();
      }
    }
  }

  std::mutex job_mu_, mu_, err_mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  size_t active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr error_;
  std::vector<std::thread> threads_;
};

}  // namespace gcr
