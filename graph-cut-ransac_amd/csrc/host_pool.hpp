// A small persistent pool of host threads for the per-pair control flow of a
// batch: the pairs' coroutines are independent (each touches only its own
// state, samplers and pending slot copies), so the engine resumes the pairs
// of one step concurrently.  One job runs at a time; a caller that finds the
// pool busy (another engine in another thread) runs its job itself.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <functional>
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

 private:
  HostPool() {
    unsigned hw = std::thread::hardware_concurrency();
    size_t want = hw > 1 ? std::min<size_t>(hw, 16) : 1;
    if (const char* e = std::getenv("GCR_HOST_THREADS")) want = (size_t)std::max(1, std::atoi(e));
    for (size_t t = 1; t < want; ++t) threads_.emplace_back([this] { worker(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }

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
