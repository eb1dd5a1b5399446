// Parallel host memory copies of the host-memory paths: staging pageable
// inputs into pinned buffers and copying hashes out of pinned landing buffers
// into pageable outputs.  One thread copies 10-20 GB/s; PCIe moves 50+ GB/s
// each way, so a single-threaded copy would bound every call whose arrays are
// pageable (numpy arrays, as the Python module passes them).
#pragma once

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace wmh {

struct CopyTask {
  void *dst;
  const void *src;
  size_t bytes;
};

// A fixed set of worker threads; copy() splits the tasks into chunks that the
// workers and the calling thread take from a shared counter, and returns when
// every chunk is copied.  Several threads may call copy() at once (the serving
// loop's slots): the workers help the oldest job that still has chunks.
class HostCopier {
 public:
  // Worker threads for the CPUs this process may run on: at most 7 besides
  // the caller, none on a single CPU (fewer if the system refuses threads:
  // the caller copies alone then).
  HostCopier() {
    cpu_set_t set;
    unsigned cpus = 1;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = (unsigned)CPU_COUNT(&set);
    const unsigned workers = std::min(7u, cpus > 1 ? cpus - 1 : 0u);
    try {
      threads_.reserve(workers);
      for (unsigned i = 0; i < workers; i++) threads_.emplace_back([this] { work(); });
    } catch (...) {
      // keep the workers that started
    }
  }
  HostCopier(const HostCopier &) = delete;
  HostCopier &operator=(const HostCopier &) = delete;
  ~HostCopier() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : threads_) t.join();
  }

  void copy(const CopyTask *tasks, size_t n) noexcept {
    std::shared_ptr<Job> job;
    try {
      job = std::make_shared<Job>();
      for (size_t i = 0; i < n; i++) {
        char *d = static_cast<char *>(tasks[i].dst);
        const char *s = static_cast<const char *>(tasks[i].src);
        for (size_t off = 0; off < tasks[i].bytes; off += kChunk)
          job->chunks.push_back(
              CopyTask{d + off, s + off, std::min(kChunk, tasks[i].bytes - off)});
      }
    } catch (...) {  // no memory for the chunk list: copy on this thread
      for (size_t i = 0; i < n; i++) std::memcpy(tasks[i].dst, tasks[i].src, tasks[i].bytes);
      return;
    }
    if (job->chunks.empty()) return;
    const bool shared = job->chunks.size() > 1 && !threads_.empty();
    if (shared) {
      {
        std::lock_guard<std::mutex> lock(mu_);
        jobs_.push_back(job);
      }
      cv_work_.notify_all();
    }
    run(*job);
    std::unique_lock<std::mutex> lock(mu_);
    cv_done_.wait(lock, [&] { return job->done.load() == job->chunks.size(); });
    if (shared) jobs_.erase(std::find(jobs_.begin(), jobs_.end(), job));
  }

 private:
  static constexpr size_t kChunk = 512u << 10;
  struct Job {
    std::vector<CopyTask> chunks;
    std::atomic<size_t> next{0}, done{0};
    bool has_chunks() const { return next.load() < chunks.size(); }
  };

  void run(Job &job) {
    const size_t n = job.chunks.size();
    for (size_t i = job.next.fetch_add(1); i < n; i = job.next.fetch_add(1)) {
      std::memcpy(job.chunks[i].dst, job.chunks[i].src, job.chunks[i].bytes);
      if (job.done.fetch_add(1) + 1 == n) {
        std::lock_guard<std::mutex> lock(mu_);
        cv_done_.notify_all();
      }
    }
  }

  // the oldest listed job with chunks left (mu_ held), or null
  std::shared_ptr<Job> open_job() const {
    for (const auto &j : jobs_)
      if (j->has_chunks()) return j;
    return nullptr;
  }

  void work() {
    for (;;) {
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_work_.wait(lock, [&] { return stop_ || (job = open_job()) != nullptr; });
        if (stop_) return;
      }
      run(*job);
    }
  }

  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::shared_ptr<Job>> jobs_;  // jobs the workers may help (mu_)
  bool stop_ = false;
};

}  // namespace wmh
