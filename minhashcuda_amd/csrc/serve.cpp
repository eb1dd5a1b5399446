// Continuous-batching serving loop over one generator (no counterpart in the
// reference, which only has the synchronous mhcuda_calc).
//
// Callers submit CSR requests held in host memory and receive tickets.  A
// scheduler thread admits pending requests (FIFO) into a free slot up to the
// slot's share of the device memory budget, uploads them, and enqueues one
// persistent hashing launch for the whole group plus the download of every
// output on the slot's stream.  Each slot has a worker thread that stages
// the uploads (pinned arrays go straight to DMA), waits for its launch and
// copies the hashes out, so slots overlap both on the device and on the
// host; a request that arrives while others are being hashed is admitted as
// soon as a slot frees up.  Each request's latency is submit -> hashes in
// host memory.  Each hash equals what a lone mhcuda_calc of that request
// gives: rows are hashed independently of whatever shares the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "minhashcuda.h"
#include "wmh_runtime.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  const float *weights = nullptr;
  const uint32_t *cols = nullptr;
  const uint32_t *rows = nullptr;
  uint32_t length = 0;
  uint32_t *output = nullptr;
  uint64_t nnz = 0;
  uint64_t bytes = 0;   // device bytes while in flight
  uint64_t ticket = 0;
  Clock::time_point submitted;
  MHCUDAResult status = mhcudaSuccess;
  bool done = false;
  double latency_ms = 0;
};

// One launch in flight: its own stream, device buffer, descriptor staging,
// pinned landing zones and worker thread.  The worker uploads the admitted
// requests, enqueues their hashing and downloads, waits for them and hands
// the hashes back, so host-side copies of different slots run in parallel
// while their launches overlap on the device.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t finished = nullptr;
  wmh::DevBuf buffer;
  wmh::PlanCtx plan;
  wmh::PinnedBuf host_in;    // staging of pageable inputs
  wmh::PinnedBuf host_out;   // landing zone of hashes bound for pageable outputs
  std::vector<Request *> reqs;
  bool busy = false;      // requests assigned and not yet retired
  std::thread worker;
  std::condition_variable cv;
};

uint64_t align256(uint64_t x) { return (x + 255u) & ~(uint64_t)255u; }

}  // namespace

struct MHCUDAServer_ {
  MinhashCudaGenerator *gen = nullptr;
  MHCUDAServeConfig cfg{};
  uint64_t slot_bytes = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work;   // scheduler: pending request or free slot
  std::condition_variable cv_done;   // waiters: a request finished
  std::deque<Request *> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_ticket = 1;
  bool stopping = false;
  bool scheduler_done = false;
  std::thread scheduler;
  // statistics since the last reset
  std::vector<double> latencies;
  uint64_t rejected = 0, batches = 0, batched_requests = 0, units = 0;
  Clock::time_point first_submit, last_done;
  bool have_window = false;

  uint64_t request_bytes(uint64_t nnz, uint32_t length) const {
    return align256(nnz * 4) * 2 + align256(((uint64_t)length + 1) * 4) +
           align256((uint64_t)length * gen->samples * 8);
  }

  // Uploads, hashes and downloads the requests of `slot`; returns when the
  // hashes are in the requests' output arrays.
  MHCUDAResult run(Slot *slot) {
    char *base = (char *)slot->buffer.ptr;
    const uint64_t S8 = (uint64_t)gen->samples * 8;
    // host staging sizes for pageable inputs and outputs
    size_t in_bytes = 0, out_bytes = 0;
    std::vector<char> in_pinned(slot->reqs.size() * 3), out_pinned(slot->reqs.size());
    for (size_t i = 0; i < slot->reqs.size(); i++) {
      Request *r = slot->reqs[i];
      const uint32_t first = r->rows[0];
      in_pinned[3 * i] = r->nnz == 0 || wmh::is_pinned(r->weights + first);
      in_pinned[3 * i + 1] = r->nnz == 0 || wmh::is_pinned(r->cols + first);
      in_pinned[3 * i + 2] = wmh::is_pinned(r->rows);
      if (!in_pinned[3 * i]) in_bytes += align256(r->nnz * 4);
      if (!in_pinned[3 * i + 1]) in_bytes += align256(r->nnz * 4);
      if (!in_pinned[3 * i + 2]) in_bytes += align256(((uint64_t)r->length + 1) * 4);
      out_pinned[i] = wmh::is_pinned(r->output);
      if (!out_pinned[i]) out_bytes += align256(r->length * S8);
    }
    if (!slot->host_in.reserve(in_bytes) || !slot->host_out.reserve(out_bytes))
      return mhcudaMemoryAllocationFailure;
    auto upload = [&](void *dst, const void *src, size_t n, bool pinned, size_t &staged) {
      if (!pinned) {
        std::memcpy(slot->host_in.ptr + staged, src, n);
        src = slot->host_in.ptr + staged;
        staged += align256(n);
      }
      return hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, slot->stream) == hipSuccess;
    };
    uint64_t off = 0;
    size_t staged = 0;
    std::vector<wmh::MatrixDesc> descs;
    descs.reserve(slot->reqs.size());
    for (size_t i = 0; i < slot->reqs.size(); i++) {
      Request *r = slot->reqs[i];
      const uint32_t first = r->rows[0];
      float *dw = (float *)(base + off);
      off += align256(r->nnz * 4);
      uint32_t *dc = (uint32_t *)(base + off);
      off += align256(r->nnz * 4);
      uint32_t *dr = (uint32_t *)(base + off);
      off += align256(((uint64_t)r->length + 1) * 4);
      uint32_t *dh = (uint32_t *)(base + off);
      off += align256(r->length * S8);
      if (r->nnz && (!upload(dw, r->weights + first, r->nnz * 4, in_pinned[3 * i], staged) ||
                     !upload(dc, r->cols + first, r->nnz * 4, in_pinned[3 * i + 1], staged)))
        return mhcudaMemoryCopyError;
      if (!upload(dr, r->rows, ((uint64_t)r->length + 1) * 4, in_pinned[3 * i + 2], staged))
        return mhcudaMemoryCopyError;
      descs.push_back(wmh::MatrixDesc{dw, dc, dr, dh, r->length, first, 0u, 0u, 0u, 0u});
    }
    MHCUDAResult res = wmh::hash_device_batch(gen, descs, mhcudaKernelPersistent, slot->stream,
                                              &slot->plan, false);
    if (res != mhcudaSuccess) return res;
    std::vector<size_t> out_off(slot->reqs.size());
    size_t landed = 0;
    for (size_t i = 0; i < slot->reqs.size(); i++) {
      Request *r = slot->reqs[i];
      void *dst = r->output;
      if (!out_pinned[i]) {
        out_off[i] = landed;
        dst = slot->host_out.ptr + landed;
        landed += align256(r->length * S8);
      }
      if (hipMemcpyAsync(dst, descs[i].output, r->length * S8, hipMemcpyDeviceToHost,
                         slot->stream) != hipSuccess)
        return mhcudaMemoryCopyError;
    }
    if (hipEventRecord(slot->finished, slot->stream) != hipSuccess ||
        hipEventSynchronize(slot->finished) != hipSuccess)
      return mhcudaRuntimeError;
    for (size_t i = 0; i < slot->reqs.size(); i++) {
      if (!out_pinned[i])
        std::memcpy(slot->reqs[i]->output, slot->host_out.ptr + out_off[i],
                    slot->reqs[i]->length * S8);
    }
    return mhcudaSuccess;
  }

  void worker_loop(Slot *slot) {
    (void)hipSetDevice(gen->device);
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
      slot->cv.wait(lock, [&] { return !slot->reqs.empty() || scheduler_done; });
      if (slot->reqs.empty()) break;  // scheduler finished and nothing assigned
      lock.unlock();
      MHCUDAResult st = run(slot);
      if (st != mhcudaSuccess) (void)hipStreamSynchronize(slot->stream);
      const Clock::time_point now = Clock::now();
      lock.lock();
      for (Request *r : slot->reqs) {
        r->status = st;
        r->done = true;
        r->latency_ms = std::chrono::duration<double, std::milli>(now - r->submitted).count();
        if (st == mhcudaSuccess) {
          latencies.push_back(r->latency_ms);
          units += r->nnz * gen->samples;
        }
      }
      if (now > last_done) last_done = now;
      slot->reqs.clear();
      slot->busy = false;
      cv_work.notify_one();
      cv_done.notify_all();
    }
  }

  void schedule_loop() {
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
      Slot *slot = nullptr;
      cv_work.wait(lock, [&] {
        if (stopping && pending.empty()) return true;
        if (pending.empty()) return false;
        for (auto &s : slots)
          if (!s->busy) return true;
        return false;
      });
      if (pending.empty()) break;  // stopping and drained
      for (auto &s : slots) {
        if (!s->busy) {
          slot = s.get();
          break;
        }
      }
      // admission: FIFO while the slot's memory budget and request cap allow
      uint64_t used = 0;
      while (!pending.empty() && slot->reqs.size() < cfg.max_batch_requests &&
             used + pending.front()->bytes <= slot_bytes) {
        used += pending.front()->bytes;
        slot->reqs.push_back(pending.front());
        pending.pop_front();
      }
      slot->busy = true;
      batches++;
      batched_requests += slot->reqs.size();
      slot->cv.notify_one();
    }
    scheduler_done = true;
    for (auto &s : slots) s->cv.notify_one();
  }
};

namespace {

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double frac = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - frac) + v[i + 1] * frac : v[i];
}

}  // namespace

extern "C" {

MHCUDAServer *mhcuda_serve_start(MinhashCudaGenerator *gen, const MHCUDAServeConfig *config,
                                 MHCUDAResult *status) {
  auto fail = [&](MHCUDAResult code) -> MHCUDAServer * {
    if (status) *status = code;
    return nullptr;
  };
  if (!gen || !config || config->slots == 0 || config->memory_budget == 0 ||
      config->max_batch_requests == 0)
    return fail(mhcudaInvalidArguments);
  if (gen->samples % 2 != 0) return fail(mhcudaInvalidArguments);
  if (hipSetDevice(gen->device) != hipSuccess) return fail(mhcudaNoSuchDevice);
  std::unique_ptr<MHCUDAServer> srv(new MHCUDAServer());
  srv->gen = gen;
  srv->cfg = *config;
  srv->slot_bytes = config->memory_budget / config->slots;
  for (uint32_t i = 0; i < config->slots; i++) {
    std::unique_ptr<Slot> s(new Slot());
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->finished, hipEventDisableTiming) != hipSuccess)
      return fail(mhcudaRuntimeError);
    if (s->buffer.reserve(srv->slot_bytes) != hipSuccess) return fail(mhcudaMemoryAllocationFailure);
    srv->slots.push_back(std::move(s));
  }
  MHCUDAServer *p = srv.get();
  for (auto &s : srv->slots) {
    Slot *sl = s.get();
    sl->worker = std::thread([p, sl] { p->worker_loop(sl); });
  }
  srv->scheduler = std::thread([p] { p->schedule_loop(); });
  return srv.release();
}

MHCUDAResult mhcuda_serve_submit(MHCUDAServer *srv, const float *weights, const uint32_t *cols,
                                 const uint32_t *rows, uint32_t length, uint32_t *output,
                                 uint64_t *ticket) {
  if (!srv || !weights || !cols || !rows || !output || !ticket || length == 0)
    return mhcudaInvalidArguments;
  for (uint32_t r = 0; r < length; r++) {
    if (rows[r + 1] < rows[r]) return mhcudaInvalidArguments;
  }
  std::unique_ptr<Request> req(new Request());
  req->weights = weights;
  req->cols = cols;
  req->rows = rows;
  req->length = length;
  req->output = output;
  req->nnz = rows[length] - rows[0];
  req->bytes = srv->request_bytes(req->nnz, length);
  req->submitted = Clock::now();
  std::lock_guard<std::mutex> lock(srv->mu);
  if (srv->stopping) return mhcudaInvalidArguments;
  req->ticket = srv->next_ticket++;
  *ticket = req->ticket;
  if (!srv->have_window) {
    srv->first_submit = req->submitted;
    srv->have_window = true;
  }
  if (req->bytes > srv->slot_bytes) {
    // larger than a slot's share of the memory budget: never admissible
    req->status = mhcudaMemoryAllocationFailure;
    req->done = true;
    srv->rejected++;
  } else {
    srv->pending.push_back(req.get());
    srv->cv_work.notify_one();
  }
  srv->requests.emplace(req->ticket, std::move(req));
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_serve_wait(MHCUDAServer *srv, uint64_t ticket, double *latency_ms) {
  if (!srv) return mhcudaInvalidArguments;
  std::unique_lock<std::mutex> lock(srv->mu);
  auto it = srv->requests.find(ticket);
  if (it == srv->requests.end()) return mhcudaInvalidArguments;
  Request *r = it->second.get();
  srv->cv_done.wait(lock, [&] { return r->done; });
  const MHCUDAResult st = r->status;
  if (latency_ms) *latency_ms = r->latency_ms;
  srv->requests.erase(it);
  return st;
}

MHCUDAResult mhcuda_serve_stats(MHCUDAServer *srv, MHCUDAServeStats *out, int reset) {
  if (!srv || !out) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(srv->mu);
  MHCUDAServeStats s{};
  s.completed = srv->latencies.size();
  s.rejected = srv->rejected;
  s.batches = srv->batches;
  s.mean_requests_per_batch = srv->batches ? (double)srv->batched_requests / srv->batches : 0;
  s.p50_ms = percentile(srv->latencies, 0.50);
  s.p99_ms = percentile(srv->latencies, 0.99);
  s.max_ms = srv->latencies.empty() ? 0 : *std::max_element(srv->latencies.begin(),
                                                            srv->latencies.end());
  double sum = 0;
  for (double x : srv->latencies) sum += x;
  s.mean_ms = s.completed ? sum / s.completed : 0;
  const double window =
      srv->have_window && s.completed
          ? std::chrono::duration<double>(srv->last_done - srv->first_submit).count()
          : 0;
  s.window_s = window;
  s.nnz_samples_per_sec = window > 0 ? (double)srv->units / window : 0;
  *out = s;
  if (reset) {
    srv->latencies.clear();
    srv->rejected = srv->batches = srv->batched_requests = srv->units = 0;
    srv->have_window = false;
  }
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_serve_stop(MHCUDAServer *srv) {
  if (!srv) return mhcudaSuccess;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stopping = true;
    srv->cv_work.notify_all();
  }
  srv->scheduler.join();
  for (auto &s : srv->slots) s->worker.join();
  (void)hipSetDevice(srv->gen->device);
  for (auto &s : srv->slots) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipEventDestroy(s->finished);
    (void)hipStreamDestroy(s->stream);
  }
  delete srv;
  return mhcudaSuccess;
}

}  // extern "C"
