// Continuous-batching serving loop over one generator (no counterpart in the
// reference, which only has the synchronous mhcuda_calc).
//
// Callers submit CSR requests held in host memory and receive tickets.  A
// scheduler thread admits pending requests (FIFO) into a free slot while
// their device bytes fit one contiguous free range of the server's device
// pool (the whole memory budget, allocated once), uploads them, and enqueues
// one persistent hashing launch for the whole group plus the download of
// every output on the slot's stream.  Each slot has a worker thread that
// stages the uploads (pinned arrays go straight to DMA, pageable ones through
// a fixed ring of pinned blocks), waits for its launch and copies the hashes
// out, so slots overlap both on the device and on the host; a request that
// arrives while others are being hashed is admitted as soon as a slot and
// pool memory free up.  Each request's latency is submit -> hashes in host
// memory.  Each hash equals what a lone mhcuda_calc of that request gives:
// rows are hashed independently of whatever shares the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
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
  Clock::time_point submitted, admitted;
  MHCUDAResult status = mhcudaSuccess;
  bool done = false;
  double latency_ms = 0;
};

uint64_t align256(uint64_t x) { return (x + 255u) & ~(uint64_t)255u; }

// First-fit allocator over the server's device pool (offsets, 256-byte
// granules), with neighbouring free ranges merged on release.
class PoolRanges {
 public:
  void reset(uint64_t bytes) {
    free_.clear();
    if (bytes) free_[0] = bytes;
  }
  uint64_t largest() const {
    uint64_t m = 0;
    for (const auto &f : free_) m = std::max(m, f.second);
    return m;
  }
  // offset of a range of `bytes`, or ~0 when no free range is large enough
  uint64_t take(uint64_t bytes) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const uint64_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      return off;
    }
    return ~0ull;
  }
  void give(uint64_t off, uint64_t bytes) {
    if (!bytes) return;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }

 private:
  std::map<uint64_t, uint64_t> free_;
};

// A ring of pinned host blocks through which a slot stages pageable inputs
// and outputs: allocated once at server start (no pinned allocation, and no
// implicit device synchronization of freeing pinned memory, on the request
// path).  Consecutive small arrays share a block.  All copies through the
// ring run on the slot's stream; block b's event is recorded after the last
// copy into or out of it, and the host touches the block again only once
// that event has completed.
constexpr size_t kStageBlock = 8u << 20;
constexpr int kStageBlocks = 4;

struct StageRing {
  wmh::PinnedBuf mem;
  wmh::HostCopier *copier = nullptr;  // the server's (copies out of the blocks)
  hipEvent_t ev[kStageBlocks] = {};
  int next = 0;      // block taken after the current one
  int cur = -1;      // block being filled (-1: none)
  size_t fill = 0;   // bytes used in the current block
  // landing zones of downloads not yet copied out, per block
  struct Landing {
    size_t at;
    char *dst;
    size_t bytes;
  };
  std::vector<Landing> owed[kStageBlocks];

  bool create() {
    if (!mem.reserve(kStageBlock * kStageBlocks)) return false;
    for (auto &e : ev)
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
    return true;
  }
  void destroy() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  char *block(int b) { return mem.ptr + (size_t)b * kStageBlock; }
  // copies out the downloads landed in block b (its event must be complete)
  void settle(int b) {
    std::vector<wmh::CopyTask> out;
    out.reserve(owed[b].size());
    for (const Landing &l : owed[b]) out.push_back(wmh::CopyTask{l.dst, block(b) + l.at, l.bytes});
    copier->copy(out.data(), out.size());
    owed[b].clear();
  }
  // true when the next room() opens another block (closing the current one)
  bool switching() const { return cur < 0 || fill == kStageBlock; }
  // closes the current block: its event marks the end of its copies
  bool close(hipStream_t s) {
    if (cur < 0) return true;
    const int b = cur;
    cur = -1;
    return hipEventRecord(ev[b], s) == hipSuccess;
  }
  // Room for up to `want` bytes in the current block (opening the next block
  // when it is full); sets *part to the bytes granted.
  char *room(size_t want, size_t *part, hipStream_t s) {
    if (cur < 0 || fill == kStageBlock) {
      if (!close(s)) return nullptr;
      const int b = next;
      next = (b + 1) % kStageBlocks;
      if (hipEventSynchronize(ev[b]) != hipSuccess) return nullptr;
      settle(b);
      cur = b;
      fill = 0;
    }
    *part = std::min(want, kStageBlock - fill);
    char *p = block(cur) + fill;
    fill = std::min(kStageBlock, (fill + *part + 63) & ~(size_t)63);
    return p;
  }
  // waits for every block and copies out all landed downloads
  bool drain(hipStream_t s) {
    if (!close(s)) return false;
    for (int b = 0; b < kStageBlocks; b++) {
      if (hipEventSynchronize(ev[b]) != hipSuccess) return false;
      settle(b);
    }
    return true;
  }
  void reset() {
    cur = -1;
    for (auto &o : owed) o.clear();
  }
};

// One launch in flight: its own stream, descriptor staging, pinned staging
// ring and worker thread.  The worker uploads the admitted requests (into its
// range of the server's device pool), enqueues their hashing and downloads,
// waits for them and hands the hashes back, so host-side copies of different
// slots run in parallel while their launches overlap on the device.
// Host-clock phases of one launch (Clock::now() at each boundary).
struct Phases {
  Clock::time_point start, uploaded, finished, drained;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t finished = nullptr;
  wmh::PlanCtx plan;
  StageRing ring;
  uint64_t pool_off = 0, pool_bytes = 0;  // this group's range of the pool
  std::vector<Request *> reqs;
  // pageable inputs staged into the current ring block, not yet copied or
  // uploaded: (device destination, staging block, host source, bytes)
  struct Staged {
    char *dev;
    char *stage;
    const char *src;
    size_t bytes;
  };
  std::vector<Staged> staged;
  Phases phases;
  bool busy = false;      // requests assigned and not yet retired
  std::thread worker;
  std::condition_variable cv;
};

}  // namespace

struct MHCUDAServer_ {
  MinhashCudaGenerator *gen = nullptr;
  MHCUDAServeConfig cfg{};
  wmh::DevBuf pool;        // the memory budget, shared by every slot
  wmh::HostCopier copier;  // staging copies of every slot, in parallel
  PoolRanges ranges;       // its free ranges (guarded by mu)
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work;   // scheduler: pending request, free slot or freed memory
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
  uint64_t timed_groups = 0;  // successful launches behind the phase sums
  double queue_sum = 0, stage_in_sum = 0, device_sum = 0, stage_out_sum = 0;
  Clock::time_point first_submit, last_done;
  bool have_window = false;

  uint64_t request_bytes(uint64_t nnz, uint32_t length) const {
    return align256(nnz * 4) * 2 + align256(((uint64_t)length + 1) * 4) +
           align256((uint64_t)length * gen->samples * 8);
  }

  // Uploads `n` bytes from host `src` to device `dst` on the slot's stream:
  // pinned sources directly, pageable ones through the staging ring (their
  // copies into a block are made together, in parallel, before the block is
  // closed or the launch is enqueued: flush_staged).
  bool upload(Slot *slot, char *dst, const char *src, size_t n, bool pinned) {
    if (pinned) return wmh::copy_checked(gen, dst, src, n, hipMemcpyHostToDevice,
                                         slot->stream) == hipSuccess;
    for (size_t done = 0; done < n;) {
      if (slot->ring.switching() && !flush_staged(slot)) return false;
      size_t part = 0;
      char *stage = slot->ring.room(n - done, &part, slot->stream);
      if (!stage) return false;
      slot->staged.push_back(Slot::Staged{dst + done, stage, src + done, part});
      done += part;
    }
    return true;
  }

  // Copies the staged pageable inputs into their ring block (in parallel) and
  // enqueues their uploads.
  bool flush_staged(Slot *slot) {
    if (slot->staged.empty()) return true;
    std::vector<wmh::CopyTask> in;
    in.reserve(slot->staged.size());
    for (const Slot::Staged &st : slot->staged)
      in.push_back(wmh::CopyTask{st.stage, st.src, st.bytes});
    copier.copy(in.data(), in.size());
    bool ok = true;
    for (const Slot::Staged &st : slot->staged)
      ok = ok && wmh::copy_checked(gen, st.dev, st.stage, st.bytes, hipMemcpyHostToDevice,
                                   slot->stream) == hipSuccess;
    slot->staged.clear();
    return ok;
  }

  // Downloads `n` bytes of hashes from device `src` to host `dst`: pinned
  // destinations directly, pageable ones into the staging ring, copied out
  // when their block is reused or the group drains.
  bool download(Slot *slot, char *dst, const char *src, size_t n, bool pinned) {
    if (pinned) return wmh::copy_checked(gen, dst, src, n, hipMemcpyDeviceToHost,
                                         slot->stream) == hipSuccess;
    StageRing &ring = slot->ring;
    for (size_t done = 0; done < n;) {
      size_t part = 0;
      char *stage = ring.room(n - done, &part, slot->stream);
      if (!stage) return false;
      if (wmh::copy_checked(gen, stage, src + done, part, hipMemcpyDeviceToHost,
                            slot->stream) != hipSuccess)
        return false;
      ring.owed[ring.cur].push_back(
          StageRing::Landing{(size_t)(stage - ring.block(ring.cur)), dst + done, part});
      done += part;
    }
    return true;
  }

  // Uploads, hashes and downloads the requests of `slot`; returns when the
  // hashes are in the requests' output arrays.
  MHCUDAResult run(Slot *slot) {
    slot->phases.start = Clock::now();
    char *base = (char *)pool.ptr + slot->pool_off;
    const uint64_t S8 = (uint64_t)gen->samples * 8;
    uint64_t off = 0;
    std::vector<wmh::MatrixDesc> descs;
    descs.reserve(slot->reqs.size());
    for (Request *r : slot->reqs) {
      const uint32_t first = r->rows[0];
      char *dw = base + off;
      off += align256(r->nnz * 4);
      char *dc = base + off;
      off += align256(r->nnz * 4);
      char *dr = base + off;
      off += align256(((uint64_t)r->length + 1) * 4);
      char *dh = base + off;
      off += align256(r->length * S8);
      if (r->nnz &&
          (!upload(slot, dw, (const char *)(r->weights + first), r->nnz * 4,
                   wmh::is_pinned(r->weights + first)) ||
           !upload(slot, dc, (const char *)(r->cols + first), r->nnz * 4,
                   wmh::is_pinned(r->cols + first))))
        return mhcudaMemoryCopyError;
      if (!upload(slot, dr, (const char *)r->rows, ((uint64_t)r->length + 1) * 4,
                  wmh::is_pinned(r->rows)))
        return mhcudaMemoryCopyError;
      descs.push_back(wmh::MatrixDesc{(const float *)dw, (const uint32_t *)dc,
                                      (const uint32_t *)dr, (uint32_t *)dh, r->length, first,
                                      0u, 0u, 0u, 0u});
    }
    if (!flush_staged(slot)) return mhcudaMemoryCopyError;
    slot->phases.uploaded = Clock::now();
    // (untimed: the generator's timing state belongs to the calls that hold
    // its lock; serving keeps its own latency statistics)
    MHCUDAResult res = wmh::hash_device_batch(gen, descs, mhcudaKernelPersistent, slot->stream,
                                              &slot->plan, false, false);
    if (res != mhcudaSuccess) return res;
    for (size_t i = 0; i < slot->reqs.size(); i++) {
      Request *r = slot->reqs[i];
      if (!download(slot, (char *)r->output, (const char *)descs[i].output, r->length * S8,
                    wmh::is_pinned(r->output)))
        return mhcudaMemoryCopyError;
    }
    if (hipEventRecord(slot->finished, slot->stream) != hipSuccess ||
        hipEventSynchronize(slot->finished) != hipSuccess)
      return mhcudaRuntimeError;
    slot->phases.finished = Clock::now();
    if (!slot->ring.drain(slot->stream)) return mhcudaRuntimeError;
    slot->phases.drained = Clock::now();
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
      // after a failure nothing of this group may still be in flight when its
      // pool range and staging blocks are reused
      if (st != mhcudaSuccess) {
        (void)hipStreamSynchronize(slot->stream);
        (void)slot->ring.close(slot->stream);
        slot->ring.reset();
        slot->staged.clear();
      }
      const Clock::time_point now = Clock::now();
      lock.lock();
      auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
      };
      for (Request *r : slot->reqs) {
        r->status = st;
        r->done = true;
        r->latency_ms = ms(r->submitted, now);
        if (st == mhcudaSuccess) {
          latencies.push_back(r->latency_ms);
          units += r->nnz * gen->samples;
          queue_sum += ms(r->submitted, r->admitted);
        }
      }
      if (st == mhcudaSuccess) {
        const Phases &ph = slot->phases;
        timed_groups++;
        stage_in_sum += ms(ph.start, ph.uploaded);
        device_sum += ms(ph.uploaded, ph.finished);
        stage_out_sum += ms(ph.finished, ph.drained);
      }
      if (now > last_done) last_done = now;
      ranges.give(slot->pool_off, slot->pool_bytes);
      slot->pool_bytes = 0;
      slot->reqs.clear();
      slot->busy = false;
      cv_work.notify_one();
      cv_done.notify_all();
    }
  }

  Slot *free_slot() {
    for (auto &s : slots)
      if (!s->busy) return s.get();
    return nullptr;
  }

  void schedule_loop() {
    std::unique_lock<std::mutex> lock(mu);
    for (;;) {
      // admissible: a free slot, and a free pool range for the oldest request
      // (every admitted request is at most the budget, so once the slots
      // drain the whole pool is one free range and it fits)
      cv_work.wait(lock, [&] {
        if (stopping && pending.empty()) return true;
        return !pending.empty() && free_slot() != nullptr &&
               pending.front()->bytes <= ranges.largest();
      });
      if (pending.empty()) break;  // stopping and drained
      Slot *slot = free_slot();
      // admission: FIFO while the largest free range and the request cap allow
      const uint64_t room = ranges.largest();
      uint64_t used = 0;
      const Clock::time_point now = Clock::now();
      while (!pending.empty() && slot->reqs.size() < cfg.max_batch_requests &&
             used + pending.front()->bytes <= room) {
        used += pending.front()->bytes;
        pending.front()->admitted = now;
        slot->reqs.push_back(pending.front());
        pending.pop_front();
      }
      slot->pool_off = ranges.take(used);
      slot->pool_bytes = used;
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
  if (!gen || !config || config->slots == 0 || config->slots > 64 ||
      config->memory_budget == 0 || config->max_batch_requests == 0)
    return fail(mhcudaInvalidArguments);
  if (gen->samples % 2 != 0) return fail(mhcudaInvalidArguments);
  if (hipSetDevice(gen->device) != hipSuccess) return fail(mhcudaNoSuchDevice);
  std::unique_ptr<MHCUDAServer> srv(new MHCUDAServer());
  auto fail_after = [&](MHCUDAResult code) -> MHCUDAServer * {
    for (auto &s : srv->slots) {  // no thread was started yet
      s->ring.destroy();
      if (s->finished) (void)hipEventDestroy(s->finished);
      if (s->stream) (void)hipStreamDestroy(s->stream);
    }
    return fail(code);
  };
  srv->gen = gen;
  srv->cfg = *config;
  const uint64_t budget = config->memory_budget & ~(uint64_t)255u;
  if (budget == 0) return fail(mhcudaInvalidArguments);
  if (wmh::reserve_checked(gen, srv->pool, budget) != hipSuccess)
    return fail(mhcudaMemoryAllocationFailure);
  srv->ranges.reset(budget);
  for (uint32_t i = 0; i < config->slots; i++) {
    srv->slots.emplace_back(new Slot());
    Slot *s = srv->slots.back().get();
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s->finished, hipEventDisableTiming) != hipSuccess)
      return fail_after(mhcudaRuntimeError);
    if (!s->ring.create()) return fail_after(mhcudaMemoryAllocationFailure);
    s->ring.copier = &srv->copier;
  }
  {
    // from here on, calls that change the tables or the tuning are refused
    std::lock_guard<std::mutex> lock(gen->mu);
    gen->servers++;
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
  if (req->bytes > (srv->cfg.memory_budget & ~(uint64_t)255u)) {
    // larger than the whole memory budget: never admissible
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
  s.queue_ms = s.completed ? srv->queue_sum / s.completed : 0;
  const double groups = (double)srv->timed_groups;
  s.stage_in_ms = groups > 0 ? srv->stage_in_sum / groups : 0;
  s.device_ms = groups > 0 ? srv->device_sum / groups : 0;
  s.stage_out_ms = groups > 0 ? srv->stage_out_sum / groups : 0;
  *out = s;
  if (reset) {
    srv->latencies.clear();
    srv->rejected = srv->batches = srv->batched_requests = srv->units = 0;
    srv->have_window = false;
    srv->timed_groups = 0;
    srv->queue_sum = srv->stage_in_sum = srv->device_sum = srv->stage_out_sum = 0;
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
    s->ring.destroy();
    (void)hipEventDestroy(s->finished);
    (void)hipStreamDestroy(s->stream);
  }
  {
    std::lock_guard<std::mutex> lock(srv->gen->mu);
    srv->gen->servers--;
  }
  delete srv;
  return mhcudaSuccess;
}

}  // extern "C"
