// Host-side runtime internals shared by the runtime sources (not part of the
// C ABI): device buffers, descriptor staging and the generator object.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "host_copy.h"
#include "minhashcuda.h"
#include "wmh_kernels.h"

namespace wmh {

// Grow-only device buffer (reference: wrappers.h unique_devptr plus the
// grow-only logic of minhashcuda.cc:396-446).  reserve() never shrinks and
// does not preserve contents when it grows.
struct DevBuf {
  void *ptr = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  hipError_t reserve(size_t bytes);
};

// Grow-only pinned host buffer.
struct PinnedBuf {
  char *ptr = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  bool reserve(size_t bytes) {
    if (bytes <= cap) return true;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    if (hipHostMalloc((void **)&ptr, bytes, hipHostMallocDefault) != hipSuccess) return false;
    cap = bytes;
    return true;
  }
  ~PinnedBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
};

// True if `p` lies in pinned host memory known to the HIP runtime (then
// copies to and from it are DMA'd asynchronously, without staging).
inline bool is_pinned(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is unknown to the runtime
    return false;
  }
  return attr.type == hipMemoryTypeHost;
}

// A copy still owed from a pinned landing zone to a caller's output array.
using PendingCopy = CopyTask;

// One device buffer set of the pipelined host-memory path, with the events
// that order its upload, hashing and download.  Pageable inputs are copied
// into `staging` (by the generator's parallel copier) and uploaded from there;
// hashes bound for pageable outputs land in `landing` (a download into
// pageable memory would block the host until the hashing finished) and are
// copied out when the set is next reused or the call ends.
struct PieceBuffers {
  DevBuf weights, cols, rows, hashes;
  PinnedBuf staging, landing;
  std::vector<PendingCopy> owed;
  hipEvent_t uploaded = nullptr, hashed = nullptr, done = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreateWithFlags(&uploaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&hashed, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    return e;
  }
  void release() {
    if (uploaded) (void)hipEventDestroy(uploaded);
    if (hashed) (void)hipEventDestroy(hashed);
    if (done) (void)hipEventDestroy(done);
    uploaded = hashed = done = nullptr;
  }
  // Completes the copies owed by the previous use of this set (its `done`
  // event must have completed).
  void settle(HostCopier &copier) {
    copier.copy(owed.data(), owed.size());
    owed.clear();
  }
};

// Descriptor staging of persistent launches: pinned host copy, device copy
// plus control block, the event marking the end of the last upload out of
// the pinned copy (so that the host does not overwrite it too early), and the
// event marking the end of the last launch that read the device copy (the
// next call's stream waits for it before overwriting the descriptors, so
// calls on different streams never share them while in flight).
struct PlanCtx {
  DevBuf plan;
  MatrixDesc *desc_host = nullptr;
  size_t cap = 0;
  hipEvent_t copied = nullptr;
  hipEvent_t released = nullptr;
  PlanCtx() = default;
  PlanCtx(const PlanCtx &) = delete;
  PlanCtx &operator=(const PlanCtx &) = delete;
  ~PlanCtx();
};

// A tuning chosen by the autotuner for one batch shape.
struct TunedChoice {
  int kernel = 0;
  Tuning tuning{2, 2048, 12, kDefaultStripeMib};
  float ms = 0;           // time of the timed prefix
  uint64_t sample_nnz = 0;
};

// Captured enqueue sequence of the device-pointer calls, replayed while the
// call (descriptors, kernel, tuning) stays the same.
struct GraphCache {
  bool enabled = false;
  std::vector<MatrixDesc> key;
  int kernel = 0;
  Tuning tuning{0, 0, 0};
  bool fast_div = false;
  CodeTable bounds{nullptr, 0.0f, 0.0f};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
  hipEvent_t launched = nullptr;  // recorded after every replay
  PlanCtx plan;
  uint64_t captures = 0;
  GraphCache() = default;
  GraphCache(const GraphCache &) = delete;
  GraphCache &operator=(const GraphCache &) = delete;
  ~GraphCache();
  void release();
};

// Grow-only device buffers of LSH candidate generation (lsh.hip).
struct LshBuffers {
  DevBuf keys_in, rows_in, keys_sorted, rows_sorted, counts, offsets, pairs, pairs_sorted, temp,
      scalars;
};

}  // namespace wmh

struct MinhashCudaGenerator_ {
  uint32_t dim = 0;
  uint16_t samples = 0;
  int verbosity = 0;
  int device = -1;
  uint32_t sp = 0;
  int num_cus = 0;
  int num_xcds = 1;
  float *params = nullptr;            // [dim][3][sp]
  // where the tables came from: nothing yet (deferred init), the on-device
  // generator with `seed`, or host tables (assign / a parameter file)
  enum { kParamsNone = 0, kParamsSeeded = 1, kParamsTables = 2 };
  int params_source = kParamsNone;
  uint32_t seed = 0;
  // every r of the table lies in [2^-60, 2^60], so the kernels may divide
  // with the short correctly rounded sequence (refresh_division_mode)
  bool fast_div = false;
  // byte-coded bound table of the skip bound ([dim][sp] bytes, wmh_bound.h;
  // null when MHCUDA_SKIP_BOUND=0 or it could not be allocated) followed in
  // the same allocation by its coarse codes (`coarse`, [dim][coarse_row_bytes]
  // bytes, null unless sp % 8 == 0), with the scale and offset of its thresholds,
  // `bounded` while it matches the parameters and they pass the bound's range
  // check, and `skip` the user's switch (mhcuda_set_skip)
  uint8_t *bounds = nullptr;
  uint8_t *coarse = nullptr;
  // the item table of the bounded stripe pieces (wmh::launch_build_items;
  // null when MHCUDA_ITEM_TABLE=0, without coarse codes, at 4 GiB or more, or
  // when it could not be allocated), rebuilt with the bound table
  float *items = nullptr;
  float code_scale = 0.0f, code_t0 = 0.0f;
  bool bounded = false;
  bool skip = true;
  wmh::DevBuf range_flag;
  hipStream_t stream = nullptr;
  // recorded on the legacy default stream by device calls given no stream,
  // so that `stream` (non-blocking) waits for work already issued there
  hipEvent_t after_null = nullptr;
  std::mutex mu;
  // grow-only caches, reference minhashcuda.cc:27-31
  wmh::DevBuf staging;
  wmh::PlanCtx plan;  // descriptors of the calls made on `stream`
  // host-memory calls: two device buffer sets used alternately, their own
  // upload and download streams, and the size limits of one piece
  wmh::PieceBuffers pieces[2];
  std::unique_ptr<wmh::HostCopier> copier;  // host-memory paths (created on first use)
  hipStream_t copy_stream = nullptr, d2h_stream = nullptr;
  uint64_t piece_nnz = 1ull << 28;
  uint64_t piece_rows = 1ull << 20;
  // 12 waves per CU by default: 3 per SIMD leave room in the register file
  // for the runtime's copy kernels of overlapped uploads and downloads (the
  // 1- and 2-samples-per-lane hashing kernels use up to 128 VGPRs, so 4 waves
  // per SIMD fill it; the 4-samples-per-lane one uses 134, so 3 waves per
  // SIMD are its full occupancy and still leave 104 registers free)
  wmh::Tuning tuning{2, 2048, 12, wmh::kDefaultStripeMib};
  bool seg_auto = true;  // let small batches use smaller work items
  wmh::GraphCache graph;  // hipGraph replay of device-pointer calls (opt-in)
  std::map<uint64_t, wmh::TunedChoice> tuned;  // autotuner cache by batch shape
  wmh::DevBuf tune_buf;
  wmh::LshBuffers lsh;  // mhcuda_lsh_candidates
  int kernel = mhcudaKernelAuto;
  // optional timing of the hashing kernel(s) of the last call
  bool timing = false;
  bool timed = false;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  // serving loops running on this generator: while any runs, calls that
  // change the parameters or the tuning are refused (the servers' launches
  // read them without taking `mu`)
  std::atomic<int> servers{0};
  // armed failures (mhcuda_inject_fault): the n-th device allocation / copy
  // of the host-memory paths from now fails; 0 = disarmed
  std::atomic<uint32_t> fail_alloc{0}, fail_copy{0};

  MinhashCudaGenerator_() = default;
  MinhashCudaGenerator_(const MinhashCudaGenerator_ &) = delete;
  MinhashCudaGenerator_ &operator=(const MinhashCudaGenerator_ &) = delete;
  ~MinhashCudaGenerator_();
};

namespace wmh {

// Enqueues the chosen hashing kernel over device-resident matrices on `s`,
// staging persistent-launch descriptors through `ctx`.  The caller owns the
// synchronization of `ctx` with other streams.
// `capturing`: the stream is being captured into a graph, so nothing may
// allocate or wait on the host (ctx's buffers must already be reserved).
// `timed`: record the generator's timing events (when timing is on) and mark
// the call as the one mhcuda_kernel_time reports; callers that do not hold
// the generator's lock (the serving loop) pass false.
MHCUDAResult hash_device_batch(MinhashCudaGenerator *gen, std::vector<MatrixDesc> &mats,
                               int kernel, hipStream_t s, PlanCtx *ctx, bool capturing,
                               bool timed = true);

// The hipGraph-replayed form of hash_device_batch for device-pointer calls.
MHCUDAResult hash_graph(MinhashCudaGenerator *gen, std::vector<MatrixDesc> &mats, int kernel,
                        hipStream_t s);

// Checks the tables on the device after the parameters change, sets
// gen->fast_div and rebuilds the bound table (gen->bounded) when the ranges
// allow (blocks until done).
MHCUDAResult refresh_division_mode(MinhashCudaGenerator *gen);

// The bound table for a persistent launch, or null when the skip bound is
// off or unusable for these tables.
inline CodeTable launch_bounds_of(const MinhashCudaGenerator *gen) {
  if (!(gen->skip && gen->bounded)) return CodeTable{nullptr, 0.0f, 0.0f};
  return CodeTable{gen->bounds, gen->code_scale, gen->code_t0, gen->coarse, gen->items};
}

inline bool same_table(const CodeTable &x, const CodeTable &y) {
  return x.codes == y.codes && x.scale == y.scale && x.t0 == y.t0 && x.coarse == y.coarse &&
         x.items == y.items;
}

// Device-pointer calls given no stream run on gen->stream, which does not
// synchronize with the legacy default stream by itself: this makes it wait
// for everything already issued there (the caller's inputs may be written
// by work on the default stream, torch's default among them).
inline hipError_t follow_default_stream(MinhashCudaGenerator *gen) {
  hipError_t e = hipEventRecord(gen->after_null, nullptr);
  if (e == hipSuccess) e = hipStreamWaitEvent(gen->stream, gen->after_null, 0);
  return e;
}

// Armed failures of the host-memory paths (mhcuda_inject_fault): each call
// counts one device allocation / copy and returns true for the one that must
// fail.
inline bool armed_failure(std::atomic<uint32_t> &left) {
  uint32_t n = left.load();
  while (n != 0) {
    if (left.compare_exchange_weak(n, n - 1)) return n == 1;
  }
  return false;
}
inline hipError_t reserve_checked(MinhashCudaGenerator *gen, DevBuf &buf, size_t bytes) {
  if (armed_failure(gen->fail_alloc)) return hipErrorOutOfMemory;
  return buf.reserve(bytes);
}
inline hipError_t copy_checked(MinhashCudaGenerator *gen, void *dst, const void *src,
                               size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  if (armed_failure(gen->fail_copy)) return hipErrorInvalidValue;
  return hipMemcpyAsync(dst, src, bytes, kind, s);
}

}  // namespace wmh
