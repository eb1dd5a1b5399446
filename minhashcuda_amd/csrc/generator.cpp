// Host runtime of the MI355X Weighted MinHash engine: the generator object,
// device selection, parameter I/O, the grow-only device buffer cache and the
// C ABI of include/minhashcuda.h.
//
// Reference counterparts: minhashcuda.cc (generator, init/calc/fini,
// buffers), private.h (error and multi-device macros) and wrappers.h (RAII
// device pointers).  Differences by design:
//   * one GPU (the first gfx950 device the bitmask selects);
//   * parameters live in the [dim][3][sp] column-block layout (see
//     wmh_kernels.h); assign/retrieve re-lay them out on the device;
//   * no host bin-packing: the kernels schedule rows themselves;
//   * every call on one generator is serialized by a mutex (the reference
//     mutates the generator through const pointers without locking);
//   * `dim * samples` is computed in 64 bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "minhashcuda.h"
#include "wmh_bound.h"
#include "wmh_kernels.h"
#include "wmh_runtime.h"

#define INFO(...) do { if (verbosity > 0) { printf(__VA_ARGS__); fflush(stdout); } } while (false)
#define DEBUG(...) do { if (verbosity > 1) { printf(__VA_ARGS__); fflush(stdout); } } while (false)

// Evaluates a HIP call; on failure logs it and returns `ret`.
#define HIPCH(call, ret)                                                      \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      INFO("%s:%d: %s -> %s\n", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return ret;                                                             \
    }                                                                         \
  } while (false)

namespace wmh {

DevBuf::~DevBuf() {
  if (ptr) (void)hipFree(ptr);
}

hipError_t DevBuf::reserve(size_t bytes) {
  if (bytes <= cap) return hipSuccess;
  if (ptr) {
    (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
  hipError_t e = hipMalloc(&ptr, bytes);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return e;
  }
  cap = bytes;
  return hipSuccess;
}

}  // namespace wmh

using wmh::DevBuf;

MinhashCudaGenerator_::~MinhashCudaGenerator_() {
  if (after_null) (void)hipEventDestroy(after_null);
  if (ev_begin) (void)hipEventDestroy(ev_begin);
  if (ev_end) (void)hipEventDestroy(ev_end);
  if (device >= 0) (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  if (copy_stream) (void)hipStreamSynchronize(copy_stream);
  if (d2h_stream) (void)hipStreamSynchronize(d2h_stream);
  for (auto &p : pieces) p.release();
  if (stream) (void)hipStreamDestroy(stream);
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
  if (d2h_stream) (void)hipStreamDestroy(d2h_stream);
  if (params) (void)hipFree(params);
  if (bounds) (void)hipFree(bounds);
  if (items) (void)hipFree(items);
}

namespace {

// Selects the device: the bitmask semantics of minhashcuda.cc:43-79 (0 = all
// devices, devices that fail validation are dropped), with the
// compute-capability check replaced by a gfx950 architecture check.  The
// engine drives one GPU: the first valid device is returned, -1 if none.
int select_device(uint32_t devices, int verbosity) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return -1;
  }
  if (devices == 0) devices = count >= 32 ? 0xffffffffu : ((1u << count) - 1u);
  int chosen = -1;
  for (int dev = 0; devices && dev < 32; dev++, devices >>= 1) {
    if (!(devices & 1u)) continue;
    if (dev >= count) {
      INFO("device %d does not exist\n", dev);
      continue;
    }
    hipDeviceProp_t props;
    if (hipGetDeviceProperties(&props, dev) != hipSuccess) {
      (void)hipGetLastError();
      INFO("failed to query device %d\n", dev);
      continue;
    }
    if (std::strncmp(props.gcnArchName, "gfx950", 6) != 0) {
      INFO("device %d is %s, not gfx950: skipped\n", dev, props.gcnArchName);
      continue;
    }
    if (chosen < 0) {
      chosen = dev;
    } else {
      INFO("device %d also selected: the engine hashes on one GPU (#%d)\n", dev, chosen);
    }
  }
  return chosen;
}

MHCUDAResult print_memory_stats(int dev) {
  size_t free_bytes = 0, total_bytes = 0;
  if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) return mhcudaRuntimeError;
  printf("GPU #%d memory: used %zu bytes (%.1f%%), free %zu bytes, total %zu bytes\n",
         dev, total_bytes - free_bytes, (total_bytes - free_bytes) * 100.0 / total_bytes,
         free_bytes, total_bytes);
  return mhcudaSuccess;
}

// Columns per re-layout chunk: the staging table is samples x chunk floats,
// at most 256 MiB.
uint32_t relayout_chunk(uint32_t dim, uint32_t samples) {
  uint64_t c = (256ull << 20) / (4ull * samples);
  c = std::max<uint64_t>(64, c & ~63ull);
  return (uint32_t)std::min<uint64_t>(c, dim);
}

}  // namespace

namespace wmh {

// Enqueues the persistent batched path on `s`: descriptors go through the
// pinned staging buffer into device memory next to the control block.
// Device plan buffer of a persistent launch: descriptors, control block and,
// with column stripes, the striped-row list (one slot per row of the batch).
struct PlanLayout {
  size_t ctl_off, list_off, bytes;
};

static PlanLayout plan_layout(const MinhashCudaGenerator *gen,
                              const std::vector<MatrixDesc> &mats) {
  PlanLayout l;
  const size_t desc_bytes = sizeof(MatrixDesc) * mats.size();
  l.ctl_off = (desc_bytes + 255) & ~(size_t)255;
  l.list_off = (l.ctl_off + sizeof(BatchCtl) + 255) & ~(size_t)255;
  uint32_t stripes = 1, min_nnz = 0;
  stripe_shape(gen->dim, gen->samples, gen->tuning.stripe_mib, &stripes, &min_nnz);
  uint64_t rows = 0;
  if (stripes > 1)
    for (const MatrixDesc &m : mats) rows += m.length;
  l.bytes = l.list_off + sizeof(StripedRow) * std::max<uint64_t>(rows, 1);
  return l;
}

static MHCUDAResult hash_persistent(MinhashCudaGenerator *gen, std::vector<MatrixDesc> &mats,
                                    hipStream_t s, PlanCtx *ctx, bool capturing, bool timed) {
  int verbosity = gen->verbosity;
  const uint32_t count = (uint32_t)mats.size();
  // the plan kernel keeps the item count below 2^31 + count (it enlarges the
  // items of huge batches), which must fit the 32-bit item indices
  if (count > (1u << 30)) return mhcudaInvalidArguments;
  uint64_t chunks = 0;
  for (MatrixDesc &m : mats) {
    m.item_begin = m.n_items = 0;
    m.chunk_begin = (uint32_t)chunks;
    chunks += (m.length + 63u) / 64u;
  }
  if (chunks >= (1ull << 32)) return mhcudaInvalidArguments;
  const size_t desc_bytes = sizeof(MatrixDesc) * count;
  const PlanLayout layout = plan_layout(gen, mats);
  const size_t ctl_off = layout.ctl_off;
  // While a graph is being captured nothing may allocate or wait here: the
  // graph's own staging is reserved and idle before capture starts.
  if (!capturing) {
    if (!ctx->copied)
      HIPCH(hipEventCreateWithFlags(&ctx->copied, hipEventDisableTiming), mhcudaRuntimeError);
    if (!ctx->released)
      HIPCH(hipEventCreateWithFlags(&ctx->released, hipEventDisableTiming), mhcudaRuntimeError);
    HIPCH(hipEventSynchronize(ctx->copied), mhcudaRuntimeError);
  }
  // a grown plan buffer is a new allocation: the old one is freed only after
  // the launches reading it have finished
  if (layout.bytes > ctx->plan.cap && !capturing)
    HIPCH(hipEventSynchronize(ctx->released), mhcudaRuntimeError);
  HIPCH(ctx->plan.reserve(layout.bytes), mhcudaMemoryAllocationFailure);
  if (ctx->cap < count) {
    if (ctx->desc_host) (void)hipHostFree(ctx->desc_host);
    ctx->desc_host = nullptr;
    ctx->cap = 0;
    HIPCH(hipHostMalloc((void **)&ctx->desc_host, desc_bytes, hipHostMallocDefault),
          mhcudaMemoryAllocationFailure);
    ctx->cap = count;
  }
  std::memcpy(ctx->desc_host, mats.data(), desc_bytes);
  char *dev = (char *)ctx->plan.ptr;
  // the previous launch through this staging (maybe on another stream) must
  // be done with the descriptors, queue heads and striped-row list
  if (!capturing) HIPCH(hipStreamWaitEvent(s, ctx->released, 0), mhcudaRuntimeError);
  HIPCH(hipMemcpyAsync(dev, ctx->desc_host, desc_bytes, hipMemcpyHostToDevice, s),
        mhcudaMemoryCopyError);
  if (!capturing) HIPCH(hipEventRecord(ctx->copied, s), mhcudaRuntimeError);
  const bool with_events = timed && gen->timing && !capturing;
  HIPCH(launch_persistent(gen->params, gen->dim, gen->samples, (MatrixDesc *)dev, count,
                          (uint32_t)chunks, (BatchCtl *)(dev + ctl_off),
                          (StripedRow *)(dev + layout.list_off), gen->tuning, gen->seg_auto,
                          gen->num_cus, gen->num_xcds, gen->fast_div, s,
                          with_events ? gen->ev_begin : nullptr,
                          with_events ? gen->ev_end : nullptr, launch_bounds_of(gen)),
        mhcudaRuntimeError);
  if (!capturing) HIPCH(hipEventRecord(ctx->released, s), mhcudaRuntimeError);
  if (timed) gen->timed = gen->timing;
  return mhcudaSuccess;
}

// Launches the chosen hashing kernel over device-resident matrices on `s`.
PlanCtx::~PlanCtx() {
  if (copied) {
    (void)hipEventSynchronize(copied);
    (void)hipEventDestroy(copied);
  }
  if (released) {
    (void)hipEventSynchronize(released);
    (void)hipEventDestroy(released);
  }
  if (desc_host) (void)hipHostFree(desc_host);
}

MHCUDAResult hash_device_batch(MinhashCudaGenerator *gen, std::vector<MatrixDesc> &mats,
                               int kernel, hipStream_t s, PlanCtx *ctx, bool capturing,
                               bool timed) {
  int verbosity = gen->verbosity;
  if (gen->samples % 2 != 0) {
    // the reference finds no even sample_delta for odd counts
    // (minhashcuda.cc:602-618) and rejects the call
    return mhcudaInvalidArguments;
  }
  if (kernel == mhcudaKernelAuto) kernel = gen->kernel;
  if (kernel == mhcudaKernelAuto) kernel = mhcudaKernelPersistent;
  if (kernel == mhcudaKernelPersistent)
    return hash_persistent(gen, mats, s, ctx, capturing, timed);
  if (kernel != mhcudaKernelPlain) return mhcudaInvalidArguments;
  const bool with_events = timed && gen->timing && !capturing;
  if (with_events) HIPCH(hipEventRecord(gen->ev_begin, s), mhcudaRuntimeError);
  for (const MatrixDesc &m : mats) {
    if (m.length == 0) continue;
    HIPCH(launch_plain(gen->params, gen->dim, gen->samples, m, gen->tuning.spl, gen->fast_div, s),
          mhcudaRuntimeError);
  }
  if (with_events) HIPCH(hipEventRecord(gen->ev_end, s), mhcudaRuntimeError);
  if (timed) gen->timed = gen->timing;
  return mhcudaSuccess;
}

MHCUDAResult refresh_division_mode(MinhashCudaGenerator *gen) {
  int verbosity = gen->verbosity;
  gen->fast_div = false;
  gen->bounded = false;
  HIPCH(gen->range_flag.reserve(3 * sizeof(uint32_t)), mhcudaMemoryAllocationFailure);
  HIPCH(hipMemsetAsync(gen->range_flag.ptr, 0, sizeof(uint32_t), gen->stream), mhcudaRuntimeError);
  HIPCH(launch_params_range(gen->params, gen->dim, gen->samples, (uint32_t *)gen->range_flag.ptr,
                            gen->stream),
        mhcudaRuntimeError);
  uint32_t bad = 1;
  HIPCH(hipMemcpyAsync(&bad, gen->range_flag.ptr, sizeof(bad), hipMemcpyDeviceToHost, gen->stream),
        mhcudaMemoryCopyError);
  HIPCH(hipStreamSynchronize(gen->stream), mhcudaRuntimeError);
  gen->fast_div = (bad & kRangeDiv) == 0;
  if (gen->bounds && bad == 0) {
    // the code range: smallest and largest entry over the table
    uint32_t *minmax = (uint32_t *)gen->range_flag.ptr + 1;
    const uint32_t init[2] = {0xffffffffu, 0u};
    HIPCH(hipMemcpyAsync(minmax, init, sizeof(init), hipMemcpyHostToDevice, gen->stream),
          mhcudaMemoryCopyError);
    HIPCH(launch_bound_range(gen->params, gen->dim, gen->samples, minmax, gen->stream),
          mhcudaRuntimeError);
    uint32_t got[2] = {0u, 0u};
    HIPCH(hipMemcpyAsync(got, minmax, sizeof(got), hipMemcpyDeviceToHost, gen->stream),
          mhcudaMemoryCopyError);
    HIPCH(hipStreamSynchronize(gen->stream), mhcudaRuntimeError);
    const float mn = wmh_order_value(got[0]), mx = wmh_order_value(got[1]);
    // an entry of -inf (|ln_c - r| beyond the fp16 range) leaves no finite
    // lower end for the codes: hash without the bound
    if (got[0] <= got[1] && std::isfinite(mn) && std::isfinite(mx)) {
      float lo = 0.0f, delta = 0.0f;
      wmh_code_range(mn, mx, &lo, &delta);
      HIPCH(launch_build_codes(gen->params, gen->dim, gen->samples, lo, delta, gen->bounds,
                               gen->coarse, gen->stream),
            mhcudaRuntimeError);
      if (gen->items)
        HIPCH(launch_build_items(gen->params, gen->dim, gen->samples, gen->items, gen->stream),
              mhcudaRuntimeError);
      HIPCH(hipStreamSynchronize(gen->stream), mhcudaRuntimeError);
      gen->code_scale = 1.0f / delta;
      gen->code_t0 = 2049.0f - lo / delta;
      gen->bounded = true;
      if (verbosity > 1)
        printf("bound codes: entries in [%g, %g], lo %g, delta %g\n", mn, mx, lo, delta);
    }
  }
  return mhcudaSuccess;
}

// hipGraph path of the device-pointer calls: the whole enqueue sequence
// (descriptor upload, plan, init, hashing, fix-up -- or the plain launches)
// is captured once per distinct (descriptors, kernel, tuning) and
// replayed with one hipGraphLaunch.  The graph owns its descriptor staging,
// so no other call can change what a replay copies; the sequence contains no
// memset nodes.
MHCUDAResult hash_graph(MinhashCudaGenerator *gen, std::vector<MatrixDesc> &mats, int kernel,
                        hipStream_t s) {
  int verbosity = gen->verbosity;
  GraphCache &gc = gen->graph;
  if (kernel == mhcudaKernelAuto) kernel = gen->kernel;
  if (kernel == mhcudaKernelAuto) kernel = mhcudaKernelPersistent;
  const bool same = gc.exec && gc.kernel == kernel &&
                    gc.tuning.spl == gen->tuning.spl && gc.tuning.seg == gen->tuning.seg &&
                    gc.tuning.waves_per_cu == gen->tuning.waves_per_cu &&
                    gc.tuning.stripe_mib == gen->tuning.stripe_mib &&
                    gc.tuning.seeds == gen->tuning.seeds && gc.fast_div == gen->fast_div &&
                    same_table(gc.bounds, launch_bounds_of(gen)) && gc.key.size() == mats.size() &&
                    std::memcmp(gc.key.data(), mats.data(), sizeof(MatrixDesc) * mats.size()) == 0;
  if (!same) {
    if (!gc.capture)
      HIPCH(hipStreamCreateWithFlags(&gc.capture, hipStreamNonBlocking), mhcudaRuntimeError);
    if (!gc.launched)
      HIPCH(hipEventCreateWithFlags(&gc.launched, hipEventDisableTiming), mhcudaRuntimeError);
    HIPCH(hipEventSynchronize(gc.launched), mhcudaRuntimeError);  // no replay in flight
    gc.release();
    // reserve the staging before capture: nothing allocates while capturing
    const size_t desc_bytes = sizeof(MatrixDesc) * mats.size();
    HIPCH(gc.plan.plan.reserve(plan_layout(gen, mats).bytes), mhcudaMemoryAllocationFailure);
    if (gc.plan.cap < mats.size()) {
      if (gc.plan.desc_host) (void)hipHostFree(gc.plan.desc_host);
      gc.plan.desc_host = nullptr;
      gc.plan.cap = 0;
      HIPCH(hipHostMalloc((void **)&gc.plan.desc_host, desc_bytes, hipHostMallocDefault),
            mhcudaMemoryAllocationFailure);
      gc.plan.cap = mats.size();
    }
    std::vector<MatrixDesc> key = mats;
    HIPCH(hipStreamBeginCapture(gc.capture, hipStreamCaptureModeRelaxed), mhcudaRuntimeError);
    MHCUDAResult r = hash_device_batch(gen, mats, kernel, gc.capture, &gc.plan, true);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(gc.capture, &graph);
    if (r != mhcudaSuccess || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return r != mhcudaSuccess ? r : mhcudaRuntimeError;
    }
    gc.graph = graph;
    HIPCH(hipGraphInstantiate(&gc.exec, gc.graph, nullptr, nullptr, 0), mhcudaRuntimeError);
    gc.key.swap(key);
    gc.kernel = kernel;
    gc.tuning = gen->tuning;
    gc.fast_div = gen->fast_div;
    gc.bounds = launch_bounds_of(gen);
    gc.captures++;
  }
  // replays share the graph's descriptor staging and control block: one on
  // another stream waits for the previous one to finish
  HIPCH(hipStreamWaitEvent(s, gc.launched, 0), mhcudaRuntimeError);
  // timing events stay outside the graph: with graphs the timed span is the
  // whole replayed sequence
  if (gen->timing) HIPCH(hipEventRecord(gen->ev_begin, s), mhcudaRuntimeError);
  HIPCH(hipGraphLaunch(gc.exec, s), mhcudaRuntimeError);
  if (gen->timing) HIPCH(hipEventRecord(gen->ev_end, s), mhcudaRuntimeError);
  HIPCH(hipEventRecord(gc.launched, s), mhcudaRuntimeError);
  gen->timed = gen->timing;
  return mhcudaSuccess;
}

GraphCache::~GraphCache() {
  if (launched) (void)hipEventSynchronize(launched);
  release();
  if (launched) (void)hipEventDestroy(launched);
  if (capture) (void)hipStreamDestroy(capture);
}

void GraphCache::release() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  key.clear();
}

}  // namespace wmh

extern "C" {

MinhashCudaGenerator *mhcuda_init(uint32_t dim, uint16_t samples, uint32_t seed,
                                  int deferred, uint32_t devices, int verbosity,
                                  MHCUDAResult *status) {
  DEBUG("mhcuda_init: %" PRIu32 " %" PRIu16 " %" PRIu32 " %d %" PRIu32 " %d %p\n",
        dim, samples, seed, deferred, devices, verbosity, (void *)status);
  if (dim == 0 || samples == 0) {
    if (status) *status = mhcudaInvalidArguments;
    return nullptr;
  }
  const int dev = select_device(devices, verbosity);
  if (dev < 0) {
    if (status) *status = mhcudaNoSuchDevice;
    return nullptr;
  }
#define FAIL(code)                   \
  do {                               \
    if (status) *status = (code);    \
    return nullptr;                  \
  } while (false)
  if (hipSetDevice(dev) != hipSuccess) FAIL(mhcudaNoSuchDevice);
  // the scratch of mhcuda_calc_device_noalloc, looked up once per device
  if (wmh::scratch_init(dev) != hipSuccess) FAIL(mhcudaRuntimeError);
  std::unique_ptr<MinhashCudaGenerator> gen(new MinhashCudaGenerator());
  gen->dim = dim;
  gen->samples = samples;
  gen->verbosity = verbosity;
  gen->device = dev;
  gen->sp = wmh::padded_samples(samples);
  gen->tuning.spl = samples <= 64 ? 1 : (samples <= 128 ? 2 : 4);
  hipDeviceProp_t props;
  if (hipGetDeviceProperties(&props, dev) != hipSuccess) FAIL(mhcudaRuntimeError);
  gen->num_cus = props.multiProcessorCount;
  // XCDs (dies with an L2 of their own): one stripe queue per XCD; 1 where the
  // runtime does not say
  int xcds = 1;
  if (hipDeviceGetAttribute(&xcds, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess ||
      xcds < 1) {
    (void)hipGetLastError();
    xcds = 1;
  }
  gen->num_xcds = xcds;
  if (hipStreamCreateWithFlags(&gen->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&gen->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&gen->d2h_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&gen->after_null, hipEventDisableTiming) != hipSuccess)
    FAIL(mhcudaRuntimeError);
  for (auto &p : gen->pieces) {
    if (p.create() != hipSuccess) FAIL(mhcudaRuntimeError);
  }
  if (const char *env = std::getenv("MHCUDA_PIECE_NNZ")) {
    const unsigned long long v = std::strtoull(env, nullptr, 10);
    if (v > 0) gen->piece_nnz = v;
  }
  if (const char *env = std::getenv("MHCUDA_PIECE_ROWS")) {
    const unsigned long long v = std::strtoull(env, nullptr, 10);
    if (v > 0) gen->piece_rows = v;
  }
  const size_t param_bytes = (size_t)dim * 3u * gen->sp * sizeof(float);
  if (hipMalloc(&gen->params, param_bytes) != hipSuccess) {
    (void)hipGetLastError();
    gen->params = nullptr;
    INFO("failed to allocate %zu bytes for the parameters\n", param_bytes);
    FAIL(mhcudaMemoryAllocationFailure);
  }
  // the bound table of the skip bound (a twelfth of the parameter bytes); the
  // engine hashes without it when it cannot be allocated or is switched off
  const char *skip_env = std::getenv("MHCUDA_SKIP_BOUND");
  if (!(skip_env && std::strcmp(skip_env, "0") == 0)) {
    if (hipMalloc((void **)&gen->bounds, (size_t)wmh::code_table_bytes(dim, samples)) !=
        hipSuccess) {
      (void)hipGetLastError();
      gen->bounds = nullptr;
      INFO("no memory for the bound table: hashing without the skip bound\n");
    } else if (gen->sp % 8u == 0u) {
      gen->coarse = gen->bounds + wmh::coarse_offset(dim, samples);
      // the item table (4/3 of the parameter bytes) where 32-bit offsets reach it
      const char *items_env = std::getenv("MHCUDA_ITEM_TABLE");
      if (!(items_env && std::strcmp(items_env, "0") == 0) &&
          wmh::item_table_fits(dim, samples) &&
          hipMalloc((void **)&gen->items, (size_t)wmh::item_table_bytes(dim, samples)) !=
              hipSuccess) {
        (void)hipGetLastError();
        gen->items = nullptr;
        INFO("no memory for the item table: the bounded pieces read the parameter table\n");
      }
    }
  }
  if (deferred) {
    if (hipMemsetAsync(gen->params, 0, param_bytes, gen->stream) != hipSuccess)
      FAIL(mhcudaRuntimeError);
  } else {
    if (wmh::launch_gen_params(gen->params, dim, samples, seed, gen->stream) != hipSuccess)
      FAIL(mhcudaRuntimeError);
    gen->params_source = MinhashCudaGenerator::kParamsSeeded;
    gen->seed = seed;
    const MHCUDAResult rr = wmh::refresh_division_mode(gen.get());
    if (rr != mhcudaSuccess) FAIL(rr);
  }
  if (hipStreamSynchronize(gen->stream) != hipSuccess) FAIL(mhcudaRuntimeError);
  if (verbosity > 1 && print_memory_stats(dev) != mhcudaSuccess) FAIL(mhcudaRuntimeError);
#undef FAIL
  return gen.release();
}

MinhashCudaGeneratorParameters mhcuda_get_parameters(const MinhashCudaGenerator *gen) {
  if (gen == nullptr) return MinhashCudaGeneratorParameters{};
  MinhashCudaGeneratorParameters p{};
  p.dim = gen->dim;
  p.samples = gen->samples;
  p.verbosity = gen->verbosity;
  return p;
}

MHCUDAResult mhcuda_retrieve_random_vars(const MinhashCudaGenerator *cgen, float *rs,
                                         float *ln_cs, float *betas) {
  if (!cgen || !rs || !ln_cs || !betas) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  const uint32_t chunk = relayout_chunk(gen->dim, gen->samples);
  HIPCH(gen->staging.reserve((size_t)chunk * gen->samples * sizeof(float)),
        mhcudaMemoryAllocationFailure);
  float *dst[3] = {rs, betas, ln_cs};  // layout order: 0 = rs, 1 = betas, 2 = ln_cs
  for (uint32_t which = 0; which < 3; which++) {
    for (uint32_t j0 = 0; j0 < gen->dim; j0 += chunk) {
      const uint32_t n = std::min(chunk, gen->dim - j0);
      HIPCH(wmh::launch_gather_params(gen->params, j0, n, which,
                                      (float *)gen->staging.ptr, gen->samples, gen->stream),
            mhcudaRuntimeError);
      HIPCH(hipMemcpy2DAsync(dst[which] + j0, (size_t)gen->dim * sizeof(float),
                             gen->staging.ptr, (size_t)n * sizeof(float),
                             (size_t)n * sizeof(float), gen->samples,
                             hipMemcpyDeviceToHost, gen->stream),
            mhcudaMemoryCopyError);
      HIPCH(hipStreamSynchronize(gen->stream), mhcudaMemoryCopyError);
    }
  }
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_assign_random_vars(const MinhashCudaGenerator *cgen, const float *rs,
                                       const float *ln_cs, const float *betas) {
  if (!cgen || !rs || !ln_cs || !betas) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tables
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  const uint32_t chunk = relayout_chunk(gen->dim, gen->samples);
  HIPCH(gen->staging.reserve((size_t)chunk * gen->samples * sizeof(float)),
        mhcudaMemoryAllocationFailure);
  const float *src[3] = {rs, betas, ln_cs};
  gen->params_source = MinhashCudaGenerator::kParamsNone;  // until every table is in
  gen->fast_div = false;                                   // until the new r are checked
  gen->bounded = false;                                    // until the bound table is rebuilt
  for (uint32_t which = 0; which < 3; which++) {
    for (uint32_t j0 = 0; j0 < gen->dim; j0 += chunk) {
      const uint32_t n = std::min(chunk, gen->dim - j0);
      HIPCH(hipMemcpy2DAsync(gen->staging.ptr, (size_t)n * sizeof(float),
                             src[which] + j0, (size_t)gen->dim * sizeof(float),
                             (size_t)n * sizeof(float), gen->samples,
                             hipMemcpyHostToDevice, gen->stream),
            mhcudaMemoryCopyError);
      HIPCH(wmh::launch_scatter_params((const float *)gen->staging.ptr, n, j0, which,
                                       gen->params, gen->samples, gen->stream),
            mhcudaRuntimeError);
      HIPCH(hipStreamSynchronize(gen->stream), mhcudaMemoryCopyError);
    }
  }
  gen->params_source = MinhashCudaGenerator::kParamsTables;
  return wmh::refresh_division_mode(gen);
}

MHCUDAResult mhcuda_calc(const MinhashCudaGenerator *cgen, const float *weights,
                         const uint32_t *cols, const uint32_t *rows, uint32_t length,
                         uint32_t *output) {
  if (!cgen || !weights || !cols || !rows || !output || length == 0)
    return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  MHCUDAMatrix m{weights, cols, rows, length, output};
  return mhcuda_calc_batch(gen, 1, &m, mhcudaKernelAuto);
}

MHCUDAResult mhcuda_fini(MinhashCudaGenerator *gen) {
  delete gen;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_calc_device(const MinhashCudaGenerator *cgen, const float *weights,
                                const uint32_t *cols, const uint32_t *rows,
                                uint32_t length, uint32_t *output, void *stream) {
  if (!cgen || !weights || !cols || !rows || !output || length == 0)
    return mhcudaInvalidArguments;
  MHCUDAMatrix m{weights, cols, rows, length, output};
  return mhcuda_calc_batch_device(cgen, 1, &m, mhcudaKernelAuto, stream);
}

MHCUDAResult mhcuda_calc_device_noalloc(const MinhashCudaGenerator *cgen, const float *weights,
                                        const uint32_t *cols, const uint32_t *rows,
                                        uint32_t length, uint32_t *output, void *stream) {
  if (!cgen || !weights || !cols || !rows || !output || length == 0)
    return mhcudaInvalidArguments;
  if (length > wmh::kNoAllocMaxRows) return mhcudaInvalidArguments;  // beyond the scratch
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  if (gen->samples % 2 != 0) return mhcudaInvalidArguments;  // as mhcuda_calc_device
  std::lock_guard<std::mutex> lock(gen->mu);
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  char *scr = static_cast<char *>(wmh::scratch_address(gen->device));
  if (!scr) return mhcudaRuntimeError;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : gen->stream;
  if (!stream) HIPCH(wmh::follow_default_stream(gen), mhcudaRuntimeError);
  // the plan kernel resets the queue heads and counters and the init pass
  // rewrites the striped-row list, so nothing left by an earlier call is read
  const wmh::MatrixDesc desc{weights, cols, rows, output, length, 0u, 0u, 0u, 0u, 0u};
  HIPCH(wmh::launch_store_desc(desc, reinterpret_cast<wmh::MatrixDesc *>(scr), s),
        mhcudaRuntimeError);
  HIPCH(wmh::launch_persistent(gen->params, gen->dim, gen->samples,
                               reinterpret_cast<wmh::MatrixDesc *>(scr), 1u, (length + 63u) / 64u,
                               reinterpret_cast<wmh::BatchCtl *>(scr + wmh::kScratchCtlOff),
                               reinterpret_cast<wmh::StripedRow *>(scr + wmh::kScratchListOff),
                               gen->tuning, gen->seg_auto, gen->num_cus, gen->num_xcds,
                               gen->fast_div, s, nullptr, nullptr, wmh::launch_bounds_of(gen)),
        mhcudaRuntimeError);
  if (!stream) HIPCH(hipStreamSynchronize(s), mhcudaRuntimeError);
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_calc_batch_device(const MinhashCudaGenerator *cgen, uint32_t count,
                                      const MHCUDAMatrix *mats, int kernel, void *stream) {
  if (!cgen || !mats || count == 0) return mhcudaInvalidArguments;
  for (uint32_t m = 0; m < count; m++) {
    if (mats[m].length && (!mats[m].weights || !mats[m].cols || !mats[m].rows || !mats[m].output))
      return mhcudaInvalidArguments;
  }
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : gen->stream;
  if (!stream) HIPCH(wmh::follow_default_stream(gen), mhcudaRuntimeError);
  std::vector<wmh::MatrixDesc> descs(count);
  for (uint32_t m = 0; m < count; m++) {
    descs[m] = wmh::MatrixDesc{mats[m].weights, mats[m].cols, mats[m].rows, mats[m].output,
                               mats[m].length, 0u, 0u, 0u, 0u, 0u};
  }
  MHCUDAResult r = gen->graph.enabled ? wmh::hash_graph(gen, descs, kernel, s)
                                      : wmh::hash_device_batch(gen, descs, kernel, s,
                                                               &gen->plan, false);
  if (r != mhcudaSuccess) return r;
  if (!stream) HIPCH(hipStreamSynchronize(s), mhcudaRuntimeError);
  return mhcudaSuccess;
}

}  // extern "C"

namespace {

// The pipelined host-memory path of mhcuda_calc_batch (called with the
// generator's lock held and its device current).
// A large host-memory call is cut into about kPipelinePieces pieces (of at
// least kMinPieceNnz nonzeros) so that the uploads, the hashing and the
// downloads of consecutive pieces overlap.
constexpr uint64_t kPipelinePieces = 4;
constexpr uint64_t kMinPieceNnz = 1ull << 21;
// Pieces whose inputs are pageable are staged through a pinned buffer per
// buffer set: they hold at most kStagedPieceNnz nonzeros (256 MiB of weights
// and columns).
constexpr uint64_t kStagedPieceNnz = 1ull << 25;

MHCUDAResult calc_batch_locked(MinhashCudaGenerator *gen, uint32_t count,
                               const MHCUDAMatrix *mats, int kernel, uint64_t max_row,
                               uint64_t total_nnz, uint64_t total_rows) {
  int verbosity = gen->verbosity;
  // Pieces: consecutive row ranges holding at most `budget` nonzeros and
  // `row_budget` rows (a row is never cut).  Piece p uses device buffer set
  // p % 2, so the upload of piece p + 1 overlaps the hashing of piece p and
  // the download of the hashes runs on a third stream.  The grow-only buffers
  // play the role of the reference's fill_buffers (minhashcuda.cc:385-455).
  const uint64_t split_nnz =
      std::max(kMinPieceNnz, (total_nnz + kPipelinePieces - 1) / kPipelinePieces);
  const uint64_t split_rows = std::max<uint64_t>(
      1, split_nnz >= total_nnz ? total_rows
                                : (total_rows + kPipelinePieces - 1) / kPipelinePieces);
  bool pageable = false;
  for (uint32_t m = 0; m < count && !pageable; m++) {
    const MHCUDAMatrix &x = mats[m];
    if (x.length == 0) continue;
    pageable = !wmh::is_pinned(x.weights + x.rows[0]) || !wmh::is_pinned(x.cols + x.rows[0]) ||
               !wmh::is_pinned(x.rows);
  }
  uint64_t budget = std::min(gen->piece_nnz, split_nnz);
  if (pageable) budget = std::min(budget, kStagedPieceNnz);
  budget = std::max<uint64_t>(budget, max_row);
  const uint64_t row_budget = std::max<uint64_t>(1, std::min(gen->piece_rows, split_rows));
  if (!gen->copier) {
    try {
      gen->copier.reset(new wmh::HostCopier());
    } catch (...) {
      return mhcudaRuntimeError;  // no threads
    }
  }
  wmh::HostCopier &copier = *gen->copier;
  struct Span {
    uint32_t m, r0, r1;
  };
  std::vector<std::vector<Span>> pieces(1);
  uint64_t cur_nnz = 0, cur_rows = 0;
  for (uint32_t m = 0; m < count; m++) {
    const MHCUDAMatrix &x = mats[m];
    uint32_t r0 = 0;
    for (uint32_t r = 0; r < x.length; r++) {
      const uint64_t n = x.rows[r + 1] - x.rows[r];
      if (cur_rows > 0 && (cur_nnz + n > budget || cur_rows + 1 > row_budget)) {
        if (r > r0) pieces.back().push_back(Span{m, r0, r});
        pieces.emplace_back();
        r0 = r;
        cur_nnz = cur_rows = 0;
      }
      cur_nnz += n;
      cur_rows++;
    }
    if (x.length > r0) pieces.back().push_back(Span{m, r0, x.length});
  }
  // Default kernel for host-memory calls when neither the caller nor the
  // autotuner chose one: one wavefront per row suits a single matrix of
  // bounded rows (no stragglers); anything else goes to the persistent queue.
  if (kernel == mhcudaKernelAuto && gen->kernel == mhcudaKernelAuto)
    kernel = (count == 1 && max_row <= 4096) ? mhcudaKernelPlain : mhcudaKernelPersistent;
  const size_t S2 = (size_t)gen->samples * 2;
  for (auto &b : gen->pieces) {
    // copies still owed here belong to a call that failed midway: drop them
    HIPCH(hipEventSynchronize(b.done), mhcudaRuntimeError);
    b.owed.clear();
  }
  for (size_t p = 0; p < pieces.size(); p++) {
    const int b = (int)(p & 1);
    wmh::PieceBuffers &buf = gen->pieces[b];
    // the set is free once its previous download has completed
    HIPCH(hipEventSynchronize(buf.done), mhcudaRuntimeError);
    buf.settle(copier);
    uint64_t nnz = 0, nrows = 0;
    for (const Span &sp : pieces[p]) {
      nnz += mats[sp.m].rows[sp.r1] - mats[sp.m].rows[sp.r0];
      nrows += sp.r1 - sp.r0;
    }
    const size_t nspans = pieces[p].size();
    HIPCH(wmh::reserve_checked(gen, buf.weights, std::max<uint64_t>(nnz, 1) * 4),
          mhcudaMemoryAllocationFailure);
    HIPCH(wmh::reserve_checked(gen, buf.cols, std::max<uint64_t>(nnz, 1) * 4),
          mhcudaMemoryAllocationFailure);
    HIPCH(wmh::reserve_checked(gen, buf.rows, (nrows + nspans) * 4), mhcudaMemoryAllocationFailure);
    HIPCH(wmh::reserve_checked(gen, buf.hashes, std::max<uint64_t>(nrows * S2, 1) * 4),
          mhcudaMemoryAllocationFailure);
    float *dw = (float *)buf.weights.ptr;
    uint32_t *dc = (uint32_t *)buf.cols.ptr, *dr = (uint32_t *)buf.rows.ptr,
             *dh = (uint32_t *)buf.hashes.ptr;
    std::vector<wmh::MatrixDesc> dev(nspans);
    // the piece's uploads: (device destination, host source, bytes); pinned
    // sources are DMA'd directly, pageable ones first copied (in parallel)
    // into the set's pinned staging buffer
    std::vector<wmh::CopyTask> uploads;
    uploads.reserve(3 * nspans);
    uint64_t nnz_off = 0, row_off = 0, out_off = 0;
    for (size_t i = 0; i < nspans; i++) {
      const Span &sp = pieces[p][i];
      const MHCUDAMatrix &x = mats[sp.m];
      const uint32_t first = x.rows[sp.r0];
      const uint64_t n = x.rows[sp.r1] - first;
      const uint32_t len = sp.r1 - sp.r0;
      if (n) {
        uploads.push_back(wmh::CopyTask{dw + nnz_off, x.weights + first, n * 4});
        uploads.push_back(wmh::CopyTask{dc + nnz_off, x.cols + first, n * 4});
      }
      uploads.push_back(wmh::CopyTask{dr + row_off, x.rows + sp.r0, ((uint64_t)len + 1) * 4});
      dev[i] = wmh::MatrixDesc{dw + nnz_off, dc + nnz_off, dr + row_off, dh + out_off,
                               len, first, 0u, 0u, 0u, 0u};
      nnz_off += n;
      row_off += (uint64_t)len + 1;
      out_off += (uint64_t)len * S2;
    }
    std::vector<char> staged(uploads.size());
    size_t stage_bytes = 0;
    for (size_t i = 0; i < uploads.size(); i++) {
      staged[i] = !wmh::is_pinned(uploads[i].src);
      if (staged[i]) stage_bytes += (uploads[i].bytes + 63) & ~(size_t)63;
    }
    if (stage_bytes) {
      // the set's previous uploads out of the staging buffer completed before
      // its `done` event
      if (!buf.staging.reserve(stage_bytes)) return mhcudaMemoryAllocationFailure;
      std::vector<wmh::CopyTask> to_stage;
      size_t at = 0;
      for (size_t i = 0; i < uploads.size(); i++) {
        if (!staged[i]) continue;
        to_stage.push_back(wmh::CopyTask{buf.staging.ptr + at, uploads[i].src, uploads[i].bytes});
        uploads[i].src = buf.staging.ptr + at;
        at += (uploads[i].bytes + 63) & ~(size_t)63;
      }
      copier.copy(to_stage.data(), to_stage.size());
    }
    for (const wmh::CopyTask &u : uploads)
      HIPCH(wmh::copy_checked(gen, u.dst, u.src, u.bytes, hipMemcpyHostToDevice,
                              gen->copy_stream),
            mhcudaMemoryCopyError);
    HIPCH(hipEventRecord(buf.uploaded, gen->copy_stream), mhcudaRuntimeError);
    HIPCH(hipStreamWaitEvent(gen->stream, buf.uploaded, 0), mhcudaRuntimeError);
    MHCUDAResult r =
        wmh::hash_device_batch(gen, dev, kernel, gen->stream, &gen->plan, false);
    if (r != mhcudaSuccess) return r;
    HIPCH(hipEventRecord(buf.hashed, gen->stream), mhcudaRuntimeError);
    HIPCH(hipStreamWaitEvent(gen->d2h_stream, buf.hashed, 0), mhcudaRuntimeError);
    // pinned outputs are written directly; pageable ones through the landing zone
    std::vector<char> direct(nspans);
    size_t land = 0;
    for (size_t i = 0; i < nspans; i++) {
      const Span &sp = pieces[p][i];
      direct[i] = wmh::is_pinned(mats[sp.m].output + (uint64_t)sp.r0 * S2);
      if (!direct[i]) land += (uint64_t)(sp.r1 - sp.r0) * S2 * 4;
    }
    if (!buf.landing.reserve(land)) return mhcudaMemoryAllocationFailure;
    land = 0;
    for (size_t i = 0; i < nspans; i++) {
      const Span &sp = pieces[p][i];
      uint32_t *dst = mats[sp.m].output + (uint64_t)sp.r0 * S2;
      const size_t bytes = (uint64_t)(sp.r1 - sp.r0) * S2 * 4;
      void *target = dst;
      if (!direct[i]) {
        target = buf.landing.ptr + land;
        buf.owed.push_back(wmh::PendingCopy{dst, target, bytes});
        land += bytes;
      }
      HIPCH(wmh::copy_checked(gen, target, dev[i].output, bytes, hipMemcpyDeviceToHost,
                              gen->d2h_stream),
            mhcudaMemoryCopyError);
    }
    HIPCH(hipEventRecord(buf.done, gen->d2h_stream), mhcudaRuntimeError);
  }
  HIPCH(hipStreamSynchronize(gen->d2h_stream), mhcudaRuntimeError);
  HIPCH(hipStreamSynchronize(gen->stream), mhcudaRuntimeError);
  for (auto &b : gen->pieces) b.settle(copier);
  INFO("mhcuda - success\n");
  return mhcudaSuccess;
}

}  // namespace

extern "C" {

MHCUDAResult mhcuda_calc_batch(const MinhashCudaGenerator *cgen, uint32_t count,
                               const MHCUDAMatrix *mats, int kernel) {
  if (!cgen || !mats || count == 0) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  int verbosity = gen->verbosity;
  uint64_t total_nnz = 0, total_rows = 0, max_row = 0;
  for (uint32_t m = 0; m < count; m++) {
    const MHCUDAMatrix &x = mats[m];
    if (x.length == 0) continue;
    if (!x.weights || !x.cols || !x.rows || !x.output) return mhcudaInvalidArguments;
    for (uint32_t r = 0; r < x.length; r++) {
      if (x.rows[r + 1] < x.rows[r]) return mhcudaInvalidArguments;  // malformed indptr
      max_row = std::max<uint64_t>(max_row, x.rows[r + 1] - x.rows[r]);
    }
    total_nnz += x.rows[x.length] - x.rows[0];
    total_rows += x.length;
  }
  if (total_rows == 0) return mhcudaInvalidArguments;
  if (gen->samples % 2 != 0) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  DEBUG("mhcuda_calc: %u matrices, %" PRIu64 " rows, %" PRIu64 " nonzeros\n", count,
        total_rows, total_nnz);
  const MHCUDAResult r =
      calc_batch_locked(gen, count, mats, kernel, max_row, total_nnz, total_rows);
  if (r != mhcudaSuccess) {
    // a call that failed midway may have left uploads, hashing and downloads
    // in flight on the generator's streams: drain them, so that the next call
    // reuses the buffers only once nothing reads or writes them any more
    (void)hipStreamSynchronize(gen->copy_stream);
    (void)hipStreamSynchronize(gen->stream);
    (void)hipStreamSynchronize(gen->d2h_stream);
    for (auto &b : gen->pieces) b.owed.clear();
  }
  return r;
}

MHCUDAResult mhcuda_set_tuning(MinhashCudaGenerator *gen, int samples_per_lane,
                               int segment, int waves_per_cu) {
  if (!gen) return mhcudaInvalidArguments;
  if (samples_per_lane && samples_per_lane != 1 && samples_per_lane != 2 && samples_per_lane != 4)
    return mhcudaInvalidArguments;
  if (segment < -1 || (segment > 0 && segment % 64 != 0) || waves_per_cu < 0 ||
      waves_per_cu > 32)
    return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tuning
  if (samples_per_lane) gen->tuning.spl = samples_per_lane;
  if (segment > 0) {
    gen->tuning.seg = segment;
    gen->seg_auto = false;
  } else if (segment == -1) {
    gen->seg_auto = true;
  }
  if (waves_per_cu) gen->tuning.waves_per_cu = waves_per_cu;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_set_timing(MinhashCudaGenerator *gen, int enable) {
  if (!gen) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  if (enable && !gen->ev_begin) {
    HIPCH(hipEventCreate(&gen->ev_begin), mhcudaRuntimeError);
    HIPCH(hipEventCreate(&gen->ev_end), mhcudaRuntimeError);
  }
  gen->timing = enable != 0;
  gen->timed = false;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_kernel_time(const MinhashCudaGenerator *cgen, float *ms) {
  if (!cgen || !ms) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  int verbosity = gen->verbosity;
  if (!gen->timed) return mhcudaInvalidArguments;
  HIPCH(hipEventSynchronize(gen->ev_end), mhcudaRuntimeError);
  HIPCH(hipEventElapsedTime(ms, gen->ev_begin, gen->ev_end), mhcudaRuntimeError);
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_jaccard_pairs(const MinhashCudaGenerator *gen, const uint32_t *hashes_a,
                                  const uint32_t *hashes_b, const uint32_t *pairs, uint32_t n,
                                  float *out, void *stream) {
  if (!gen || !hashes_a || !hashes_b || !pairs || !out) return mhcudaInvalidArguments;
  if (n == 0) return mhcudaSuccess;
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : gen->stream;
  if (!stream)
    HIPCH(wmh::follow_default_stream(const_cast<MinhashCudaGenerator *>(gen)), mhcudaRuntimeError);
  HIPCH(wmh::launch_jaccard_pairs(hashes_a, hashes_b, pairs, n, gen->samples, out, s),
        mhcudaRuntimeError);
  if (!stream) HIPCH(hipStreamSynchronize(s), mhcudaRuntimeError);
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_lsh_band_keys(const MinhashCudaGenerator *gen, const uint32_t *hashes,
                                  uint32_t rows, uint32_t bands, uint64_t *keys, void *stream) {
  if (!gen || !hashes || !keys || bands == 0 || gen->samples % bands != 0)
    return mhcudaInvalidArguments;
  if (rows == 0) return mhcudaSuccess;
  int verbosity = gen->verbosity;
  HIPCH(hipSetDevice(gen->device), mhcudaNoSuchDevice);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : gen->stream;
  if (!stream)
    HIPCH(wmh::follow_default_stream(const_cast<MinhashCudaGenerator *>(gen)), mhcudaRuntimeError);
  HIPCH(wmh::launch_band_keys(hashes, rows, gen->samples, bands,
                              reinterpret_cast<unsigned long long *>(keys), s),
        mhcudaRuntimeError);
  if (!stream) HIPCH(hipStreamSynchronize(s), mhcudaRuntimeError);
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_set_graph(MinhashCudaGenerator *gen, int enable) {
  if (!gen) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  gen->graph.enabled = enable != 0;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_graph_captures(const MinhashCudaGenerator *gen, uint64_t *captures) {
  if (!gen || !captures) return mhcudaInvalidArguments;
  *captures = gen->graph.captures;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_get_tuning(const MinhashCudaGenerator *gen, int *samples_per_lane,
                               int *segment, int *waves_per_cu) {
  if (!gen) return mhcudaInvalidArguments;
  if (samples_per_lane) *samples_per_lane = gen->tuning.spl;
  if (segment) *segment = gen->tuning.seg;
  if (waves_per_cu) *waves_per_cu = gen->tuning.waves_per_cu;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_set_stripes(MinhashCudaGenerator *gen, int stripe_mib) {
  if (!gen || stripe_mib < 0) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tuning
  gen->tuning.stripe_mib = stripe_mib;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_set_seeds(MinhashCudaGenerator *gen, int enable) {
  if (!gen) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tuning
  gen->tuning.seeds = enable != 0;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_get_seeds(const MinhashCudaGenerator *gen, int *enabled) {
  if (!gen || !enabled) return mhcudaInvalidArguments;
  *enabled = gen->tuning.seeds;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_set_skip(MinhashCudaGenerator *gen, int enable) {
  if (!gen) return mhcudaInvalidArguments;
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads it
  gen->skip = enable != 0;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_get_skip(const MinhashCudaGenerator *gen, int *enabled, int *active) {
  if (!gen) return mhcudaInvalidArguments;
  if (enabled) *enabled = gen->skip ? 1 : 0;
  if (active) *active = wmh::launch_bounds_of(gen).codes != nullptr;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_inject_fault(MinhashCudaGenerator *gen, int what, uint32_t nth) {
  if (!gen) return mhcudaInvalidArguments;
  switch (what) {
    case 0:
      gen->fail_alloc.store(0);
      gen->fail_copy.store(0);
      return mhcudaSuccess;
    case 1:
      gen->fail_alloc.store(nth);
      return mhcudaSuccess;
    case 2:
      gen->fail_copy.store(nth);
      return mhcudaSuccess;
    default:
      return mhcudaInvalidArguments;
  }
}

MHCUDAResult mhcuda_get_stripes(const MinhashCudaGenerator *gen, int *stripe_mib,
                                uint32_t *stripes, uint32_t *min_nnz) {
  if (!gen) return mhcudaInvalidArguments;
  uint32_t k = 1, n = 0;
  wmh::stripe_shape(gen->dim, gen->samples, gen->tuning.stripe_mib, &k, &n);
  if (stripe_mib) *stripe_mib = gen->tuning.stripe_mib;
  if (stripes) *stripes = k;
  if (min_nnz) *min_nnz = k > 1 ? n : 0u;
  return mhcudaSuccess;
}

}  // extern "C"
