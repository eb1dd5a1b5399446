// Autotuner of the hashing kernels (no counterpart in the reference, whose
// launch shape is fixed by sample_delta and binpack, minhashcuda.cc:457-528
// and :602-620).
//
// The key of a batch is its sample count plus the shape of its nonzero-
// weighted row-length histogram (the log2 bins holding the median and the
// 90th percentile of nonzeros, and a size class of the total), measured on the
// device.  On a cache miss the tuner times candidates on a bounded prefix of
// the batch with HIP events, by coordinate descent: kernel (plain or
// persistent), samples per lane, then -- for the persistent kernel -- the
// work-item size, the waves per CU and the column-stripe size.  The prefix
// holds up to 2^32 nonzeros (the whole batch when smaller): the gain of column
// stripes depends on how often the batch reuses each stripe of the table,
// which a small prefix understates.  The winner becomes the generator's tuning
// and is cached under the key, so a later batch of the same shape is tuned
// without timing anything.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "minhashcuda.h"
#include "wmh_runtime.h"

namespace {

// Nonzeros timed per candidate: a prefix of the batch of at most this size.
// It is large on purpose: column stripes and the lane width interact with how
// often the batch reuses each part of the parameter table, which a small
// prefix understates.
constexpr uint64_t kSampleNnz = 1ull << 32;

struct Candidate {
  int kernel;
  wmh::Tuning tuning;
};

}  // namespace

extern "C" MHCUDAResult mhcuda_autotune(MinhashCudaGenerator *gen, uint32_t count,
                                        const MHCUDAMatrix *mats, void *stream,
                                        MHCUDATuneResult *result) {
  if (!gen || !mats || count == 0) return mhcudaInvalidArguments;
  if (gen->samples % 2 != 0) return mhcudaInvalidArguments;
  for (uint32_t m = 0; m < count; m++) {
    if (mats[m].length && (!mats[m].weights || !mats[m].cols || !mats[m].rows || !mats[m].output))
      return mhcudaInvalidArguments;
  }
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tuning
  const int verbosity = gen->verbosity;
  if (hipSetDevice(gen->device) != hipSuccess) return mhcudaNoSuchDevice;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : gen->stream;
  if (!stream && wmh::follow_default_stream(gen) != hipSuccess) return mhcudaRuntimeError;

  // descriptors (with the 64-row chunk prefix the row scans use) + histogram
  std::vector<wmh::MatrixDesc> descs(count);
  uint64_t chunks = 0;
  for (uint32_t m = 0; m < count; m++) {
    descs[m] = wmh::MatrixDesc{mats[m].weights, mats[m].cols, mats[m].rows, mats[m].output,
                               mats[m].length, 0u, 0u, 0u, (uint32_t)chunks, 0u};
    chunks += (mats[m].length + 63u) / 64u;
  }
  if (chunks >= (1ull << 32)) return mhcudaInvalidArguments;
  const size_t desc_bytes = sizeof(wmh::MatrixDesc) * count;
  const size_t hist_off = (desc_bytes + 255) & ~(size_t)255;
  if (gen->tune_buf.reserve(hist_off + 66 * sizeof(unsigned long long)) != hipSuccess)
    return mhcudaMemoryAllocationFailure;
  char *dev = (char *)gen->tune_buf.ptr;
  unsigned long long hist[66];
  if (hipMemcpyAsync(dev, descs.data(), desc_bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(dev + hist_off, 0, sizeof(hist), s) != hipSuccess)
    return mhcudaMemoryCopyError;
  if (wmh::launch_histogram((const wmh::MatrixDesc *)dev, count, (uint32_t)chunks,
                            (unsigned long long *)(dev + hist_off), s) != hipSuccess)
    return mhcudaRuntimeError;
  if (hipMemcpyAsync(hist, dev + hist_off, sizeof(hist), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return mhcudaMemoryCopyError;

  uint64_t total = 0, rows = 0;
  for (int b = 0; b < 33; b++) {
    rows += hist[b];
    total += hist[33 + b];
  }
  MHCUDATuneResult res{};
  if (total == 0) {  // nothing to time: keep the current tuning
    res.kernel = gen->kernel;
    res.samples_per_lane = gen->tuning.spl;
    res.segment = gen->tuning.seg;
    res.waves_per_cu = gen->tuning.waves_per_cu;
    res.stripe_mib = gen->tuning.stripe_mib;
    res.seeds = gen->tuning.seeds;
    res.cached = 1;
    if (result) *result = res;
    return mhcudaSuccess;
  }
  int p50 = 0, p90 = 0;
  uint64_t acc = 0;
  for (int b = 0; b < 33; b++) {
    acc += hist[33 + b];
    if (!p50 && acc * 2 >= total) p50 = b;
    if (!p90 && acc * 10 >= total * 9) {
      p90 = b;
      break;
    }
  }
  const int size_class = (int)(std::log2((double)total) / 2.0);
  const uint64_t key = (uint64_t)gen->samples | ((uint64_t)p50 << 16) | ((uint64_t)p90 << 24) |
                       ((uint64_t)size_class << 32);
  res.key = key;
  res.histogram_rows = rows;
  res.histogram_nnz = total;

  auto it = gen->tuned.find(key);
  if (it == gen->tuned.end()) {
    // nonzeros of every matrix
    std::vector<uint64_t> nnz_of(count, 0);
    for (uint32_t m = 0; m < count; m++) {
      if (descs[m].length == 0) continue;
      uint32_t r0 = 0, r1 = 0;
      if (hipMemcpy(&r0, descs[m].rows, 4, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(&r1, descs[m].rows + descs[m].length, 4, hipMemcpyDeviceToHost) != hipSuccess)
        return mhcudaMemoryCopyError;
      nnz_of[m] = r1 >= r0 ? r1 - r0 : 0;
    }
    // the timed prefix: whole matrices, the last one cut to a row prefix
    std::vector<wmh::MatrixDesc> sample;
    uint64_t budget = kSampleNnz, have = 0;
    for (uint32_t m = 0; m < count && have < budget; m++) {
      if (descs[m].length == 0) continue;
      const uint64_t n = nnz_of[m];
      wmh::MatrixDesc d = descs[m];
      d.chunk_begin = 0;
      if (have + n > budget && n > 0) {
        d.length = (uint32_t)std::max<uint64_t>(1, (uint64_t)d.length * (budget - have) / n);
        have = budget;
      } else {
        have += n;
      }
      sample.push_back(d);
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      return mhcudaRuntimeError;
    const wmh::Tuning saved = gen->tuning;
    const bool saved_timing = gen->timing;
    gen->timing = false;
    auto time_on = [&](const std::vector<wmh::MatrixDesc> &batch, const Candidate &c,
                       float *ms) -> MHCUDAResult {
      gen->tuning = c.tuning;
      std::vector<wmh::MatrixDesc> d = batch;
      if (hipEventRecord(e0, s) != hipSuccess) return mhcudaRuntimeError;
      MHCUDAResult r = wmh::hash_device_batch(gen, d, c.kernel, s, &gen->plan, false);
      if (r != mhcudaSuccess) return r;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        return mhcudaRuntimeError;
      return mhcudaSuccess;
    };
    auto time_one = [&](const Candidate &c, float *ms) { return time_on(sample, c, ms); };
    Candidate best{mhcudaKernelPersistent, saved};
    float best_ms = 0;
    MHCUDAResult r = time_one(best, &best_ms);  // warm-up
    if (r == mhcudaSuccess) r = time_one(best, &best_ms);
    auto consider = [&](Candidate c) {
      float ms = 0;
      if (r != mhcudaSuccess) return;
      r = time_one(c, &ms);
      if (r == mhcudaSuccess && ms < best_ms) {
        best = c;
        best_ms = ms;
      }
    };
    consider(Candidate{mhcudaKernelPlain, saved});
    const Candidate after_kernel = best;
    for (int spl : {1, 2, 4}) {
      if (spl == after_kernel.tuning.spl) continue;
      Candidate c = after_kernel;
      c.tuning.spl = spl;
      consider(c);
    }
    if (best.kernel == mhcudaKernelPersistent) {
      const Candidate after_spl = best;
      for (int seg : {512, 1024, 2048, 4096}) {
        if (seg == after_spl.tuning.seg) continue;
        Candidate c = after_spl;
        c.tuning.seg = seg;
        consider(c);
      }
      const Candidate after_seg = best;
      // (the hashing kernels hold at most 4 waves per SIMD, 16 per CU)
      for (int wpc : {8, 12, 16}) {
        if (wpc == after_seg.tuning.waves_per_cu) continue;
        Candidate c = after_seg;
        c.tuning.waves_per_cu = wpc;
        consider(c);
      }
      // column stripes (only where the table spans several stripes), with
      // and without heavy seeds where the stripe pieces can use them (the
      // best stripe size differs: seeded pieces gather less)
      const Candidate after_wpc = best;
      const bool seedable = wmh::launch_bounds_of(gen).codes != nullptr &&
                            gen->samples % 8 == 0 && after_wpc.tuning.spl == 2;
      uint32_t k = 1, min_nnz = 0;
      for (int seeds : {after_wpc.tuning.seeds, 1 - after_wpc.tuning.seeds}) {
        if (seeds != after_wpc.tuning.seeds && !seedable) continue;
        for (int mib : {0, 4, 8, 16, 32, 64, 256}) {
          if (mib == after_wpc.tuning.stripe_mib && seeds == after_wpc.tuning.seeds) continue;
          wmh::stripe_shape(gen->dim, gen->samples, mib, &k, &min_nnz);
          if (mib != 0 && k < 2) continue;
          if (mib == 0 && seeds != after_wpc.tuning.seeds) continue;  // seeds need stripes
          Candidate c = after_wpc;
          c.tuning.stripe_mib = mib;
          c.tuning.seeds = seeds;
          consider(c);
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    gen->timing = saved_timing;
    gen->tuning = saved;
    if (r != mhcudaSuccess) return r;
    if (verbosity > 0)
      printf("autotune key %llx: kernel %d spl %d seg %d waves/CU %d stripes %d MiB "
             "seeds %d (%.3f ms on %llu nnz)\n",
             (unsigned long long)key, best.kernel, best.tuning.spl, best.tuning.seg,
             best.tuning.waves_per_cu, best.tuning.stripe_mib, best.tuning.seeds, best_ms,
             (unsigned long long)have);
    wmh::TunedChoice choice;
    choice.kernel = best.kernel;
    choice.tuning = best.tuning;
    choice.ms = best_ms;
    choice.sample_nnz = have;
    it = gen->tuned.emplace(key, choice).first;
  } else {
    res.cached = 1;
  }
  gen->kernel = it->second.kernel;
  gen->tuning = it->second.tuning;
  res.kernel = it->second.kernel;
  res.samples_per_lane = it->second.tuning.spl;
  res.segment = it->second.tuning.seg;
  res.waves_per_cu = it->second.tuning.waves_per_cu;
  res.stripe_mib = it->second.tuning.stripe_mib;
  res.seeds = it->second.tuning.seeds;
  res.sample_ms = it->second.ms;
  res.sample_nnz = it->second.sample_nnz;
  if (result) *result = res;
  return mhcudaSuccess;
}
