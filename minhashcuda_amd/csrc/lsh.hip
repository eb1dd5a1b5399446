// LSH candidate generation on the GPU (no counterpart in the reference; the
// reference's users compare hash pairs with datasketch's WeightedMinHash.jaccard,
// test.py:104-106, and find candidates on the host).
//
// Input: the band keys of mhcuda_lsh_band_keys, keys[row * bands + g].  Rows
// whose keys agree in some band are candidate pairs.  Per band the (key, row)
// pairs are radix-sorted by key; inside every run of equal keys holding 2 to
// max_bucket rows (the reserved key 2^64 - 1 of bands with an empty row's
// sample excluded) every pair of rows is emitted as the 64-bit code
// min << 32 | max; the codes of all bands are then sorted and deduplicated.
// Each element finds its run by looking at most max_bucket positions either
// way, so runs longer than max_bucket (skipped) cost O(max_bucket) per
// element, never more.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "minhashcuda.h"
#include "wmh_runtime.h"

namespace {

constexpr int kBlock = 256;
constexpr unsigned long long kReservedKey = ~0ull;

unsigned grid_for(uint64_t n) {
  const uint64_t blocks = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 1u << 20));
}

// keys [rows][bands] -> per band contiguous keys and row ids [bands][rows]
__global__ __launch_bounds__(kBlock) void lsh_extract_kernel(
    const unsigned long long *__restrict__ keys, uint32_t rows, uint32_t bands,
    unsigned long long *__restrict__ band_keys, uint32_t *__restrict__ band_rows) {
  const uint64_t total = (uint64_t)rows * bands;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint32_t g = (uint32_t)(i / rows);
    const uint32_t r = (uint32_t)(i - (uint64_t)g * rows);
    band_keys[i] = keys[(uint64_t)r * bands + g];
    band_rows[i] = r;
  }
}

// The run of equal keys around sorted position p of its band (positions
// [base, base + rows)): [*start, *end), or an empty run when the key is
// reserved or the run holds more than max_bucket rows.
__device__ __forceinline__ void run_of(const unsigned long long *__restrict__ k, uint64_t base,
                                       uint64_t end_of_band, uint64_t p, uint32_t max_bucket,
                                       uint64_t *start, uint64_t *end) {
  const unsigned long long key = k[p];
  *start = *end = p;
  if (key == kReservedKey) return;
  uint64_t s = p, e = p + 1;
  while (s > base && k[s - 1] == key) {
    if (e - s >= max_bucket) return;  // longer than max_bucket: skipped
    s--;
  }
  while (e < end_of_band && k[e] == key) {
    if (e - s >= max_bucket) return;
    e++;
  }
  *start = s;
  *end = e;
}

// counts[p] = pairs position p forms with the later positions of its run
__global__ __launch_bounds__(kBlock) void lsh_count_kernel(
    const unsigned long long *__restrict__ sorted_keys, uint32_t rows, uint32_t bands,
    uint32_t max_bucket, unsigned long long *__restrict__ counts) {
  const uint64_t total = (uint64_t)rows * bands;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t base = i / rows * rows;
    uint64_t s, e;
    run_of(sorted_keys, base, base + rows, i, max_bucket, &s, &e);
    counts[i] = e - s >= 2 ? e - 1 - i : 0;
  }
}

__global__ __launch_bounds__(kBlock) void lsh_emit_kernel(
    const unsigned long long *__restrict__ sorted_keys, const uint32_t *__restrict__ sorted_rows,
    uint32_t rows, uint32_t bands, uint32_t max_bucket,
    const unsigned long long *__restrict__ offsets, unsigned long long *__restrict__ pairs) {
  const uint64_t total = (uint64_t)rows * bands;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t base = i / rows * rows;
    uint64_t s, e;
    run_of(sorted_keys, base, base + rows, i, max_bucket, &s, &e);
    if (e - s < 2) continue;
    const uint32_t a = sorted_rows[i];
    unsigned long long *out = pairs + offsets[i];
    for (uint64_t q = i + 1; q < e; q++) {
      const uint32_t b = sorted_rows[q];
      const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
      *out++ = ((unsigned long long)lo << 32) | hi;
    }
  }
}

__global__ void lsh_total_kernel(const unsigned long long *__restrict__ offsets,
                                 const unsigned long long *__restrict__ counts, uint64_t n,
                                 unsigned long long *__restrict__ total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? offsets[n - 1] + counts[n - 1] : 0ull;
}

}  // namespace

extern "C" MHCUDAResult mhcuda_lsh_candidates(const MinhashCudaGenerator *gen,
                                              const uint64_t *keys, uint32_t rows,
                                              uint32_t bands, uint32_t max_bucket,
                                              uint64_t *pairs, uint64_t capacity,
                                              uint64_t *count, void *stream) {
  if (!gen || !keys || !count || bands == 0 || max_bucket < 2 || (capacity && !pairs))
    return mhcudaInvalidArguments;
  *count = 0;
  if (rows < 2) return mhcudaSuccess;
  auto *g = const_cast<MinhashCudaGenerator *>(gen);
  // the generator's grow-only buffers (gen->lsh) are used under its lock
  std::lock_guard<std::mutex> lock(g->mu);
  if (hipSetDevice(g->device) != hipSuccess) return mhcudaNoSuchDevice;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g->stream;
  if (!stream && wmh::follow_default_stream(g) != hipSuccess) return mhcudaRuntimeError;
  wmh::LshBuffers &B = g->lsh;
  const uint64_t n = (uint64_t)rows * bands;
  using u64 = unsigned long long;
  auto reserve = [&](wmh::DevBuf &b, size_t bytes) {
    return b.reserve(std::max<size_t>(bytes, 256)) == hipSuccess;
  };
  if (!reserve(B.keys_in, n * 8) || !reserve(B.rows_in, n * 4) || !reserve(B.keys_sorted, n * 8) ||
      !reserve(B.rows_sorted, n * 4) || !reserve(B.counts, n * 8) || !reserve(B.offsets, n * 8) ||
      !reserve(B.scalars, 256))
    return mhcudaMemoryAllocationFailure;
  u64 *k_in = (u64 *)B.keys_in.ptr, *k_sorted = (u64 *)B.keys_sorted.ptr;
  uint32_t *r_in = (uint32_t *)B.rows_in.ptr, *r_sorted = (uint32_t *)B.rows_sorted.ptr;
  u64 *counts = (u64 *)B.counts.ptr, *offsets = (u64 *)B.offsets.ptr;
  u64 *scalars = (u64 *)B.scalars.ptr;
  hipLaunchKernelGGL(lsh_extract_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s,
                     reinterpret_cast<const u64 *>(keys), rows, bands, k_in, r_in);
  if (hipGetLastError() != hipSuccess) return mhcudaRuntimeError;
  // temporary storage: the largest any of the library calls below needs
  size_t t_sort = 0, t_scan = 0;
  if (rocprim::radix_sort_pairs(nullptr, t_sort, k_in, k_sorted, r_in, r_sorted, rows, 0, 64,
                                s) != hipSuccess ||
      rocprim::exclusive_scan(nullptr, t_scan, counts, offsets, 0ull, (size_t)n,
                              rocprim::plus<u64>(), s) != hipSuccess)
    return mhcudaRuntimeError;
  if (!reserve(B.temp, std::max(t_sort, t_scan))) return mhcudaMemoryAllocationFailure;
  for (uint32_t g = 0; g < bands; g++) {
    size_t bytes = B.temp.cap;
    const uint64_t o = (uint64_t)g * rows;
    if (rocprim::radix_sort_pairs(B.temp.ptr, bytes, k_in + o, k_sorted + o, r_in + o,
                                  r_sorted + o, rows, 0, 64, s) != hipSuccess)
      return mhcudaRuntimeError;
  }
  hipLaunchKernelGGL(lsh_count_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, k_sorted, rows,
                     bands, max_bucket, counts);
  if (hipGetLastError() != hipSuccess) return mhcudaRuntimeError;
  {
    size_t bytes = B.temp.cap;
    if (rocprim::exclusive_scan(B.temp.ptr, bytes, counts, offsets, 0ull, (size_t)n,
                                rocprim::plus<u64>(), s) != hipSuccess)
      return mhcudaRuntimeError;
  }
  hipLaunchKernelGGL(lsh_total_kernel, dim3(1), dim3(64), 0, s, offsets, counts, n, scalars);
  u64 total = 0;
  if (hipMemcpyAsync(&total, scalars, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return mhcudaMemoryCopyError;
  if (total == 0) return mhcudaSuccess;
  if (!reserve(B.pairs, total * 8) || !reserve(B.pairs_sorted, total * 8))
    return mhcudaMemoryAllocationFailure;
  u64 *pr = (u64 *)B.pairs.ptr, *pr_sorted = (u64 *)B.pairs_sorted.ptr;
  hipLaunchKernelGGL(lsh_emit_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, k_sorted, r_sorted,
                     rows, bands, max_bucket, offsets, pr);
  if (hipGetLastError() != hipSuccess) return mhcudaRuntimeError;
  size_t t_sort2 = 0, t_uniq = 0;
  if (rocprim::radix_sort_keys(nullptr, t_sort2, pr, pr_sorted, (size_t)total, 0, 64, s) !=
          hipSuccess ||
      rocprim::unique(nullptr, t_uniq, pr_sorted, pr, scalars + 1, (size_t)total,
                      rocprim::equal_to<u64>(), s) != hipSuccess)
    return mhcudaRuntimeError;
  if (!reserve(B.temp, std::max(t_sort2, t_uniq))) return mhcudaMemoryAllocationFailure;
  {
    size_t bytes = B.temp.cap;
    if (rocprim::radix_sort_keys(B.temp.ptr, bytes, pr, pr_sorted, (size_t)total, 0, 64, s) !=
        hipSuccess)
      return mhcudaRuntimeError;
    bytes = B.temp.cap;
    if (rocprim::unique(B.temp.ptr, bytes, pr_sorted, pr, scalars + 1, (size_t)total,
                        rocprim::equal_to<u64>(), s) != hipSuccess)
      return mhcudaRuntimeError;
  }
  u64 unique_count = 0;
  if (hipMemcpyAsync(&unique_count, scalars + 1, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return mhcudaMemoryCopyError;
  const uint64_t n_out = std::min<uint64_t>(unique_count, capacity);
  if (n_out && (hipMemcpyAsync(pairs, pr, n_out * 8, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
    return mhcudaMemoryCopyError;
  *count = unique_count;
  return mhcudaSuccess;
}
