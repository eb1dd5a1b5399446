// Test infrastructure: exhaustive checks of the engine's short division
// (wmh::recip_rn, wmh::div_rn_fast in minhashcuda_amd/csrc/wmh_div.h)
// against IEEE binary32 division on the GPU.  Built by tests/native/build_helpers.py
// into libfastdiv_check.so and driven through ctypes by
// tests/test_gpu_fastdiv.py.  No memory is read by the kernels; mismatch
// counts go to a small device buffer.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wmh_div.h"

namespace {

// res[0]: reciprocals that differ from 1.0f / r; res[1]: the first such bits
__global__ void rcp_kernel(uint32_t lo, uint32_t n, unsigned long long *res) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float r = __uint_as_float(lo + i);
  if (__float_as_uint(wmh::recip_rn(r)) != __float_as_uint(1.0f / r)) {
    if (atomicAdd(&res[0], 1ull) == 0ull) res[1] = lo + i;
  }
}

// One thread per b significand (b = 2^eb * [1, 2)); each loops over `na`
// significands of a = 2^ea * [1, 2) from a0.  res[2]: quotients that differ
// from a / b, res[3]: the first (b significand << 32 | a significand).
__global__ void div_kernel(int ea, int eb, uint32_t a0, uint32_t na, unsigned long long *res) {
  const uint32_t ib = blockIdx.x * 256u + threadIdx.x;  // < 2^23
  const float b = __uint_as_float(((uint32_t)(127 + eb) << 23) | ib);
  uint32_t bad = 0, first = 0;
  for (uint32_t k = 0; k < na; k++) {
    const uint32_t ia = a0 + k;
    const float a = __uint_as_float(((uint32_t)(127 + ea) << 23) | ia);
    const bool differ = __float_as_uint(wmh::div_rn_fast(a, b)) != __float_as_uint(a / b);
    first = (differ && bad == 0) ? ia : first;
    bad += differ;
  }
  if (bad) {
    if (atomicAdd(&res[2], (unsigned long long)bad) == 0ull)
      res[3] = ((unsigned long long)ib << 32) | first;
  }
}

int finish(unsigned long long *dev, unsigned long long *host) {
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  if (hipMemcpy(host, dev, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
    return 1;
  return hipFree(dev) == hipSuccess ? 0 : 1;
}

}  // namespace

extern "C" {

// Checks recip_rn(r) == 1.0f / r for every float whose bit pattern is in
// [lo, hi).  out[0] = mismatches, out[1] = bits of the first one.
int fastdiv_rcp_check(uint32_t lo, uint32_t hi, unsigned long long *out) {
  if (hi <= lo || out == nullptr) return 2;
  unsigned long long *dev = nullptr;
  if (hipMalloc(&dev, 4 * sizeof(unsigned long long)) != hipSuccess) return 1;
  if (hipMemset(dev, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return 1;
  for (uint64_t s = lo; s < hi; s += 1ull << 26) {
    const uint32_t n = (uint32_t)((hi - s) < (1ull << 26) ? (hi - s) : (1ull << 26));
    rcp_kernel<<<(n + 255) / 256, 256>>>((uint32_t)s, n, dev);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  unsigned long long h[4];
  if (finish(dev, h)) return 1;
  out[0] = h[0];
  out[1] = h[1];
  return 0;
}

// Checks div_rn_fast(a, b) == a / b for a = 2^ea * (1 + ia 2^-23), ia in
// [a0, a0 + na), and every b = 2^eb * [1, 2) (2^23 values), in launches of
// at most 2^17 a-values.  Exponents must keep both operands normal.
// out[0] = mismatches, out[1] = the first (b significand << 32 | a significand).
int fastdiv_pair_check(int ea, int eb, uint32_t a0, uint32_t na, unsigned long long *out) {
  if (out == nullptr || ea < -126 || ea > 127 || eb < -126 || eb > 127) return 2;
  if (a0 > (1u << 23) || na > (1u << 23) - a0) return 2;
  unsigned long long *dev = nullptr;
  if (hipMalloc(&dev, 4 * sizeof(unsigned long long)) != hipSuccess) return 1;
  if (hipMemset(dev, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return 1;
  const uint32_t step = 1u << 17;
  for (uint32_t s = 0; s < na; s += step) {
    const uint32_t n = na - s < step ? na - s : step;
    div_kernel<<<(1u << 23) / 256, 256>>>(ea, eb, a0 + s, n, dev);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  unsigned long long h[4];
  if (finish(dev, h)) return 1;
  out[0] = h[2];
  out[1] = h[3];
  return 0;
}

}  // extern "C"
