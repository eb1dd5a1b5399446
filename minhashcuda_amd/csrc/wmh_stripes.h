// Bounds of one column-stripe piece of a CSR row.  Shared by the persistent
// kernel and a host-side check of the partition property
// (tests/native/stripe_check.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wmh {

// Stripe piece k (of K) of the CSR positions [s, e): from the first position
// whose column is >= k * width to the first whose column is >= (k + 1) *
// width, or to e for the last stripe.  The two binary searches run together
// (a wider search needs fewer rounds but more loads, and the gathers that
// bound the kernel compete with them for the texture addresser; units of
// several pieces hide the rounds' latency instead).  Whatever the column
// order, a search's result never decreases as its key grows (two keys take
// the same path until the first probed column lying between them, where the
// smaller key turns left and the larger right), so the end of piece k is the
// start of piece k + 1 and the K pieces partition [s, e).
__host__ __device__ inline void stripe_piece(const uint32_t *__restrict__ cols, uint32_t s,
                                             uint32_t e, uint32_t k, uint32_t K,
                                             uint32_t width, uint32_t &a, uint32_t &b) {
  const uint64_t x0 = (uint64_t)k * width, x1 = x0 + width;
  uint32_t lo0 = s, hi0 = k == 0 ? s : e;
  uint32_t lo1 = k + 1 == K ? e : s, hi1 = e;
  while (lo0 < hi0 || lo1 < hi1) {
    const uint32_t m0 = lo0 + ((hi0 - lo0) >> 1), m1 = lo1 + ((hi1 - lo1) >> 1);
    const uint32_t c0 = lo0 < hi0 ? cols[m0] : 0u;
    const uint32_t c1 = lo1 < hi1 ? cols[m1] : 0u;
    if (lo0 < hi0) {
      if (c0 < x0) lo0 = m0 + 1;
      else hi0 = m0;
    }
    if (lo1 < hi1) {
      if (c1 < x1) lo1 = m1 + 1;
      else hi1 = m1;
    }
  }
  a = lo0;
  b = lo1;
}

}  // namespace wmh
