// Host build of the persistent kernel's stripe_piece (wmh_stripes.h) for the
// CPU test of the partition property (tests/test_stripe_pieces.py): the same
// source, compiled as host C++.
#include <stdint.h>

#include "wmh_stripes.h"

// bounds[2k], bounds[2k + 1] = piece k of the CSR positions [s, e), k < K
extern "C" void wmh_stripe_pieces(const uint32_t *cols, uint32_t s, uint32_t e, uint32_t K,
                                  uint32_t width, uint32_t *bounds) {
  for (uint32_t k = 0; k < K; k++)
    wmh::stripe_piece(cols, s, e, k, K, width, bounds[2 * k], bounds[2 * k + 1]);
}
