// Host-side helpers shared by the runtime sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

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

}  // namespace wmh
