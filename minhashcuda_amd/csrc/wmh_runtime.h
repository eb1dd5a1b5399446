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

// One device buffer set of the pipelined host-memory path, with the events
// that order its upload, hashing and download.
struct PieceBuffers {
  DevBuf weights, cols, rows, hashes;
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
};

}  // namespace wmh
