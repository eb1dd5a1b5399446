// Parameter files: the generator's three tables on disk, so that hashes can be
// reproduced in a later process (no counterpart in the reference, whose only
// persistence is retrieve_random_vars / assign_random_vars into caller arrays,
// minhashcuda.cc:254-285, driven from Python as in test.py:154-182).
//
// Format, version 1 (little-endian):
//
//   offset  size  field
//        0     8  magic "WMHPARAM"
//        8     4  version (1)
//       12     4  kind: 1 = tables follow, 2 = seeded (regenerate from seed)
//       16     4  dim
//       20     4  samples
//       24     4  seed (kind 2; 0 otherwise)
//       28     4  recipe of the on-device generator (kind 2: 1; 0 otherwise)
//       32     8  checksum of rs     (kind 1; 0 otherwise)
//       40     8  checksum of ln_cs
//       48     8  checksum of betas
//       56     8  header checksum: h = 0; h = mix64(h ^ w) over the seven
//                 64-bit words at offsets 0..48
//       64        kind 1: rs, ln_cs, betas, each [samples][dim] float32
//                 (the layout of retrieve_random_vars / assign_random_vars)
//
// Table checksum: sum over positions p = k * dim + j (mod 2^64) of
// ((bits(table[k][j]) ^ p * 0x9E3779B97F4A7C15) * 0xBF58476D1CE4E5B9); the
// multiplier is odd, so any single flipped bit changes it.  mix64 is the
// splitmix64 finalizer.
//
// Files are written to a temporary name next to the target and renamed into
// place once complete.  The tables move through two pinned buffers of a few
// tens of MiB, so a table set of any size is streamed: device relayout and
// DMA of one column chunk overlap the file I/O of the previous one.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "minhashcuda.h"
#include "wmh_kernels.h"
#include "wmh_runtime.h"

namespace {

constexpr char kMagic[8] = {'W', 'M', 'H', 'P', 'A', 'R', 'A', 'M'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kKindTables = 1;
constexpr uint32_t kKindSeeded = 2;
// Identifies the on-device generator (Philox4x32-10 streams, the Gamma(2,1)
// sampler and the uniform conversions); a seeded file is only loaded by a
// library with the same recipe.
constexpr uint32_t kRecipe = 1;
constexpr uint64_t kHeaderBytes = 64;
constexpr uint64_t kPosMul = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kSumMul = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t kChunkBytes = 32ull << 20;  // per pinned buffer

struct Header {
  char magic[8];
  uint32_t version, kind, dim, samples, seed, recipe;
  uint64_t checksum[3];  // rs, ln_cs, betas
  uint64_t header_checksum;
};
static_assert(sizeof(Header) == kHeaderBytes, "packed 64-byte header");

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t header_checksum(const Header &h) {
  uint64_t w[7];
  std::memcpy(w, &h, sizeof(w));
  uint64_t x = 0;
  for (uint64_t v : w) x = mix64(x ^ v);
  return x;
}

// Checksum terms of `n` consecutive values of one table starting at position p.
uint64_t checksum_run(const uint32_t *bits, uint64_t n, uint64_t p) {
  uint64_t sum = 0, x = p * kPosMul;
  for (uint64_t i = 0; i < n; i++, x += kPosMul) sum += (bits[i] ^ x) * kSumMul;
  return sum;
}

// Columns per chunk: a multiple of 64 whose [samples][n] float block fits one
// pinned buffer (MHCUDA_PARAMS_CHUNK_BYTES overrides the buffer size).
uint32_t file_chunk(uint32_t dim, uint32_t samples) {
  uint64_t bytes = kChunkBytes;
  if (const char *env = std::getenv("MHCUDA_PARAMS_CHUNK_BYTES")) {
    const unsigned long long v = std::strtoull(env, nullptr, 10);
    if (v > 0) bytes = v;
  }
  uint64_t c = bytes / (4ull * samples);
  c = std::max<uint64_t>(64, c & ~63ull);
  return (uint32_t)std::min<uint64_t>(c, dim);
}

bool write_all(int fd, const void *buf, uint64_t n, uint64_t off) {
  const char *p = static_cast<const char *>(buf);
  while (n > 0) {
    const ssize_t w = pwrite(fd, p, (size_t)std::min<uint64_t>(n, 1ull << 30), (off_t)off);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return false;
    p += w;
    off += (uint64_t)w;
    n -= (uint64_t)w;
  }
  return true;
}

bool read_all(int fd, void *buf, uint64_t n, uint64_t off) {
  char *p = static_cast<char *>(buf);
  while (n > 0) {
    const ssize_t r = pread(fd, p, (size_t)std::min<uint64_t>(n, 1ull << 30), (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    p += r;
    off += (uint64_t)r;
    n -= (uint64_t)r;
  }
  return true;
}

// Reads and validates the header; mhcudaMemoryCopyError if the file cannot be
// read, mhcudaInvalidArguments if it is not a well-formed version-1 file.
MHCUDAResult read_header(int fd, Header *h) {
  struct stat st;
  if (fstat(fd, &st) != 0) return mhcudaMemoryCopyError;
  if ((uint64_t)st.st_size < kHeaderBytes) return mhcudaInvalidArguments;
  if (!read_all(fd, h, kHeaderBytes, 0)) return mhcudaMemoryCopyError;
  if (std::memcmp(h->magic, kMagic, 8) != 0 || h->version != kVersion ||
      h->header_checksum != header_checksum(*h) || h->dim == 0 || h->samples == 0 ||
      h->samples > 0xffffu)
    return mhcudaInvalidArguments;
  uint64_t expect = kHeaderBytes;
  if (h->kind == kKindTables) {
    expect += 12ull * h->dim * h->samples;
  } else if (h->kind != kKindSeeded) {
    return mhcudaInvalidArguments;
  }
  if ((uint64_t)st.st_size != expect) return mhcudaInvalidArguments;
  return mhcudaSuccess;
}

// Tables in file order (rs, ln_cs, betas) -> their slot in the column blocks
// of the device layout (0 = rs, 1 = betas, 2 = ln_cs).
constexpr uint32_t kSlot[3] = {0, 2, 1};

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

struct Events {
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool create() {
    return hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
  }
  ~Events() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace

extern "C" {

MHCUDAResult mhcuda_params_file_info(const char *path, uint32_t *dim, uint16_t *samples) {
  if (!path) return mhcudaInvalidArguments;
  Fd f;
  f.fd = open(path, O_RDONLY | O_CLOEXEC);
  if (f.fd < 0) return mhcudaMemoryCopyError;
  Header h;
  const MHCUDAResult r = read_header(f.fd, &h);
  if (r != mhcudaSuccess) return r;
  if (dim) *dim = h.dim;
  if (samples) *samples = (uint16_t)h.samples;
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_save_random_vars(const MinhashCudaGenerator *cgen, const char *path,
                                     int allow_seed_only) {
  if (!cgen || !path) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  const int verbosity = gen->verbosity;
  if (gen->params_source == MinhashCudaGenerator::kParamsNone) {
    if (verbosity > 0) printf("save_random_vars: the generator holds no tables yet\n");
    return mhcudaInvalidArguments;
  }
  if (hipSetDevice(gen->device) != hipSuccess) return mhcudaNoSuchDevice;
  const uint32_t dim = gen->dim, samples = gen->samples;
  const bool seeded = allow_seed_only && gen->params_source == MinhashCudaGenerator::kParamsSeeded;

  Header h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.kind = seeded ? kKindSeeded : kKindTables;
  h.dim = dim;
  h.samples = samples;
  h.seed = seeded ? gen->seed : 0u;
  h.recipe = seeded ? kRecipe : 0u;

  const std::string tmp = std::string(path) + ".tmp." + std::to_string((long)getpid());
  Fd f;
  f.fd = open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (f.fd < 0) return mhcudaMemoryCopyError;
  auto fail = [&](MHCUDAResult r) {
    (void)hipStreamSynchronize(gen->stream);  // no DMA into the pinned buffers after return
    close(f.fd);
    f.fd = -1;
    unlink(tmp.c_str());
    return r;
  };

  if (!seeded) {
    // stream the tables out: relayout + DMA of chunk i overlap the write of chunk i-1
    const uint32_t chunk = file_chunk(dim, samples);
    const uint64_t block = (uint64_t)chunk * samples * sizeof(float);
    wmh::PinnedBuf pin[2];
    Events evs;
    if (!pin[0].reserve(block) || !pin[1].reserve(block) ||
        gen->staging.reserve(2 * block) != hipSuccess)
      return fail(mhcudaMemoryAllocationFailure);
    if (!evs.create()) return fail(mhcudaRuntimeError);
    const uint32_t nchunks = (dim + chunk - 1) / chunk;
    const uint64_t steps = 3ull * nchunks;
    const uint64_t table_bytes = 4ull * dim * samples;
    auto drain = [&](uint64_t i) -> MHCUDAResult {  // write step i from its pinned buffer
      const uint32_t t = (uint32_t)(i / nchunks);
      const uint32_t j0 = (uint32_t)(i % nchunks) * chunk;
      const uint32_t n = std::min(chunk, dim - j0);
      if (hipEventSynchronize(evs.ev[i & 1]) != hipSuccess) return mhcudaMemoryCopyError;
      const uint32_t *bits = reinterpret_cast<const uint32_t *>(pin[i & 1].ptr);
      for (uint32_t k = 0; k < samples; k++) {
        const uint64_t p = (uint64_t)k * dim + j0;
        h.checksum[t] += checksum_run(bits + (uint64_t)k * n, n, p);
        if (!write_all(f.fd, bits + (uint64_t)k * n, 4ull * n,
                       kHeaderBytes + t * table_bytes + 4ull * p))
          return mhcudaMemoryCopyError;
      }
      return mhcudaSuccess;
    };
    for (uint64_t i = 0; i < steps; i++) {
      const uint32_t t = (uint32_t)(i / nchunks);
      const uint32_t j0 = (uint32_t)(i % nchunks) * chunk;
      const uint32_t n = std::min(chunk, dim - j0);
      float *dstage = reinterpret_cast<float *>(static_cast<char *>(gen->staging.ptr) +
                                                (i & 1) * block);
      if (wmh::launch_gather_params(gen->params, j0, n, kSlot[t], dstage, samples,
                                    gen->stream) != hipSuccess ||
          hipMemcpyAsync(pin[i & 1].ptr, dstage, 4ull * n * samples, hipMemcpyDeviceToHost,
                         gen->stream) != hipSuccess ||
          hipEventRecord(evs.ev[i & 1], gen->stream) != hipSuccess)
        return fail(mhcudaRuntimeError);
      if (i > 0) {
        const MHCUDAResult r = drain(i - 1);
        if (r != mhcudaSuccess) return fail(r);
      }
    }
    const MHCUDAResult r = drain(steps - 1);
    if (r != mhcudaSuccess) return fail(r);
  }
  h.header_checksum = header_checksum(h);
  if (!write_all(f.fd, &h, kHeaderBytes, 0) || fsync(f.fd) != 0) return fail(mhcudaMemoryCopyError);
  if (close(f.fd) != 0) {
    f.fd = -1;
    unlink(tmp.c_str());
    return mhcudaMemoryCopyError;
  }
  f.fd = -1;
  if (rename(tmp.c_str(), path) != 0) {
    unlink(tmp.c_str());
    return mhcudaMemoryCopyError;
  }
  return mhcudaSuccess;
}

MHCUDAResult mhcuda_load_random_vars(const MinhashCudaGenerator *cgen, const char *path) {
  if (!cgen || !path) return mhcudaInvalidArguments;
  auto *gen = const_cast<MinhashCudaGenerator *>(cgen);
  std::lock_guard<std::mutex> lock(gen->mu);
  if (gen->servers.load() > 0) return mhcudaInvalidArguments;  // a server reads the tables
  const int verbosity = gen->verbosity;
  Fd f;
  f.fd = open(path, O_RDONLY | O_CLOEXEC);
  if (f.fd < 0) return mhcudaMemoryCopyError;
  Header h;
  MHCUDAResult r = read_header(f.fd, &h);
  if (r != mhcudaSuccess) return r;
  const uint32_t dim = gen->dim, samples = gen->samples;
  if (h.dim != dim || h.samples != samples) {
    if (verbosity > 0)
      printf("load_random_vars: the file holds %u x %u tables, the generator needs %u x %u\n",
             h.samples, h.dim, samples, dim);
    return mhcudaInvalidArguments;
  }
  if (h.kind == kKindSeeded && h.recipe != kRecipe) return mhcudaInvalidArguments;
  if (hipSetDevice(gen->device) != hipSuccess) return mhcudaNoSuchDevice;
  gen->params_source = MinhashCudaGenerator::kParamsNone;  // until every table is in
  gen->fast_div = false;                                   // until the new r are checked
  gen->bounded = false;                                    // until the bound table is rebuilt

  if (h.kind == kKindSeeded) {
    if (wmh::launch_gen_params(gen->params, dim, samples, h.seed, gen->stream) != hipSuccess ||
        hipStreamSynchronize(gen->stream) != hipSuccess)
      return mhcudaRuntimeError;
    gen->params_source = MinhashCudaGenerator::kParamsSeeded;
    gen->seed = h.seed;
    return wmh::refresh_division_mode(gen);
  }

  // stream the tables in: the read of chunk i overlaps DMA + relayout of chunk i-1
  const uint32_t chunk = file_chunk(dim, samples);
  const uint64_t block = (uint64_t)chunk * samples * sizeof(float);
  wmh::PinnedBuf pin[2];
  Events evs;
  if (!pin[0].reserve(block) || !pin[1].reserve(block) ||
      gen->staging.reserve(2 * block) != hipSuccess)
    return mhcudaMemoryAllocationFailure;
  if (!evs.create()) return mhcudaRuntimeError;
  const uint32_t nchunks = (dim + chunk - 1) / chunk;
  const uint64_t steps = 3ull * nchunks;
  const uint64_t table_bytes = 4ull * dim * samples;
  uint64_t sums[3] = {0, 0, 0};
  bool recorded[2] = {false, false};
  for (uint64_t i = 0; i < steps && r == mhcudaSuccess; i++) {
    const uint32_t t = (uint32_t)(i / nchunks);
    const uint32_t j0 = (uint32_t)(i % nchunks) * chunk;
    const uint32_t n = std::min(chunk, dim - j0);
    // the DMA of step i-2 read from this pinned buffer
    if (recorded[i & 1] && hipEventSynchronize(evs.ev[i & 1]) != hipSuccess) {
      r = mhcudaMemoryCopyError;
      break;
    }
    uint32_t *bits = reinterpret_cast<uint32_t *>(pin[i & 1].ptr);
    for (uint32_t k = 0; k < samples; k++) {
      const uint64_t p = (uint64_t)k * dim + j0;
      if (!read_all(f.fd, bits + (uint64_t)k * n, 4ull * n, kHeaderBytes + t * table_bytes + 4ull * p)) {
        r = mhcudaMemoryCopyError;
        break;
      }
      sums[t] += checksum_run(bits + (uint64_t)k * n, n, p);
    }
    if (r != mhcudaSuccess) break;
    float *dstage = reinterpret_cast<float *>(static_cast<char *>(gen->staging.ptr) +
                                              (i & 1) * block);
    if (hipMemcpyAsync(dstage, bits, 4ull * n * samples, hipMemcpyHostToDevice, gen->stream) !=
            hipSuccess ||
        wmh::launch_scatter_params(dstage, n, j0, kSlot[t], gen->params, samples, gen->stream) !=
            hipSuccess ||
        hipEventRecord(evs.ev[i & 1], gen->stream) != hipSuccess) {
      r = mhcudaRuntimeError;
      break;
    }
    recorded[i & 1] = true;
  }
  if (hipStreamSynchronize(gen->stream) != hipSuccess && r == mhcudaSuccess) r = mhcudaRuntimeError;
  if (r == mhcudaSuccess &&
      (sums[0] != h.checksum[0] || sums[1] != h.checksum[1] || sums[2] != h.checksum[2])) {
    if (verbosity > 0) printf("load_random_vars: table checksum mismatch in %s\n", path);
    r = mhcudaInvalidArguments;
  }
  if (r != mhcudaSuccess) {
    // the tables are partly overwritten: leave them as a deferred init does
    (void)hipMemsetAsync(gen->params, 0, (size_t)dim * 3u * gen->sp * sizeof(float), gen->stream);
    (void)hipStreamSynchronize(gen->stream);
    gen->fast_div = false;
    gen->bounded = false;
    return r;
  }
  gen->params_source = MinhashCudaGenerator::kParamsTables;
  return wmh::refresh_division_mode(gen);
}

}  // extern "C"
