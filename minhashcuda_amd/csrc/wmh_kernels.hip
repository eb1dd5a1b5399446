// gfx950 (MI355X, CDNA4) kernels of the Weighted MinHash engine.
//
// Hot path (reference: kernel.cu:56-112, weighted_minhash_cuda).  For a CSR
// row with nonzeros (j_p, w_p) in CSR order and every sample k:
//   lw   = ln(w_p)                       (once per nonzero, wmh_logf)
//   t    = floor(lw / r[k,j] + beta[k,j])
//   ln_a = (ln_c[k,j] - (t - beta[k,j]) * r[k,j]) - r[k,j]
//   keep the first p with the smallest ln_a (strict <, initial FLT_MAX),
//   output (j_p, (uint32)(int32)t).
// All arithmetic is binary32, unfused (this file is compiled with
// -ffp-contract=off), with IEEE division and preserved denormals.
//
// Mapping to CDNA4: one wavefront owns one (row, sample tile).  Lane l holds
// V consecutive samples, so a wavefront covers 64*V samples with running
// argmins in VGPRs and never needs a cross-lane reduction.  The row's
// (col, w) pairs are read 64 at a time with one coalesced load per lane, each
// lane takes the logarithm of its own weight, and the pairs are then
// broadcast to SGPRs with v_readlane: the column becomes a scalar base address
// and each nonzero costs three coalesced 64*V*4-byte loads from the column's
// contiguous parameter block.  No MFMA: this is a gather + compare, not a
// contraction; the kernel is bound by HBM bandwidth on those gathers.
#include <float.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "wmh_bound.h"
#include "wmh_div.h"
#include "wmh_kernels.h"
#include "wmh_math.h"
#include "wmh_stripes.h"

namespace wmh {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // 4 wavefronts per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t i) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)i);
}

__device__ __forceinline__ float readlane_f32(float v, uint32_t i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)i));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  return ((unsigned long long)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v);
}

// ---------------------------------------------------------------------------
// Parameter generation: Philox4x32-10 (Salmon et al., SC'11) + Gamma(2,1).
// Replaces the reference's cuRAND + gamma_cuda/log_cuda (kernel.cu:18-36,
// minhashcuda.cc:176-200) with a counter-based generator whose stream is a
// pure function of (seed, k, j), so any device -- or the CPU -- regenerates
// the same table without communication.

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; round++) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

constexpr uint32_t kPhiloxKey1 = 0x4d484344u;  // "MHCD"

__device__ __forceinline__ float u_open(uint32_t x) {
  return (float)((x >> 9) * 2u + 1u) * 0x1p-24f;  // odd multiples of 2^-24 in (0,1)
}
__device__ __forceinline__ float u_half_open(uint32_t x) {
  return (float)(x >> 8) * 0x1p-24f;  // multiples of 2^-24 in [0,1)
}

__global__ __launch_bounds__(kBlock) void gen_params_kernel(
    float *__restrict__ params, uint32_t dim, uint32_t samples, uint32_t sp,
    uint32_t seed) {
  WMH_NO_CONTRACT
  const uint64_t total = (uint64_t)dim * sp;
  for (uint64_t idx = (uint64_t)blockIdx.x * kBlock + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * kBlock) {
    const uint32_t j = (uint32_t)(idx / sp);
    const uint32_t k = (uint32_t)(idx - (uint64_t)j * sp);
    float r = 0.0f, beta = 0.0f, lc = 0.0f;
    if (k < samples) {
      const uint64_t n = (uint64_t)k * dim + j;  // logical [samples][dim] index
      const U4 a = philox4x32_10(U4{(uint32_t)n, (uint32_t)(n >> 32), 0u, 0u}, seed, kPhiloxKey1);
      const U4 b = philox4x32_10(U4{(uint32_t)n, (uint32_t)(n >> 32), 1u, 0u}, seed, kPhiloxKey1);
      r = -wmh_logf(u_open(a.x) * u_open(a.y));
      lc = wmh_logf(-wmh_logf(u_open(a.z) * u_open(a.w)));
      beta = u_half_open(b.x);
    }
    float *col = params + (uint64_t)j * 3u * sp;
    col[k] = r;
    col[sp + k] = beta;
    col[2u * sp + k] = lc;
  }
}

// ---------------------------------------------------------------------------
// Re-layout between the reference's [samples][dim] tables (kernel.cu:97-99,
// python.cc:157) and the internal [dim][3][sp] blocks: 64x64 LDS tiles so that
// both the read and the write side are coalesced.

__global__ __launch_bounds__(kBlock) void scatter_params_kernel(
    const float *__restrict__ staging, uint32_t ncols, uint32_t j0,
    uint32_t which, float *__restrict__ params, uint32_t samples, uint32_t sp) {
  __shared__ float tile[64][65];
  const uint32_t jj0 = blockIdx.x * 64u, k0 = blockIdx.y * 64u;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  for (uint32_t r = ty; r < 64u; r += kWavesPerBlock) {
    const uint32_t k = k0 + r, jj = jj0 + tx;
    tile[r][tx] = (k < samples && jj < ncols) ? staging[(uint64_t)k * ncols + jj] : 0.0f;
  }
  __syncthreads();
  for (uint32_t r = ty; r < 64u; r += kWavesPerBlock) {
    const uint32_t jj = jj0 + r, k = k0 + tx;
    if (jj < ncols && k < sp)
      params[(uint64_t)(j0 + jj) * 3u * sp + which * sp + k] = tile[tx][r];
  }
}

__global__ __launch_bounds__(kBlock) void gather_params_kernel(
    const float *__restrict__ params, uint32_t j0, uint32_t ncols,
    uint32_t which, float *__restrict__ staging, uint32_t samples, uint32_t sp) {
  __shared__ float tile[64][65];
  const uint32_t jj0 = blockIdx.x * 64u, k0 = blockIdx.y * 64u;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  for (uint32_t r = ty; r < 64u; r += kWavesPerBlock) {
    const uint32_t jj = jj0 + r, k = k0 + tx;
    tile[r][tx] = (jj < ncols && k < samples)
                      ? params[(uint64_t)(j0 + jj) * 3u * sp + which * sp + k]
                      : 0.0f;
  }
  __syncthreads();
  for (uint32_t r = ty; r < 64u; r += kWavesPerBlock) {
    const uint32_t k = k0 + r, jj = jj0 + tx;
    if (k < samples && jj < ncols) staging[(uint64_t)k * ncols + jj] = tile[tx][r];
  }
}

// Range checks of the tables, one bit each in *bad (one ballot and at most
// one atomic per wavefront and bit):
//   kRangeDiv:   some r of a real sample is outside [2^-60, 2^60] or NaN (the
//                short division div_rn_fast needs it inside);
//   kRangeBound: some beta is outside [0, 1) or some ln_c is not finite or
//                beyond 2^60 in magnitude (the skip bound of wmh_bound.h
//                needs it inside, and r inside the division range).
__global__ __launch_bounds__(kBlock) void params_range_kernel(const float *__restrict__ params,
                                                              uint32_t dim, uint32_t samples,
                                                              uint32_t sp,
                                                              uint32_t *__restrict__ bad) {
  const uint64_t total = (uint64_t)dim * samples;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  // every lane of a wavefront runs the same number of iterations, so the
  // ballots are taken with the whole wavefront present
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock + (threadIdx.x & ~(uint64_t)(kWave - 1));
       base < total; base += stride) {
    const uint64_t idx = base + lane_id();
    bool out_div = false, out_bound = false;
    if (idx < total) {
      const uint32_t j = (uint32_t)(idx / samples);
      const uint32_t k = (uint32_t)(idx - (uint64_t)j * samples);
      const float *col = params + (uint64_t)j * 3u * sp;
      const float r = col[k], beta = col[sp + k], lc = col[2u * sp + k];
      out_div = !(r >= kFastDivMinB && r <= kFastDivMaxB);
      out_bound = !(beta >= 0.0f && beta < 1.0f) ||
                  !(lc >= -WMH_BOUND_MAX_ABS_LC && lc <= WMH_BOUND_MAX_ABS_LC);
    }
    const unsigned long long md = __ballot(out_div), mb = __ballot(out_bound);
    if (lane_id() == 0 && (md | mb) != 0ull)
      atomicOr(bad, (md ? kRangeDiv : 0u) | (mb ? kRangeBound : 0u));
  }
}

// Smallest and largest bound entry of the real samples (order-preserving
// keys of their float values), one atomic pair per wavefront.
__global__ __launch_bounds__(kBlock) void bound_range_kernel(const float *__restrict__ params,
                                                             uint32_t dim, uint32_t samples,
                                                             uint32_t sp,
                                                             uint32_t *__restrict__ minmax) {
  const uint64_t total = (uint64_t)dim * samples;
  uint32_t lo = 0xffffffffu, hi = 0u;
  for (uint64_t idx = (uint64_t)blockIdx.x * kBlock + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * kBlock) {
    const uint32_t j = (uint32_t)(idx / samples);
    const uint32_t k = (uint32_t)(idx - (uint64_t)j * samples);
    const float *col = params + (uint64_t)j * 3u * sp;
    const uint32_t key = wmh_order_bits(wmh_f16_to_f32(wmh_bound_entry(col[k], col[2u * sp + k])));
    lo = min(lo, key);
    hi = max(hi, key);
  }
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, d));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, d));
  }
  if (lane_id() == 0u) {
    atomicMin(minmax, lo);
    atomicMax(minmax + 1, hi);
  }
}

// The byte-coded bound table: entry (j, k) = wmh_bound_code of column j,
// sample k, laid out [dim][sp] (one sp-byte run per column); padded samples
// get 255.
__global__ __launch_bounds__(kBlock) void build_codes_kernel(const float *__restrict__ params,
                                                             uint32_t dim, uint32_t samples,
                                                             uint32_t sp, float lo, float delta,
                                                             uint8_t *__restrict__ codes) {
  const uint64_t total = (uint64_t)dim * sp;
  const float inv_delta = 1.0f / delta;  // a power of two: exact
  for (uint64_t idx = (uint64_t)blockIdx.x * kBlock + threadIdx.x; idx < total;
       idx += (uint64_t)gridDim.x * kBlock) {
    const uint32_t j = (uint32_t)(idx / sp);
    const uint32_t k = (uint32_t)(idx - (uint64_t)j * sp);
    const float *col = params + (uint64_t)j * 3u * sp;
    codes[idx] = (uint8_t)(k < samples ? wmh_bound_code(wmh_bound_entry(col[k], col[2u * sp + k]),
                                                        lo, delta, inv_delta)
                                       : 255u);
  }
}

// coarse[j][g] = the smallest of codes[j][8g .. 8g + 7]: per column, the
// smallest code of each 8-sample group, in rows of rb bytes (padding: 255).
// A group whose smallest code is skipped against the group's largest
// threshold T' has every code skipped against its own threshold: the skip
// test (wmh_code_skips) only turns on as c grows or T' falls.
__global__ __launch_bounds__(kBlock) void build_coarse_kernel(const uint8_t *__restrict__ codes,
                                                              uint32_t dim, uint32_t sp,
                                                              uint32_t rb,
                                                              uint8_t *__restrict__ coarse) {
  const uint64_t total = (uint64_t)dim * rb;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t j = i / rb;
    const uint32_t g = (uint32_t)(i - j * rb);
    uint32_t m = 255u;
    if (g < sp / 8u) {
      const uint2 w = *reinterpret_cast<const uint2 *>(codes + j * sp + 8u * g);
#pragma unroll
      for (int b = 0; b < 4; b++) {
        m = min(m, (w.x >> (8 * b)) & 0xffu);
        m = min(m, (w.y >> (8 * b)) & 0xffu);
      }
    }
    coarse[i] = (uint8_t)m;
  }
}

// The item table: (r, beta, ln_c, 0) of every (column, sample) of the
// column-block parameter table side by side, [dim][sp] entries of 16 bytes,
// so that an exact evaluation reads one memory line instead of three.
__global__ __launch_bounds__(kBlock) void build_items_kernel(const float *__restrict__ params,
                                                             uint32_t dim, uint32_t sp,
                                                             float4 *__restrict__ items) {
  const uint64_t total = (uint64_t)dim * sp;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t j = i / sp;
    const uint32_t k = (uint32_t)(i - j * sp);
    const float *blk = params + j * 3u * sp;
    items[i] = make_float4(blk[k], blk[sp + k], blk[2u * sp + k], 0.0f);
  }
}

// ---------------------------------------------------------------------------
// The hashing core.

template <int V>
struct Argmin {
  float best[V];   // smallest ln_a so far (FLT_MAX: nothing selected)
  float t[V];      // t of the selected nonzero
  uint32_t j[V];   // its column (WMH_SENTINEL: nothing selected)

  __device__ __forceinline__ void init() {
#pragma unroll
    for (int v = 0; v < V; v++) {
      best[v] = FLT_MAX;
      t[v] = 0.0f;
      j[v] = WMH_SENTINEL;
    }
  }
};

// Nonzeros x samples-per-lane whose gathers are issued together per group
// (16: 24 loads in flight per wavefront at 2 samples per lane; doubling it
// measured no change on MI355X).
#ifndef WMH_GROUP_NNZ
#define WMH_GROUP_NNZ 16
#endif

// lw / r: the short sequence (Fast: every r of the table is in
// [kFastDivMinB, kFastDivMaxB] and no lw of the chunk is +inf) or IEEE.
template <bool Fast>
__device__ __forceinline__ float div_lw(float lw, float r) {
  if constexpr (Fast) return div_rn_fast(lw, r);
  else return lw / r;
}

// Runs body(std::true_type) when `fast` (wave-uniform) holds, else
// body(std::false_type): the division mode is chosen once per chunk of
// nonzeros, outside the loops that keep gathers in flight (a branch per fold
// would join the two division sequences before every fold and make the
// compiler wait there for every outstanding load and store).
template <typename Body>
__device__ __forceinline__ void with_division(bool fast, Body &&body) {
  if (fast) body(std::true_type{});
  else body(std::false_type{});
}

// Folds one nonzero (column j, log-weight lw, parameters r/beta/lc for the
// lane's V samples) into the running argmins.  `ok` (wave-uniform) is false
// for columns >= dim, which are never selected.
template <int V, bool Fast>
__device__ __forceinline__ void fold(const float (&r)[V], const float (&beta)[V],
                                     const float (&lc)[V], uint32_t j, float lw,
                                     bool ok, Argmin<V> &acc) {
  WMH_NO_CONTRACT
#pragma unroll
  for (int v = 0; v < V; v++) {
    const float t = __builtin_floorf(div_lw<Fast>(lw, r[v]) + beta[v]);  // kernel.cu:103
    const float ln_y = (t - beta[v]) * r[v];                 // kernel.cu:104
    const float ln_a = (lc[v] - ln_y) - r[v];                // kernel.cu:105
    const bool upd = ok && ln_a < acc.best[v];               // kernel.cu:106
    acc.best[v] = upd ? ln_a : acc.best[v];
    acc.t[v] = upd ? t : acc.t[v];
    acc.j[v] = upd ? j : acc.j[v];
  }
}

// Byte offsets of the lane's r / beta / ln_c values inside a column block.
struct LaneOffsets {
  uint32_t r, beta, lc;
};

// A column block is read through a raw buffer resource held in SGPRs (the
// block's address, its size as the range, dword 3 = 32-bit data format as
// gfx9 requires), so each gather is one buffer load with the lane's byte
// offset: no per-lane 64-bit address arithmetic.  Offsets past the block's
// end read zero instead of memory.
constexpr int kRawBufferWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const char *blk, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(blk), (short)0, (int)bytes,
                                           kRawBufferWord3);
}

// Loads V consecutive floats at byte offset off of a block.  Streaming
// loads carry the non-temporal cache policy (kept out of the way of the L2
// lines that other wavefronts re-read).
constexpr int kLoadNonTemporal = 2;  // cache-policy operand: nt

template <int V, bool Streaming = false>
__device__ __forceinline__ void load_b(__amdgpu_buffer_rsrc_t rs, uint32_t off,
                                       float (&out)[V]) {
  constexpr int aux = Streaming ? kLoadNonTemporal : 0;
  if constexpr (V == 1) {
    out[0] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, aux));
  } else if constexpr (V == 2) {
    const float2 x = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, aux));
    out[0] = x.x;
    out[1] = x.y;
  } else {
    const float4 x = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, aux));
    out[0] = x.x;
    out[1] = x.y;
    out[2] = x.z;
    out[3] = x.w;
  }
}

// Hashes CSR positions [begin, end) of one row into `acc` (CSR order, so a
// tie keeps the earliest position).  Nonzeros are folded in groups of U whose
// 3*U gathers are all issued before the first one is consumed.
template <int V, int U>
__device__ __forceinline__ void hash_range(const float *__restrict__ params,
                                           uint32_t dim, uint64_t block_bytes,
                                           LaneOffsets lo,
                                           const float *__restrict__ weights,
                                           const uint32_t *__restrict__ cols,
                                           uint32_t begin, uint32_t end, bool fast_div,
                                           Argmin<V> &acc) {
  static_assert(kWave % U == 0, "groups tile a 64-nonzero chunk");
  const uint32_t lane = lane_id();
  const char *pbase = reinterpret_cast<const char *>(params);
  for (uint32_t base = begin; base < end; base += kWave) {
    const uint32_t n = uniform(min((uint32_t)kWave, end - base));
    uint32_t c = WMH_SENTINEL;
    float lw = 0.0f;
    if (lane < n) {
      c = cols[base + lane];
      lw = wmh_logf(weights[base + lane]);  // once per nonzero (kernel.cu:95)
    }
    const bool fast = fast_div && __ballot(lw == __builtin_inff()) == 0ull;
    // The last group of a chunk may run past n: lanes >= n hold the column
    // WMH_SENTINEL (>= dim), so those slots read column 0's block and are
    // never selected, and the group keeps all of its gathers in flight.
    with_division(fast, [&](auto mode) {
      for (uint32_t i = 0; i < n; i += U) {
        uint32_t j[U];
        float lwu[U];
        float r[U][V], beta[U][V], lc[U][V];
#pragma unroll
        for (int u = 0; u < U; u++) {
          j[u] = readlane_u32(c, i + u);
          lwu[u] = readlane_f32(lw, i + u);
          const __amdgpu_buffer_rsrc_t rs = block_rsrc(
              pbase + (uint64_t)(j[u] < dim ? j[u] : 0u) * block_bytes, block_bytes);
          load_b<V>(rs, lo.r, r[u]);
          load_b<V>(rs, lo.beta, beta[u]);
          load_b<V>(rs, lo.lc, lc[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          fold<V, decltype(mode)::value>(r[u], beta[u], lc[u], j[u], lwu[u], j[u] < dim, acc);
      }
    });
  }
}

// Writes the lane's V hashes of `row` (kernel.cu:83-85: out[(row*S + k)*2] =
// column, out[... + 1] = (uint32)(int32)t).
template <int V>
__device__ __forceinline__ void write_hashes(uint32_t *__restrict__ output,
                                             uint64_t row, uint32_t samples,
                                             uint32_t k0, const Argmin<V> &acc) {
  uint32_t h[2 * V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    const bool sel = acc.j[v] != WMH_SENTINEL;
    h[2 * v] = acc.j[v];
    h[2 * v + 1] = sel ? wmh_t_bits(acc.t[v]) : WMH_SENTINEL;
  }
  uint32_t *dst = output + (row * samples + k0) * 2u;
  if (k0 + V <= samples) {
    if constexpr (V == 1) {
      *reinterpret_cast<uint2 *>(dst) = make_uint2(h[0], h[1]);
    } else if constexpr (V == 2) {
      *reinterpret_cast<uint4 *>(dst) = make_uint4(h[0], h[1], h[2], h[3]);
    } else {
      reinterpret_cast<uint4 *>(dst)[0] = make_uint4(h[0], h[1], h[2], h[3]);
      reinterpret_cast<uint4 *>(dst)[1] = make_uint4(h[4], h[5], h[6], h[7]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) {
      if (k0 + v < samples) {
        dst[2 * v] = h[2 * v];
        dst[2 * v + 1] = h[2 * v + 1];
      }
    }
  }
}

// Plain kernel: wavefront w of the grid handles (row, tile) = divmod(w, ntiles)
// for the whole row.  Reference: kernel.cu:56-112.
template <int V>
__global__ __launch_bounds__(kBlock) void wmh_plain_kernel(
    const float *__restrict__ params, uint32_t dim, uint32_t samples,
    uint32_t sp, uint32_t ntiles, const float *__restrict__ weights,
    const uint32_t *__restrict__ cols, const uint32_t *__restrict__ rows,
    uint32_t length, uint32_t base0, uint32_t *__restrict__ output, uint32_t fast_div) {
  const uint64_t total = (uint64_t)length * ntiles;
  const uint32_t lane = lane_id();
  // CSR positions are clamped to [rows[0], rows[length]], so a malformed
  // indptr can give wrong hashes but never reads outside the arrays
  const uint32_t first = rows[0] - base0;
  const uint32_t last = max(first, rows[length] - base0);
  for (uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
       w < total; w += (uint64_t)gridDim.x * kWavesPerBlock) {
    const uint32_t row = uniform((uint32_t)(w / ntiles));
    const uint32_t tile = uniform((uint32_t)(w - (uint64_t)row * ntiles));
    const uint32_t begin = uniform(min(max(rows[row] - base0, first), last));
    const uint32_t end = uniform(min(max(rows[row + 1] - base0, first), last));
    const uint32_t k0 = tile * (kWave * V) + lane * V;
    const uint32_t koff = (k0 < samples) ? k0 : 0u;
    const LaneOffsets lo{koff * 4u, (sp + koff) * 4u, (2u * sp + koff) * 4u};
    Argmin<V> acc;
    acc.init();
    hash_range<V, WMH_GROUP_NNZ / V>(params, dim, 12ull * sp, lo, weights, cols, begin, end,
                                     fast_div != 0u, acc);
    if (k0 < samples) write_hashes<V>(output, row, samples, k0, acc);
  }
}

// ---------------------------------------------------------------------------
// Batched, nnz-balanced persistent hashing.
//
// The nonzeros of each matrix are cut into items of `seg` consecutive CSR
// positions (the last item of a matrix may be shorter); a work unit is an
// (item, sample tile) pair.  Waves of a fixed-size grid pull units from an
// atomic queue until it is drained.  Inside an item a wave streams the
// nonzeros in chunks of 64 regardless of row boundaries, so many short rows
// share one wavefront, and flushes a row's argmins whenever the row ends.
//
// A row that crosses an item boundary ("split") is merged across items with a
// 64-bit atomicMin on key = order(ln_a) << 32 | position, written into the
// row's own 8-byte output slots: the key orders by ln_a, then by CSR position,
// which is exactly the strict-< first-minimum of the sequential walk.  A
// fix-up pass then turns each winning key back into (column, t), recomputing
// t for the winning position.  Empty and split rows are pre-filled with all
// ones (the sentinel pair == the "no key" value) by an init pass.
//
// Launch sequence: plan (item prefix over matrices, queue reset) -> init ->
// persistent -> fixup.  No memsets, so the sequence can be graph-captured.

// Maps a binary32 value that can be selected (not NaN, < FLT_MAX) to an
// unsigned integer with the same order; -0 and +0 share a key.
__device__ __forceinline__ uint32_t order_key(float f) {
  const uint32_t u = wmh_f2u(f == 0.0f ? 0.0f : f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr unsigned long long kNoKey = ~0ull;

// The row containing CSR position a: the largest r in [0, length] with
// rows[r] <= a, given rows[0] <= a < rows[length].  64-ary search: one
// wavefront-wide load and one ballot per level.
__device__ __forceinline__ uint32_t find_row(const uint32_t *__restrict__ rows,
                                             uint32_t length, uint32_t a) {
  const uint32_t lane = lane_id();
  uint32_t lo = 0, hi = length;  // rows[lo] <= a; the answer is in [lo, hi]
  while (hi - lo >= (uint32_t)kWave) {
    const uint32_t step = (hi - lo + kWave) / kWave;  // ceil((hi - lo + 1) / 64)
    const uint32_t idx = lo + lane * step;
    const bool le = idx <= hi && rows[idx] <= a;
    const unsigned long long mask = __ballot(le);  // lane 0 is always set
    lo = uniform(lo + (uint32_t)(63 - __builtin_clzll(mask)) * step);
    hi = uniform(min(hi, lo + step - 1u));
  }
  const uint32_t idx = lo + lane;
  const bool le = idx <= hi && rows[idx] <= a;
  const unsigned long long mask = __ballot(le);
  return uniform(lo + (uint32_t)(63 - __builtin_clzll(mask)));
}

// Flushes the argmins of one row: final hashes for a row that lies inside
// the item, keyed atomic minima for a split row.
template <int V>
__device__ __forceinline__ void flush_row(uint32_t *__restrict__ output, uint32_t row,
                                          uint32_t samples, uint32_t k0, bool split,
                                          const Argmin<V> &acc) {
  if (k0 >= samples) return;
  if (split) {
    unsigned long long *slot =
        reinterpret_cast<unsigned long long *>(output) + ((uint64_t)row * samples + k0);
    // A slot only decreases, so a value read here is >= its current one: when
    // it is already <= key the atomic would change nothing and is skipped
    // (most pieces of a striped row do not improve on the earlier ones).
    unsigned long long seen[V];
#pragma unroll
    for (int v = 0; v < V; v++) seen[v] = k0 + v < samples ? slot[v] : 0ull;
#pragma unroll
    for (int v = 0; v < V; v++) {
      if (k0 + v < samples && acc.j[v] != WMH_SENTINEL) {
        const unsigned long long key =
            ((unsigned long long)order_key(acc.best[v]) << 32) | acc.j[v];
        if (key < seen[v]) atomicMin(slot + v, key);
      }
    }
  } else {
    write_hashes<V>(output, row, samples, k0, acc);
  }
}

// Skip bound inside a row walk (wmh_bound.h; the byte-coded bound table is
// laid out [dim][sp]).  A lane holds V samples k0 .. k0 + V - 1 of every
// nonzero, so its V codes of one nonzero are one V-byte load.  Per block of
// kCodeGroups groups of positions the lane tests each (nonzero, sample) against the
// threshold T' of its running minimum for that sample and marks the nonzeros
// for which some of its samples are not skipped; only marked (nonzero, lane)
// pairs gather parameters and fold.  An unmarked pair's computed ln_a is
// strictly above the lane's running minimum at the start of the block, so it
// can never be selected (strict <, ties keep the earlier position) -- nor win
// the keyed merge of a split row, whose minimum it is compared with belongs to
// a real candidate of the same row.  Positions past the end of the row that
// is current at the start of the block are marked for every lane.
// (one group per block keeps the hashing kernels within 128 VGPRs, 4 waves
// per SIMD; two groups need 141 at 4 samples per lane).  At 4 samples per
// lane the codes of the next block are loaded while a block's parameters are
// gathered, so a block waits for one round trip instead of two (+7% on a
// serving batch at 256 samples, although its 134 VGPRs leave 3 waves per
// SIMD; at 2 samples per lane the same would cost the stripe pieces, which
// share that kernel, their fourth wave).
#ifndef WMH_CODE_GROUPS
#define WMH_CODE_GROUPS 1
#endif
constexpr uint32_t kCodeGroups = WMH_CODE_GROUPS;

// Buffer offsets with the top bit set lie past every column block (block
// sizes are far below 2^31): an unmarked lane's gather returns zero and reads
// no memory.

// Packed 0x4800 - T' of the lane's V running minima, two samples per word
// (wmh_bound.h).  At one sample per lane the high half is never tested.
template <int V>
__device__ __forceinline__ void lane_thresholds(const float (&best)[V], float scale, float t0,
                                                uint32_t (&tp)[(V + 1) / 2]) {
#pragma unroll
  for (int q = 0; q < (V + 1) / 2; q++) {
    const uint32_t lo16 = 0x4800u - wmh_code_threshold(best[2 * q], scale, t0);
    const uint32_t hi16 =
        2 * q + 1 < V ? 0x4800u - wmh_code_threshold(best[(2 * q + 1) % V], scale, t0) : 0u;
    tp[q] = lo16 | (hi16 << 16);
  }
}

// The lane's V codes of one nonzero (bytes of a b32 / b16 / b8 load).
template <int V>
__device__ __forceinline__ uint32_t load_codes(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  if constexpr (V == 4) return __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0);
  else if constexpr (V == 2) return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rs, off, 0, 0);
  else return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(rs, off, 0, 0);
}

// 1 when some of the lane's V samples is not skipped for the nonzero whose
// codes are x and whose term (0x4000 - cB, both halves) is lw_pair.
template <int V>
__device__ __forceinline__ bool codes_mark(uint32_t x, const uint32_t (&tp)[(V + 1) / 2],
                                           uint32_t lw_pair) {
  uint32_t h;
  if constexpr (V == 4) {
    h = (__builtin_amdgcn_perm(0u, x, 0x0c010c00u) + tp[0] + lw_pair) &
        (__builtin_amdgcn_perm(0u, x, 0x0c030c02u) + tp[1] + lw_pair);
  } else if constexpr (V == 2) {
    h = __builtin_amdgcn_perm(0u, x, 0x0c010c00u) + tp[0] + lw_pair;
  } else {
    h = (x + tp[0] + lw_pair) | 0x80000000u;
  }
  return (h & 0x80008000u) != 0x80008000u;
}

// The code table and its scale (1 / delta) and t0 (wmh_code_threshold).
struct BoundCodes {
  const uint8_t *codes;
  float scale, t0;
};

// Hashes CSR positions [a, b) of matrix M (one item or one stripe piece) for
// the lane's samples.  In a split row the argmin tracks the CSR position
// instead of the column.  `row` is the row containing a, or ~0u to search
// for it; for a stripe piece (`piece_end` != 0) [a, b) lies inside `row`,
// whose positions are [piece_start, piece_end), and no row bound is read.
// Unless skip_nnz is 0, rows with at least `skip_nnz` nonzeros are skipped
// (their stripe pieces hash them); skip_nnz >= 64.  `streaming` gathers with
// non-temporal loads (items hashed while other wavefronts work through a
// cache-resident stripe).  Bounded: the skip bound above, on bc's codes.
template <int V, int U, bool Bounded>
__device__ __forceinline__ void hash_item(const float *__restrict__ params, uint32_t dim,
                                          uint32_t sp, uint32_t samples, uint32_t k0,
                                          LaneOffsets lo, const MatrixDesc &M,
                                          uint32_t a, uint32_t b, uint32_t row,
                                          uint32_t piece_start, uint32_t piece_end,
                                          uint32_t skip_nnz, bool fast_div, bool streaming,
                                          BoundCodes bc) {
  static_assert(kWave % U == 0, "groups tile a 64-nonzero chunk");
  // positions whose codes are loaded together; codes loaded a block ahead
  constexpr uint32_t CB = kCodeGroups * (uint32_t)U;
  constexpr bool kPrefetch = V == 4;
  // A block uses the codes once it starts kBoundAfter positions into its
  // row's part of [a, b): closer to the row's start its minima rarely skip
  // enough to repay the codes' round trip.  At 1 or 2 samples per lane a
  // 128-byte line of a parameter array serves 32 or 16 lanes, so the bound
  // saves HBM traffic only late in a row (measured on 100,000 rows of
  // Poisson(100) nonzeros: +6.5% at 64 against 0); at 4 samples per lane 16
  // (serving rows: 64 cost 1%, 128 cost 6%).
  constexpr uint32_t kBoundAfter = V == 4 ? 16u : 64u;
  static_assert(kWave % CB == 0, "blocks tile a chunk");
  const uint32_t lane = lane_id();
  const char *pbase = reinterpret_cast<const char *>(params);
  const uint64_t block_bytes = 12ull * sp;
  const uint32_t *__restrict__ rows = M.rows;
  const uint32_t *__restrict__ cols = M.cols - M.base;
  const float *__restrict__ weights = M.weights - M.base;
  const bool in_row = piece_end != 0u;
  // the lane's codes sit at byte koff of a column's sp-byte run
  const uint32_t koff = k0 < samples ? k0 : 0u;
  const bool lane_ok = k0 < samples;
  uint32_t r = row != ~0u ? row : find_row(rows, M.length, a);
  uint32_t row_start = in_row ? piece_start : uniform(rows[r]);
  // rows[rp_base + lane]: the next 64 row ends, read with one vector load
  // (never needed inside one row: no position of [a, b) ends it before b)
  uint32_t rp_base = r + 1;
  uint32_t rp = (!in_row && rp_base + lane <= M.length) ? rows[rp_base + lane] : 0xffffffffu;
  uint32_t rp_idx = 0;
  uint32_t row_end = in_row ? piece_end : readlane_u32(rp, 0);
  // a stripe piece always merges through the keyed minima, even when it
  // holds the whole row (the fix-up pass reads every striped row as keys)
  bool split = in_row || row_start < a || row_end > b;
  bool skip = skip_nnz != 0u && row_end - row_start >= skip_nnz;
  // the current row's first position in [a, b): its minima start there
  uint32_t row_first = row_start > a ? row_start : a;
  Argmin<V> acc;
  acc.init();

  // Called after position p_next - 1 ended row r and p_next < b: moves to the
  // next non-empty row (it contains p_next).
  auto next_row = [&]() {
    do {
      r++;
      row_start = row_end;
      if (++rp_idx == (uint32_t)kWave) {
        rp_base += kWave;
        rp = (rp_base + lane <= M.length) ? rows[rp_base + lane] : 0xffffffffu;
        rp_idx = 0;
      }
      row_end = readlane_u32(rp, rp_idx);
    } while (row_end == row_start);
    row_first = row_start;
    split = row_end > b;
    skip = skip_nnz != 0u && row_end - row_start >= skip_nnz;
    acc.init();
  };

  uint32_t cb = a;
  while (cb < b) {
    if (skip) {  // a striped row: jump past it
      if (row_end >= b) break;
      cb = row_end;
      next_row();
      continue;
    }
    const uint32_t n = uniform(min((uint32_t)kWave, b - cb));
    uint32_t c = WMH_SENTINEL;
    float lw = 0.0f;
    if (lane < n) {
      c = cols[cb + lane];
      lw = wmh_logf(weights[cb + lane]);
    }
    // the nonzero's term of the bound test, 0x4000 - cB (lane = nonzero)
    uint32_t lwt = 0u;
    if constexpr (Bounded) lwt = wmh_code_lw(wmh_bound_lw(lw), bc.scale);
    const bool fast = fast_div && __ballot(lw == __builtin_inff()) == 0ull;
    // As in hash_range, the last group may run past n (slots >= n hold the
    // column WMH_SENTINEL and are never selected); only real positions can
    // end a row.  A striped row starting inside the chunk runs past its end
    // (it has at least 64 nonzeros): its slots are not folded.
    auto groups = [&](auto mode, auto stream) {
      constexpr bool kFast = decltype(mode)::value;
      constexpr bool kStream = decltype(stream)::value;
      // the lane's codes of positions cb + i0 .. cb + i0 + CB - 1
      auto load_block = [&](uint32_t i0, uint32_t (&x)[CB]) {
#pragma unroll
        for (int u = 0; u < (int)CB; u++) {
          const uint32_t j = readlane_u32(c, i0 + u);
          const __amdgpu_buffer_rsrc_t rs =
              block_rsrc(reinterpret_cast<const char *>(bc.codes) +
                             (uint64_t)(j < dim ? j : 0u) * sp, sp);
          x[u] = load_codes<V>(rs, koff);
        }
      };
      auto codes_at = [&](uint32_t i0) { return cb + i0 - row_first >= kBoundAfter; };
      uint32_t xn[kPrefetch ? CB : 1];
      bool have_xn = false;
      if constexpr (Bounded && kPrefetch) {
        have_xn = codes_at(0);
        if (have_xn) load_block(0, xn);
      }
      for (uint32_t i0 = 0; i0 < n && !skip; i0 += CB) {
        // bit u: this lane folds position cb + i0 + u
        uint32_t marks = 0xffffffffu;
        bool use_codes = false;
        uint32_t x[CB];
        if constexpr (Bounded) {
          if constexpr (kPrefetch) {
            use_codes = have_xn && codes_at(i0);
#pragma unroll
            for (int u = 0; u < (int)CB; u++) x[u] = xn[u];
            have_xn = i0 + CB < n && codes_at(i0 + CB);
            if (have_xn) load_block(i0 + CB, xn);
          } else {
            use_codes = codes_at(i0);
            if (use_codes) load_block(i0, x);
          }
        }
        if (Bounded && use_codes) {
          uint32_t tp[(V + 1) / 2];
          lane_thresholds<V>(acc.best, bc.scale, bc.t0, tp);
          marks = 0u;
#pragma unroll
          for (int u = 0; u < (int)CB; u++) {
            const uint32_t src = i0 + u;
            const uint32_t lt = readlane_u32(lwt, src);
            const uint32_t j = readlane_u32(c, src);
            // past the current row: another row's minima do not apply
            const bool other_row = cb + src >= row_end;
            const bool ev = j < dim && lane_ok &&
                            (other_row || codes_mark<V>(x[u], tp, lt | (lt << 16)));
            marks |= ev ? 1u << u : 0u;
          }
        }
        for (uint32_t i = i0; i < i0 + CB && i < n && !skip; i += U) {
          const uint32_t gm = marks >> (i - i0);
          const uint32_t p0 = cb + i;
          // the current row runs past the group: no row checks needed
          const bool inside = row_end > p0 + min((uint32_t)U, n - i);
          if (Bounded && inside && __ballot((gm & ((1u << U) - 1u)) != 0u) == 0ull) continue;
          uint32_t j[U];
          float lwu[U];
          float rr[U][V], beta[U][V], lc[U][V];
#pragma unroll
          for (int u = 0; u < U; u++) {
            j[u] = readlane_u32(c, i + u);
            lwu[u] = readlane_f32(lw, i + u);
            const __amdgpu_buffer_rsrc_t rs = block_rsrc(
                pbase + (uint64_t)(j[u] < dim ? j[u] : 0u) * block_bytes, block_bytes);
            // an unmarked lane's offsets get the top bit (arithmetic, not a
            // select: the compiler turns a select of load offsets into a
            // branch around each load, with a wait for every load in flight
            // at the join)
            const uint32_t nog = Bounded ? (~(gm >> u) & 1u) << 31 : 0u;
            load_b<V, kStream>(rs, lo.r | nog, rr[u]);
            load_b<V, kStream>(rs, lo.beta | nog, beta[u]);
            load_b<V, kStream>(rs, lo.lc | nog, lc[u]);
          }
          if (inside) {
#pragma unroll
            for (int u = 0; u < U; u++)
              fold<V, kFast>(rr[u], beta[u], lc[u], split ? p0 + u : j[u], lwu[u],
                             j[u] < dim && (!Bounded || ((gm >> u) & 1u) != 0u), acc);
          } else {
#pragma unroll
            for (int u = 0; u < U; u++) {
              const uint32_t p = p0 + u;
              fold<V, kFast>(rr[u], beta[u], lc[u], split ? p : j[u], lwu[u],
                             j[u] < dim && !skip && (!Bounded || ((gm >> u) & 1u) != 0u), acc);
              if (i + u < n && p + 1 == row_end) {
                flush_row<V>(M.output, r, samples, k0, split, acc);
                if (p + 1 < b) next_row();
              }
            }
          }
        }
      }
    };
    with_division(fast, [&](auto mode) {
      // (only the 2-samples-per-lane kernel, which hashes stripe pieces in
      // pairs, streams: a second copy of this loop slowed the others)
      if constexpr (V == 2) {
        if (streaming) groups(mode, std::true_type{});
        else groups(mode, std::false_type{});
      } else {
        groups(mode, std::false_type{});
      }
    });
    cb += n;
  }
  if (row_end > b && !skip) flush_row<V>(M.output, r, samples, k0, true, acc);  // continues later
}

// Stripe pieces of a 128-sample tile, two nonzeros per step: lanes 0-31
// take the even offsets of [a, b) and lanes 32-63 the odd ones, each lane 4
// samples (k0 = 128 tile + 4 (lane % 32)), so every gather is one full-wave
// 16-byte-per-lane load (3 per step of two nonzeros instead of 3 per
// nonzero).  Within a half the positions are folded in CSR order (strict <);
// at the end the halves merge by (ln_a, position), which keeps the earliest
// of equal minima, and lanes 0-31 merge the result into the row's keyed
// minima like any piece of a split row.  A chunk's column offsets and
// log-weights reach the lanes of each half through ds_bpermute; a column >=
// dim gets the log-weight NaN, which is never selected.  Gathers go through
// a buffer resource at the stripe's first column block with 32-bit offsets;
// a chunk holding a column too far from the stripe for that (an unsorted row
// in a table of 4 GiB or more) uses 64-bit addresses instead.
constexpr int kPairSteps = 3;  // pairs of nonzeros whose gathers are issued together
constexpr uint32_t kItemsFirstEvery = 8;

__device__ __forceinline__ uint32_t bperm_u32(uint32_t v, uint32_t src_lane) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src_lane << 2), (int)v);
}

// 16 bytes at voffset + soffset (soffset wave-uniform, in an SGPR)
__device__ __forceinline__ float4 load4(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
}

template <bool Fast>
__device__ __forceinline__ void fold4(const float4 &r4, const float4 &b4, const float4 &c4,
                                      float lw, uint32_t p, float (&best)[4],
                                      uint32_t (&pos)[4]) {
  WMH_NO_CONTRACT
  const float r[4] = {r4.x, r4.y, r4.z, r4.w};
  const float beta[4] = {b4.x, b4.y, b4.z, b4.w};
  const float lc[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const float t = __builtin_floorf(div_lw<Fast>(lw, r[v]) + beta[v]);  // kernel.cu:103
    const float ln_y = (t - beta[v]) * r[v];              // kernel.cu:104
    const float ln_a = (lc[v] - ln_y) - r[v];             // kernel.cu:105
    const bool upd = ln_a < best[v];                      // kernel.cu:106
    best[v] = upd ? ln_a : best[v];
    pos[v] = upd ? p : pos[v];
  }
}

__device__ __forceinline__ void hash_piece_pairs(const float *__restrict__ params, uint32_t dim,
                                                 uint32_t sp, uint32_t samples, uint32_t tile,
                                                 const MatrixDesc &M, uint32_t a, uint32_t b,
                                                 uint32_t row, uint32_t col0_in, bool fast_div) {
  const uint32_t lane = lane_id();
  const uint32_t half = lane >> 5;
  const uint32_t col0 = uniform(col0_in);
  const uint32_t k0 = tile * 128u + 4u * (lane & 31u);
  const uint32_t koff = k0 < samples ? k0 : 0u;
  const uint32_t block_bytes = 12u * sp;
  const uint32_t off_r = koff * 4u;              // lane part of the offset
  const uint32_t s_b = uniform(sp * 4u), s_c = uniform(sp * 8u);  // beta, ln_c arrays
  // column offsets c - col0 up to `span` keep every lane's 16 bytes below 2^32
  const uint32_t span = (0xffffffffu - 12u * sp) / block_bytes;
  // the resource must live in SGPRs: its base is made wave-uniform explicitly
  const uint64_t sb = reinterpret_cast<uint64_t>(params) + (uint64_t)col0 * block_bytes;
  char *sbase = reinterpret_cast<char *>(((uint64_t)uniform((uint32_t)(sb >> 32)) << 32) |
                                         uniform((uint32_t)sb));
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(sbase, (short)0, (int)0xffffffff, kRawBufferWord3);
  const uint32_t *__restrict__ cols = M.cols - M.base;
  const float *__restrict__ weights = M.weights - M.base;
  float best[4];
  uint32_t pos[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    best[v] = FLT_MAX;
    pos[v] = WMH_SENTINEL;
  }
  for (uint32_t cb = a; cb < b; cb += kWave) {
    const uint32_t n = uniform(min((uint32_t)kWave, b - cb));
    // lanes without a selectable nonzero read a valid block (the stripe's
    // first, or column 0 for a stripe past dim) and are never selected
    uint32_t c = col0 < dim ? col0 : 0u;
    float lw = __builtin_nanf("");
    if (lane < n) {
      const uint32_t cc = cols[cb + lane];
      if (cc < dim) {
        c = cc;
        lw = wmh_logf(weights[cb + lane]);  // once per nonzero (kernel.cu:95)
      }
    }
    const bool fast = fast_div && __ballot(lw == __builtin_inff()) == 0ull;
    const uint32_t rel = c - col0;
    const bool fits = __ballot(rel > span) == 0ull;
    const uint32_t o = rel * block_bytes;
    // One loop per (division mode, addressing), so that no branch sits
    // between a group's gathers and its folds.  Full groups of kPairSteps
    // pairs first, then the chunk's remaining positions one pair at a time
    // (a piece averages a few dozen nonzeros: padding its last group to a
    // full one would spend folds and gathers on empty slots).  A pair never
    // reaches past lane 63 (n is even when it is 64); lanes past n hold NaN
    // log-weights.
    auto pairs = [&](auto mode, auto near) {
      auto group = [&](auto steps, uint32_t i) {
        constexpr int P = decltype(steps)::value;
        float4 r4[P], b4[P], c4[P];
        float lws[P];
#pragma unroll
        for (int u = 0; u < P; u++) {
          const uint32_t src = i + 2u * u + half;
          lws[u] = __uint_as_float(bperm_u32(__float_as_uint(lw), src));
          if constexpr (decltype(near)::value) {
            const uint32_t ou = bperm_u32(o, src) + off_r;
            r4[u] = load4(rs, ou, 0u);
            b4[u] = load4(rs, ou, s_b);
            c4[u] = load4(rs, ou, s_c);
          } else {
            const char *blk = reinterpret_cast<const char *>(params) +
                              (uint64_t)bperm_u32(c, src) * block_bytes + off_r;
            r4[u] = *reinterpret_cast<const float4 *>(blk);
            b4[u] = *reinterpret_cast<const float4 *>(blk + s_b);
            c4[u] = *reinterpret_cast<const float4 *>(blk + s_c);
          }
        }
#pragma unroll
        for (int u = 0; u < P; u++)
          fold4<decltype(mode)::value>(r4[u], b4[u], c4[u], lws[u], cb + i + 2u * u + half,
                                       best, pos);
      };
      uint32_t i = 0;
      for (; i + 2 * kPairSteps <= n; i += 2 * kPairSteps)
        group(std::integral_constant<int, kPairSteps>{}, i);
      for (; i < n; i += 2) group(std::integral_constant<int, 1>{}, i);
    };
    with_division(fast, [&](auto mode) {
      if (fits) pairs(mode, std::true_type{});
      else pairs(mode, std::false_type{});
    });
  }
  // merge the odd positions (lanes 32-63) into the even ones (lanes 0-31)
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const float ob = __uint_as_float(bperm_u32(__float_as_uint(best[v]), lane ^ 32u));
    const uint32_t op = bperm_u32(pos[v], lane ^ 32u);
    const bool take = ob < best[v] || (ob == best[v] && op < pos[v]);
    best[v] = take ? ob : best[v];
    pos[v] = take ? op : pos[v];
  }
  if (half != 0u || k0 >= samples) return;
  unsigned long long *slot =
      reinterpret_cast<unsigned long long *>(M.output) + ((uint64_t)row * samples + k0);
  unsigned long long seen[4];
#pragma unroll
  for (int v = 0; v < 4; v++) seen[v] = k0 + v < samples ? slot[v] : 0ull;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    if (k0 + v < samples && pos[v] != WMH_SENTINEL) {
      const unsigned long long key = ((unsigned long long)order_key(best[v]) << 32) | pos[v];
      if (key < seen[v]) atomicMin(slot + v, key);
    }
  }
}

// Stripe pieces of a 128-sample tile with the skip bound (wmh_bound.h), for
// pieces whose row already holds a keyed minimum for every sample of the
// tile (heavy seeds or the earlier stripes' pieces wrote them).  Quarter-wave
// layout: lane l hashes samples k0 = 128 tile + 8 (l % 16) .. k0 + 7 at the
// positions cb + 4m + l / 16 (m = 0..15) of each 64-position chunk.
//   Phase 1: for every (nonzero, sample) the code c is tested against the
//   sample's threshold T' -- from the tile's minimum: the row's keyed minimum
//   at the start, then the piece's own -- and the nonzero's term
//   (wmh_code_skips, two samples per 32-bit add).  First the coarse test, one
//   byte per (nonzero, lane): the smallest of the lane's 8 codes against the
//   largest of its 8 thresholds; a group it skips has every sample skipped.
//   Then the 8 codes of the groups it kept (a few percent once the row holds
//   good minima), sample by sample.  A skipped (nonzero, sample) has a
//   computed ln_a strictly above a minimum that a real candidate of the row
//   reached, so it cannot change the merged result.
//   Phase 2: every (nonzero, sample) pair the fine test keeps becomes an
//   item of an LDS list; the wavefront evaluates the list one item per lane
//   (the formula of kernel.cu:103-106 for one sample) and merges each result
//   into the tile's minima in LDS by an atomic minimum of the key
//   (order(ln_a) << 32 | position), which keeps the earliest of equal
//   values.  Every kThresholdEvery-th chunk the lanes' thresholds then follow
//   the tile's minima, which the four quarters share.
// At the end lanes merge the tile's minima into the row's keyed minima like
// every piece of a split row.  Requires samples % 8 == 0 (every lane's 8
// samples lie in its column's runs).

// The value of a keyed minimum (order_key(ln_a) << 32 | position).
__device__ __forceinline__ float key_value(unsigned long long key) {
  const uint32_t h = (uint32_t)(key >> 32);
  return __uint_as_float((h & 0x80000000u) ? (h & 0x7fffffffu) : ~h);
}


// Phase-1 steps whose coarse code loads are issued together (the loop over
// the two groups of a chunk stays rolled: unrolled, it needs 151 VGPRs).
constexpr int kBoundGroup = 8;

// Two byte codes of a 32-bit word (bytes 2i, 2i + 1) as 16-bit halves.
template <int I>
__device__ __forceinline__ uint32_t code_pair(uint32_t w) {
  return __builtin_amdgcn_perm(0u, w, I == 0 ? 0x0c010c00u : 0x0c030c02u);
}

// Packed 0x4800 - T' of the lane's 8 samples, two per word (wmh_bound.h),
// and 0x4800 - max T' in both halves of tmax2 (the coarse test).
__device__ __forceinline__ void code_thresholds(const float (&thr)[8], float scale, float t0,
                                                uint32_t (&tpair)[4], uint32_t &tmax2) {
  uint32_t tmax = 0u;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t lo = wmh_code_threshold(thr[2 * i], scale, t0);
    const uint32_t hi = wmh_code_threshold(thr[2 * i + 1], scale, t0);
    tpair[i] = (0x4800u - lo) | ((0x4800u - hi) << 16);
    tmax = max(tmax, max(lo, hi));
  }
  tmax2 = (0x4800u - tmax) * 0x10001u;
}

// Folds samples H*4 .. H*4 + 3 of the lane's 8 (one 16-byte load of each
// parameter array) for the nonzero at position p into the running minima;
// true when one of them changed.
template <bool Fast, int H>
__device__ __forceinline__ bool fold_half(const float4 &r4, const float4 &b4, const float4 &c4,
                                          float lw, uint32_t p, float (&best)[8],
                                          uint32_t (&pos)[8]) {
  WMH_NO_CONTRACT
  const float r[4] = {r4.x, r4.y, r4.z, r4.w};
  const float beta[4] = {b4.x, b4.y, b4.z, b4.w};
  const float lc[4] = {c4.x, c4.y, c4.z, c4.w};
  bool any = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int v = 4 * H + i;
    const float t = __builtin_floorf(div_lw<Fast>(lw, r[i]) + beta[i]);  // kernel.cu:103
    const float ln_y = (t - beta[i]) * r[i];              // kernel.cu:104
    const float ln_a = (lc[i] - ln_y) - r[i];             // kernel.cu:105
    const bool upd = ln_a < best[v];                      // kernel.cu:106
    best[v] = upd ? ln_a : best[v];
    pos[v] = upd ? p : pos[v];
    any |= upd;
  }
  return any;
}

// Chunks between refreshes of the code thresholds from the running minima.
// Thresholds of older minima are higher: the test then skips a little less,
// never a pair that could win (a skipped pair stays above an existing
// candidate).  Minima change in few chunks once the row holds good ones, and
// a refresh costs about as much as a quarter of the chunk's tests.
constexpr uint32_t kThresholdEvery = 8;

// Working memory of one wavefront in hash_piece_bounded (LDS): the tile's
// 128 keyed minima, the packed code thresholds of its samples (two per
// word), the chunk's (nonzero, group) pairs that passed the coarse test (at
// most 16 per lane) and the (nonzero, sample) items of one fine-test pass
// (at most 8 per lane).
struct PieceLds {
  unsigned long long key[128];
  uint32_t tq[64];
  uint32_t pair[kWave * 16];
  uint32_t item[kWave * 8];
  uint32_t npair, nitem;
};

// Base0: the gathers address the table from column 0 with 32-bit offsets
// (the whole table below 4 GiB: no column is out of reach, whatever the row's
// order), else from the stripe's first column (a piece holding a column out
// of reach then returns false before writing anything).  `items`: the item
// table (build_items_kernel) when Base0 and it lies below 4 GiB, else null
// (the items read the column-block table).
template <bool Base0>
__device__ __forceinline__ bool hash_piece_bounded(const float *__restrict__ params,
                                                   const uint8_t *__restrict__ codes,
                                                   const uint8_t *__restrict__ coarse,
                                                   const float4 *__restrict__ items,
                                                   float code_scale, float code_t0,
                                                   uint32_t dim, uint32_t sp, uint32_t samples,
                                                   uint32_t tile, const MatrixDesc &M, uint32_t a,
                                                   uint32_t b, uint32_t row, uint32_t col0_in,
                                                   bool fast_div,
                                                   const unsigned long long (&skey)[8]) {
  __shared__ PieceLds piece_lds[kWavesPerBlock];
  PieceLds &L = piece_lds[threadIdx.x >> 6];
  const uint32_t lane = lane_id();
  const uint32_t col0 = Base0 ? 0u : uniform(col0_in);
  const uint32_t t0k = tile * 128u;  // the tile's first sample
  const uint32_t block_bytes = 12u * sp;  // the code table holds sp bytes per column
  const uint32_t s_b = uniform(sp * 4u), s_c = uniform(sp * 8u);  // beta, ln_c arrays
  // column offsets c - col0 up to `span` keep every lane's bytes below 2^32
  const uint32_t span = (0xffffffffu - 12u * sp) / block_bytes;
  // the resources must live in SGPRs: their bases are made wave-uniform
  const uint64_t pb = reinterpret_cast<uint64_t>(params) + (uint64_t)col0 * block_bytes;
  const uint64_t bb = reinterpret_cast<uint64_t>(codes) + (uint64_t)col0 * sp;
  char *pbase = reinterpret_cast<char *>(((uint64_t)uniform((uint32_t)(pb >> 32)) << 32) |
                                         uniform((uint32_t)pb));
  char *bbase = reinterpret_cast<char *>(((uint64_t)uniform((uint32_t)(bb >> 32)) << 32) |
                                         uniform((uint32_t)bb));
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(pbase, (short)0, (int)0xffffffff, kRawBufferWord3);
  const __amdgpu_buffer_rsrc_t rb =
      __builtin_amdgcn_make_buffer_rsrc(bbase, (short)0, (int)0xffffffff, kRawBufferWord3);
  // coarse codes: coarse_row_bytes(sp) per column, the tile's 16 groups as
  // one aligned 16-byte load
  const uint32_t spc = uniform(coarse_row_bytes(sp));
  const uint64_t cbb = reinterpret_cast<uint64_t>(coarse) + (uint64_t)col0 * spc;
  char *cbase = reinterpret_cast<char *>(((uint64_t)uniform((uint32_t)(cbb >> 32)) << 32) |
                                         uniform((uint32_t)cbb));
  const __amdgpu_buffer_rsrc_t rc =
      __builtin_amdgcn_make_buffer_rsrc(cbase, (short)0, (int)0xffffffff, kRawBufferWord3);
  const uint64_t ib = reinterpret_cast<uint64_t>(items);
  char *ibase = reinterpret_cast<char *>(((uint64_t)uniform((uint32_t)(ib >> 32)) << 32) |
                                         uniform((uint32_t)ib));
  const __amdgpu_buffer_rsrc_t ri =
      __builtin_amdgcn_make_buffer_rsrc(ibase, (short)0, (int)0xffffffff, kRawBufferWord3);
  const uint32_t *__restrict__ cols = M.cols - M.base;
  const float *__restrict__ weights = M.weights - M.base;
  // groups of the tile that hold samples (samples % 8 == 0: whole groups)
  const uint32_t ngroups = uniform(min(16u, (samples - t0k) / 8u));
  const uint32_t group_mask = ngroups >= 16u ? 0xffffu : (1u << ngroups) - 1u;
  // The tile's minima start as the row's keyed minima (the lanes of the
  // first quarter hold them, 8 samples each; samples past the end: none) and
  // take every item's key by an LDS atomic minimum.
  if (lane < 16u) {
#pragma unroll
    for (int v = 0; v < 8; v++) L.key[8u * lane + v] = lane < ngroups ? skey[v] : kNoKey;
  }
  if (lane == 0u) {
    L.npair = 0u;
    L.nitem = 0u;
  }
  // Code thresholds from the tile's minima: tq[l] packs 0x4800 - T' of
  // samples 2l and 2l + 1 (the fine test), tc[i] that of the largest T' of
  // groups 2i and 2i + 1 (the coarse test, wave-uniform).
  uint32_t tc[8];
  // (a sample without a minimum yet: FLT_MAX, whose T' skips nothing)
  auto thr_of = [&](unsigned long long key) {
    return wmh_code_threshold(key == kNoKey ? FLT_MAX : key_value(key), code_scale, code_t0);
  };
  auto refresh = [&]() {
    const uint32_t ta = thr_of(L.key[2u * lane]);
    const uint32_t tb = thr_of(L.key[2u * lane + 1u]);
    L.tq[lane] = (0x4800u - ta) | ((0x4800u - tb) << 16);
    uint32_t m = max(ta, tb);
    m = max(m, (uint32_t)__shfl_xor((int)m, 1));
    m = max(m, (uint32_t)__shfl_xor((int)m, 2));  // lanes 4g .. 4g + 3: group g's largest
    const uint32_t cg = 0x4800u - m;
#pragma unroll
    for (int i = 0; i < 8; i++)
      tc[i] = readlane_u32(cg, 8u * i) | (readlane_u32(cg, 8u * i + 4u) << 16);
  };
  refresh();
  bool stale = false;  // the tile's minima moved below the thresholds
  uint32_t chunk_no = 0;
  // the next chunk's columns and weights are loaded while this one is tested
  uint32_t next_c = dim;
  float next_w = 0.0f;
  if (lane < b - a) {
    next_c = cols[a + lane];
    next_w = weights[a + lane];
  }
  for (uint32_t cb = a; cb < b; cb += kWave) {
    const uint32_t n = uniform(min((uint32_t)kWave, b - cb));
    const uint32_t cc = next_c;
    const float wv = next_w;
    next_c = dim;
    if (cb + kWave < b && lane < b - cb - kWave) {
      next_c = cols[cb + kWave + lane];
      next_w = weights[cb + kWave + lane];
    }
    // lane l holds the nonzero at position cb + l; lanes without a
    // selectable one point at a valid block (the stripe's first, or column 0
    // for a stripe past dim) and list no pairs
    const bool own = lane < n && cc < dim;
    uint32_t c = col0 < dim ? col0 : 0u;
    float lw = __builtin_nanf("");
    float lwb = -__builtin_inff();
    if (own) {
      c = cc;
      lw = wmh_logf(wv);  // once per nonzero (kernel.cu:95)
      lwb = wmh_bound_lw(lw);
    }
    // every gather goes through the resources with 32-bit offsets; a piece
    // holding a column too far from the stripe for that (an unsorted row in
    // a table of 4 GiB or more, or a stripe starting at dim, whose lanes
    // without a column point at column 0) is left to hash_piece_pairs,
    // before anything of it is written
    if (!Base0 && __ballot(c - col0 > span) != 0ull) return false;
    const bool fast = fast_div && __ballot(lw == __builtin_inff()) == 0ull;
    const uint32_t rel = c - col0;
    const uint32_t o = rel * block_bytes, ob = rel * sp;
    const uint32_t lw_term = wmh_code_lw(lwb, code_scale);
    const uint32_t lw_pair = lw_term | (lw_term << 16);
    // Phase 1a, the coarse test of the lane's own nonzero for the tile's 16
    // groups: the smallest code of each group against the group's largest
    // threshold; a group it skips has all of its samples skipped.  Two groups
    // per 32-bit add (16-bit halves); bit g of `kept` for group g.
    const uint4 cw = __builtin_bit_cast(
        uint4, __builtin_amdgcn_raw_buffer_load_b128(rc, rel * spc + 16u * tile, 0, 0));
    const uint32_t h[8] = {code_pair<0>(cw.x) + tc[0] + lw_pair, code_pair<1>(cw.x) + tc[1] + lw_pair,
                           code_pair<0>(cw.y) + tc[2] + lw_pair, code_pair<1>(cw.y) + tc[3] + lw_pair,
                           code_pair<0>(cw.z) + tc[4] + lw_pair, code_pair<1>(cw.z) + tc[5] + lw_pair,
                           code_pair<0>(cw.w) + tc[6] + lw_pair, code_pair<1>(cw.w) + tc[7] + lw_pair};
    // the skip bits (bit 7 of bytes 1 and 3 of each sum) gathered four at a
    // time: the bytes side by side, then a multiply that adds the four bits
    // of a word into its top byte
    uint32_t skip = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const uint32_t w = (__builtin_amdgcn_perm(h[2 * i + 1], h[2 * i], 0x07050301u) >> 7) & 0x01010101u;
      skip |= ((w * 0x01020408u) >> 24) << (4 * i);
    }
    uint32_t kept = own ? ~skip & group_mask : 0u;
    // Phase 1b, the fine test of the (nonzero, group) pairs the coarse test
    // kept (a few percent), spread over the wavefront through an LDS list:
    // each lane tests one pair against its 8 codes and the 8 samples'
    // thresholds, and every (nonzero, sample) it does not skip becomes an
    // item.  Phase 2 evaluates the items, one per lane (the formula of
    // kernel.cu:103-106 for one sample), and merges each into the tile's
    // minimum of its sample with an LDS atomic minimum of the key
    // (order(ln_a) << 32 | position), which keeps the earliest of equal
    // values.
    if (__ballot(kept != 0u) != 0ull) {
      const uint32_t np = (uint32_t)__builtin_popcount(kept);
      uint32_t at = 0u;
      if (np != 0u) at = atomicAdd(&L.npair, np);
      while (kept != 0u) {
        const uint32_t g = (uint32_t)__builtin_ctz(kept);
        kept &= kept - 1u;
        L.pair[at++] = lane | (g << 6);
      }
      const uint32_t pairs = uniform(L.npair);
      if (lane == 0u) L.npair = 0u;
      stale = true;
      for (uint32_t p0 = 0; p0 < pairs; p0 += kWave) {
        const bool valid = p0 + lane < pairs;
        const uint32_t pr = valid ? L.pair[p0 + lane] : 0u;
        const uint32_t src = pr & 63u, g = pr >> 6;
        const uint32_t ob_m = bperm_u32(ob, src);
        const uint32_t lp = bperm_u32(lw_pair, src);
        const uint4 tp = *reinterpret_cast<const uint4 *>(&L.tq[4u * g]);
        const uint2 x = __builtin_bit_cast(
            uint2, __builtin_amdgcn_raw_buffer_load_b64(
                       rb, (ob_m + t0k + 8u * g) & (valid ? ~0u : 0u), 0, 0));
        // bit 15 of each half: that sample skipped (wmh_code_skips)
        const uint32_t f0 = code_pair<0>(x.x) + tp.x + lp;
        const uint32_t f1 = code_pair<1>(x.x) + tp.y + lp;
        const uint32_t f2 = code_pair<0>(x.y) + tp.z + lp;
        const uint32_t f3 = code_pair<1>(x.y) + tp.w + lp;
        const uint32_t w0 = (__builtin_amdgcn_perm(f1, f0, 0x07050301u) >> 7) & 0x01010101u;
        const uint32_t w1 = (__builtin_amdgcn_perm(f3, f2, 0x07050301u) >> 7) & 0x01010101u;
        const uint32_t fskip = ((w0 * 0x01020408u) >> 24) | (((w1 * 0x01020408u) >> 24) << 4);
        uint32_t kv = valid ? ~fskip & 0xffu : 0u;  // bit v: sample 8g + v kept
        const uint32_t ni = (uint32_t)__builtin_popcount(kv);
        uint32_t ia = 0u;
        if (ni != 0u) ia = atomicAdd(&L.nitem, ni);
        while (kv != 0u) {
          const uint32_t v = (uint32_t)__builtin_ctz(kv);
          kv &= kv - 1u;
          L.item[ia++] = src | ((8u * g + v) << 6);
        }
        const uint32_t total = uniform(L.nitem);
        if (lane == 0u) L.nitem = 0u;
        for (uint32_t i0 = 0; i0 < total; i0 += kWave) {
          const bool ivalid = i0 + lane < total;
          const uint32_t it = ivalid ? L.item[i0 + lane] : 0u;
          const uint32_t isrc = it & 63u;
          const uint32_t k = it >> 6;  // the sample's place in the tile
          const float lws = __uint_as_float(bperm_u32(__float_as_uint(lw), isrc));
          float r, beta, lc;
          if (Base0 && items != nullptr) {
            // the item's (r, beta, ln_c) in one 16-byte load (below 2^32:
            // the item table lies below 4 GiB)
            const uint32_t off =
                ((bperm_u32(ob, isrc) + t0k + k) << 4) & (ivalid ? ~0u : 0u);
            const float4 q = __builtin_bit_cast(
                float4, __builtin_amdgcn_raw_buffer_load_b128(ri, off, 0u, 0));
            r = q.x;
            beta = q.y;
            lc = q.z;
          } else {
            const uint32_t off = (bperm_u32(o, isrc) + 4u * (t0k + k)) & (ivalid ? ~0u : 0u);
            r = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0u, 0));
            beta =
                __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, s_b, 0));
            lc = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, s_c, 0));
          }
          float ln_a;
          {
            WMH_NO_CONTRACT
            const float t = __builtin_floorf(
                (fast ? div_lw<true>(lws, r) : div_lw<false>(lws, r)) + beta);  // kernel.cu:103
            const float ln_y = (t - beta) * r;                                   // kernel.cu:104
            ln_a = (lc - ln_y) - r;                                              // kernel.cu:105
          }
          // kernel.cu:106: below FLT_MAX, and the smallest (ln_a, position)
          if (ivalid && ln_a < FLT_MAX)
            atomicMin(&L.key[k], ((unsigned long long)order_key(ln_a) << 32) | (cb + isrc));
        }
      }
    }
    if (++chunk_no % kThresholdEvery == 0u && stale) {
      refresh();
      stale = false;
    }
  }
  // the tile's minima into the row's keyed minima (lane l: samples l and
  // l + 64 of the tile; a slot only decreases)
  unsigned long long *slot =
      reinterpret_cast<unsigned long long *>(M.output) + ((uint64_t)row * samples + t0k);
#pragma unroll
  for (uint32_t hh = 0; hh < 2u; hh++) {
    const uint32_t k = lane + 64u * hh;
    if (t0k + k < samples) {
      const unsigned long long key = L.key[k];
      if (key != kNoKey && key < slot[k]) atomicMin(slot + k, key);
    }
  }
  return true;
}

// Matrix whose range of `field` values contains x: the largest m with
// mats[m].field <= x (mats sorted by field).
template <uint32_t MatrixDesc::*Field>
__device__ __forceinline__ uint32_t find_matrix(const MatrixDesc *__restrict__ mats,
                                                uint32_t count, uint32_t x) {
  uint32_t lo = 0, hi = count - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (mats[mid].*Field <= x) lo = mid;
    else hi = mid - 1;
  }
  return uniform(lo);
}

// Work of a persistent launch: every queue, with stripe pieces through the
// skip bound or two nonzeros per step (kAllWork); every queue with every
// stripe piece through the skip bound from column 0 (kBoundedPieces: without
// the paired pieces' code the 2-samples kernel needs 98 VGPRs instead of 128,
// 5 waves per SIMD, which the LDS of the bounded pieces allows); or the items
// of the unstriped rows alone (kItemsOnly).
enum PersistentMode { kAllWork = 0, kBoundedPieces = 1, kItemsOnly = 2 };
constexpr int kPieceWavesPerCu = 20;

// 4 waves per SIMD (128 VGPRs) at 1 and 2 samples per lane: left alone the
// compiler takes 167 VGPRs for the 2-samples kernel and drops to 3 waves; the
// 4-samples kernel (131 VGPRs) keeps 3.
template <int V, int Mode>
__global__ __launch_bounds__(kBlock) __attribute__((
    amdgpu_waves_per_eu(Mode == kBoundedPieces ? 5 : (V == 4 ? 3 : 4))))
void wmh_persistent_kernel(
    const float *__restrict__ params, uint32_t dim, uint32_t samples, uint32_t sp,
    const MatrixDesc *__restrict__ mats, uint32_t count, BatchCtl *__restrict__ ctl,
    StripePlan stripes, uint32_t fast_div, const uint8_t *__restrict__ codes, float code_scale,
    float code_t0, const uint8_t *__restrict__ coarse, const float4 *__restrict__ items) {
  const uint32_t lane = lane_id();
  if (uniform(ctl->invalid)) return;  // malformed indptr: hash nothing
  const uint32_t seg = uniform(ctl->seg);
  const uint32_t ntiles = uniform(ctl->ntiles);
  // Stripe pieces go through Q = stripes.queues queues (1 or one per XCD):
  // queue q holds the stripes k with k % Q == q, in increasing k, a unit being
  // `per_unit` consecutive pieces of one stripe.  Workgroups are dispatched to
  // the XCDs round-robin, so with one queue per XCD a wave of workgroup w
  // serves queue w % Q first: each XCD then gathers from its own stripe, which
  // its L2 can hold, rather than every XCD caching the same one (stripes larger than
  // an L2 are better shared, Q = 1, so that together they stay in the
  // Infinity Cache).  A wave then takes items of the other rows from the
  // shared queue, and finally helps drain the other stripe queues, so every
  // queue is emptied whatever the dispatch order.
  const uint64_t pieces = (uint64_t)uniform(ctl->n_striped) * ntiles;  // per stripe
  const uint32_t K = stripes.stripes > 1 ? stripes.stripes : 0u;
  const uint64_t waves = (uint64_t)gridDim.x * kWavesPerBlock;
  const uint32_t Q = stripes.queues;
  // pieces per unit: at most stripes.per_unit, and few enough that every
  // stripe has several units per wavefront of its XCD, but at least 8 when
  // stripes.per_unit allows (each unit starts with the searches for its
  // pieces' bounds, whose latency fewer pieces do not hide)
  const uint32_t per_unit = (uint32_t)max(
      (unsigned long long)min(8u, stripes.per_unit),
      min((unsigned long long)stripes.per_unit, pieces * Q / (4ull * waves)));
  const uint64_t per_stripe = (pieces + per_unit - 1) / per_unit;
  const uint64_t item_units = uniform_u64(ctl->total_units);
  const uint32_t skip_nnz = K ? stripes.min_nnz : 0u;
  const uint32_t home = blockIdx.x % Q;
  // At 2 samples per lane, when striped rows hold most of the batch (3/4 of
  // its nonzeros or more), one workgroup in kItemsFirstEvery (spread over the
  // XCDs) takes the items of the unstriped rows before its stripe queue, with
  // non-temporal gathers: their random gathers, bound by HBM, then run beside
  // the striped ones, bound by the texture path and the VALU, without
  // evicting the stripes from L2 (+1.5% on the uneven batch; with ordinary
  // loads they cost 2%, one in 2 or 16 workgroups gained less).  Batches of
  // mostly unstriped rows keep ordinary loads everywhere (non-temporal random
  // gathers are slower).
  const bool items_first =
      V == 2 && K != 0 &&
      uniform_u64(ctl->striped_nnz) * 4ull >= uniform_u64(ctl->total_nnz) * 3ull &&
      (blockIdx.x / stripes.xcds) % kItemsFirstEvery == 0;
  for (uint32_t step = 0; step <= Q; step++) {
    const uint32_t qid = step == 0 ? (items_first ? kMaxXcds : home)
                                   : (step == 1 ? (items_first ? home : kMaxXcds)
                                                : (home + step - 1) % Q);
    const uint32_t nk = qid < Q && qid < K ? (K - qid + Q - 1) / Q : 0u;
    uint64_t qtotal = qid == kMaxXcds ? item_units : per_stripe * nk;
    if (Mode == kItemsOnly && qid < kMaxXcds) qtotal = 0;
    if (qtotal == 0) continue;
    for (;;) {
      uint64_t u = 0;
      if (lane == 0) u = atomicAdd(&ctl->next[qid], 1ull);
      u = ((uint64_t)readlane_u32((uint32_t)(u >> 32), 0) << 32) | readlane_u32((uint32_t)u, 0);
      if (u >= qtotal) break;  // this queue is drained
      const bool piece = qid < kMaxXcds;
      uint32_t k = 0;
      uint64_t q0 = 0, q_end = 1;
      // lane l holds the row of the unit's piece l and the piece's bounds
      // (one vector load of the list and of the row bounds for the whole
      // unit, the lanes' binary searches side by side)
      uint32_t lm = 0, lrow = 0, ls = 0, le = 0, la = 0, lb = 0;
      if (piece) {
        const uint32_t jj = (uint32_t)(u / per_stripe);
        k = qid + jj * Q;
        q0 = (u - (uint64_t)jj * per_stripe) * per_unit;
        q_end = min(q0 + per_unit, pieces);
        if (q0 + lane < q_end) {
          const StripedRow sr = stripes.list[(uint32_t)((q0 + lane) / ntiles)];
          lm = sr.matrix;
          lrow = sr.row;
          const uint32_t *rws = mats[lm].rows;
          ls = rws[lrow];
          le = rws[lrow + 1];
          stripe_piece(mats[lm].cols - mats[lm].base, ls, le, k, K, stripes.width, la, lb);
        }
      }
      for (uint64_t q = q0; q < q_end; q++) {
        uint32_t tile, m, a, b, row, ps = 0, pe = 0;
        if (piece) {
          const uint32_t l = (uint32_t)(q - q0);
          tile = (uint32_t)(q % ntiles);
          m = readlane_u32(lm, l);
          row = readlane_u32(lrow, l);
          ps = readlane_u32(ls, l);
          pe = readlane_u32(le, l);
          a = readlane_u32(la, l);
          b = readlane_u32(lb, l);
          if (a == b) continue;
        } else {
          const uint32_t item = (uint32_t)(u / ntiles);
          tile = (uint32_t)(u - (uint64_t)item * ntiles);
          m = find_matrix<&MatrixDesc::item_begin>(mats, count, item);
          const uint32_t rows0 = uniform(mats[m].rows[0]);
          const uint32_t rend = uniform(mats[m].rows[mats[m].length]);
          a = rows0 + (item - mats[m].item_begin) * seg;
          b = a + min(seg, rend - a);
          row = ~0u;
        }
        const MatrixDesc M = mats[m];
        if (V == 2 && Mode == kBoundedPieces && piece) {
          // every piece through the skip bound (its row's minima from heavy
          // seeds; a sample without one is not skipped)
          const uint32_t kq = tile * 128u + 8u * (lane & 15u);
          const unsigned long long *slot = reinterpret_cast<const unsigned long long *>(M.output) +
                                           ((uint64_t)row * samples + (kq < samples ? kq : 0u));
          unsigned long long skey[8];
#pragma unroll
          for (int v = 0; v < 8; v++) skey[v] = slot[v];
          hash_piece_bounded<true>(params, codes, coarse, items, code_scale, code_t0, dim, sp,
                                   samples, tile, M, a, b, row, 0u, fast_div != 0u, skey);
          continue;
        } else if constexpr (V == 2 && Mode != kBoundedPieces) {
          if (piece) {
            const uint64_t col0 = (uint64_t)k * stripes.width;
            const uint32_t c0 = col0 < dim ? (uint32_t)col0 : dim;
            if (codes != nullptr && coarse != nullptr && samples % 8u == 0u) {
              // the row's keyed minima of the lane's 8 samples (quarter
              // layout); once the earlier pieces left one for every sample
              // of the tile, the bound can skip most exact evaluations
              const uint32_t kq = tile * 128u + 8u * (lane & 15u);
              unsigned long long skey[8];
              bool none = false;
              const unsigned long long *slot =
                  reinterpret_cast<const unsigned long long *>(M.output) +
                  ((uint64_t)row * samples + (kq < samples ? kq : 0u));
#pragma unroll
              for (int v = 0; v < 8; v++) {
                skey[v] = slot[v];
                none |= kq < samples && skey[v] == kNoKey;
              }
              if (__ballot(none) == 0ull &&
                  hash_piece_bounded<false>(params, codes, coarse, nullptr, code_scale, code_t0,
                                            dim, sp, samples, tile, M, a, b, row, c0,
                                            fast_div != 0u, skey))
                continue;
            }
            hash_piece_pairs(params, dim, sp, samples, tile, M, a, b, row, c0, fast_div != 0u);
            continue;
          }
        }
        const uint32_t k0 = tile * (kWave * V) + lane * V;
        const uint32_t koff = (k0 < samples) ? k0 : 0u;
        const LaneOffsets lo{koff * 4u, (sp + koff) * 4u, (2u * sp + koff) * 4u};
        const BoundCodes bc{codes, code_scale, code_t0};
        if (codes != nullptr)
          hash_item<V, WMH_GROUP_NNZ / V, true>(params, dim, sp, samples, k0, lo, M, a, b, row,
                                                 ps, pe, row == ~0u ? skip_nnz : 0u,
                                                 fast_div != 0u, items_first && step == 0, bc);
        else
          hash_item<V, WMH_GROUP_NNZ / V, false>(params, dim, sp, samples, k0, lo, M, a, b, row,
                                                  ps, pe, row == ~0u ? skip_nnz : 0u,
                                                  fast_div != 0u, items_first && step == 0, bc);
      }
    }
  }
}

// Weights in flight per lane of the seed scan (16-byte runs) and seeds whose
// gathers are issued together (8 runs, 16 seeds, or 6 workgroups per CU
// instead of 4 measured the same: the kernel is bound by HBM).
constexpr int kSeedRuns = 4;
constexpr int kSeedGroup = 8;
constexpr int kSeedBlocksPerCu = 4;

// Heavy seeds of the striped rows: one wavefront per (striped row, 128-sample
// tile), after the init pass and before the persistent kernel, when the
// stripe pieces use the skip bound (2 samples per lane, samples % 8 == 0).
// Lane l takes the heaviest nonzero of its share of the row (largest weight,
// the earliest of equal ones, NaN never; a column >= dim leaves the lane
// without a seed), so the 64 seeds are heavy nonzeros spread over the whole
// row.  They are hashed exactly for the tile's samples (2 per lane,
// kernel.cu:103-106), keeping per sample the smallest (ln_a, position) -- the
// seeds are not in CSR order, so equal values keep the earlier position
// explicitly -- and stored as keyed minima in the row's slots, which the init
// pass set to kNoKey and which only this wavefront writes before the
// persistent kernel.  Every stripe piece of the row then finds minima for its
// tile and goes through the skip bound from its first nonzero on: a
// nonzero's chance to hold a sample's minimum grows with its weight, so heavy
// seeds start the bound near the row's final minima.  The seeds are real
// candidates of the row and the merge orders by (ln_a, position), so no hash
// changes (a seed's own position is evaluated again by its piece, with the
// same key).  Costs one read of the striped rows' weights and 64 nonzeros per
// row and tile (bound by HBM: the seeds' parameter gathers move about twice
// the weights' bytes).
__global__ __launch_bounds__(kBlock) void batch_seed_kernel(
    const float *__restrict__ params, uint32_t dim, uint32_t samples, uint32_t sp,
    const MatrixDesc *__restrict__ mats, const BatchCtl *__restrict__ ctl,
    StripePlan stripes, uint32_t fast_div) {
  WMH_NO_CONTRACT
  if (uniform(ctl->invalid)) return;
  const uint32_t lane = lane_id();
  const uint32_t ntiles = (samples + 127u) / 128u;
  const uint64_t units = (uint64_t)uniform(ctl->n_striped) * ntiles;
  const uint32_t block_bytes = 12u * sp;
  const char *pbase = reinterpret_cast<const char *>(params);
  for (uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); w < units;
       w += (uint64_t)gridDim.x * kWavesPerBlock) {
    const uint32_t idx = uniform((uint32_t)(w / ntiles));
    const uint32_t tile = uniform((uint32_t)(w - (uint64_t)idx * ntiles));
    const StripedRow sr = stripes.list[idx];
    const uint32_t m = uniform(sr.matrix), row = uniform(sr.row);
    const MatrixDesc M = mats[m];
    const uint32_t s = uniform(M.rows[row]), e = uniform(M.rows[row + 1]);
    const float *__restrict__ weights = M.weights - M.base;
    const uint32_t *__restrict__ cols = M.cols - M.base;
    if (s >= e) continue;
    // The lane's heaviest nonzero of its runs of 4 positions: the runs are
    // 16-byte aligned from the one holding position s on, lane l takes runs
    // l, l + 64, ..., kSeedRuns of them in flight.  Positions outside [s, e)
    // that share a run with s or with a position of the row are read (the
    // same 16 bytes, so mapped memory) but never taken.  Runs and positions
    // go up, so the strict > keeps the earliest of equal weights.
    float bw = -__builtin_inff();
    uint32_t bp = WMH_SENTINEL;
    const int64_t a0 = (int64_t)s - (int64_t)((reinterpret_cast<uintptr_t>(weights + s) >> 2) & 3u);
    for (int64_t q0 = a0; q0 < (int64_t)e; q0 += 4 * kSeedRuns * kWave) {
      float4 x[kSeedRuns];
#pragma unroll
      for (int u = 0; u < kSeedRuns; u++) {
        const int64_t q = q0 + 4 * (int64_t)(lane + u * kWave);
        x[u] = q < (int64_t)e ? *reinterpret_cast<const float4 *>(weights + q)
                              : make_float4(-__builtin_inff(), -__builtin_inff(),
                                            -__builtin_inff(), -__builtin_inff());
      }
#pragma unroll
      for (int u = 0; u < kSeedRuns; u++) {
        const int64_t q = q0 + 4 * (int64_t)(lane + u * kWave);
        const float xs[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const bool take = q + k >= (int64_t)s && q + k < (int64_t)e && xs[k] > bw;
          bw = take ? xs[k] : bw;
          bp = take ? (uint32_t)(q + k) : bp;
        }
      }
    }
    uint32_t c = dim;
    if (bp != WMH_SENTINEL) c = cols[bp];
    const bool ok = c < dim;
    const float lw = ok ? wmh_logf(bw) : __builtin_nanf("");
    const bool fast = fast_div != 0u && __ballot(lw == __builtin_inff()) == 0ull;
    const uint32_t k0 = tile * 128u + 2u * lane;
    const uint32_t koff = k0 < samples ? k0 : 0u;
    const LaneOffsets lo{koff * 4u, (sp + koff) * 4u, (2u * sp + koff) * 4u};
    float best[2] = {FLT_MAX, FLT_MAX};
    uint32_t pos[2] = {WMH_SENTINEL, WMH_SENTINEL};
    const unsigned long long have = __ballot(ok);
    with_division(fast, [&](auto mode) {
      constexpr bool kFast = decltype(mode)::value;
      constexpr int U = kSeedGroup;
      for (uint32_t i = 0; i < (uint32_t)kWave; i += U) {
        if (((have >> i) & ((1ull << U) - 1ull)) == 0ull) continue;
        float rr[U][2], beta[U][2], lc[U][2];
        float lwu[U];
        uint32_t pu[U];
        bool oku[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const uint32_t j = readlane_u32(c, i + u);
          oku[u] = ((have >> (i + u)) & 1ull) != 0ull;
          lwu[u] = readlane_f32(lw, i + u);
          pu[u] = readlane_u32(bp, i + u);
          const __amdgpu_buffer_rsrc_t rs =
              block_rsrc(pbase + (uint64_t)(oku[u] ? j : 0u) * block_bytes, block_bytes);
          load_b<2>(rs, lo.r, rr[u]);
          load_b<2>(rs, lo.beta, beta[u]);
          load_b<2>(rs, lo.lc, lc[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
          for (int v = 0; v < 2; v++) {
            const float t = __builtin_floorf(div_lw<kFast>(lwu[u], rr[u][v]) + beta[u][v]);
            const float ln_y = (t - beta[u][v]) * rr[u][v];
            const float ln_a = (lc[u][v] - ln_y) - rr[u][v];
            const bool upd = oku[u] && (ln_a < best[v] || (ln_a == best[v] &&
                                                           pos[v] != WMH_SENTINEL &&
                                                           pu[u] < pos[v]));
            best[v] = upd ? ln_a : best[v];
            pos[v] = upd ? pu[u] : pos[v];
          }
        }
      }
    });
    unsigned long long *slot =
        reinterpret_cast<unsigned long long *>(M.output) + ((uint64_t)row * samples + k0);
#pragma unroll
    for (int v = 0; v < 2; v++) {
      if (k0 + v < samples && pos[v] != WMH_SENTINEL)
        slot[v] = ((unsigned long long)order_key(best[v]) << 32) | pos[v];
    }
  }
}

// Item plan (one workgroup): the item size, the item prefix over the
// matrices, and the reset of the work queue.  With `target_items` > 0 the
// item size is the batch's nonzeros / target_items rounded down to a multiple
// of 64, within [64, max_seg], so that a small batch is still spread over the
// CUs; otherwise it is max_seg.  Either way it is at
// least ceil(nonzeros / 2^31) rounded up to 64, so that the item count stays
// below 2^31 + count and fits the 32-bit item indices.
__global__ __launch_bounds__(kBlock) void batch_plan_kernel(MatrixDesc *__restrict__ mats,
                                                            uint32_t count, uint32_t max_seg,
                                                            uint32_t target_items,
                                                            uint32_t ntiles,
                                                            BatchCtl *__restrict__ ctl) {
  __shared__ uint32_t part[kBlock];
  __shared__ uint32_t carry;
  __shared__ uint32_t bad;
  __shared__ unsigned long long total;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    carry = 0;
    bad = 0;
    total = 0;
  }
  __syncthreads();
  // pass 1: total nonzeros (and indptr sanity)
  for (uint32_t m = tid; m < count; m += kBlock) {
    if (mats[m].length > 0) {
      const uint32_t first = mats[m].rows[0], last = mats[m].rows[mats[m].length];
      if (last < first) bad = 1;
      else atomicAdd(&total, (unsigned long long)(last - first));
    }
  }
  __syncthreads();
  uint32_t seg = max_seg;
  if (target_items > 0) {
    const unsigned long long want = (total / target_items) & ~63ull;
    seg = (uint32_t)min((unsigned long long)max_seg, max(64ull, want));
  }
  const unsigned long long min_seg = (((total + (1ull << 31) - 1) >> 31) + 63ull) & ~63ull;
  if (seg < min_seg) seg = (uint32_t)min_seg;
  // pass 2: item prefix
  for (uint32_t base = 0; base < count; base += kBlock) {
    const uint32_t m = base + tid;
    uint32_t n = 0;
    if (m < count && mats[m].length > 0) {
      const uint32_t first = mats[m].rows[0], last = mats[m].rows[mats[m].length];
      if (last >= first) {
        const uint32_t nnz = last - first;
        n = nnz / seg + (nnz % seg != 0u);
      }
    }
    part[tid] = n;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kBlock; off <<= 1) {
      const uint32_t v = tid >= off ? part[tid - off] : 0u;
      __syncthreads();
      part[tid] += v;
      __syncthreads();
    }
    if (m < count) {
      mats[m].item_begin = carry + part[tid] - n;
      mats[m].n_items = n;
    }
    __syncthreads();
    if (tid == kBlock - 1) carry += part[kBlock - 1];
    __syncthreads();
  }
  if (tid <= kMaxXcds) ctl->next[tid] = 0;
  if (tid == 0) {
    ctl->n_striped = 0;
    ctl->total_nnz = total;
    ctl->striped_nnz = 0;
    ctl->invalid = bad;
    ctl->seg = seg;
    ctl->total_items = carry;
    ctl->total_units = (unsigned long long)carry * ntiles;
    ctl->ntiles = ntiles;
  }
}

// Row scan over every matrix, 64 rows per wavefront.
//   kInit:  empty rows and split rows get all-ones in every sample slot, and
//           rows to stripe are listed in stripes.list;
//   kFixup: each split row's winning keys become (column, t).
// A row is split when it crosses an item boundary or is striped.
enum RowPass { kInit = 0, kFixup = 1 };

template <int PASS>
__global__ __launch_bounds__(kBlock) void batch_rows_kernel(
    const float *__restrict__ params, uint32_t samples, uint32_t sp,
    const MatrixDesc *__restrict__ mats, uint32_t count, uint32_t total_chunks,
    BatchCtl *__restrict__ ctl, StripePlan stripes) {
  WMH_NO_CONTRACT
  const uint32_t lane = lane_id();
  if (PASS == kFixup && uniform(ctl->invalid)) return;
  const uint32_t seg = uniform(ctl->seg);
  const uint32_t stripe_nnz = stripes.stripes > 1 ? stripes.min_nnz : 0xffffffffu;
  for (uint32_t w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); w < total_chunks;
       w += gridDim.x * kWavesPerBlock) {
    const uint32_t wv = uniform(w);
    const uint32_t m = find_matrix<&MatrixDesc::chunk_begin>(mats, count, wv);
    const MatrixDesc M = mats[m];
    const uint32_t r0 = (wv - M.chunk_begin) * kWave;
    const uint32_t row = r0 + lane;
    bool flag = false, striped = false;
    uint32_t len = 0;  // nonzeros of a striped row
    if (row < M.length) {
      const uint32_t rows0 = M.rows[0];
      const uint32_t s = M.rows[row], e = M.rows[row + 1];
      if (PASS == kInit && (e < s || s < rows0)) atomicOr(&ctl->invalid, 1u);
      const bool empty = s == e;
      striped = e > s && e - s >= stripe_nnz;
      len = striped ? e - s : 0u;
      const bool split = !empty && (striped || (s - rows0) / seg != (e - 1u - rows0) / seg);
      flag = PASS == kInit ? (empty || split) : split;
    }
    if (PASS == kInit) {
      const unsigned long long smask = __ballot(striped);
      if (smask) {
        // the wave's striped nonzeros, summed over the ballot's lanes
        unsigned long long snnz = 0;
        for (unsigned long long mm = smask; mm; mm &= mm - 1)
          snnz += readlane_u32(len, (uint32_t)__builtin_ctzll(mm));
        uint32_t base = 0;
        if (lane == 0) {
          base = atomicAdd(&ctl->n_striped, (uint32_t)__builtin_popcountll(smask));
          atomicAdd(&ctl->striped_nnz, snnz);
        }
        base = readlane_u32(base, 0);
        if (striped) {
          const uint32_t rank = (uint32_t)__builtin_popcountll(smask & ((1ull << lane) - 1ull));
          stripes.list[base + rank] = StripedRow{m, row};
        }
      }
    }
    unsigned long long mask = __ballot(flag);
    while (mask) {
      const uint32_t l = (uint32_t)__builtin_ctzll(mask);
      mask &= mask - 1;
      const uint32_t rr = uniform(r0 + l);
      unsigned long long *slot =
          reinterpret_cast<unsigned long long *>(M.output) + (uint64_t)rr * samples;
      for (uint32_t k = lane; k < samples; k += kWave) {
        if (PASS == kInit) {
          slot[k] = kNoKey;
        } else {
          const unsigned long long key = slot[k];
          if (key == kNoKey) continue;
          const uint32_t pos = (uint32_t)key - M.base;
          const uint32_t j = M.cols[pos];
          const float lw = wmh_logf(M.weights[pos]);
          const float *blk = params + (uint64_t)j * 3u * sp;
          const float t = __builtin_floorf(lw / blk[k] + blk[sp + k]);
          reinterpret_cast<uint2 *>(slot)[k] = make_uint2(j, wmh_t_bits(t));
        }
      }
    }
  }
}

// Row-length histogram of a batch for the autotuner: hist[b] counts rows and
// hist[33 + b] sums nonzeros over rows whose nonzero count n has bit length
// b (b = 0 for empty rows, 1 for n = 1, 2 for 2..3, ...).
__global__ __launch_bounds__(kBlock) void batch_histogram_kernel(
    const MatrixDesc *__restrict__ mats, uint32_t count, uint32_t total_chunks,
    unsigned long long *__restrict__ hist) {
  __shared__ unsigned long long local[66];
  for (uint32_t i = threadIdx.x; i < 66; i += kBlock) local[i] = 0;
  __syncthreads();
  const uint32_t lane = lane_id();
  for (uint32_t w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); w < total_chunks;
       w += gridDim.x * kWavesPerBlock) {
    const uint32_t wv = uniform(w);
    const uint32_t m = find_matrix<&MatrixDesc::chunk_begin>(mats, count, wv);
    const MatrixDesc M = mats[m];
    const uint32_t row = (wv - M.chunk_begin) * kWave + lane;
    if (row < M.length) {
      const uint32_t s = M.rows[row], e = M.rows[row + 1];
      const uint32_t n = e >= s ? e - s : 0u;
      const uint32_t b = n ? 32u - (uint32_t)__builtin_clz(n) : 0u;
      atomicAdd(&local[b], 1ull);
      atomicAdd(&local[33 + b], (unsigned long long)n);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < 66; i += kBlock)
    if (local[i]) atomicAdd(&hist[i], local[i]);
}

// ---------------------------------------------------------------------------
// Consumers of the hashes: weighted-Jaccard estimates and LSH band keys.

// Estimated weighted Jaccard similarity of row pairs: the fraction of samples
// whose (column, t) pairs are equal (consistent weighted sampling, Ioffe
// 2010), counting sentinel samples (rows without a selectable nonzero) as
// unequal.  One wavefront per pair; lanes compare samples, a ballot counts.
__global__ __launch_bounds__(kBlock) void jaccard_pairs_kernel(
    const uint2 *__restrict__ a, const uint2 *__restrict__ b, const uint32_t *__restrict__ pairs,
    uint32_t n, uint32_t samples, float *__restrict__ out) {
  const uint32_t lane = lane_id();
  for (uint32_t w = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); w < n;
       w += gridDim.x * kWavesPerBlock) {
    const uint32_t p = uniform(w);
    const uint64_t ra = (uint64_t)uniform(pairs[2 * p]) * samples;
    const uint64_t rb = (uint64_t)uniform(pairs[2 * p + 1]) * samples;
    uint32_t equal = 0;
    for (uint32_t k0 = 0; k0 < samples; k0 += kWave) {
      bool eq = false;
      if (k0 + lane < samples) {
        const uint2 x = a[ra + k0 + lane], y = b[rb + k0 + lane];
        eq = x.x == y.x && x.y == y.y && !(x.x == WMH_SENTINEL && x.y == WMH_SENTINEL);
      }
      equal += (uint32_t)__builtin_popcountll(__ballot(eq));
    }
    if (lane == 0) out[p] = (float)equal / (float)samples;
  }
}

// 64-bit mixing step (the splitmix64 finalizer constants of Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// LSH band keys: band g of a row covers samples [g*r, (g+1)*r), r =
// samples / bands; its key folds the band's (column, t) pairs in sample
// order, h = mix64(h ^ (column << 32 | t) + g), starting from h = mix64(g).
// A band that holds a sentinel sample gets the reserved key 2^64 - 1.  One
// thread per (row, band).
__global__ __launch_bounds__(kBlock) void band_keys_kernel(const uint2 *__restrict__ hashes,
                                                           uint32_t rows, uint32_t samples,
                                                           uint32_t bands,
                                                           unsigned long long *__restrict__ keys) {
  const uint32_t r = samples / bands;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < (uint64_t)rows * bands;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t row = i / bands;
    const uint32_t g = (uint32_t)(i - row * bands);
    const uint2 *h = hashes + row * samples + (uint64_t)g * r;
    uint64_t acc = mix64(g);
    bool sentinel = false;
    for (uint32_t k = 0; k < r; k++) {
      const uint2 x = h[k];
      sentinel |= x.x == WMH_SENTINEL && x.y == WMH_SENTINEL;
      acc = mix64((acc ^ (((uint64_t)x.x << 32) | x.y)) + g);
    }
    keys[i] = sentinel ? ~0ull : acc;
  }
}

// Plan of the allocation-free entry point (see scratch_init).  Host code
// never uses this name as an address: it takes the per-device address that
// hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void store_desc_kernel(MatrixDesc desc, MatrixDesc *__restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = desc;
}

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};

unsigned grid_for(uint64_t waves) {
  const uint64_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  return (unsigned)(blocks < (1u << 20) ? (blocks ? blocks : 1) : (1u << 20));
}

}  // namespace

hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) return e != hipSuccess ? e : hipErrorInvalidSymbol;
  scratch_ptr[device] = p;
  return hipSuccess;
}

void *scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_ptr[device];
}

hipError_t launch_store_desc(const MatrixDesc &desc, MatrixDesc *dst, hipStream_t stream) {
  hipLaunchKernelGGL(store_desc_kernel, dim3(1), dim3(64), 0, stream, desc, dst);
  return hipGetLastError();
}

hipError_t launch_gen_params(float *params, uint32_t dim, uint32_t samples,
                             uint32_t seed, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  const uint64_t total = (uint64_t)dim * sp;
  uint64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  hipLaunchKernelGGL(gen_params_kernel, dim3((unsigned)blocks), dim3(kBlock), 0,
                     stream, params, dim, samples, sp, seed);
  return hipGetLastError();
}

hipError_t launch_params_range(const float *params, uint32_t dim, uint32_t samples,
                               uint32_t *bad, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  const uint64_t total = (uint64_t)dim * samples;
  uint64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536u) blocks = 65536u;
  hipLaunchKernelGGL(params_range_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)),
                     dim3(kBlock), 0, stream, params, dim, samples, sp, bad);
  return hipGetLastError();
}

hipError_t launch_scatter_params(const float *staging, uint32_t ncols,
                                 uint32_t j0, uint32_t which, float *params,
                                 uint32_t samples, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  dim3 grid((ncols + 63) / 64, (sp + 63) / 64);
  hipLaunchKernelGGL(scatter_params_kernel, grid, dim3(kBlock), 0, stream,
                     staging, ncols, j0, which, params, samples, sp);
  return hipGetLastError();
}

hipError_t launch_gather_params(const float *params, uint32_t j0, uint32_t ncols,
                                uint32_t which, float *staging, uint32_t samples,
                                hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  dim3 grid((ncols + 63) / 64, (samples + 63) / 64);
  hipLaunchKernelGGL(gather_params_kernel, grid, dim3(kBlock), 0, stream,
                     params, j0, ncols, which, staging, samples, sp);
  return hipGetLastError();
}

hipError_t launch_plain(const float *params, uint32_t dim, uint32_t samples,
                        const MatrixDesc &m, int spl, bool fast_div, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  const uint32_t per_wave = (uint32_t)(kWave * spl);
  const uint32_t ntiles = (samples + per_wave - 1) / per_wave;
  const unsigned grid = grid_for((uint64_t)m.length * ntiles);
  switch (spl) {
    case 1:
      hipLaunchKernelGGL(wmh_plain_kernel<1>, dim3(grid), dim3(kBlock), 0, stream,
                         params, dim, samples, sp, ntiles, m.weights, m.cols, m.rows,
                         m.length, m.base, m.output, (uint32_t)fast_div);
      break;
    case 2:
      hipLaunchKernelGGL(wmh_plain_kernel<2>, dim3(grid), dim3(kBlock), 0, stream,
                         params, dim, samples, sp, ntiles, m.weights, m.cols, m.rows,
                         m.length, m.base, m.output, (uint32_t)fast_div);
      break;
    case 4:
      hipLaunchKernelGGL(wmh_plain_kernel<4>, dim3(grid), dim3(kBlock), 0, stream,
                         params, dim, samples, sp, ntiles, m.weights, m.cols, m.rows,
                         m.length, m.base, m.output, (uint32_t)fast_div);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_jaccard_pairs(const uint32_t *a, const uint32_t *b, const uint32_t *pairs,
                                uint32_t n, uint32_t samples, float *out, hipStream_t stream) {
  hipLaunchKernelGGL(jaccard_pairs_kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream,
                     reinterpret_cast<const uint2 *>(a), reinterpret_cast<const uint2 *>(b),
                     pairs, n, samples, out);
  return hipGetLastError();
}

hipError_t launch_band_keys(const uint32_t *hashes, uint32_t rows, uint32_t samples,
                            uint32_t bands, unsigned long long *keys, hipStream_t stream) {
  const uint64_t n = (uint64_t)rows * bands;
  uint64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > (1u << 20)) blocks = 1u << 20;
  hipLaunchKernelGGL(band_keys_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)),
                     dim3(kBlock), 0, stream, reinterpret_cast<const uint2 *>(hashes), rows,
                     samples, bands, keys);
  return hipGetLastError();
}

hipError_t launch_bound_range(const float *params, uint32_t dim, uint32_t samples,
                              uint32_t *minmax, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  const uint64_t total = (uint64_t)dim * samples;
  uint64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536u) blocks = 65536u;
  hipLaunchKernelGGL(bound_range_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)),
                     dim3(kBlock), 0, stream, params, dim, samples, sp, minmax);
  return hipGetLastError();
}

hipError_t launch_build_codes(const float *params, uint32_t dim, uint32_t samples, float lo,
                              float delta, uint8_t *codes, uint8_t *coarse, hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  const uint64_t total = (uint64_t)dim * sp;
  uint64_t blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  hipLaunchKernelGGL(build_codes_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)),
                     dim3(kBlock), 0, stream, params, dim, samples, sp, lo, delta, codes);
  if (coarse != nullptr) {
    if (sp % 8u != 0u) return hipErrorInvalidValue;
    const uint32_t rb = coarse_row_bytes(sp);
    uint64_t cblocks = ((uint64_t)dim * rb + kBlock - 1) / kBlock;
    if (cblocks > 65536u * 4u) cblocks = 65536u * 4u;
    hipLaunchKernelGGL(build_coarse_kernel, dim3((unsigned)std::max<uint64_t>(1, cblocks)),
                       dim3(kBlock), 0, stream, codes, dim, sp, rb, coarse);
  }
  return hipGetLastError();
}

hipError_t launch_build_items(const float *params, uint32_t dim, uint32_t samples, float *items,
                              hipStream_t stream) {
  const uint32_t sp = padded_samples(samples);
  uint64_t blocks = ((uint64_t)dim * sp + kBlock - 1) / kBlock;
  if (blocks > 65536u * 4u) blocks = 65536u * 4u;
  hipLaunchKernelGGL(build_items_kernel, dim3((unsigned)std::max<uint64_t>(1, blocks)),
                     dim3(kBlock), 0, stream, params, dim, sp, reinterpret_cast<float4 *>(items));
  return hipGetLastError();
}

hipError_t launch_histogram(const MatrixDesc *mats, uint32_t count, uint32_t total_chunks,
                            unsigned long long *hist, hipStream_t stream) {
  const unsigned grid = std::min<unsigned>(grid_for(total_chunks), 1024u);
  hipLaunchKernelGGL(batch_histogram_kernel, dim3(grid), dim3(kBlock), 0, stream, mats, count,
                     total_chunks, hist);
  return hipGetLastError();
}

hipError_t launch_persistent(const float *params, uint32_t dim, uint32_t samples,
                             MatrixDesc *mats, uint32_t count, uint32_t total_chunks,
                             BatchCtl *ctl, StripedRow *striped, const Tuning &tuning,
                             bool auto_seg, int num_cus, int num_xcds, bool fast_div,
                             hipStream_t stream, hipEvent_t ev_begin, hipEvent_t ev_end,
                             CodeTable table) {
  const uint32_t sp = padded_samples(samples);
  // (the bounded stripe pieces also need 8 whole samples per lane: the
  // kernel checks samples % 8)
  const uint8_t *codes = table.codes;
  const int spl = tuning.spl;
  if (spl != 1 && spl != 2 && spl != 4) return hipErrorInvalidValue;
  if (tuning.seg <= 0 || tuning.seg % kWave != 0 || count == 0) return hipErrorInvalidValue;
  const uint32_t xcds = (uint32_t)std::min<int>(std::max(num_xcds, 1), (int)kMaxXcds);
  StripePlan stripes{1u, 0xffffffffu, 1u, 1u, 0u, striped, xcds};
  stripe_shape(dim, samples, tuning.stripe_mib, &stripes.stripes, &stripes.min_nnz);
  stripes.width = (uint32_t)(((uint64_t)dim + stripes.stripes - 1) / stripes.stripes);
  // stripes that fit an XCD's L2 (twice over) get one queue per XCD
  const uint64_t stripe_bytes = (uint64_t)dim * 12u * sp / stripes.stripes;
  stripes.queues = stripe_bytes <= kXcdStripeBytes ? xcds : 1u;
  // about half an item's nonzeros of the shortest pieces per queue unit (at
  // most one per lane: a unit's pieces are read with one vector load)
  stripes.per_unit =
      std::min<uint32_t>(kWave, std::max(1u, (uint32_t)tuning.seg / (2u * kStripePieceNnz)));
  if (stripes.stripes > 1 && striped == nullptr) return hipErrorInvalidValue;
  const uint32_t seg = (uint32_t)tuning.seg;
  // Automatic item size: about one item per two waves of the grid.  Fewer,
  // longer items split fewer rows across items (each split row costs keyed
  // atomics and the fix-up pass); measured on batches of 4-64 serving
  // requests (1.3M-12M nonzeros), whole-call times against four items per
  // wave: 1.43 -> 1.0 ms, 2.15 -> 1.8 ms, 4.9 -> 4.5 ms.
  const uint32_t target = auto_seg ? (uint32_t)std::max(1, num_cus * tuning.waves_per_cu / 2) : 0u;
  const uint32_t per_wave = (uint32_t)(kWave * spl);
  const uint32_t ntiles = (samples + per_wave - 1) / per_wave;
  hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(kBlock), 0, stream, mats, count, seg,
                     target, ntiles, ctl);
  const unsigned row_grid = grid_for(total_chunks);
  hipLaunchKernelGGL(batch_rows_kernel<kInit>, dim3(row_grid), dim3(kBlock), 0, stream,
                     params, samples, sp, mats, count, total_chunks, ctl, stripes);
  // heavy seeds where the stripe pieces use the skip bound
  const bool seeded = tuning.seeds && codes != nullptr && table.coarse != nullptr &&
                      samples % 8u == 0u && spl == 2 && stripes.stripes > 1;
  // With seeds and a table below 4 GiB, every stripe piece takes the skip
  // bound from column 0 and the kernel leaves out the paired pieces (5 waves
  // per SIMD instead of 4).
  const bool split = seeded && (uint64_t)dim * 12u * sp < (1ull << 32);
  if (seeded)
    hipLaunchKernelGGL(batch_seed_kernel,
                       dim3((unsigned)std::max(1, num_cus * kSeedBlocksPerCu)),
                       dim3(kBlock), 0, stream, params, dim, samples, sp, mats, ctl, stripes,
                       (uint32_t)fast_div);
  const unsigned grid = (unsigned)std::max(1, num_cus * tuning.waves_per_cu / kWavesPerBlock);
  if (ev_begin) (void)hipEventRecord(ev_begin, stream);
  switch (spl) {
    case 1:
      hipLaunchKernelGGL((wmh_persistent_kernel<1, kAllWork>), dim3(grid), dim3(kBlock), 0, stream,
                         params, dim, samples, sp, mats, count, ctl, stripes,
                         (uint32_t)fast_div, codes, table.scale, table.t0, table.coarse, nullptr);
      break;
    case 2:
      if (split) {
        hipLaunchKernelGGL((wmh_persistent_kernel<2, kBoundedPieces>),
                           dim3((unsigned)std::max(1, num_cus * kPieceWavesPerCu / kWavesPerBlock)),
                           dim3(kBlock), 0, stream, params, dim, samples, sp, mats, count, ctl,
                           stripes, (uint32_t)fast_div, codes, table.scale, table.t0,
                           table.coarse, reinterpret_cast<const float4 *>(table.items));
      } else {
        hipLaunchKernelGGL((wmh_persistent_kernel<2, kAllWork>), dim3(grid), dim3(kBlock), 0,
                           stream, params, dim, samples, sp, mats, count, ctl, stripes,
                           (uint32_t)fast_div, codes, table.scale, table.t0, table.coarse, nullptr);
      }
      break;
    default:
      hipLaunchKernelGGL((wmh_persistent_kernel<4, kAllWork>), dim3(grid), dim3(kBlock), 0, stream,
                         params, dim, samples, sp, mats, count, ctl, stripes,
                         (uint32_t)fast_div, codes, table.scale, table.t0, table.coarse, nullptr);
      break;
  }
  if (ev_end) (void)hipEventRecord(ev_end, stream);
  hipLaunchKernelGGL(batch_rows_kernel<kFixup>, dim3(row_grid), dim3(kBlock), 0, stream,
                     params, samples, sp, mats, count, total_chunks, ctl, stripes);
  return hipGetLastError();
}

}  // namespace wmh
