// Internal launcher interface between the host runtime (generator.cpp) and the
// gfx950 kernels (wmh_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wmh {

// Parameter table layout in HBM: one contiguous block per column j,
//   params[j * 3 * sp + which * sp + k],  which: 0 = rs, 1 = betas, 2 = ln_cs,
// with sp = samples rounded up to a multiple of 4 (pad entries are zero).
// A nonzero's parameters for every sample are therefore one 3*sp*4-byte run
// (1.5 KiB at 128 samples) read with fully coalesced wavefront loads.
inline uint32_t padded_samples(uint32_t samples) { return (samples + 3u) & ~3u; }

// One CSR matrix of a batch (all device pointers).  rows has length + 1
// entries; CSR position p lives at weights[p - base] / cols[p - base].
struct MatrixDesc {
  const float *weights;
  const uint32_t *cols;
  const uint32_t *rows;
  uint32_t *output;       // [length][samples][2]
  uint32_t length;
  uint32_t base;
  uint32_t item_begin;    // first work item of this matrix (set on the device)
  uint32_t n_items;       // work items covering this matrix (set on the device)
  uint32_t chunk_begin;   // first 64-row chunk of this matrix in the row scans
  uint32_t pad_;
};

// Control block of one persistent launch (device memory).
// Work queues of the persistent kernel: up to one per XCD for stripe pieces
// (with Q of them, the stripes k with k % Q == q go to queue q), then the
// items' queue at index kMaxXcds.
constexpr uint32_t kMaxXcds = 8;
struct BatchCtl {
  unsigned long long next[kMaxXcds + 1];  // queue heads
  unsigned long long total_units;      // items * sample tiles
  uint32_t total_items;   // below 2^31 + count (see batch_plan_kernel)
  uint32_t ntiles;
  uint32_t invalid;       // nonzero: some indptr decreases; nothing is hashed
  uint32_t seg;           // item size chosen by the plan kernel
  uint32_t n_striped;     // striped rows listed by the init pass
  unsigned long long total_nnz;    // nonzeros of the batch (plan kernel)
  unsigned long long striped_nnz;  // nonzeros of the striped rows (init pass)
};

// Kernel tuning knobs (chosen by the autotuner).
struct Tuning {
  int spl;          // samples per lane: 1, 2 or 4
  int seg;          // nonzeros per persistent work item (multiple of 64)
  int waves_per_cu; // persistent grid size = CUs * waves_per_cu
  int stripe_mib;   // parameter bytes per column stripe in MiB (0: no stripes)
  int seeds = 0;    // 1: heavy seeds give striped rows their first minima
};

// Column stripes of the persistent launch.  When the parameter table is
// larger than one stripe, the columns are cut into `stripes` ranges of
// `width` columns, and every row with at least `min_nnz` nonzeros is cut into
// one piece per range: piece k holds the row's positions from the first
// whose column is >= k * width to the first whose column is >= (k + 1) *
// width (found by binary search; the last piece runs to the row's end).  The
// work queue hands out all first pieces, then all second pieces, and so on,
// so for rows whose columns are sorted the waves in flight gather from one
// stripe of parameters, small enough to stay in the caches and address
// translation caches.  A binary search is monotone in its key whatever the
// column order, so the pieces of any row partition it exactly and are merged
// like rows split across work items.  Other rows are packed into items as
// before.
struct StripedRow {
  uint32_t matrix;
  uint32_t row;
};
struct StripePlan {
  uint32_t stripes;    // 1: no striping
  uint32_t min_nnz;    // rows with at least this many nonzeros are striped
  uint32_t per_unit;   // pieces handed out per queue unit
  uint32_t queues;     // stripe queues: `xcds` (one per XCD) or 1
  uint32_t width;      // columns per stripe: ceil(dim / stripes)
  StripedRow *list;    // capacity: every row of the batch
  uint32_t xcds;       // the device's XCDs (workgroups go to them round-robin)
};
// Stripes up to this size get one queue per XCD (each XCD's L2 holds 4 MiB).
constexpr uint64_t kXcdStripeBytes = 8ull << 20;

// Stripe count and row threshold for a table of dim x samples parameters:
// stripes of about `stripe_mib` MiB (at most kMaxStripes of them), rows of at
// least 1024 nonzeros and of kStripePieceNnz nonzeros per stripe.
constexpr uint32_t kMaxStripes = 1024;
constexpr int kDefaultStripeMib = 4;
constexpr uint32_t kStripePieceNnz = 8;
inline void stripe_shape(uint32_t dim, uint32_t samples, int stripe_mib, uint32_t *stripes,
                         uint32_t *min_nnz) {
  const uint64_t table = (uint64_t)dim * 12u * padded_samples(samples);
  const uint64_t stripe = (uint64_t)(stripe_mib > 0 ? stripe_mib : 0) << 20;
  uint64_t k = stripe ? (table + stripe - 1) / stripe : 1;
  if (k > kMaxStripes) k = kMaxStripes;
  *stripes = k < 2 ? 1u : (uint32_t)k;
  const uint32_t rows_min = *stripes * kStripePieceNnz;
  *min_nnz = *stripes > 1 ? (rows_min > 1024u ? rows_min : 1024u) : 0xffffffffu;
}

// Device scratch of the allocation-free entry point
// (mhcuda_calc_device_noalloc): the plan of one persistent launch over one
// matrix -- its descriptor at offset 0, the control block at kScratchCtlOff
// and the striped-row list (8 bytes per row) at kScratchListOff -- in a
// file-scope __device__ array of kScratchBytes, one per GPU, kept until the
// process ends.
constexpr uint32_t kNoAllocMaxRows = 1u << 20;
constexpr size_t kScratchCtlOff = 256;
constexpr size_t kScratchListOff = 512;
constexpr size_t kScratchBytes = kScratchListOff + sizeof(StripedRow) * kNoAllocMaxRows;
static_assert(sizeof(MatrixDesc) <= kScratchCtlOff, "descriptor fits its scratch slot");
static_assert(kScratchCtlOff + sizeof(BatchCtl) <= kScratchListOff, "control block fits");

// Looks up the scratch array's address on `device` (which must be current)
// once, with hipGetSymbolAddress; later calls return at once.
hipError_t scratch_init(int device);
// The scratch array's device address on `device` (null before scratch_init).
void *scratch_address(int device);
// Writes one descriptor into device memory with a one-thread kernel (the
// descriptor travels as a kernel argument: no copy from host memory).
hipError_t launch_store_desc(const MatrixDesc &desc, MatrixDesc *dst, hipStream_t stream);

hipError_t launch_gen_params(float *params, uint32_t dim, uint32_t samples,
                             uint32_t seed, hipStream_t stream);

// Range checks of the tables: ORs into *bad (device memory, zeroed by the
// caller) kRangeDiv unless every r of a real sample lies in [2^-60, 2^60]
// (NaNs count as outside), and kRangeBound unless every beta lies in [0, 1)
// and every ln_c in [-2^60, 2^60].
constexpr uint32_t kRangeDiv = 1u, kRangeBound = 2u;
hipError_t launch_params_range(const float *params, uint32_t dim, uint32_t samples,
                               uint32_t *bad, hipStream_t stream);

// The byte-coded bound table of the skip bound (wmh_bound.h): `codes`
// [dim][sp] bytes, and the scale 1 / delta and offset t0 = 2049 - lo / delta
// of its thresholds.  codes == null: the skip bound is not used.  `coarse`
// [dim][coarse_row_bytes(sp)] holds the smallest code of every 8-sample group
// (null unless sp % 8 == 0); the bounded stripe pieces test it before the 8
// codes.
struct CodeTable {
  const uint8_t *codes;
  float scale;
  float t0;
  const uint8_t *coarse = nullptr;
  // the item table (launch_build_items) or null
  const float *items = nullptr;
};

// Bytes of a column's coarse codes: one per 8-sample group, padded to a
// multiple of 16 so that the 16 groups of a 128-sample tile are one aligned
// 16-byte load.
__host__ __device__ inline uint32_t coarse_row_bytes(uint32_t sp) {
  return ((sp / 8u) + 15u) & ~15u;
}

// Offset of the coarse codes in the bound table's allocation (after the
// codes, 256-byte aligned) and the allocation's bytes for dim x samples.
inline uint64_t coarse_offset(uint32_t dim, uint32_t samples) {
  return ((uint64_t)dim * padded_samples(samples) + 255u) & ~255ull;
}
inline uint64_t code_table_bytes(uint32_t dim, uint32_t samples) {
  const uint32_t sp = padded_samples(samples);
  return sp % 8u == 0u ? coarse_offset(dim, samples) + (uint64_t)dim * coarse_row_bytes(sp)
                       : (uint64_t)dim * sp;
}

// Smallest and largest bound entry (wmh_bound_entry) over the real samples
// of the table, as order-preserving keys (wmh_order_bits) into minmax[0]
// (atomic minimum) and minmax[1] (atomic maximum); the caller sets them to
// 0xffffffff and 0 first.  Valid only while both range checks pass.
hipError_t launch_bound_range(const float *params, uint32_t dim, uint32_t samples,
                              uint32_t *minmax, hipStream_t stream);

// Fills codes [dim][sp] with wmh_bound_code of every entry for the given lo
// and delta (padded samples: 255), then, when coarse is not null (sp % 8 ==
// 0), coarse [dim][coarse_row_bytes(sp)] with the smallest of every 8
// consecutive codes (padding bytes: 255).
hipError_t launch_build_codes(const float *params, uint32_t dim, uint32_t samples, float lo,
                              float delta, uint8_t *codes, uint8_t *coarse, hipStream_t stream);

// The item table of the bounded stripe pieces: items [dim][sp][4] holds the
// (r, beta, ln_c, 0) of every (column, sample) side by side (one memory line
// per exact evaluation instead of three).  Used only where its bytes stay
// below 2^32 (item_table_fits), which the 32-bit offsets of the gathers need.
inline uint64_t item_table_bytes(uint32_t dim, uint32_t samples) {
  return (uint64_t)dim * padded_samples(samples) * 16u;
}
inline bool item_table_fits(uint32_t dim, uint32_t samples) {
  return item_table_bytes(dim, samples) <= (1ull << 32);
}
hipError_t launch_build_items(const float *params, uint32_t dim, uint32_t samples, float *items,
                              hipStream_t stream);

// Scatter one [samples][ncols] FP32 table (rows of pitch ncols) into
// params[(j0 + jj) * 3 * sp + which * sp + k].
hipError_t launch_scatter_params(const float *staging, uint32_t ncols,
                                 uint32_t j0, uint32_t which, float *params,
                                 uint32_t samples, hipStream_t stream);
// The inverse: gather params back into a [samples][ncols] staging table.
hipError_t launch_gather_params(const float *params, uint32_t j0, uint32_t ncols,
                                uint32_t which, float *staging, uint32_t samples,
                                hipStream_t stream);

// One launch hashing every row of one matrix: one wavefront per
// (row, sample tile).
// `fast_div`: every r of the table lies in [2^-60, 2^60] (launch_params_range),
// so the kernels may divide with the short correctly rounded sequence.
hipError_t launch_plain(const float *params, uint32_t dim, uint32_t samples,
                        const MatrixDesc &m, int spl, bool fast_div, hipStream_t stream);

// Batched hashing of `count` matrices in one persistent launch (plus the
// plan, init and fix-up passes).  `mats` and `ctl` are device memory; the
// host fills every MatrixDesc field except item_begin / n_items, and
// chunk_begin must be the exclusive prefix of ceil(length / 64).
// tuning.seg is the item size, or with `auto_seg` its upper bound (the plan
// kernel then sizes items for about four per resident wavefront).
// `ev_begin` / `ev_end` (may be null) are recorded right around the
// persistent hashing kernel.
// `striped` (device memory, one entry per row of the batch) receives the
// striped rows when tuning.stripe_mib asks for stripes; it may be null
// otherwise.
// `table`: the byte-coded bound table when the skip bound may be used (both
// range checks passed), else codes = null; it is used only when samples is
// a multiple of 8, and only by the 2-samples-per-lane kernel's stripe pieces.
hipError_t launch_persistent(const float *params, uint32_t dim, uint32_t samples,
                             MatrixDesc *mats, uint32_t count, uint32_t total_chunks,
                             BatchCtl *ctl, StripedRow *striped, const Tuning &tuning,
                             bool auto_seg, int num_cus, int num_xcds, bool fast_div,
                             hipStream_t stream,
                             hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr,
                             CodeTable table = CodeTable{nullptr, 0.0f, 0.0f});

// Weighted-Jaccard estimates of row pairs (pairs[2i], pairs[2i+1] index the
// rows of hash tables a and b, each [rows][samples][2]).
hipError_t launch_jaccard_pairs(const uint32_t *a, const uint32_t *b, const uint32_t *pairs,
                                uint32_t n, uint32_t samples, float *out, hipStream_t stream);

// LSH band keys [rows][bands] of a hash table [rows][samples][2]; samples
// must be a multiple of bands.
hipError_t launch_band_keys(const uint32_t *hashes, uint32_t rows, uint32_t samples,
                            uint32_t bands, unsigned long long *keys, hipStream_t stream);

// Row-length histogram of a batch (device MatrixDesc array with chunk_begin
// filled in): hist[0..32] row counts and hist[33..65] nonzero sums by the bit
// length of the row's nonzero count.  `hist` must be zeroed beforehand.
hipError_t launch_histogram(const MatrixDesc *mats, uint32_t count, uint32_t total_chunks,
                            unsigned long long *hist, hipStream_t stream);

}  // namespace wmh
