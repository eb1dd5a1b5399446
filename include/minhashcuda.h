/*
 * Public C ABI of the MI355X Weighted MinHash engine (libMHCUDA.so).
 *
 * The first block is the drop-in boundary: the same six entry points, the same
 * result codes and the same parameter struct as the reference's minhashcuda.h,
 * with the same argument meaning and error behaviour.  Each declaration cites
 * the reference interface it replaces.  Unlike the reference header this one
 * also compiles as C (guards added, `enum` spelled out through a typedef).
 *
 * The second block adds entry points the reference does not have: hashing of
 * device-resident CSR data on a caller stream, batched hashing of many
 * matrices in one persistent launch, kernel selection, and a serving loop.
 */
#ifndef MHCUDA_MINHASHCUDA_H
#define MHCUDA_MINHASHCUDA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef __GNUC__
#define MALLOC __attribute__((malloc))
#else
#define MALLOC
#endif

/* Opaque generator: parameter tables resident in HBM plus grow-only device
 * buffers.  Reference: minhashcuda.h:41, minhashcuda.cc:13-38. */
typedef struct MinhashCudaGenerator_ MinhashCudaGenerator;

/* Reference: minhashcuda.h:44-48 (12-byte layout: dim@0, samples@4, verbosity@8). */
typedef struct {
  uint32_t dim;
  uint16_t samples;
  int verbosity;
} MinhashCudaGeneratorParameters;

/* Reference: minhashcuda.h:51-58, same values. */
enum MHCUDAResult {
  mhcudaSuccess,
  mhcudaInvalidArguments,
  mhcudaNoSuchDevice,
  mhcudaMemoryAllocationFailure,
  mhcudaRuntimeError,
  mhcudaMemoryCopyError
};
typedef enum MHCUDAResult MHCUDAResult;

/* Creates a generator for `dim`-dimensional rows and `samples` hashes per row.
 * Unless `deferred`, the parameter tables are generated on the device with
 * Philox4x32-10 from `seed` (rs ~ Gamma(2,1), ln_cs = ln Gamma(2,1),
 * betas ~ U[0,1)).  `devices` is a bitmask (0 = all); devices that are not
 * gfx950 are dropped and the first remaining device is used.  On error
 * returns NULL and stores the code in *status (if non-NULL); on success
 * *status is left untouched.
 * Reference: minhashcuda.h:74-76, minhashcuda.cc:212-242. */
MinhashCudaGenerator *mhcuda_init(uint32_t dim, uint16_t samples, uint32_t seed,
                                  int deferred, uint32_t devices, int verbosity,
                                  MHCUDAResult *status) MALLOC;

/* {dim, samples, verbosity}; all zero for NULL.
 * Reference: minhashcuda.h:81-82, minhashcuda.cc:244-252. */
MinhashCudaGeneratorParameters mhcuda_get_parameters(const MinhashCudaGenerator *gen);

/* Copies the three [samples][dim] float32 tables to host memory.
 * Reference: minhashcuda.h:91-92, minhashcuda.cc:254-270. */
MHCUDAResult mhcuda_retrieve_random_vars(const MinhashCudaGenerator *gen,
                                         float *rs, float *ln_cs, float *betas);

/* Uploads three [samples][dim] float32 tables (host memory).
 * Reference: minhashcuda.h:101-103, minhashcuda.cc:272-285. */
MHCUDAResult mhcuda_assign_random_vars(const MinhashCudaGenerator *gen,
                                       const float *rs, const float *ln_cs,
                                       const float *betas);

/* Hashes `length` CSR rows held in host memory.  rows has length + 1 entries
 * that index weights/cols absolutely (rows[0] need not be 0).  output receives
 * length * samples * 2 uint32: (column, (uint32)(int32)t) per sample.  Returns
 * after the output is in host memory.  An odd sample count is rejected with
 * mhcudaInvalidArguments, as in the reference.
 * Reference: minhashcuda.h:115-118, minhashcuda.cc:583-642. */
MHCUDAResult mhcuda_calc(const MinhashCudaGenerator *gen, const float *weights,
                         const uint32_t *cols, const uint32_t *rows,
                         uint32_t length, uint32_t *output);

/* Frees the generator and every device buffer; NULL is accepted.
 * Reference: minhashcuda.h:124, minhashcuda.cc:644-649. */
MHCUDAResult mhcuda_fini(MinhashCudaGenerator *gen);

/* ------------------------------------------------------------------------ */
/* Extensions (no counterpart in the reference).                            */

/* One CSR matrix for the batched and device entry points.  All pointers are
 * device pointers for the *_device calls and host pointers for
 * mhcuda_calc_batch. */
typedef struct {
  const float *weights;
  const uint32_t *cols;
  const uint32_t *rows;   /* length + 1 entries, absolute positions */
  uint32_t length;
  uint32_t *output;       /* length * samples * 2 */
} MHCUDAMatrix;

/* Kernel choice for the hashing calls. */
enum MHCUDAKernel {
  mhcudaKernelAuto = 0,        /* autotuned choice */
  mhcudaKernelPlain = 1,       /* one launch per matrix, one wavefront per row */
  mhcudaKernelPersistent = 2   /* nnz-balanced persistent work queue */
};

/* Hashes device-resident CSR data.  `stream` is a hipStream_t; NULL means the
 * generator's own stream, ordered after the work already issued to the
 * legacy default stream, and the call returns after completion; otherwise
 * the work is only enqueued on `stream`.  (The same holds for every `stream`
 * argument below.) */
MHCUDAResult mhcuda_calc_device(const MinhashCudaGenerator *gen,
                                const float *weights, const uint32_t *cols,
                                const uint32_t *rows, uint32_t length,
                                uint32_t *output, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device or
 * host memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same as mhcuda_calc_device with the persistent
 * kernel, but the launch plan lives in a fixed device scratch array of
 * 8 MiB + 512 bytes per GPU (looked up at mhcuda_init, kept until the process
 * ends) instead of the generator's grow-only buffers: nothing is allocated.
 * At most 1,048,576 rows per call (more: mhcudaInvalidArguments).  Every call
 * on one GPU shares that GPU's scratch, so calls that overlap on one GPU (on
 * different streams, or from generators on the same device) race: order them
 * on one stream or synchronize between them. */
MHCUDAResult mhcuda_calc_device_noalloc(const MinhashCudaGenerator *gen,
                                        const float *weights, const uint32_t *cols,
                                        const uint32_t *rows, uint32_t length,
                                        uint32_t *output, void *stream);

/* Hashes `count` device-resident matrices with the chosen kernel (plain: one
 * launch per matrix; persistent: one launch for the whole batch). */
MHCUDAResult mhcuda_calc_batch_device(const MinhashCudaGenerator *gen,
                                      uint32_t count, const MHCUDAMatrix *mats,
                                      int kernel, void *stream);

/* Host-memory version of mhcuda_calc_batch_device (uploads, hashes, downloads). */
MHCUDAResult mhcuda_calc_batch(const MinhashCudaGenerator *gen, uint32_t count,
                               const MHCUDAMatrix *mats, int kernel);

/* Tuning knobs of the generator (samples per lane 1/2/4, nonzeros per
 * persistent work item, persistent waves per CU); 0 keeps the current value.
 * An explicit segment fixes the item size; segment = -1 restores the default
 * behaviour, where calls whose nonzero count the host knows (host-memory
 * calls, serving) shrink it for small batches. */
MHCUDAResult mhcuda_set_tuning(MinhashCudaGenerator *gen, int samples_per_lane,
                               int segment, int waves_per_cu);
MHCUDAResult mhcuda_get_tuning(const MinhashCudaGenerator *gen,
                               int *samples_per_lane, int *segment,
                               int *waves_per_cu);

/* Column stripes of the persistent kernel: with a parameter table larger
 * than `stripe_mib` MiB, the columns are cut into as many ranges as the table
 * has stripes of that size, rows of many nonzeros are cut into one piece per
 * column range, and all first pieces are hashed before all second pieces, and
 * so on, so that concurrent gathers share one stripe of the table (for rows
 * with sorted columns; any column order is hashed exactly).  0 turns stripes
 * off.
 * The hashes do not depend on it. */
MHCUDAResult mhcuda_set_stripes(MinhashCudaGenerator *gen, int stripe_mib);
/* The stripe size, the resulting stripe count for this generator's table
 * (1: off) and the smallest striped row (nonzeros; 0 when off). */
MHCUDAResult mhcuda_get_stripes(const MinhashCudaGenerator *gen, int *stripe_mib,
                                uint32_t *stripes, uint32_t *min_nnz);

/* The skip bound of the persistent kernel (on by default): a table of one
 * byte per (column, sample), a code of a lower bound of the hash value, lets
 * the kernel skip the exact evaluation (and the parameter gathers) of every
 * (nonzero, sample) that provably cannot beat the row's current minimum, in
 * its row walks and in its column-stripe pieces.  The hashes are the same
 * either way.  It needs tables with every r in [2^-60, 2^60], beta in
 * [0, 1) and |ln_c| <= 2^60 whose every |ln_c - r| stays within the fp16 range
 * (generated tables always qualify; assigned or loaded ones are checked),
 * and dim * roundup(samples, 4) bytes of device memory (plus dim * roundup(samples
 * / 8, 16) bytes of coarse codes when samples % 8 == 0, and then, while
 * 16 * dim * roundup(samples, 4) bytes stay within 4 GiB, that many bytes of
 * an item table of interleaved parameters; allocated by mhcuda_init;
 * MHCUDA_SKIP_BOUND=0 in the environment skips all of them,
 * MHCUDA_ITEM_TABLE=0 the item table alone); the stripe pieces also need a
 * sample count divisible by 8.
 * mhcuda_set_skip(0) turns it off; mhcuda_get_skip reports the switch and
 * whether the bound is in use for the current tables. */
MHCUDAResult mhcuda_set_skip(MinhashCudaGenerator *gen, int enable);
MHCUDAResult mhcuda_get_skip(const MinhashCudaGenerator *gen, int *enabled, int *active);

/* Heavy seeds of the column-stripe pieces (off by default; the autotuner may
 * turn them on): before the persistent kernel, one wavefront per (striped
 * row, 128-sample tile) hashes 64 heavy nonzeros of the row (the heaviest of
 * each 64th position) and stores their minima, so that every stripe piece,
 * the row's first included, takes the skip bound from its first nonzero on.
 * Used when the skip bound is active, at 2 samples per lane, with a sample
 * count divisible by 8 and more than one stripe.  Pays on rows whose weights
 * spread widely (the heavy nonzeros then hold most minima); the hashes do not
 * depend on it.  No counterpart in the reference. */
MHCUDAResult mhcuda_set_seeds(MinhashCudaGenerator *gen, int enable);
MHCUDAResult mhcuda_get_seeds(const MinhashCudaGenerator *gen, int *enabled);

/* Result of mhcuda_autotune. */
typedef struct {
  uint64_t key;              /* samples | p50 bin << 16 | p90 bin << 24 | size class << 32 */
  uint64_t histogram_rows;   /* rows in the batch */
  uint64_t histogram_nnz;    /* nonzeros in the batch */
  int kernel;                /* chosen MHCUDAKernel */
  int samples_per_lane;
  int segment;
  int waves_per_cu;
  float sample_ms;           /* time of the chosen configuration on the timed prefix */
  uint64_t sample_nnz;       /* nonzeros of the timed prefix */
  int cached;                /* 1: taken from the cache, nothing was timed */
  int stripe_mib;            /* column-stripe size (0: no stripes) */
  int seeds;                 /* heavy seeds of the striped rows (mhcuda_set_seeds) */
} MHCUDATuneResult;

/* Autotunes the kernel choice and tuning for a batch of device-resident
 * matrices: measures the batch's nonzero-weighted row-length histogram on the
 * device, and unless a batch of the same shape (histogram bins of the median
 * and 90th-percentile nonzero, size class, samples) was tuned before, times
 * candidates (kernel, samples per lane, item size, waves per CU, column-stripe
 * size, heavy seeds) on a prefix of the batch holding up to 2^32 nonzeros.  The choice
 * becomes the generator's tuning (kernel mhcudaKernelAuto then uses it).  The
 * timed runs write the prefix's hashes into its outputs. */
MHCUDAResult mhcuda_autotune(MinhashCudaGenerator *gen, uint32_t count,
                             const MHCUDAMatrix *mats, void *stream,
                             MHCUDATuneResult *result);

/* Enables (1) or disables (0) hipGraph replay of mhcuda_calc_device and
 * mhcuda_calc_batch_device: the first call with a given set of descriptors,
 * kernel and tuning is captured into a graph, later identical calls replay
 * it with a single launch. */
MHCUDAResult mhcuda_set_graph(MinhashCudaGenerator *gen, int enable);

/* Number of graphs captured so far (each distinct call shape captures once). */
MHCUDAResult mhcuda_graph_captures(const MinhashCudaGenerator *gen, uint64_t *captures);

/* Enables (1) or disables (0) HIP-event timing of the hashing kernel(s) of
 * each call: the persistent kernel alone, or the span of the plain launches;
 * for a graph replay, the whole replayed sequence. */
MHCUDAResult mhcuda_set_timing(MinhashCudaGenerator *gen, int enable);

/* Milliseconds taken by the timed kernel(s) of the last hashing call (waits
 * for them to finish); mhcudaInvalidArguments if that call was not timed. */
MHCUDAResult mhcuda_kernel_time(const MinhashCudaGenerator *gen, float *ms);

/* Test hook: arms a failure of the host-memory paths (mhcuda_calc,
 * mhcuda_calc_batch, serving) so that callers can check how errors
 * propagate.  what = 1: the nth device-buffer reservation from now fails
 * (mhcudaMemoryAllocationFailure); what = 2: the nth host<->device copy from
 * now fails (mhcudaMemoryCopyError); what = 0 disarms both.  A failed call
 * leaves nothing in flight, and the generator (or server) stays usable. */
MHCUDAResult mhcuda_inject_fault(MinhashCudaGenerator *gen, int what, uint32_t nth);

/* ------------------------------------------------------------------------ */
/* Parameter files: the three tables (or, for tables generated on the device,
 * just the seed) in a versioned, checksummed file, so that a later process
 * reproduces the same hashes.  They extend the retrieve / assign pair
 * (reference minhashcuda.cc:254-285) and the deferred-init flow of the
 * reference's test.py:154-182 to a file.  Format version 1: a 64-byte header
 * (magic "WMHPARAM", version, kind, dim, samples, seed, generator recipe,
 * three table checksums, header checksum), then for kind 1 the tables rs,
 * ln_cs, betas, each [samples][dim] float32 little-endian.
 * File errors return mhcudaMemoryCopyError; a malformed file, a checksum
 * mismatch or a file for another dim / sample count returns
 * mhcudaInvalidArguments. */

/* Writes the generator's tables to `path` (through a temporary file renamed
 * into place).  With allow_seed_only != 0 and tables generated on the device
 * from a seed, only the seed is written (kind 2, 64 bytes).  A deferred
 * generator whose tables were never assigned is rejected. */
MHCUDAResult mhcuda_save_random_vars(const MinhashCudaGenerator *gen, const char *path,
                                     int allow_seed_only);

/* Loads a parameter file into a generator of the same dim and sample count
 * (typically created with deferred = 1).  If the tables fail their checksum
 * the generator's tables are zeroed, as after a deferred init, and
 * mhcudaInvalidArguments is returned. */
MHCUDAResult mhcuda_load_random_vars(const MinhashCudaGenerator *gen, const char *path);

/* Reads dim and samples from a parameter file's header (no device work). */
MHCUDAResult mhcuda_params_file_info(const char *path, uint32_t *dim, uint16_t *samples);

/* ------------------------------------------------------------------------ */
/* Consumers of the hashes (device memory; hash tables are [rows][samples][2]
 * uint32 as written by the hashing calls of this generator).               */

/* Weighted-Jaccard estimates: out[i] = fraction of samples whose (column, t)
 * pairs are equal between row pairs[2i] of hashes_a and row pairs[2i+1] of
 * hashes_b (the estimator datasketch's WeightedMinHash.jaccard computes for
 * the reference's test.py:104-106).  Samples of rows without a selectable
 * nonzero (the sentinel pair) never count as equal.  NULL stream: generator
 * stream and wait; otherwise enqueued on `stream`. */
MHCUDAResult mhcuda_jaccard_pairs(const MinhashCudaGenerator *gen, const uint32_t *hashes_a,
                                  const uint32_t *hashes_b, const uint32_t *pairs, uint32_t n,
                                  float *out, void *stream);

/* LSH band keys for candidate generation: keys[row * bands + g] is a 64-bit
 * key of the (column, t) pairs of samples [g*r, (g+1)*r), r = samples / bands
 * (bands must divide samples); rows sharing a key in any band are candidate
 * pairs.  h = mix(g); for each sample h = mix((h ^ (column << 32 | t)) + g),
 * with mix the splitmix64 finalizer.  A band holding a sentinel sample gets
 * the reserved key 2^64 - 1, which consumers skip. */
MHCUDAResult mhcuda_lsh_band_keys(const MinhashCudaGenerator *gen, const uint32_t *hashes,
                                  uint32_t rows, uint32_t bands, uint64_t *keys, void *stream);

/* LSH candidate pairs from band keys (device memory, [rows][bands] as written
 * by mhcuda_lsh_band_keys): the distinct row pairs (i < j) whose keys agree in
 * at least one band, skipping the reserved key and buckets of more than
 * max_bucket rows (max_bucket >= 2), as pair codes i << 32 | j sorted
 * ascending.  Writes the first min(*count, capacity) codes to `pairs` (device
 * memory; may be NULL when capacity is 0) and the number of distinct pairs to
 * *count (host memory): call again with a larger buffer when *count exceeds
 * capacity.  Returns after completion; `stream` orders the work after the
 * caller's (NULL: after the legacy default stream). */
MHCUDAResult mhcuda_lsh_candidates(const MinhashCudaGenerator *gen, const uint64_t *keys,
                                   uint32_t rows, uint32_t bands, uint32_t max_bucket,
                                   uint64_t *pairs, uint64_t capacity, uint64_t *count,
                                   void *stream);

/* ------------------------------------------------------------------------ */
/* Serving loop: continuous batching of CSR requests over one generator.     */

typedef struct MHCUDAServer_ MHCUDAServer;

typedef struct {
  uint64_t memory_budget;       /* device bytes for requests in flight: one pool, shared by
                                   every slot (a request of up to the whole budget is admitted) */
  uint32_t slots;               /* launches in flight at once (one stream each) */
  uint32_t max_batch_requests;  /* most requests coalesced into one launch */
} MHCUDAServeConfig;

typedef struct {
  uint64_t completed;           /* requests hashed since the last reset */
  uint64_t rejected;            /* requests larger than the whole memory budget */
  uint64_t batches;             /* launches */
  double mean_requests_per_batch;
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> hashes in host memory */
  double window_s;              /* first submit -> last completion */
  double nnz_samples_per_sec;   /* completed nonzeros * samples / window_s */
  /* where the time goes (means): per request, submit -> admitted into a slot;
     per launch, the host side of the uploads (staging pageable arrays through
     pinned blocks, enqueueing the copies), then the device side (the hashing
     launch and the downloads, until they complete), then copying hashes out
     of the staging blocks into pageable outputs */
  double queue_ms, stage_in_ms, device_ms, stage_out_ms;
} MHCUDAServeStats;

/* Starts the scheduler, slot workers and their device pool (memory_budget
 * bytes, allocated here; at most 64 slots).  The generator must outlive the
 * server.  Hashing calls on the generator may run meanwhile; calls that
 * change its parameters or tuning (assign / load random vars, set_tuning,
 * set_stripes, set_skip, autotune) return mhcudaInvalidArguments until
 * mhcuda_serve_stop, since the server's launches read them without taking
 * the generator's lock. */
MHCUDAServer *mhcuda_serve_start(MinhashCudaGenerator *gen, const MHCUDAServeConfig *config,
                                 MHCUDAResult *status);

/* Queues one CSR request (host memory, same conventions as mhcuda_calc);
 * the arrays must stay valid and `output` untouched until the request is
 * waited for.  Requests are admitted in arrival order as slots and pool
 * memory free up; one larger than the whole budget completes at once with
 * mhcudaMemoryAllocationFailure. */
MHCUDAResult mhcuda_serve_submit(MHCUDAServer *server, const float *weights,
                                 const uint32_t *cols, const uint32_t *rows,
                                 uint32_t length, uint32_t *output, uint64_t *ticket);

/* Blocks until the request is done; returns its status and (optionally) its
 * latency.  Each ticket can be waited for once. */
MHCUDAResult mhcuda_serve_wait(MHCUDAServer *server, uint64_t ticket, double *latency_ms);

/* Latency percentiles and throughput since the last reset (reset != 0 resets). */
MHCUDAResult mhcuda_serve_stats(MHCUDAServer *server, MHCUDAServeStats *stats, int reset);

/* Hashes everything still queued, then stops the threads and frees the server. */
MHCUDAResult mhcuda_serve_stop(MHCUDAServer *server);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* MHCUDA_MINHASHCUDA_H */
