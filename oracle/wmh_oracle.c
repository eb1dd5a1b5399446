// TEST INFRASTRUCTURE -- the CPU oracle for the Weighted MinHash hashing path.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only to check or time against it; the product
// (libMHCUDA.so) never links or calls it.
//
// It restates, on the CPU and in plain C, what the reference computes:
//   * wmh_oracle_calc      -- the per-row/per-sample argmin of kernel.cu:76-111
//                             (formula kernel.cu:95-109, output layout and the
//                             (uint32)(int32)t cast kernel.cu:83-85,108,
//                             parameter layout [samples][dim] kernel.cu:97-99),
//                             single thread.
//   * wmh_oracle_calc_omp  -- the same, rows spread over OpenMP threads (the
//                             CPU baseline).
//   * wmh_oracle_calc_f64  -- an FP64 twin (double log/div) used only to report
//                             how often FP32 rounding flips the argmin.
//   * wmh_oracle_philox    -- Philox4x32-10 (Salmon et al., SC'11).
//   * wmh_oracle_gen_params-- the parameter distributions of the reference's
//                             init path (minhashcuda.cc:176-200, kernel.cu:18-36):
//                             rs = -ln(U1*U2) ~ Gamma(2,1), ln_cs = ln(-ln(U3*U4)),
//                             betas = U5, with uniforms drawn from Philox.
//
// Semantics fixed where the reference is undefined (documented divergences):
//   * an empty row, or a row none of whose nonzeros is selectable (all weights
//     zero / negative / NaN), yields (0xffffffff, 0xffffffff) for every sample;
//   * a nonzero whose column is >= dim is ignored;
//   * (int32)t saturates, NaN -> 0.
// Selection is `ln_a < best` with best starting at FLT_MAX (kernel.cu:79,106):
// ties keep the earliest CSR position, -0 == +0, NaN and +inf never win.
//
// Build: see oracle/Makefile (-O2 -ffp-contract=off, no fast-math).
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../minhashcuda_amd/csrc/wmh_bound.h"
#include "../minhashcuda_amd/csrc/wmh_math.h"

float wmh_oracle_logf(float x) { return wmh_logf(x); }

// The skip bound's building blocks (wmh_bound.h), vectorised for the tests.
void wmh_oracle_f16_round_down(const float *x, uint16_t *h, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) h[i] = wmh_f16_round_down(x[i]);
}
void wmh_oracle_f16_to_f32(const uint16_t *h, float *x, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) x[i] = wmh_f16_to_f32(h[i]);
}
void wmh_oracle_bound_entry(const float *r, const float *lc, uint16_t *h, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) h[i] = wmh_bound_entry(r[i], lc[i]);
}
void wmh_oracle_bound_lw(const float *lw, float *out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) out[i] = wmh_bound_lw(lw[i]);
}
void wmh_oracle_code_range(float mn, float mx, float *lo, float *delta) {
  wmh_code_range(mn, mx, lo, delta);
}
void wmh_oracle_bound_code(const uint16_t *h, uint8_t *c, uint64_t n, float lo, float delta) {
  for (uint64_t i = 0; i < n; i++) c[i] = (uint8_t)wmh_bound_code(h[i], lo, delta, 1.0f / delta);
}
void wmh_oracle_code_threshold(const float *thr, uint32_t *t, uint64_t n, float scale, float t0) {
  for (uint64_t i = 0; i < n; i++) t[i] = wmh_code_threshold(thr[i], scale, t0);
}
void wmh_oracle_code_lw(const float *lwb, uint32_t *b, uint64_t n, float scale) {
  for (uint64_t i = 0; i < n; i++) b[i] = wmh_code_lw(lwb[i], scale);
}

void wmh_oracle_logf_vec(const float *x, float *y, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) y[i] = wmh_logf(x[i]);
}

// One row: kernel.cu:76-111 restated for a single row and all samples.
static void oracle_row(const float *rs, const float *ln_cs, const float *betas,
                       uint32_t dim, uint32_t samples, const float *w,
                       const uint32_t *cols, uint64_t begin, uint64_t end,
                       uint32_t *out, float *best, float *best_t,
                       uint32_t *best_j) {
  for (uint32_t k = 0; k < samples; k++) {
    best[k] = FLT_MAX;  // kernel.cu:79
    best_j[k] = WMH_SENTINEL;
  }
  for (uint64_t p = begin; p < end; p++) {
    const uint32_t j = cols[p];
    if (j >= dim) continue;
    const float lw = wmh_logf(w[p]);  // kernel.cu:95
    for (uint32_t k = 0; k < samples; k++) {
      const uint64_t ci = (uint64_t)k * dim + j;  // kernel.cu:97
      const float r = rs[ci];
      const float beta = betas[ci];
      const float t = floorf(lw / r + beta);        // kernel.cu:103
      const float ln_y = (t - beta) * r;            // kernel.cu:104
      const float ln_a = ln_cs[ci] - ln_y - r;      // kernel.cu:105
      if (ln_a < best[k]) {                         // kernel.cu:106-109
        best[k] = ln_a;
        best_j[k] = j;
        best_t[k] = t;
      }
    }
  }
  for (uint32_t k = 0; k < samples; k++) {
    if (best_j[k] == WMH_SENTINEL) {
      out[2 * k] = WMH_SENTINEL;
      out[2 * k + 1] = WMH_SENTINEL;
    } else {
      out[2 * k] = best_j[k];
      out[2 * k + 1] = wmh_t_bits(best_t[k]);
    }
  }
}

// rows: uint64 indptr of length+1 entries (absolute positions in w/cols).
void wmh_oracle_calc(const float *rs, const float *ln_cs, const float *betas,
                     uint32_t dim, uint32_t samples, const float *w,
                     const uint32_t *cols, const uint64_t *rows,
                     uint64_t length, uint32_t *out) {
  float *best = (float *)malloc(sizeof(float) * samples * 2);
  uint32_t *best_j = (uint32_t *)malloc(sizeof(uint32_t) * samples);
  for (uint64_t r = 0; r < length; r++) {
    oracle_row(rs, ln_cs, betas, dim, samples, w, cols, rows[r], rows[r + 1],
               out + r * samples * 2, best, best + samples, best_j);
  }
  free(best);
  free(best_j);
}

// The same computation with rows spread over OpenMP threads; returns the
// number of threads used.
int wmh_oracle_calc_omp(const float *rs, const float *ln_cs, const float *betas,
                        uint32_t dim, uint32_t samples, const float *w,
                        const uint32_t *cols, const uint64_t *rows,
                        uint64_t length, uint32_t *out, int threads) {
  int used = 1;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
#ifdef _OPENMP
#pragma omp single
    used = omp_get_num_threads();
#endif
    float *best = (float *)malloc(sizeof(float) * samples * 2);
    uint32_t *best_j = (uint32_t *)malloc(sizeof(uint32_t) * samples);
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < (int64_t)length; r++) {
      oracle_row(rs, ln_cs, betas, dim, samples, w, cols, rows[r], rows[r + 1],
                 out + (uint64_t)r * samples * 2, best, best + samples, best_j);
    }
    free(best);
    free(best_j);
  }
  return used;
}

// FP64 twin of wmh_oracle_calc (information only: argmin flip rate).
void wmh_oracle_calc_f64(const float *rs, const float *ln_cs, const float *betas,
                         uint32_t dim, uint32_t samples, const float *w,
                         const uint32_t *cols, const uint64_t *rows,
                         uint64_t length, uint32_t *out) {
  double *best = (double *)malloc(sizeof(double) * samples * 2);
  uint32_t *best_j = (uint32_t *)malloc(sizeof(uint32_t) * samples);
  for (uint64_t row = 0; row < length; row++) {
    for (uint32_t k = 0; k < samples; k++) {
      best[k] = DBL_MAX;
      best_j[k] = WMH_SENTINEL;
    }
    for (uint64_t p = rows[row]; p < rows[row + 1]; p++) {
      const uint32_t j = cols[p];
      if (j >= dim) continue;
      const double lw = log((double)w[p]);
      for (uint32_t k = 0; k < samples; k++) {
        const uint64_t ci = (uint64_t)k * dim + j;
        const double r = rs[ci], beta = betas[ci];
        const double t = floor(lw / r + beta);
        const double ln_a = (double)ln_cs[ci] - (t - beta) * r - r;
        if (ln_a < best[k]) {
          best[k] = ln_a;
          best_j[k] = j;
          best[samples + k] = t;
        }
      }
    }
    uint32_t *o = out + row * samples * 2;
    for (uint32_t k = 0; k < samples; k++) {
      if (best_j[k] == WMH_SENTINEL) {
        o[2 * k] = o[2 * k + 1] = WMH_SENTINEL;
      } else {
        o[2 * k] = best_j[k];
        o[2 * k + 1] = wmh_t_bits((float)best[samples + k]);
      }
    }
  }
  free(best);
  free(best_j);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the Weyl key
// schedule.
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void wmh_oracle_philox(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; round++) {
    if (round > 0) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Uniforms used by the parameter generator.
//   open (0,1): odd multiples of 2^-24, so products never reach 1 and logs of
//               them are finite and negative;
//   [0,1):      multiples of 2^-24.
static float u_open(uint32_t x) { return (float)((x >> 9) * 2u + 1u) * 0x1p-24f; }
static float u_half_open(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

#define WMH_PHILOX_KEY1 0x4d484344u  // second key word: "MHCD"

// Parameters for logical element (k, j), n = k*dim + j:
//   A = Philox({n_lo, n_hi, 0, 0}, {seed, KEY1}), B = Philox({n_lo, n_hi, 1, 0}, ...)
//   rs    = -ln(u(A0) * u(A1))
//   ln_cs =  ln(-ln(u(A2) * u(A3)))
//   betas =  u'(B0)
// Outputs are in the reference layout [samples][dim].
// One parameter triple of the generated table: sample k of column j of a
// table `dim` columns wide (counter k * dim + j).
static void gen_one(uint32_t seed, uint64_t n, float *r, float *lc, float *beta) {
  const uint32_t key[2] = {seed, WMH_PHILOX_KEY1};
  uint32_t a[4], b[4];
  uint32_t ca[4] = {(uint32_t)n, (uint32_t)(n >> 32), 0u, 0u};
  uint32_t cb[4] = {(uint32_t)n, (uint32_t)(n >> 32), 1u, 0u};
  wmh_oracle_philox(ca, key, a);
  wmh_oracle_philox(cb, key, b);
  *r = -wmh_logf(u_open(a[0]) * u_open(a[1]));
  *lc = wmh_logf(-wmh_logf(u_open(a[2]) * u_open(a[3])));
  *beta = u_half_open(b[0]);
}

void wmh_oracle_gen_params(uint32_t seed, uint32_t dim, uint32_t samples,
                           float *rs, float *ln_cs, float *betas) {
  const uint64_t total = (uint64_t)dim * samples;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < (int64_t)total; n++) gen_one(seed, (uint64_t)n, rs + n, ln_cs + n, betas + n);
}

// The columns cols[0..ncols) of the table wmh_oracle_gen_params would produce
// for `dim` columns, as [samples][ncols] tables (column i of the result is
// table column cols[i]): the counter-based generator needs no other column.
void wmh_oracle_gen_params_cols(uint32_t seed, uint32_t dim, uint32_t samples,
                                const uint32_t *cols, uint64_t ncols, float *rs,
                                float *ln_cs, float *betas) {
  const uint64_t total = ncols * samples;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)total; i++) {
    const uint64_t k = (uint64_t)i / ncols, c = (uint64_t)i % ncols;
    gen_one(seed, k * dim + cols[c], rs + i, ln_cs + i, betas + i);
  }
}
