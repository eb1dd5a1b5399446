// TEST INFRASTRUCTURE -- the timed CPU baseline of bench.py (never part of
// the product): the same hashing formula as wmh_oracle.c (kernel.cu:95-109,
// binary32, the same operation order, contraction off, IEEE division), with
// the parameter tables transposed to one contiguous run per column
// ([dim][samples] for each table) so that the loop over samples is
// contiguous and vectorises (x86-64-v3: 8 floats per AVX2 instruction), and
// rows spread over OpenMP threads.  Its hashes equal the oracle's bit for
// bit (checked by tests/test_oracle.py and by bench.py on every sample it
// times).
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../minhashcuda_amd/csrc/wmh_math.h"

// rs_t, lc_t, b_t: [dim][samples]; rows: length + 1 absolute positions.
int wmh_cpu_calc_fast(const float *rs_t, const float *lc_t, const float *b_t, uint32_t dim,
                      uint32_t samples, const float *w, const uint32_t *cols,
                      const uint64_t *rows, uint64_t length, uint32_t *out, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  int used = 1;
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
    float *restrict best = (float *)malloc(sizeof(float) * samples);
    float *restrict best_t = (float *)malloc(sizeof(float) * samples);
    uint32_t *restrict best_j = (uint32_t *)malloc(sizeof(uint32_t) * samples);
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < (int64_t)length; r++) {
      for (uint32_t k = 0; k < samples; k++) {
        best[k] = FLT_MAX;
        best_t[k] = 0.0f;
        best_j[k] = WMH_SENTINEL;
      }
      for (uint64_t p = rows[r]; p < rows[r + 1]; p++) {
        const uint32_t j = cols[p];
        if (j >= dim) continue;
        const float lw = wmh_logf(w[p]);
        const float *restrict R = rs_t + (uint64_t)j * samples;
        const float *restrict B = b_t + (uint64_t)j * samples;
        const float *restrict C = lc_t + (uint64_t)j * samples;
        for (uint32_t k = 0; k < samples; k++) {
          const float t = floorf(lw / R[k] + B[k]);
          const float ln_y = (t - B[k]) * R[k];
          const float ln_a = C[k] - ln_y - R[k];
          const int upd = ln_a < best[k];
          best[k] = upd ? ln_a : best[k];
          best_t[k] = upd ? t : best_t[k];
          best_j[k] = upd ? j : best_j[k];
        }
      }
      uint32_t *o = out + (uint64_t)r * samples * 2;
      for (uint32_t k = 0; k < samples; k++) {
        const int sel = best_j[k] != WMH_SENTINEL;
        o[2 * k] = best_j[k];
        o[2 * k + 1] = sel ? wmh_t_bits(best_t[k]) : WMH_SENTINEL;
      }
    }
    free(best);
    free(best_t);
    free(best_j);
  }
  return used;
}
