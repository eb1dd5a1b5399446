/* Test infrastructure: runs the C oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer (make -C oracle sanitize) on inputs that reach
 * its edge paths: empty rows, columns >= dim, zero / negative / NaN /
 * infinite / subnormal / huge weights, sample counts that are not multiples
 * of 4, dim = 1; the single-thread, OpenMP and FP64 walks, the parameter
 * generator, and wmh_logf on random bit patterns.  Exits non-zero on a
 * mismatch between the single-thread and OpenMP walks; the sanitizers abort
 * on any invalid access or undefined operation. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void wmh_oracle_calc(const float *, const float *, const float *, uint32_t, uint32_t,
                     const float *, const uint32_t *, const uint64_t *, uint64_t, uint32_t *);
int wmh_oracle_calc_omp(const float *, const float *, const float *, uint32_t, uint32_t,
                        const float *, const uint32_t *, const uint64_t *, uint64_t,
                        uint32_t *, int);
void wmh_oracle_calc_f64(const float *, const float *, const float *, uint32_t, uint32_t,
                         const float *, const uint32_t *, const uint64_t *, uint64_t,
                         uint32_t *);
void wmh_oracle_gen_params(uint32_t, uint32_t, uint32_t, float *, float *, float *);
float wmh_oracle_logf(float);

static uint32_t state = 12345u;
static uint32_t next(void) {  /* xorshift32 */
  state ^= state << 13;
  state ^= state >> 17;
  state ^= state << 5;
  return state;
}

int main(void) {
  const uint32_t dims[] = {1, 7, 130, 1000}, samples[] = {2, 6, 128, 130};
  const float specials[] = {0.0f, -0.0f, -1.0f, INFINITY, -INFINITY, NAN,
                            1e-45f, 1.17549435e-38f, 3.4e38f, 1.0f};
  for (int di = 0; di < 4; di++) {
    for (int si = 0; si < 4; si++) {
      const uint32_t dim = dims[di], S = samples[si];
      const size_t n = (size_t)dim * S;
      float *rs = malloc(n * 4), *lc = malloc(n * 4), *bt = malloc(n * 4);
      wmh_oracle_gen_params(7u + (uint32_t)di, dim, S, rs, lc, bt);
      const uint64_t R = 50;
      uint64_t rows[51];
      rows[0] = 0;
      for (uint64_t r = 0; r < R; r++) rows[r + 1] = rows[r] + (r % 7 == 0 ? 0 : next() % 40);
      const uint64_t nnz = rows[R];
      float *w = malloc((nnz + 1) * 4);
      uint32_t *c = malloc((nnz + 1) * 4);
      for (uint64_t p = 0; p < nnz; p++) {
        w[p] = next() % 5 == 0 ? specials[next() % 10] : (float)(next() % 1000) / 7.0f;
        c[p] = next() % 11 == 0 ? dim + next() % 5 : next() % dim;
      }
      uint32_t *o1 = malloc(R * S * 8), *o2 = malloc(R * S * 8), *o3 = malloc(R * S * 8);
      wmh_oracle_calc(rs, lc, bt, dim, S, w, c, rows, R, o1);
      wmh_oracle_calc_omp(rs, lc, bt, dim, S, w, c, rows, R, o2, 4);
      wmh_oracle_calc_f64(rs, lc, bt, dim, S, w, c, rows, R, o3);
      if (memcmp(o1, o2, R * S * 8) != 0) {
        printf("OpenMP walk differs: dim %u samples %u\n", dim, S);
        return 1;
      }
      free(rs); free(lc); free(bt); free(w); free(c); free(o1); free(o2); free(o3);
    }
  }
  unsigned zeros = 0;
  for (uint32_t i = 0; i < 2000000u; i++) {
    const uint32_t b = next();
    float x;
    memcpy(&x, &b, 4);
    zeros += wmh_oracle_logf(x) == 0.0f;
  }
  for (int i = 0; i < 10; i++) zeros += wmh_oracle_logf(specials[i]) == 0.0f;
  printf("sanitize_check: ok (%u zero logarithms)\n", zeros);
  return 0;
}
