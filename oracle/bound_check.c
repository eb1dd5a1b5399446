// TEST INFRASTRUCTURE -- brute-force check of the skip bound of
// minhashcuda_amd/csrc/wmh_bound.h against the hashing formula evaluated
// exactly as the oracle and the kernels do (kernel.cu:103-105, binary32,
// round to nearest, unfused, IEEE division).
//
//   bound_check [variant] [trials]
//
// Variants 0-4 count trials where the bound L = X - lwb exceeds the computed
// ln_a (a skip that could drop a selectable nonzero).  Variant 0 is the
// bound the engine uses and must give no violation; the planted mistakes
// must give some, which shows the trials reach the cases that matter:
//   1: the "- r" term dropped from the table entry;
//   2: the sign of lw flipped;
//   3: no margins, table entry rounded to nearest instead of down;
//   4: no margins (table entry rounded down, lwb = lw): only the binary32
//      rounding of ln_a below lc - r - lw, a few ulps, remains to be caught.
// Variants 5-8 test the byte codes the engine stores (wmh_code_*): over a
// table range around the entry and running minima thr at or just above
// ln_a (ties included), a violation is a skip (c >= T + cB) with
// ln_a <= thr.  Variant 5 is the engine's test and must give none; planted:
//   6: T without its + 1 (floor((thr - lo) / delta));
//   7: cB rounded down instead of up;
//   8: codes rounded up instead of down.
// Inputs: r, lc, beta and lw drawn over the bound's whole range (generated
// table distributions, log-uniform extremes, tiny and large |lw|), half of
// them with beta placed so that lw / r + beta lands on or just above an
// integer, where ln_a meets its lower end lc - r - lw.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../minhashcuda_amd/csrc/wmh_bound.h"

static uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static double unit(uint64_t *s) {  // (0, 1)
  *s += 0x9e3779b97f4a7c15ull;
  return ((mix(*s) >> 11) + 0.5) * 0x1p-53;
}

static float log_uniform(uint64_t *s, double lo_exp, double hi_exp) {
  return (float)exp2(lo_exp + (hi_exp - lo_exp) * unit(s));
}

// kernel.cu:103-105 in the engine's operation order
static float ln_a_of(float lw, float r, float beta, float lc) {
  const float t = floorf(lw / r + beta);
  const float ln_y = (t - beta) * r;
  return (lc - ln_y) - r;
}

int main(int argc, char **argv) {
  const int variant = argc > 1 ? atoi(argv[1]) : 0;
  const long long trials = argc > 2 ? atoll(argv[2]) : 20000000LL;
  long long violations = 0, near_low_end = 0;
#pragma omp parallel for reduction(+ : violations, near_low_end) schedule(static)
  for (long long i = 0; i < trials; i++) {
    uint64_t s = mix((uint64_t)i * 0x2545f4914f6cdd1dull + (uint64_t)variant);
    float r, lc, lw, beta;
    switch ((int)(unit(&s) * 4)) {
      case 0:  // generated-table distributions (Gamma(2,1), ln Gamma(2,1))
        r = (float)-log(unit(&s) * unit(&s));
        lc = (float)log(-log(unit(&s) * unit(&s)));
        break;
      case 1:  // r at both ends of its range
        r = log_uniform(&s, -60, 60);
        lc = (float)((unit(&s) - 0.5) * 40);
        break;
      case 2:  // large |lc|
        r = log_uniform(&s, -30, 30);
        lc = log_uniform(&s, -20, 60) * (unit(&s) < 0.5 ? -1.0f : 1.0f);
        break;
      default:  // tiny r, moderate lc
        r = log_uniform(&s, -60, -10);
        lc = (float)((unit(&s) - 0.5) * 10);
        break;
    }
    const double pick = unit(&s);
    if (pick < 0.5) lw = (float)(-103.3 + 192.0 * unit(&s));
    else if (pick < 0.9) lw = log_uniform(&s, -25, 7) * (unit(&s) < 0.5 ? -1.0f : 1.0f);
    else lw = 0.0f;
    if (unit(&s) < 0.5) {
      beta = (float)unit(&s);
      if (beta >= 1.0f) beta = 0.0f;
    } else {
      // d + beta on (or one ulp above) an integer: ln_a at its lower end
      const float d = lw / r;
      double b = ceil((double)d) - (double)d;
      if (unit(&s) < 0.5) b += fabs((double)d) * 0x1p-24 * unit(&s);
      b -= floor(b);
      beta = (float)b;
      if (beta >= 1.0f || beta < 0.0f) beta = 0.0f;
    }
    if (variant >= 5 && unit(&s) < 0.02) lw = INFINITY;  // ln_a = -inf
    const float ln_a = ln_a_of(lw, r, beta, lc);
    uint16_t x16;
    float lwb;
    switch (variant) {
      case 1:
        x16 = wmh_bound_entry(0.0f, lc);
        lwb = wmh_bound_lw(lw);
        break;
      case 2:
        x16 = wmh_bound_entry(r, lc);
        lwb = -wmh_bound_lw(lw);
        break;
      case 3: {
        // nearest-rounded fp16 of lc - r through the round-down routine:
        // round down x + half an fp16 ulp of x
        const float base = lc - r;
        const uint16_t lo = wmh_f16_round_down(base);
        const uint16_t hi = (uint16_t)(lo + ((lo & 0x8000u) ? -1 : 1));
        const float vlo = wmh_f16_to_f32(lo), vhi = wmh_f16_to_f32(hi);
        x16 = (fabsf(base - vhi) < fabsf(base - vlo)) ? hi : lo;
        lwb = lw;
        break;
      }
      case 4:
        x16 = wmh_f16_round_down(lc - r);
        lwb = lw;
        break;
      default:
        x16 = wmh_bound_entry(r, lc);
        lwb = wmh_bound_lw(lw);
        break;
    }
    if (variant < 5) {
      const float L = wmh_f16_to_f32(x16) - lwb;
      if (L > ln_a) violations++;
      if (ln_a - (lc - r - lw) <= 0x1p-10f * (fabsf(lc) + r + fabsf(lw))) near_low_end++;
      continue;
    }
    // byte codes over a table range [mn, mx] holding X (an entry of -inf
    // leaves the bound off)
    const float x = wmh_f16_to_f32(x16);
    if (!(x > -INFINITY) || !(ln_a == ln_a)) continue;
    const float mn = unit(&s) < 0.2 ? x : x - (float)(unit(&s) * 64.0);
    const float mx = unit(&s) < 0.2 ? x : x + (float)(unit(&s) * 64.0);
    float lo, delta;
    wmh_code_range(mn, mx, &lo, &delta);
    const float scale = 1.0f / delta;
    uint32_t code = wmh_bound_code(x16, lo, delta, scale);
    if (variant == 8) {
      const float q = ceilf((x - lo) * scale);
      code = q > 255.0f ? 255u : (uint32_t)q;
    }
    // thr: ln_a itself (a tie), the next float up, a random step above, or
    // the grid point of the codes at or above ln_a
    float thr = ln_a;
    const double how = unit(&s);
    if (how < 0.25) thr = nextafterf(ln_a, INFINITY);
    else if (how < 0.5) thr = ln_a + (float)(unit(&s) * 2.0) * delta;
    else if (how < 0.75 && ln_a > -INFINITY) thr = lo + ceilf((ln_a - lo) * scale) * delta;
    if (!(thr >= ln_a)) thr = ln_a;
    if (!(thr < FLT_MAX)) continue;  // running minima are below FLT_MAX
    const float t0 = 2049.0f - lo * scale;
    const uint32_t tp = wmh_code_threshold(thr, scale, variant == 6 ? t0 - 1.0f : t0);
    uint32_t b = wmh_code_lw(lwb, scale);
    if (variant == 7 && lwb > -INFINITY && lwb < INFINITY) {
      const float q = floorf(lwb * scale);
      b = (uint32_t)(0x4000 - (int32_t)q);
    }
    if (wmh_code_skips(code, tp, b) && !(ln_a > thr)) violations++;
    if (ln_a - (lc - r - lw) <= 0x1p-10f * (fabsf(lc) + r + fabsf(lw))) near_low_end++;
  }
  printf("variant %d trials %lld violations %lld near_low_end %lld\n", variant, trials,
         violations, near_low_end);
  return (variant == 0 || variant == 5) && violations != 0 ? 1 : 0;
}
