// Test infrastructure (oracle/): exhaustive accuracy check of wmh_logf.
//
// For every non-negative binary32 input (0 .. +inf, 2^31 - 2^23 + 1 values)
// compares wmh_logf(x) with ln(x) evaluated by glibc's logl on the x87 80-bit
// format (64-bit mantissa) and reports:
//   * how many results equal the correctly rounded binary32 value,
//   * how many differ (and by how many ulps at most),
//   * the maximum error in ulps measured against the logl value,
//   * "hard" inputs whose logl value is so close to a binary32 rounding
//     boundary (within 2^-40 ulp) that its own rounding cannot be trusted.
// Usage: logf_check [first_bits last_bits]   (default: the whole range)
// Prints one JSON object.  Compiled with -ffp-contract=off, no fast-math.
//
//        logf_check --emit-exceptions
// (built with -DWMH_LOGF_NO_EXCEPTIONS) prints wmh_log_exceptions.h: every
// input the float-float evaluation misrounds, with its correctly rounded
// result taken from the 80-bit logl value.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../minhashcuda_amd/csrc/wmh_math.h"

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char **argv) {
  uint32_t first = 0, last = 0x7f800000u;
  int emit = argc == 2 && strcmp(argv[1], "--emit-exceptions") == 0;
  static uint32_t exc[4096][2];
  int n_exc = 0;
  if (argc == 3) {
    first = (uint32_t)strtoul(argv[1], NULL, 0);
    last = (uint32_t)strtoul(argv[2], NULL, 0);
  }
  long long n = 0, cr = 0, wrong = 0, hard = 0, specials_bad = 0;
  long double max_ulp = 0;
  uint32_t worst = 0;
  uint32_t first_wrong[16];
  int n_first_wrong = 0;
#pragma omp parallel reduction(+:n,cr,wrong,hard,specials_bad)
  {
  long double t_max = 0;
  uint32_t t_worst = 0;
#pragma omp for schedule(dynamic, 1)
  for (long long blk = first >> 16; blk <= (long long)(last >> 16); blk++) {
    uint32_t lo = (uint32_t)blk << 16, hi = lo + 0xffffu;
    if (lo < first) lo = first;
    if (hi > last) hi = last;
    for (uint64_t b = lo; b <= hi; b++) {
      float x = u2f((uint32_t)b);
      float y = wmh_logf(x);
      n++;
      if (b == 0) { if (!(isinf(y) && y < 0)) specials_bad++; continue; }
      if (b == 0x7f800000u) { if (!(isinf(y) && y > 0)) specials_bad++; continue; }
      if (b == 0x3f800000u) { if (f2u(y) != 0) specials_bad++; else cr++; continue; }
      long double L = logl((long double)x);
      float yr = (float)L;  // correctly rounded from the 80-bit value
      // distance of L to the nearest binary32 rounding boundary, in ulps of yr
      float up = nextafterf(yr, INFINITY), dn = nextafterf(yr, -INFINITY);
      long double ulp = ((long double)up - (long double)dn) / 2;
      long double mid_up = ((long double)yr + (long double)up) / 2;
      long double mid_dn = ((long double)yr + (long double)dn) / 2;
      long double dist = fminl(fabsl(L - mid_up), fabsl(L - mid_dn)) / ulp;
      if (dist < 0x1p-40L) hard++;
      int ex;
      frexpl(L, &ex);
      long double err = fabsl((long double)y - L) / ldexpl(1.0L, ex - 24);
      if (y == yr) {
        cr++;
      } else {
        wrong++;
#pragma omp critical
        {
          if (n_first_wrong < 16) first_wrong[n_first_wrong++] = (uint32_t)b;
          if (n_exc < 4096) { exc[n_exc][0] = (uint32_t)b; exc[n_exc][1] = f2u(yr); n_exc++; }
        }
      }
      if (err > t_max) { t_max = err; t_worst = (uint32_t)b; }
    }
  }
#pragma omp critical
  {
    if (t_max > max_ulp) { max_ulp = t_max; worst = t_worst; }
  }
  }
  if (emit) {
    // sort by input bits (the device code binary-searches the list)
    for (int a = 1; a < n_exc; a++)
      for (int c = a; c > 0 && exc[c - 1][0] > exc[c][0]; c--) {
        uint32_t t0 = exc[c][0], t1 = exc[c][1];
        exc[c][0] = exc[c - 1][0]; exc[c][1] = exc[c - 1][1];
        exc[c - 1][0] = t0; exc[c - 1][1] = t1;
      }
    printf("// Generated by oracle/logf_check --emit-exceptions -- do not edit by hand.\n");
    printf("// Inputs (binary32 bits) whose float-float logarithm lands within 2^-22 ulp\n");
    printf("// of a rounding tie on the wrong side, with the correctly rounded result.\n");
    printf("#pragma once\n#define WMH_LOG_EXCEPTIONS %d\n", n_exc);
    printf("WMH_TABLE_QUALIFIER const uint32_t wmh_log_exceptions[%d][2] = {\n", n_exc);
    for (int a = 0; a < n_exc; a++) printf("  {0x%08xu, 0x%08xu},\n", exc[a][0], exc[a][1]);
    printf("};\n");
    return 0;
  }
  printf("{\"first\": %u, \"last\": %u, \"n\": %lld, \"correctly_rounded\": %lld, "
         "\"misrounded\": %lld, \"hard\": %lld, \"specials_bad\": %lld, "
         "\"max_ulp\": %.6Lf, \"worst_bits\": %u, \"first_misrounded\": [",
         first, last, n, cr, wrong, hard, specials_bad, max_ulp, worst);
  for (int i = 0; i < n_first_wrong; i++) printf("%s%u", i ? ", " : "", first_wrong[i]);
  printf("]}\n");
  return 0;
}
