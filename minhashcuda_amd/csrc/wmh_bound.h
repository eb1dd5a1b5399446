// A lower bound on a nonzero's hash value, shared bit for bit by the gfx950
// kernels and the host (table build, CPU checks).
//
// The hashing formula (kernel.cu:103-105, evaluated in binary32 with round to
// nearest, unfused):
//   d = lw / r;  t = floor(d + beta);  y = (t - beta) * r;  ln_a = (lc - y) - r
// In exact arithmetic t - beta lies in (lw/r - 1, lw/r], so y lies in
// (lw - r, lw] and ln_a in [lc - r - lw, lc - lw).  The rounded evaluation
// stays above lc - r - lw - 2^-24 (6|lw| + 10r + 2|lc|) (plus at most 6 * 2^-150
// from results in the subnormal range) for r in [2^-60, 2^60], beta in
// [0, 1), |lc| <= 2^60 and |lw| <= 2^7 (every logarithm of a positive finite
// float).  The bound below is lower than that by a wide margin:
//   X   = fp16 value of (lc - r) - 2^-16 (|lc| + r), rounded toward -inf
//         (one 2-byte entry per (column, sample): the bound table);
//   lwb = lw + 2^-16 |lw|                  (per nonzero);
//   L   = X - lwb <= lc - r - lw - (2^-16 - 3.2 * 2^-24)(|lc| + r + |lw|).
// So a computed ln_a is never below L, and a nonzero whose L exceeds the
// running minimum of a sample (strictly: ties keep the earlier position)
// cannot be selected for it: its exact evaluation can be skipped without
// changing any hash.  Non-finite cases: lw = -inf (w = 0) and NaN (w < 0 or
// NaN) give ln_a = +inf or NaN, never selected, and lwb = -inf (always
// skipped); lw = +inf gives ln_a = -inf and lwb = +inf (never skipped).
#pragma once

#include "wmh_math.h"

// fp16 bit pattern of x rounded toward -inf (finite x below -65504 gives
// -inf, x above 65504 gives 65504; NaN gives -inf, the bound that never
// skips).  Integer arithmetic only, so host and device agree.
WMH_HD uint16_t wmh_f16_round_down(float x) {
  const uint32_t bits = wmh_f2u(x);
  const uint32_t sign = bits >> 31;
  const uint32_t exp = (bits >> 23) & 0xffu;
  const uint32_t mant = bits & 0x7fffffu;
  if (exp == 0xffu) {
    if (mant != 0u) return 0xfc00u;                 // NaN: -inf
    return sign ? 0xfc00u : 0x7c00u;                // +-inf
  }
  uint32_t h = 0u;
  uint32_t inexact = 0u;
  const int e = (int)exp - 127;
  if (e > 15) {                                     // |x| >= 65536
    h = 0x7c00u;                                    // inf (only reached as the
    inexact = 1u;                                   // rounded-up magnitude)
    if (!sign) return 0x7bffu;                      // largest finite: round down
  } else if (exp != 0u && e >= -14) {               // normal half
    h = ((uint32_t)(e + 15) << 10) | (mant >> 13);
    inexact = (mant & 0x1fffu) != 0u;
  } else if (exp != 0u && e >= -25) {               // subnormal half (m * 2^-24)
    const uint32_t sig = mant | 0x800000u;
    const uint32_t shift = (uint32_t)(-14 - e) + 13u;  // 14 .. 24
    h = sig >> shift;
    inexact = (sig & ((1u << shift) - 1u)) != 0u;
  } else {                                          // |x| < 2^-25 or zero
    h = 0u;
    inexact = (exp | mant) != 0u;
  }
  if (!sign) return (uint16_t)h;                    // truncation rounds down
  if (inexact && h < 0x7c00u) h += 1u;              // magnitude up (may reach inf)
  return (uint16_t)(0x8000u | h);
}

// The float value of fp16 bits (exact).
WMH_HD float wmh_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h >> 15) << 31;
  const uint32_t exp = (h >> 10) & 0x1fu;
  const uint32_t mant = h & 0x3ffu;
  if (exp == 0x1fu) return wmh_u2f(sign | 0x7f800000u | (mant << 13));
  if (exp != 0u) return wmh_u2f(sign | ((exp + 112u) << 23) | (mant << 13));
  // subnormal: mant * 2^-24, exact in binary32
  const float v = (float)mant * 0x1p-24f;
  return sign ? -v : v;
}

// Bound-table entry of one (column, sample) with parameters r and lc.
WMH_HD uint16_t wmh_bound_entry(float r, float lc) {
  WMH_NO_CONTRACT
  const float base = lc - r;
  const float margin = ((lc < 0.0f ? -lc : lc) + r) * 0x1p-16f;
  return wmh_f16_round_down(base - margin);
}

// The per-nonzero term lwb of the bound for a log-weight lw.
WMH_HD float wmh_bound_lw(float lw) {
  WMH_NO_CONTRACT
  if (!(lw == lw) || lw == -__builtin_inff()) return -__builtin_inff();
  if (lw == __builtin_inff()) return lw;
  return lw + (lw < 0.0f ? -lw : lw) * 0x1p-16f;
}

// Range of the parameter tables within which the bound holds (checked on the
// device whenever the tables change; generated tables always pass).
#define WMH_BOUND_MAX_ABS_LC 0x1p60f

// ---------------------------------------------------------------------------
// Byte codes of the bound table.
//
// The engine keeps entry X (the fp16 value above) of every (column, sample)
// as one byte c = min(255, floor((X - lo) / delta)), with delta = 2^e >= 1/16
// and lo a multiple of delta at or below every X of the table, so that
// lo + c delta <= X.  For a sample whose running minimum is thr and a nonzero
// whose term is lwb, (nonzero, sample) is skipped when c >= T + cB, where
//   T  = floor((thr - lo) / delta) + 1     (so lo + T delta > thr),
//   cB = ceil(lwb / delta)                 (so cB delta >= lwb):
// then X >= lo + (T + cB) delta > thr + lwb, and the computed ln_a, above
// X - lwb, is above thr -- strictly, so ties keep the earlier position.
// Clamps, each toward evaluating more: T to [-2048, 2303] and cB to
// [-2048, 2304] (lwb = -inf: -2048, never selectable; lwb = +inf: 2304,
// never skipped).  A finite lwb is below 104 in magnitude (the logarithms of
// positive finite floats), so |cB| <= 1664 never reaches the clamps, and
// T = 2303 skips nothing such a nonzero brings (T + cB >= 639 > 255).  With
// T' = T + 2048 in [0, 4351] the 16-bit sum
//   H = c + (0x4800 - T') + (0x4000 - cB)
// lies in [0x6E01, 0x90FF], and its bit 15 is set exactly when c >= T + cB:
// the kernel tests two samples per 32-bit add, one per 16-bit half.
#define WMH_CODE_T_MAX 4351u

// Byte code of an entry (lo, delta = 1 / inv_delta as above).  The rounded
// subtraction may land on the next integer; the exact check lo + c delta <= X
// (a multiple of delta below 2^21 delta: exact) takes it back.
WMH_HD uint32_t wmh_bound_code(uint16_t x16, float lo, float delta, float inv_delta) {
  WMH_NO_CONTRACT
  const float x = wmh_f16_to_f32(x16);
  float q = __builtin_floorf((x - lo) * inv_delta);
  q = q > 0.0f ? q : 0.0f;
  q = q < 255.0f ? q : 255.0f;
  uint32_t c = (uint32_t)q;
  if (c > 0u && lo + (float)c * delta > x) c--;
  return c;
}

// T' = T + 2048 of a running minimum thr, with scale = 1 / delta and
// t0 = 2049 - lo / delta (both exact): the floor of the correctly rounded
// fma is never below the exact T' (rounding does not cross the integer
// below), then clamped to [0, 4351] (NaN: 0).
WMH_HD uint32_t wmh_code_threshold(float thr, float scale, float t0) {
  float q = __builtin_fmaf(thr, scale, t0);
  q = q > 0.0f ? q : 0.0f;
  q = q < 4351.0f ? q : 4351.0f;
  return (uint32_t)__builtin_floorf(q);
}

// 0x4000 - cB of a nonzero's term lwb (scale = 1 / delta <= 16: lwb * scale
// is an exact power-of-two scaling).
WMH_HD uint32_t wmh_code_lw(float lwb, float scale) {
  WMH_NO_CONTRACT
  float q = __builtin_ceilf(lwb * scale);
  q = q > -2048.0f ? q : -2048.0f;  // -inf (NaN is never produced): -2048
  q = q < 2304.0f ? q : 2304.0f;
  return (uint32_t)(0x4000 - (int32_t)q);
}

// 1 when (code c, threshold T', nonzero term) is skipped.
WMH_HD uint32_t wmh_code_skips(uint32_t c, uint32_t tprime, uint32_t lw_term) {
  return ((c + (0x4800u - tprime) + lw_term) >> 15) & 1u;
}

// Order-preserving unsigned key of a non-NaN float, and back.
WMH_HD uint32_t wmh_order_bits(float x) {
  const uint32_t u = wmh_f2u(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
WMH_HD float wmh_order_value(uint32_t key) {
  return wmh_u2f((key & 0x80000000u) ? (key & 0x7fffffffu) : ~key);
}

// delta and lo of a table whose entries lie in [mn, mx] (finite): delta the
// smallest power of two from 1/16 up whose 255 steps cover 90% of the range
// (entries above lo + 255 delta take code 255), lo = mn rounded down to a
// multiple of delta; |lo| / delta <= 2^20 (|mn| <= 65504).
WMH_HD void wmh_code_range(float mn, float mx, float *lo, float *delta) {
  WMH_NO_CONTRACT
  float d = 0.0625f;
  while (255.0f * d < 0.9f * (mx - mn) && d < 65536.0f) d *= 2.0f;
  *delta = d;
  *lo = __builtin_floorf(mn / d) * d;
}
