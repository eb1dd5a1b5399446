// Scalar FP32 math shared bit-for-bit by the gfx950 kernels and the host code.
//
// Everything here is built from IEEE-754 binary32 operations that round to
// nearest-even (add, sub, mul, div, floor, explicit fma) and integer bit
// manipulation.  With FP contraction disabled and denormals preserved, the
// same source evaluates to the same bits on the host (SSE) and on CDNA4
// (VALU), which is what makes the device hashes comparable to a CPU
// restatement bit for bit.
//
// The logarithm is this project's own: table-driven range reduction
// (x = 2^e * m, m = c * (1 + r)) followed by a degree-6 log1p series carried in
// float-float (pairs hi + lo) and one final rounding.  Its accuracy is pinned
// by an exhaustive comparison over every binary32 input against a 64-bit
// mantissa reference logarithm (see the tests).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WMH_HD __host__ __device__ __forceinline__
#define WMH_TABLE_QUALIFIER static __constant__ __device__
#else
#define WMH_HD static inline
#define WMH_TABLE_QUALIFIER static
#endif

#if defined(__clang__)
#define WMH_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define WMH_NO_CONTRACT
#endif

#include "wmh_log_table.h"
#if !defined(WMH_LOGF_NO_EXCEPTIONS)
#include "wmh_log_exceptions.h"
#endif

WMH_HD uint32_t wmh_f2u(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, sizeof(u));
  return u;
}
WMH_HD float wmh_u2f(uint32_t u) {
  float x;
  __builtin_memcpy(&x, &u, sizeof(x));
  return x;
}

// Error-free transformation: s + e == a + b exactly (Knuth's 2Sum, 6 flops,
// no precondition on the magnitudes).
WMH_HD void wmh_two_sum(float a, float b, float *s, float *e) {
  WMH_NO_CONTRACT
  float sum = a + b;
  float bv = sum - a;
  float av = sum - bv;
  float eb = b - bv;
  float ea = a - av;
  *s = sum;
  *e = ea + eb;
}

// ln(2) = LN2_HI + LN2_LO with LN2_HI = RN(ln 2); |error| < 2^-49.
#define WMH_LN2_HI 0x1.62e430p-1f
#define WMH_LN2_LO -0x1.05c610p-29f

// Natural logarithm of a binary32 value.
//   x == +-0      -> -inf
//   x <  0 or NaN -> canonical quiet NaN (0x7fc00000)
//   x == +inf     -> +inf
//   subnormals are handled exactly (scaled by 2^23 first).
WMH_HD float wmh_logf(float x) {
  WMH_NO_CONTRACT
  const float x0 = x;
  uint32_t ix = wmh_f2u(x);
  if (ix - 0x00800000u >= 0x7f000000u) {
    // zero, subnormal, negative, inf or NaN
    if ((ix & 0x7fffffffu) == 0) return wmh_u2f(0xff800000u);  // -inf
    if (ix == 0x7f800000u) return x;                            // +inf
    if (ix >= 0x7f800000u) return wmh_u2f(0x7fc00000u);         // NaN / negative
    // positive subnormal: make it normal
    x = x * 0x1p23f;
    ix = wmh_f2u(x) - (23u << 23);
  }
  // x = 2^e * m with m in [0.75, 1.5)
  uint32_t tmp = ix - 0x3f400000u;
  int32_t e = (int32_t)tmp >> 23;
  uint32_t i = (tmp >> 14) & 511u;
  float m = wmh_u2f(ix - (tmp & 0xff800000u));
  float c = wmh_log_table[i][0];
  float inv_c = wmh_log_table[i][1];
  float logc_hi = wmh_log_table[i][2];
  float logc_lo = wmh_log_table[i][3];

  // r = (m - c) / c as r_hi + r_lo; m - c is exact (Sterbenz).
  float d = m - c;
  float r_hi = d * inv_c;
  float r_lo = __builtin_fmaf(-r_hi, c, d) * inv_c;

  // log1p(r) = r - r^2/2 + r^3/3 - r^4/4 + r^5/5 - r^6/6, |r| <= 2^-9.
  float s_hi = r_hi * r_hi;                           // r_hi^2 exactly as
  float s_lo = __builtin_fmaf(r_hi, r_hi, -s_hi);     // s_hi + s_lo
  float c_hi = s_hi * r_hi;                           // r_hi^3 as c_hi + c_lo
  float c_lo = __builtin_fmaf(s_hi, r_hi, -c_hi) + s_lo * r_hi;
  float t3_hi = c_hi * 0x1.555556p-2f;                // (c_hi + c_lo) / 3
  float t3_lo = (__builtin_fmaf(-t3_hi, 3.0f, c_hi) + c_lo) * 0x1.555556p-2f;
  float r4 = s_hi * s_hi;
  float tail = r4 * (-0.25f + r_hi * (0x1.99999ap-3f - r_hi * 0x1.555556p-3f));

  float ef = (float)e;
  float p_hi = ef * WMH_LN2_HI;                       // e * LN2_HI exactly as
  float p_lo = __builtin_fmaf(ef, WMH_LN2_HI, -p_hi); // p_hi + p_lo

  // small terms: everything below the leading part of the result
  float small = tail + t3_lo;
  small = small + (s_hi * r_lo - r_hi * r_lo);
  small = small + (r_lo - 0.5f * s_lo);
  small = small + (logc_lo + ef * WMH_LN2_LO);
  small = small + p_lo;

  // leading terms, summed without rounding error
  float s, err, acc = 0.0f;
  wmh_two_sum(p_hi, logc_hi, &s, &err);
  acc = acc + err;
  wmh_two_sum(s, r_hi, &s, &err);
  acc = acc + err;
  wmh_two_sum(s, -0.5f * s_hi, &s, &err);
  acc = acc + err;
  wmh_two_sum(s, t3_hi, &s, &err);
  acc = acc + err;
  float z, res;
  wmh_two_sum(s, acc + small, &z, &res);
#if !defined(WMH_LOGF_NO_EXCEPTIONS)
  // Rounding test: the float-float value above carries a relative error of
  // about 2^-46, so only results whose exact residual |res| lies within
  // 2^-16 ulp of a half-ulp tie can be misrounded.  Those few inputs consult
  // the exception list of correctly rounded results (found exhaustively).
  uint32_t ez = (wmh_f2u(z) >> 23) & 0xffu;           // |z| >= 2^-24: ez >= 103
  float half = wmh_u2f((ez - 24u) << 23);
  float tol = wmh_u2f((ez - 40u) << 23);
  float dres = __builtin_fabsf(res);
  int near_tie = __builtin_fabsf(dres - half) <= tol;
  if ((wmh_f2u(z) & 0x7fffffu) == 0u)                 // power of two: the tie
    near_tie = near_tie || __builtin_fabsf(dres - 0.5f * half) <= tol;  // below is closer
  if (near_tie) {
    uint32_t key = wmh_f2u(x0);
    int lo = 0, hi = WMH_LOG_EXCEPTIONS - 1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      uint32_t k = wmh_log_exceptions[mid][0];
      if (k == key) return wmh_u2f(wmh_log_exceptions[mid][1]);
      if (k < key) lo = mid + 1; else hi = mid - 1;
    }
  }
#endif
  return z;
}

// (uint32)(int32)t with the conversion made total: NaN -> 0, values beyond the
// int32 range saturate.  t is always integral here (it comes from floorf).
WMH_HD uint32_t wmh_t_bits(float t) {
  if (!(t == t)) return 0u;
  if (t >= 2147483648.0f) return 0x7fffffffu;
  if (t < -2147483648.0f) return 0x80000000u;
  return (uint32_t)(int32_t)t;
}

// The empty / unselectable row sentinel written for both halves of a hash.
#define WMH_SENTINEL 0xffffffffu
