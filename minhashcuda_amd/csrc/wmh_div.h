// Correctly rounded binary32 division in 6 gfx950 instructions, for the
// operand range of the hashing path (a = ln w, b = r).  Shared by the
// kernels and the exhaustive checker in tests/native/fastdiv_check.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace wmh {

// Range of b for which div_rn_fast is exact (see below).
constexpr float kFastDivMinB = 0x1p-60f;
constexpr float kFastDivMaxB = 0x1p60f;

// One Newton step from v_rcp_f32: RN(1/b) for every b in
// [kFastDivMinB, kFastDivMaxB] (checked exhaustively on gfx950 over every
// float in [2^-60, 2^61): 1,015,021,568 values).
__device__ __forceinline__ float recip_rn(float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
}

// a / b rounded to nearest even, in 6 instructions instead of the 11 of the
// general IEEE sequence (which also scales operands near the ends of the
// exponent range and fixes up zeros, infinities and NaNs).  Exact for every
// b in [kFastDivMinB, kFastDivMaxB] and every a = +0 or finite a with |a| in
// [2^-25, 2^7], which covers a = ln(w) for every positive finite float w:
//   y = recip_rn(b) = RN(1/b);
//   q0 = RN(a * y) is within 1 ulp of a / b, so rem = a - b * q0 is exact
//   (no underflow in this range) and RN(q0 + rem * y) = RN(a / b)
//   (P. Markstein, IBM J. Res. Dev. 34(1), 1990; J.-M. Muller et al.,
//   Handbook of Floating-Point Arithmetic, on FMA-based division).
// Also checked exhaustively on gfx950 against IEEE division for all 2^46
// significand pairs at the exponent pairs (a, b) in ([1, 2), [1, 2)),
// ([2^-20, 2^-19), [2^30, 2^31)), ([64, 128), [2^-60, 2^-59)) and
// ([2^-25, 2^-24), [2^59, 2^60)).  For a = -inf or NaN the IEEE quotient
// (-inf or NaN) and this one (NaN) both give a hash candidate that is never
// selected; a = +inf must take the IEEE path (+inf here is NaN).
__device__ __forceinline__ float div_rn_fast(float a, float b) {
  const float y = recip_rn(b);
  const float q0 = a * y;
  return __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
}

}  // namespace wmh
