// mh_device.h -- device-side arithmetic shared by the step kernels.
//
// Everything here is a fixed sequence of IEEE single-precision operations
// (the library is compiled with -ffp-contract=off; fused multiply-adds only
// where written as fmaf), so results are reproducible bit for bit by any
// conforming implementation of the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

// Sine used for the rest-length actuation.  k = rint(x * 2/pi) picks the
// quadrant; x - k*pi/2 is reduced with pi/2 split into three floats
// (Cody & Waite 1980) using fused multiply-adds; the reduced argument
// |r| <= pi/4 goes through the Taylor polynomial of sin (to r^9) or cos (to
// r^10), selected and signed by k mod 4.  Coefficients are 1/n! rounded to
// float.  Absolute error below 2e-7 for |x| <= 300.
__device__ __forceinline__ float act_sin(float x) {
  const float kInvHalfPi = 0x1.45f306p-1f;
  const float kHalfPiHi = 0x1.921fb6p+0f;
  const float kHalfPiMid = -0x1.777a5cp-25f;
  const float kHalfPiLo = -0x1.ee59dap-50f;
  const float k = rintf(x * kInvHalfPi);
  float r = fmaf(-k, kHalfPiHi, x);
  r = fmaf(-k, kHalfPiMid, r);
  r = fmaf(-k, kHalfPiLo, r);
  const int quadrant = static_cast<int>(k) & 3;
  const float r2 = r * r;
  // sin: r + r^3 (-1/3! + r^2 (1/5! + r^2 (-1/7! + r^2 / 9!)))
  float ps = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
  ps = fmaf(r2, ps, 0x1.111112p-7f);
  ps = fmaf(r2, ps, -0x1.555556p-3f);
  const float r3 = r * r2;
  const float sin_r = fmaf(r3, ps, r);
  // cos: 1 + r^2 (-1/2 + r^2 (1/4! + r^2 (-1/6! + r^2 (1/8! - r^2/10!))))
  float pc = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
  pc = fmaf(r2, pc, -0x1.6c16c2p-10f);
  pc = fmaf(r2, pc, 0x1.555556p-5f);
  pc = fmaf(r2, pc, -0x1.0p-1f);
  const float cos_r = fmaf(r2, pc, 1.0f);
  const float v = (quadrant & 1) ? cos_r : sin_r;
  return (quadrant & 2) ? -v : v;
}


// Correctly rounded sqrt for x in [2^-96, 2^128) without the range
// scaling and special-value handling of the general sequence: with the
// hardware reciprocal square root y, s = x y refined once by Newton's step
// in residual form, s + (x - s s) (y / 2), the residual computed exactly by
// a fused multiply-add (Markstein 1990, "Computation of elementary functions
// on the IBM RISC System/6000 processor").  On gfx950 this equals sqrtf for
// every float in the range (checked exhaustively by mh_check_arith); the
// general sequence in the error analysis is not guaranteed to round
// correctly, so the check is what admits it.
__device__ __forceinline__ float sqrt_rn_inrange(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float s = x * y;
  const float r = fmaf(-s, s, x);
  return fmaf(r, 0.5f * y, s);
}

// The previous short sequence, kept as a cross-check: the hardware square
// root (within 1 ulp) corrected to the nearest float by the signs of the
// residuals x - s*(s -+ 1 ulp), Tuckerman's rounding test.
__device__ __forceinline__ float sqrt_rn_tuckerman_inrange(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
  const float r_dn = fmaf(-s_dn, s, x);
  const float r_up = fmaf(-s_up, s, x);
  float r = (r_dn <= 0.f) ? s_dn : s;
  r = (r_up > 0.f) ? s_up : r;
  return r;
}
constexpr float kSqrtFastLo = 0x1p-96f;
constexpr float kSqrtFastHi = 0x1p+126f;

// Correctly rounded 1/y for y in [2^-64, 2^64]: the hardware reciprocal
// (within 1 ulp) refined by one Newton step with a fused residual,
// r + r (1 - y r).  Checked against 1.0f / y for every float in range by
// mh_check_arith.
__device__ __forceinline__ float rcp_rn_inrange(float y) {
  const float r = __builtin_amdgcn_rcpf(y);
  const float e = fmaf(-y, r, 1.0f);
  return fmaf(e, r, r);
}
constexpr float kRcpFastLo = 0x1p-64f;
constexpr float kRcpFastHi = 0x1p+64f;

// Sum of squares x*x + y*y + z*z as fmaf(z, z, fmaf(y, y, x * x)): three
// operations, two of them fused.
__device__ __forceinline__ float sum_sq3(float x, float y, float z) {
  return fmaf(z, z, fmaf(y, y, x * x));
}

// |(x, y, z)| as sqrtf(sum_sq3), correctly rounded sqrt.  When every lane
// of the wave has its sum of squares in the fast range (the usual case:
// spring lengths are far from 2^-48 and 2^63) the short sequence runs;
// otherwise the whole wave takes the general one.  The branch is
// wave-uniform, so no lane pays for both.
__device__ __forceinline__ float norm3(float x, float y, float z) {
  const float s = sum_sq3(x, y, z);
  // s is +0, positive or NaN: one unsigned compare of the bit pattern
  // tests kSqrtFastLo <= s <= kSqrtFastHi
  const bool fast = __float_as_uint(s) - __float_as_uint(kSqrtFastLo) <=
                    __float_as_uint(kSqrtFastHi) - __float_as_uint(kSqrtFastLo);
  if (__builtin_amdgcn_ballot_w64(fast) == __builtin_amdgcn_read_exec())
    return sqrt_rn_inrange(s);
  return sqrtf(s);
}

// Sums of squares from which a spring takes the short sequences below:
// s >= 2^-30 gives d = RN(sqrt(s)) >= 2^-15, whose half ulp is at least
// 2^-39, so adding any eps below 2^-39 (the step's 1e-12 is 0x1.19799ap-40)
// rounds back to d and the sum need not be formed.
constexpr float kLenFastLo = 0x1p-30f;

// d = |(x, y, z)| and inv = 1 / (d + eps) for 0 <= eps < 2^-39, both
// correctly rounded.  For a sum of squares in [2^-30, 2^128), d lies in
// [2^-15, 2^64), d + eps == d, and d is in the fast reciprocal's range, so
// one wave-uniform test covers both short sequences.  That test is the
// single compare s >= 2^-30: NaN fails it, and +inf passes it and gives
// d = inv = NaN where the general sequences give d = +inf, inv = 0; the
// force solver's only use of the pair, (k (L - d)) inv, is NaN either way.
__device__ __forceinline__ void length_and_inverse(float x, float y, float z, float eps,
                                                   float& d, float& inv) {
  const float s = sum_sq3(x, y, z);
  const bool fast = s >= kLenFastLo;
  if (__builtin_amdgcn_ballot_w64(fast) == __builtin_amdgcn_read_exec()) {
    d = sqrt_rn_inrange(s);
    inv = rcp_rn_inrange(d);
  } else {
    d = sqrtf(s);
    inv = 1.0f / (d + eps);
  }
}

// Correctly rounded a / b from y = RN(1/b): q0 = a y, then two fused
// corrections q + (a - b q) y.  The first leaves q faithful, so its
// remainder is exact and the second rounds correctly (Markstein's theorem
// on fused multiply-add division).  Valid when a is 0 or |a| >= 2^-79, b and
// a / b are normal and below 2^64, so no remainder falls below 2^-126.
__device__ __forceinline__ float div_rn_from_rcp(float a, float b, float y) {
  const float q0 = a * y;
  const float r0 = fmaf(-b, q0, a);
  const float q1 = fmaf(r0, y, q0);
  const float r1 = fmaf(-b, q1, a);
  return fmaf(r1, y, q1);
}

// Rest-length factor A * sin(omega * t + phi) of one material at time t.
__device__ __forceinline__ float actuation(float dL0, float omega, float phi, float t) {
  return dL0 * act_sin(fmaf(omega, t, phi));
}

}  // namespace mh
