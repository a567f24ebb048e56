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

// |(x, y, z)| as sqrtf((x*x + y*y) + z*z), correctly rounded sqrt.
__device__ __forceinline__ float norm3(float x, float y, float z) {
  float s = x * x + y * y;
  s = s + z * z;
  return sqrtf(s);
}

// Rest-length factor A * sin(omega * t + phi) of one material at time t.
__device__ __forceinline__ float actuation(float dL0, float omega, float phi, float t) {
  return dL0 * act_sin(fmaf(omega, t, phi));
}

}  // namespace mh
