/*
 * fitness.c -- oracle restatement of the reference fitness and the
 * project's float sine.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Fitness follows common/evolvables/SoftBody.cpp:
 *   updateLength 98-110  (extremes start at the zero vector; non-air masses)
 *   updateCOM    112-122 (mean += ((p - mean) * 1.0f) / (i + 1), in float)
 *   updateFitness 124-137 (max((COM.x - base.x) / L, 0))
 * Pinned by tests/golden/fitness_reference.bin (compiled reference
 * SoftBody::updateBaseline / Update / fitness) and hand-computed 2-3 mass
 * cases.
 *
 * orc_sinf is the project's sine spec, written independently of the HIP
 * copy: k = rint(x * 2/pi); Cody-Waite reduction of x - k*pi/2 with a
 * three-float split of pi/2 (Cody & Waite, "Software Manual for the
 * Elementary Functions", 1980) evaluated with fused multiply-adds; then the
 * Taylor polynomials of sin (to r^9) or cos (to r^10) on |r| <= pi/4,
 * chosen and signed by k mod 4.  Every step is an IEEE single operation,
 * so any conforming implementation of this sequence gives the same bits.
 * Pinned by: |orc_sinf(x) - sin(x)| <= 2e-7 over a dense sweep of
 * [-300, 300] plus exact special points (tests/test_oracle_math.py).
 */
#include <math.h>
#include "oracle.h"

void orc_com(const float* pos, const uint8_t* is_air, int n, float* com3) {
  float mx = 0.0f, my = 0.0f, mz = 0.0f;
  float i = 0;
  for (int k = 0; k < n; k++) {
    if (is_air[k]) continue;
    float d = i + 1;
    mx = mx + ((pos[3 * k] - mx) * 1.0f) / d;
    my = my + ((pos[3 * k + 1] - my) * 1.0f) / d;
    mz = mz + ((pos[3 * k + 2] - mz) * 1.0f) / d;
    i++;
  }
  com3[0] = mx; com3[1] = my; com3[2] = mz;
}

float orc_length(const float* pos, const uint8_t* is_air, int n) {
  float cx = 0.0f, fx = 0.0f;
  for (int k = 0; k < n; k++) {
    if (is_air[k]) continue;
    float x = pos[3 * k];
    if (x < cx) cx = x;
    if (x > fx) fx = x;
  }
  float l = fabsf(fx - cx);
  return l > 1.0f ? l : 1.0f;
}

float orc_fitness(float com_x, float base_com_x, float length) {
  float f = (com_x - base_com_x) / length;
  return f > 0.0f ? f : 0.0f;
}

float orc_sinf(float x) {
  const float two_over_pi = 0x1.45f306p-1f;
  const float p1 = 0x1.921fb6p+0f, p2 = -0x1.777a5cp-25f,
              p3 = -0x1.ee59dap-50f;
  const float s1 = -0x1.555556p-3f, s2 = 0x1.111112p-7f,
              s3 = -0x1.a01a02p-13f, s4 = 0x1.71de3ap-19f;
  const float c1 = -0x1.0p-1f, c2 = 0x1.555556p-5f, c3 = -0x1.6c16c2p-10f,
              c4 = 0x1.a01a02p-16f, c5 = -0x1.27e4fcp-22f;
  float prod = x * two_over_pi;
  float k = rintf(prod);
  float r = fmaf(-k, p1, x);
  r = fmaf(-k, p2, r);
  r = fmaf(-k, p3, r);
  int q = ((int)k) & 3;
  float r2 = r * r;
  float sp = fmaf(r2, s4, s3);
  sp = fmaf(r2, sp, s2);
  sp = fmaf(r2, sp, s1);
  float r3 = r * r2;
  float s = fmaf(r3, sp, r);
  float cp = fmaf(r2, c5, c4);
  cp = fmaf(r2, cp, c3);
  cp = fmaf(r2, cp, c2);
  cp = fmaf(r2, cp, c1);
  float c = fmaf(r2, cp, 1.0f);
  float v = (q & 1) ? c : s;
  return (q & 2) ? -v : v;
}
