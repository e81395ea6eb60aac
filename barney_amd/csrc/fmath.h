// Sine, cosine and x^5 with one fixed sequence of IEEE float operations, for
// the shading step's sampling (cosine-hemisphere, concentric-disk, GGX and
// phase-function directions) and the Schlick weight.
//
// The kernels and the CPU oracle both evaluate these instead of their
// platform's libm, so a path's arithmetic rounds alike on both and its
// random stream -- re-seeded from the bits of each hit distance
// (barney/kernels/shadeRays.cu:610) -- stays the same. Only correctly
// rounded operations are used (+, *, fmaf, rintf), in code built with
// floating-point contraction off on both sides.
//
// sin/cos: Cody-Waite reduction by pi/2 with a three-part constant (Cody and
// Waite, Software Manual for the Elementary Functions, 1980), then
//   sin r = r + r^3 P(r^2),  cos r = 1 + r^2 Q(r^2)  on [-pi/4, pi/4],
// P of degree 3 and Q of degree 4 fitted by weighted least squares on
// Chebyshev nodes (fmath_fit.py prints the coefficients below). Error: within
// 2 ulp of sin and cos for |x| < 2^16 (tests/test_oracle_fmath.py); NaN
// outside that domain and for non-finite x.
//
// pow5: (x^2)^2 * x, within 3 ulp of x^5 (powf(x, 5) in barney's
// schlick_weight, NVisii.h).
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define BNF_HD __host__ __device__ inline
#else
#define BNF_HD inline
#endif

namespace bnf {

// float32 constants written exactly (hexadecimal floating literals)
constexpr float kTwoOverPi = 0x1.45f306p-1f;
constexpr float kPiO2Hi = 0x1.921fb6p+0f;
constexpr float kPiO2Mid = -0x1.777a5cp-25f;
constexpr float kPiO2Lo = -0x1.ee59dap-50f;
constexpr float kS1 = -0x1.555556p-3f;
constexpr float kS2 = 0x1.111108p-7f;
constexpr float kS3 = -0x1.a00f6cp-13f;
constexpr float kS4 = 0x1.6cce0cp-19f;
constexpr float kC1 = -0x1.0p-1f;
constexpr float kC2 = 0x1.555556p-5f;
constexpr float kC3 = -0x1.6c16b4p-10f;
constexpr float kC4 = 0x1.a00eaap-16f;
constexpr float kC5 = -0x1.23c718p-22f;

BNF_HD void sincos(float x, float &s, float &c) {
  if (!(fabsf(x) < 65536.f)) {
    s = c = __builtin_nanf("");
    return;
  }
  const float k = rintf(x * kTwoOverPi);
  float r = fmaf(-k, kPiO2Hi, x);
  r = fmaf(-k, kPiO2Mid, r);
  r = fmaf(-k, kPiO2Lo, r);
  const float u = r * r;
  const float p = fmaf(fmaf(fmaf(kS4, u, kS3), u, kS2), u, kS1);
  const float sr = fmaf(r * u, p, r);
  const float q = fmaf(fmaf(fmaf(fmaf(kC5, u, kC4), u, kC3), u, kC2), u, kC1);
  const float cr = fmaf(u, q, 1.f);
  // x = k pi/2 + r: the quadrant k mod 4 (two's complement for k < 0)
  switch ((int)k & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
  }
}

BNF_HD float sin(float x) {
  float s, c;
  sincos(x, s, c);
  return s;
}

BNF_HD float cos(float x) {
  float s, c;
  sincos(x, s, c);
  return c;
}

BNF_HD float pow5(float x) {
  const float x2 = x * x;
  return (x2 * x2) * x;
}

}  // namespace bnf
