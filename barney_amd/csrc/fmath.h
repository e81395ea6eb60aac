// Sine, cosine, arc cosine, two-argument arctangent and x^5 with one fixed
// sequence of IEEE float operations, for the shading step's sampling
// (cosine-hemisphere, concentric-disk, GGX and phase-function directions),
// the Schlick weight and the environment map's spherical coordinates.
//
// The kernels and the CPU oracle both evaluate these instead of their
// platform's libm, so a path's arithmetic rounds alike on both and its
// random stream -- re-seeded from the bits of each hit distance
// (barney/kernels/shadeRays.cu:610) -- stays the same. Only correctly
// rounded operations are used (+, -, *, /, sqrtf, fmaf, rintf), in code
// built with floating-point contraction off on both sides.
//
// sin/cos: Cody-Waite reduction by pi/2 with a three-part constant (Cody and
// Waite, Software Manual for the Elementary Functions, 1980), then
//   sin r = r + r^3 P(r^2),  cos r = 1 + r^2 Q(r^2)  on [-pi/4, pi/4],
// P of degree 3 and Q of degree 4 fitted by weighted least squares on
// Chebyshev nodes (fmath_fit.py prints the coefficients below). Error: within
// 2 ulp of sin and cos for |x| < 2^16 (tests/test_oracle_fmath.py); NaN
// outside that domain and for non-finite x.
//
// atan2: |y|/|x| or |x|/|y| reduced to [0, 1], then to [-tan(pi/8),
// tan(pi/8)] by atan a = pi/4 + atan((a - 1)/(a + 1)) above tan(pi/8);
// atan t = t + t^3 A(t^2), A of degree 4; quadrants by pi/2 -+ r, pi - r and
// the sign of y, with atan2's signed-zero cases (atan2(+-0, -0) = +-pi).
// acos: pi/2 - asin x for |x| <= 1/2, else 2 asin(sqrt((1 - |x|)/2))
// (reflected for x < 0); asin t = t + t^3 B(t^2), B of degree 4. Both
// within 3 ulp of float64 (tests/test_oracle_fmath.py); NaN for NaN input
// and for |x| > 1 in acos.
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
constexpr float kA1 = -0x1.555548p-2f;
constexpr float kA2 = 0x1.999156p-3f;
constexpr float kA3 = -0x1.23b1p-3f;
constexpr float kA4 = 0x1.b1b53ap-4f;
constexpr float kA5 = -0x1.f0eebcp-5f;
constexpr float kB1 = 0x1.5555c8p-3f;
constexpr float kB2 = 0x1.33027ap-4f;
constexpr float kB3 = 0x1.746dc6p-5f;
constexpr float kB4 = 0x1.8cd42ep-6f;
constexpr float kB5 = 0x1.58d7b2p-5f;
constexpr float kTanPiO8 = 0x1.a8279ap-2f;
// pi/4 and pi in two parts: halving / doubling pi/2's parts is exact
constexpr float kPiO4Hi = 0x1.921fb6p-1f;
constexpr float kPiO4Mid = -0x1.777a5cp-26f;
constexpr float kPiHi = 0x1.921fb6p+1f;
constexpr float kPiMid = -0x1.777a5cp-24f;

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

// asin of t in [-1/2, 1/2] (odd in t)
BNF_HD float asin_core(float t) {
  const float u = t * t;
  const float p = fmaf(fmaf(fmaf(fmaf(kB5, u, kB4), u, kB3), u, kB2), u, kB1);
  return fmaf(t * u, p, t);
}

BNF_HD float acos(float x) {
  const float ax = fabsf(x);
  if (!(ax <= 1.f)) return __builtin_nanf("");
  if (ax <= 0.5f) return (kPiO2Mid - asin_core(x)) + kPiO2Hi;
  // 1 - ax is exact for ax in [1/2, 1], and so is the halving
  const float r = 2.f * asin_core(sqrtf((1.f - ax) * 0.5f));
  return x > 0.f ? r : (kPiMid - r) + kPiHi;
}

BNF_HD float atan2(float y, float x) {
  if (x != x || y != y) return __builtin_nanf("");
  const float ax = fabsf(x), ay = fabsf(y);
  const bool swap = ay > ax;
  const float num = swap ? ax : ay, den = swap ? ay : ax;
  float r;
  if (den == 0.f) {
    r = 0.f;
  } else if (!(den < __builtin_inff())) {
    r = (num < __builtin_inff()) ? 0.f : kPiO4Hi;
  } else {
    const float a = num / den;
    const bool big = a > kTanPiO8;
    const float t = big ? (a - 1.f) / (a + 1.f) : a;
    const float u = t * t;
    const float p = fmaf(fmaf(fmaf(fmaf(kA5, u, kA4), u, kA3), u, kA2), u, kA1);
    r = fmaf(t * u, p, t);
    if (big) r = (r + kPiO4Mid) + kPiO4Hi;
  }
  // small parts first: r = 0 leaves exactly the float nearest pi/2 or pi
  const bool neg = __builtin_signbit(x);
  if (swap)
    r = neg ? (kPiO2Mid + r) + kPiO2Hi : (kPiO2Mid - r) + kPiO2Hi;
  else if (neg)
    r = (kPiMid - r) + kPiHi;
  return __builtin_copysignf(r, y);
}

BNF_HD float pow5(float x) {
  const float x2 = x * x;
  return (x2 * x2) * x;
}

}  // namespace bnf
