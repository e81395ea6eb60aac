// barney-amd: shared device/host definitions for the gfx950 ray-tracing core.
//
// Small self-contained vector math (barney uses OWL's vec3f/box3f/affine3f,
// which is not available here), barney's RNG and hash (restated from
// barney/common/barney-common.h:65-152), and the record layouts the kernels
// exchange through HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define BN_HD __host__ __device__ __forceinline__
#define BN_D __device__ __forceinline__

namespace bnamd {

// barney/common/math.h:12-18
#define BN_PI ((float)M_PI)
#define BN_TWO_PI (2.f * BN_PI)
#define BN_FOUR_PI (4.f * BN_PI)
#define BN_ONE_OVER_PI (1.f / BN_PI)
#define BN_ONE_OVER_TWO_PI (1.f / BN_TWO_PI)
#define BN_ONE_OVER_FOUR_PI (1.f / BN_FOUR_PI)
#define BN_INF __builtin_huge_valf()

struct vec3f {
  float x, y, z;
  BN_HD vec3f() = default;
  BN_HD constexpr vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
  BN_HD explicit constexpr vec3f(float a) : x(a), y(a), z(a) {}
  BN_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct vec3i {
  int x, y, z;
};
struct vec2i {
  int x, y;
};

BN_HD vec3f operator+(vec3f a, vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
BN_HD vec3f operator-(vec3f a, vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
BN_HD vec3f operator-(vec3f a) { return {-a.x, -a.y, -a.z}; }
BN_HD vec3f operator*(vec3f a, vec3f b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
BN_HD vec3f operator*(float s, vec3f a) { return {s * a.x, s * a.y, s * a.z}; }
BN_HD vec3f operator*(vec3f a, float s) { return {s * a.x, s * a.y, s * a.z}; }
BN_HD vec3f operator/(vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
BN_HD vec3f &operator+=(vec3f &a, vec3f b) { a = a + b; return a; }
BN_HD vec3f &operator*=(vec3f &a, vec3f b) { a = a * b; return a; }
BN_HD vec3f &operator*=(vec3f &a, float s) { a = a * s; return a; }
BN_HD bool operator==(vec3f a, vec3f b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
BN_HD float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BN_HD vec3f cross(vec3f a, vec3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
BN_HD float length(vec3f a) { return sqrtf(dot(a, a)); }
BN_HD vec3f normalize(vec3f a) { return a * (1.f / sqrtf(dot(a, a))); }
BN_HD vec3f vmin(vec3f a, vec3f b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
BN_HD vec3f vmax(vec3f a, vec3f b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
BN_HD vec3f vmin(vec3f a, float b) { return {fminf(a.x, b), fminf(a.y, b), fminf(a.z, b)}; }
BN_HD vec3f vabs(vec3f a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
BN_HD float reduce_max(vec3f a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
BN_HD float reduce_min(vec3f a) { return fminf(a.x, fminf(a.y, a.z)); }
BN_HD vec3f rcp(vec3f a) { return {1.f / a.x, 1.f / a.y, 1.f / a.z}; }
BN_HD vec3f load3(const float *p) { return {p[0], p[1], p[2]}; }

// Affine transform as barney's affine3f: linear part columns vx,vy,vz + p.
struct affine3f {
  vec3f vx, vy, vz, p;
};
BN_HD vec3f xfmVector(const affine3f &a, vec3f v) { return v.x * a.vx + v.y * a.vy + v.z * a.vz; }
BN_HD vec3f xfmPoint(const affine3f &a, vec3f v) { return xfmVector(a, v) + a.p; }

// Orthonormal frame around N, following the published embree/ospray
// 'frame(N)' construction that OWL re-exports (used by
// barney/packedBSDF/Lambertian.h:81).
struct frame3f {
  vec3f vx, vy, vz;
};
BN_HD frame3f make_frame(vec3f N) {
  const vec3f dx0 = {0.f, N.z, -N.y};
  const vec3f dx1 = {-N.z, 0.f, N.x};
  const vec3f dx = normalize(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  const vec3f dy = normalize(cross(N, dx));
  return {dx, dy, N};
}
BN_HD vec3f xfmVector(const frame3f &f, vec3f v) { return v.x * f.vx + v.y * f.vy + v.z * f.vz; }

BN_HD uint32_t float_as_uint(float f) {
  union { float f; uint32_t u; } c;
  c.f = f;
  return c.u;
}
BN_HD float uint_as_float(uint32_t u) {
  union { float f; uint32_t u; } c;
  c.u = u;
  return c.f;
}

// ---------------------------------------------------------------------------
// RNG: barney's RNGSeed / Random2 (barney/common/barney-common.h:65-111).
// A 64-bit LCG whose output is a PCG-style xorshift + variable shift.
// ---------------------------------------------------------------------------
struct RNGSeed {
  uint64_t state = 0xcafef00dd15ea5e5ull;
  BN_HD uint32_t next() {
    uint64_t x = state;
    const unsigned count = (unsigned)(x >> 61);
    state = x * 6364136223846793005ull;
    x ^= x >> 22;
    return (uint32_t)(x >> (22 + count));
  }
  BN_HD uint32_t next_hash(uint64_t h) {
    state ^= (h << 1);
    return next();
  }
  BN_HD void seed(uint32_t a, uint32_t b) {
    state ^= 2ull * a + 1ull;
    next();
    state ^= 2ull * b;
    next();
  }
};

struct Random2 : RNGSeed {
  BN_HD Random2() = default;
  BN_HD Random2(uint32_t a, uint32_t b) { seed(a, b); }
  // Random2(RNGSeed&, b): copies the parent's state, advances the parent,
  // then mixes in b (290374 when b == 0).
  BN_HD Random2(RNGSeed &parent, uint64_t b) {
    state = parent.state;
    parent.next();
    next_hash(b ? b : 290374ull);
  }
  BN_HD float operator()() { return (float)(next() & 0x00FFFFFFu) * (1.f / (float)0x01000000); }
};

// FNV-style hash (barney/common/barney-common.h:132-152)
BN_HD uint64_t bn_hash(uint32_t v) { return 0xcbf29ce484222325ull ^ (0x100000001b3ull * v); }
BN_HD uint64_t bn_hash(uint64_t h, uint32_t v) { return h * 0x100000001b3ull ^ v; }
BN_HD uint64_t bn_hash3(uint32_t a, uint32_t b, uint32_t c) { return bn_hash(bn_hash(bn_hash(a), b), c); }

// Kernels that give a pixel L lanes run its paths in rounds: in round r lane
// 'sub' of the pixel's group traces path r*L + sub and holds v = (rgb sum,
// alpha). This adds the round's first 'count' values to acc in path order
// (lane base+0, base+1, ...), the order in which a sequential loop over the
// pixel's paths adds them, so the float sums round alike. Every lane of the
// wavefront must call it (it shuffles); each ends with its group's acc.
template <int L>
BN_D void add_round_in_path_order(float4 &acc, float4 v, int count) {
  const int base = (int)(threadIdx.x & 63) & ~(L - 1);
#pragma unroll
  for (int j = 0; j < L; ++j) {
    const float x = __shfl(v.x, base + j), y = __shfl(v.y, base + j);
    const float z = __shfl(v.z, base + j), w = __shfl(v.w, base + j);
    if (j < count) {
      acc.x += x;
      acc.y += y;
      acc.z += z;
      acc.w += w;
    }
  }
}

}  // namespace bnamd
