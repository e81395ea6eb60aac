// 2D textures and the HDRI environment light.
//
//   tex2D            the embree backend's sampler (rtcore/embree/Texture.cpp:
//                    computeAddress :26-84, getTexel :166-214, nearest
//                    :232-240, bilinear :266-289), texels widened to float4
//                    when the texture is created
//   EnvMapLight::DD  barney/light/EnvMap.h:97-246 (sampleCDF, pdf,
//                    worldToPixel, uvToWorld, sample)
//   radianceFromEnv  barney/kernels/shadeRays.cu:437-456
//
// Spherical angles use the shared acos / atan2 and sines / cosines of
// fmath.h (the same float operations in the oracle).
#pragma once
#include "common.h"
#include "fmath.h"

namespace bnamd {

// rtc::AddressMode and rtc::FilterMode values (rtcore/common/rtcore-common.h)
enum TexAddress : int { TEX_WRAP = 0, TEX_CLAMP = 1, TEX_BORDER = 2, TEX_MIRROR = 3 };
enum TexFilter : int { TEX_POINT = 0, TEX_LINEAR = 1 };

// A 2D texture on the device: float4 texels, row-major.
struct Tex2DDD {
  const float4 *texels;
  int w, h;
  int filter;
  int addr[2];
  int pad;
};

struct TexLerp {
  int i0, i1;
  float f;
};

// computeAddress (rtcore/embree/Texture.cpp:26-84) for one axis of N texels.
BN_HD TexLerp tex_address(int mode, float tc, int N) {
#pragma clang fp contract(off)
  TexLerp o{0, 0, 0.f};
  if (mode == TEX_WRAP) {
    tc = tc * (float)N - .5f;
    tc = fmodf(tc, (float)N);
    if (tc < 0.f) tc += (float)N;
    const int ic = (int)floorf(tc);
    o.f = tc - ic;
    // tc + N can round to N itself: that texel wraps to 0 (the embree
    // sampler reads one past the row there)
    o.i0 = ic >= N ? ic - N : ic;
    o.i1 = (int)fmodf(tc + 1.f, (float)N);
  } else if (mode == TEX_MIRROR) {
    tc = tc * N + .5f;
    const int ic = (int)floorf(tc);
    o.f = tc - ic;
    int i0 = ic - 1, i1 = ic;
    if (i0 < 0) i0 = -i0;
    if (i1 < 0) i1 = -i1;
    i0 %= 2 * N;
    i1 %= 2 * N;
    if (i0 >= N) i0 = 2 * N - 1 - i0;
    if (i1 >= N) i1 = 2 * N - 1 - i1;
    o.i0 = i0;
    o.i1 = i1;
  } else if (mode == TEX_CLAMP) {
    tc = tc * N;
    if (tc - .5f <= 0.f) {
      o.i0 = o.i1 = 0;
    } else if (tc - .5f >= N - 1) {
      o.i0 = o.i1 = N - 1;
    } else {
      o.f = (tc - .5f) - (int)(tc - .5f);
      o.i0 = (int)(tc - .5f);
      o.i1 = o.i0 + 1;
    }
  }
  // border: the embree sampler leaves the record zeroed (texel 0, weight 0)
  return o;
}

BN_HD float4 lerp4(float f, float4 a, float4 b) {
#pragma clang fp contract(off)
  return make_float4((1.f - f) * a.x + f * b.x, (1.f - f) * a.y + f * b.y, (1.f - f) * a.z + f * b.z,
                     (1.f - f) * a.w + f * b.w);
}

// tex2D<float4> with normalized coordinates
BN_HD float4 tex2d(const Tex2DDD &t, float u, float v) {
#pragma clang fp contract(off)
  if (t.filter == TEX_POINT) {
    const int ix = (int)((uint32_t)(fabsf(u) * t.w) % (uint32_t)t.w);
    const int iy = (int)((uint32_t)(fabsf(v) * t.h) % (uint32_t)t.h);
    return t.texels[ix + iy * t.w];
  }
  const TexLerp lx = tex_address(t.addr[0], u, t.w), ly = tex_address(t.addr[1], v, t.h);
  const float4 v00 = t.texels[lx.i0 + ly.i0 * t.w], v01 = t.texels[lx.i1 + ly.i0 * t.w];
  const float4 v10 = t.texels[lx.i0 + ly.i1 * t.w], v11 = t.texels[lx.i1 + ly.i1 * t.w];
  return lerp4(ly.f, lerp4(lx.f, v00, v01), lerp4(lx.f, v10, v11));
}

// EnvMapLight::DD. texels == nullptr: no map (uniform ambient radiance).
struct EnvDD {
  Tex2DDD tex;
  vec3f toWorld[3];  // columns
  vec3f toLocal[3];
  float scale;
  const float *cdfY;  // [h]
  const float *cdfX;  // [h][w]
};

// linear3f * vector (columns)
BN_HD vec3f xfm_lin(const vec3f *m, vec3f v) {
#pragma clang fp contract(off)
  return v.x * m[0] + v.y * m[1] + v.z * m[2];
}

// render/DG.h:97-106
BN_HD float spherical_theta(vec3f v) { return bnf::acos(fmaxf(fminf(v.z, 1.f), -1.f)); }
BN_HD float spherical_phi(vec3f v) {
#pragma clang fp contract(off)
  const float p = bnf::atan2(v.y, v.x);
  return (p < 0.f) ? (p + float(2.f * M_PI)) : p;
}

// cdfGetPDF / sampleCDF (EnvMap.h:97-133)
BN_HD float cdf_pdf(int pos, const float *cdf, int N) {
#pragma clang fp contract(off)
  return N * (cdf[pos] - (pos ? cdf[pos - 1] : 0.f));
}
BN_HD int sample_cdf(const float *cdf, int N, float v, float &pdf) {
  int begin = 0, end = N;
  while (end - begin > 1) {
    const int mid = (begin + end) / 2;
    if (v < cdf[mid - 1])
      end = mid;
    else
      begin = mid;
  }
  pdf = cdf_pdf(begin, cdf, N);
  return begin;
}

// EnvMapLight::DD::pdf with a texture (EnvMap.h:135-159, worldToPixel :161-177)
BN_HD float env_pdf(const EnvDD &e, vec3f dir) {
#pragma clang fp contract(off)
  const vec3f l = xfm_lin(e.toLocal, dir);
  const float u = spherical_phi(l) * BN_ONE_OVER_TWO_PI, v = spherical_theta(l) * BN_ONE_OVER_PI;
  const int ix = (int)fminf(fmaxf(u * (float)e.tex.w, 0.f), e.tex.w - 1.f);
  const int iy = (int)fminf(fmaxf(v * (float)e.tex.h, 0.f), e.tex.h - 1.f);
  const float py = cdf_pdf(iy, e.cdfY, e.tex.h);
  const float px = cdf_pdf(ix, e.cdfX + (size_t)iy * e.tex.w, e.tex.w);
  const float theta = BN_PI * ((iy + .5f) / e.tex.h);
  return px * py * 1.f / (BN_TWO_PI * BN_PI * bnf::sin(theta));
}

// EnvMapLight::DD::sample (EnvMap.h:208-246; uvToWorld :195-206)
template <class RNG>
BN_HD void env_sample(const EnvDD &e, RNG &r, vec3f &dir, vec3f &rad, float &pdf) {
#pragma clang fp contract(off)
  const float ry = r();
  const float rx = r();
  float py, px;
  const int iy = sample_cdf(e.cdfY, e.tex.h, ry, py);
  const int ix = sample_cdf(e.cdfX + (size_t)e.tex.w * iy, e.tex.w, rx, px);
  const float sx = (ix + r()) / e.tex.w;
  const float sy = (iy + r()) / e.tex.h;
  const float4 c = tex2d(e.tex, sx, sy);
  rad = e.scale * vec3f(c.x, c.y, c.z);
  float sp, cp, st, ct;
  bnf::sincos(BN_TWO_PI * sx, sp, cp);
  bnf::sincos(BN_PI * sy, st, ct);
  dir = xfm_lin(e.toWorld, vec3f(cp * st, sp * st, ct));
  pdf = px * py * 1.f / (BN_TWO_PI * BN_PI * st);
}

// radianceFromEnv with a texture (shadeRays.cu:437-456)
BN_HD vec3f env_radiance(const EnvDD &e, vec3f rayDir) {
#pragma clang fp contract(off)
  const vec3f d = xfm_lin(e.toLocal, normalize(rayDir));
  const float theta = spherical_theta(d), phi = spherical_phi(d);
  const float4 c = tex2d(e.tex, phi * (1.f / (2.f * (float)M_PI)), theta * (1.f / (float)M_PI));
  return e.scale * vec3f(c.x, c.y, c.z);
}

}  // namespace bnamd

// The environment-light handle of the C ABI (envmap.hip creates it).
struct bnamd_envmap_s {
  bnamd::EnvDD dd;
};
