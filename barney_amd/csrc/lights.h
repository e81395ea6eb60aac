// Quad, directional and point lights and their samplers:
//   QuadLight::DD     barney/light/QuadLight.h:15-30, getDD QuadLight.cu:9-18
//   DirLight::DD      light/DirLight.h, getDD DirLight.cu:10-20
//   PointLight::DD    light/PointLight.h:12-25 (radianceTowards), getDD
//                     PointLight.cu:9-17
//   sampleAreaLights  kernels/shadeRays.cu:46-139 (8-candidate reservoir)
//   sampleDirLights   shadeRays.cu:141-193 (2 candidates)
//   samplePointLights shadeRays.cu:195-265 (2 candidates, 1/d^2 falloff)
// The float operations are barney's, in its order (no contraction), so the
// random streams and the sampled radiance match the oracle bit for bit.
#pragma once
#include "common.h"

namespace bnamd {

// QuadLight::DD. barney's getDD computes the area but leaves 'normal'
// unset; here it is normalize(edge0 x edge1). The samplers weigh a point by
// -dot(toLight, normal), so the quad lights the side its normal points to.
struct QuadLightDev {
  vec3f corner, edge0, edge1;
  vec3f emission;
  vec3f normal;
  float area;
};

// DirLight::DD with color * radiance folded (the samplers use only the
// product): direction = normalize(direction), the way the light travels.
struct DirLightDev {
  vec3f direction;
  vec3f radiance;
};

// PointLight::DD with radianceTowards folded: color * power / (4 pi) when
// intensity is NaN, else color * intensity.
struct PointLightDev {
  vec3f position;
  vec3f radiance;
};

struct LightsDD {
  const QuadLightDev *quad;
  const DirLightDev *dir;
  const PointLightDev *point;
  int numQuad, numDir, numPoint;
  int pad;
};

BN_HD bool has_light_lists(const LightsDD &l) { return (l.numQuad | l.numDir | l.numPoint) != 0; }

// Light::Sample (light/Light.h): pdf 0 marks "no sample".
struct LightSample {
  vec3f dir;
  vec3f rad;
  float dist;
  float pdf;
};

// a reservoir draw of a light index: min(int(random() * n), n - 1)
template <class RNG>
BN_HD int pick_light(RNG &random, int n) {
#pragma clang fp contract(off)
  const int i = (int)(random() * (float)n);
  return i < n - 1 ? i : n - 1;
}

// the scan that picks candidate i with probability w[i] / sum: r runs down
// the weights until it falls below one (kR when it never does)
template <int kR>
BN_HD int reservoir_pick(const float (&w)[kR], float r) {
#pragma clang fp contract(off)
  int sel = kR;
#pragma unroll
  for (int i = 0; i < kR; ++i) {
    if (sel == kR) {
      if (r >= w[i])
        r -= w[i];
      else
        sel = i;
    }
  }
  return sel;
}

// sampleAreaLights (shadeRays.cu:46-139). N is zero for volume samples.
// Each candidate takes exactly three draws (index, u, v), so the chosen
// one's index and position are drawn again from a copy of the generator
// instead of being kept for all eight candidates.
template <class RNG>
BN_HD bool sample_quad_lights(const LightsDD &L, vec3f P, vec3f N, RNG &random, LightSample &ls) {
#pragma clang fp contract(off)
  constexpr int kR = 8;
  if (L.numQuad == 0) return false;
  const RNG start = random;
  float w[kR];
  float sum = 0.f;
  const bool noN = (N == vec3f(0.f));
#pragma unroll
  for (int i = 0; i < kR; ++i) {
    const QuadLightDev &q = L.quad[pick_light(random, L.numQuad)];
    w[i] = 0.f;
    const float u = random();
    const float v = random();
    const vec3f LP = q.corner + u * q.edge0 + v * q.edge1;
    vec3f d = LP - P;
    const float dist = length(d);
    if (dist < 1e-3f) continue;
    d = d * (1.f / dist);
    float weight = noN ? 1.f : dot(d, N);
    if (weight <= 1e-3f) continue;
    weight *= -dot(d, q.normal);
    if (weight <= 1e-3f) continue;
    weight *= (1.f / (dist * dist)) * q.area * reduce_max(q.emission);
    sum += weight;
    w[i] = weight;
  }
  if (sum == 0.f) return false;
  const int sel = reservoir_pick<kR>(w, random() * sum);
  if (sel == kR) return false;
  float sw = 0.f;
#pragma unroll
  for (int i = 0; i < kR; ++i)
    if (i == sel) sw = w[i];
  RNG again = start;
  for (int k = 0; k < 3 * sel; ++k) (void)again();
  const QuadLightDev &q = L.quad[pick_light(again, L.numQuad)];
  const float su = again();
  const float sv = again();
  const vec3f LD = (q.corner + su * q.edge0 + sv * q.edge1) - P;
  ls.dir = normalize(LD);
  ls.dist = length(LD);
  ls.rad = q.emission * (q.area * -dot(q.normal, ls.dir) / (ls.dist * ls.dist));
  ls.pdf = sw / sum * ((float)kR / (float)L.numQuad);
  return true;
}

// sampleDirLights (shadeRays.cu:141-193): pdf = inf, the reservoir's
// probability divided into the radiance.
template <class RNG>
BN_HD bool sample_dir_lights(const LightsDD &L, vec3f N, RNG &random, LightSample &ls) {
#pragma clang fp contract(off)
  constexpr int kR = 2;
  if (L.numDir == 0) return false;
  int id[kR];
  float w[kR];
  float sum = 0.f;
  const bool noN = (N == vec3f(0.f));
#pragma unroll
  for (int i = 0; i < kR; ++i) {
    id[i] = pick_light(random, L.numDir);
    w[i] = 0.f;
    const DirLightDev &d = L.dir[id[i]];
    float weight = noN ? 1.f : dot(-d.direction, N);
    if (weight <= 1e-3f) continue;
    weight *= reduce_max(d.radiance);
    if (weight <= 1e-3f) continue;
    w[i] = weight;
    sum += weight;
  }
  if (sum == 0.f) return false;
  const int sel = reservoir_pick<kR>(w, random() * sum);
  if (sel == kR) return false;
  const int sid = sel == 0 ? id[0] : id[1];
  const float sw = sel == 0 ? w[0] : w[1];
  const DirLightDev &d = L.dir[sid];
  ls.dir = -d.direction;
  ls.dist = BN_INF;
  ls.rad = d.radiance * 1.f / (sw / sum * ((float)kR / (float)L.numDir));
  ls.pdf = BN_INF;
  return sw != 0.f;
}

// samplePointLights (shadeRays.cu:195-265): falloff 0 + 0 d + 1 d^2, the
// shadow ray stops at 0.9999 of the distance.
template <class RNG>
BN_HD bool sample_point_lights(const LightsDD &L, vec3f P, vec3f N, RNG &random, LightSample &ls) {
#pragma clang fp contract(off)
  constexpr int kR = 2;
  if (L.numPoint == 0) return false;
  int id[kR];
  float w[kR];
  float sum = 0.f;
  const bool noN = (N == vec3f(0.f));
#pragma unroll
  for (int i = 0; i < kR; ++i) {
    id[i] = pick_light(random, L.numPoint);
    w[i] = 0.f;
    const PointLightDev &p = L.point[id[i]];
    vec3f d = p.position - P;
    const float dist = length(d);
    const float falloff = dist * dist;
    d = normalize(d);
    const vec3f rad = p.radiance * (1.f / falloff);
    float weight = noN ? 1.f : dot(d, N);
    if (weight <= 1e-3f) continue;
    weight *= reduce_max(rad);
    if (weight <= 1e-3f) continue;
    w[i] = weight;
    sum += weight;
  }
  if (sum == 0.f) return false;
  const int sel = reservoir_pick<kR>(w, random() * sum);
  if (sel == kR) return false;
  const int sid = sel == 0 ? id[0] : id[1];
  const float sw = sel == 0 ? w[0] : w[1];
  const PointLightDev &p = L.point[sid];
  vec3f d = p.position - P;
  const float dist = length(d);
  ls.dir = normalize(d);
  ls.dist = dist * .9999f;
  ls.rad = p.radiance * (1.f / (dist * dist)) * 1.f / (sw / sum * ((float)kR / (float)L.numPoint));
  ls.pdf = BN_INF;
  return sw != 0.f;
}

}  // namespace bnamd

// The light-list handle of the C ABI (lights.hip creates it).
struct bnamd_lights_s {
  bnamd::LightsDD dd;
};
