// Camera rays and the shade/bounce step of barney's path tracer with the
// environment (ambient or HDRI) light, quad / directional / point lights and
// single scattering:
//   camera           barney/Camera.cu:81-100
//   ray generation   barney/kernels/generateRays.cu:73-129, 199-210
//   bounce()         barney/kernels/shadeRays.cu:476-878
//   env light        shadeRays.cu:267-293 (sampleEnvLight), 295-404 (sampleLights),
//                    437-470 (radianceFromEnv, primaryRayMissColor); the
//                    HDRI map in envmap.h; the light lists in lights.h
//   BSDFs            packedBSDF/Lambertian.h:55-87, Phase.h:120-172
//   materials        material/AnariMatte.h:51-63
//   accumulation     shadeRays.cu:970-1038
#pragma once
#include <hip/hip_fp16.h>
#include <cstring>
#include "envmap.h"
#include "fmath.h"
#include "lights.h"
#include "world.h"

namespace bnamd {

// Camera in the form PerspectiveCamera::commit produces (Camera.cu:81-100).
struct CameraDD {
  vec3f org;
  vec3f dir_00, dir_du, dir_dv;
};

inline vec3f host_normalize(vec3f v) { return v * (1.f / sqrtf(dot(v, v))); }

inline CameraDD make_camera(const bnamd_camera &c) {
  CameraDD d;
  d.org = {c.position[0], c.position[1], c.position[2]};
  // Camera::set3f("direction") normalizes and commit() normalizes again.
  vec3f dir00 = host_normalize(host_normalize({c.direction[0], c.direction[1], c.direction[2]}));
  const vec3f up = {c.up[0], c.up[1], c.up[2]};
  d.dir_du = host_normalize(cross(dir00, up));
  d.dir_dv = host_normalize(cross(d.dir_du, dir00));
  dir00 = dir00 * (float)(1.f / (2.0f * tanf((0.5f * c.fovy) * (float)M_PI / 180.0f)));
  d.dir_00 = dir00;
  return d;
}

struct RenderDD {
  int width, height, accumID;
  float ambient;
  float bg[4];
  int maxPathDepth, maxDiffuseBounces, rrDepth;
  int cutOn;        // cut plane enabled
  float4 cut;       // Renderer "cutPlane" (nx, ny, nz, d)
  EnvDD env;        // env.tex.texels == nullptr: uniform ambient radiance
  LightsDD lights;  // quad, directional and point lights (counts 0: none)
};

inline RenderDD make_render_dd(const bnamd_render_params &p) {
  RenderDD r;
  memset(&r, 0, sizeof r);  // padding too: graph cache keys compare these bytes
  r.width = p.width;
  r.height = p.height;
  r.accumID = p.accumID;
  r.ambient = p.ambientRadiance;
  for (int k = 0; k < 4; ++k) r.bg[k] = p.bgColor[k];
  r.maxPathDepth = p.maxPathDepth > 0 ? p.maxPathDepth : 10;
  r.maxDiffuseBounces = p.maxDiffuseBounces >= 0 ? p.maxDiffuseBounces : 1;
  r.rrDepth = p.russianRoulette;
  r.cut = make_float4(p.cutPlane[0], p.cutPlane[1], p.cutPlane[2], p.cutPlane[3]);
  // disabled when d <= -1e28 (Renderer.h:59); an all-zero plane rejects
  // nothing either (0 < 0 never holds), so it is skipped the same way
  r.cutOn = (p.cutPlane[3] > -1e28f) &&
            !(p.cutPlane[0] == 0.f && p.cutPlane[1] == 0.f && p.cutPlane[2] == 0.f && p.cutPlane[3] == 0.f);
  if (p.envMap)
    r.env = p.envMap->dd;
  else
    r.env = EnvDD{};
  r.lights = p.lights ? p.lights->dd : LightsDD{};
  return r;
}

BN_HD bool has_map(const RenderDD &rp) { return rp.env.tex.texels != nullptr; }
// Kernels are instantiated with kLights = has_lights(rp): an HDRI map or a
// light list, so that scenes lit by the uniform ambient radiance alone run
// code without the map's and the lists' samplers.
BN_HD bool has_lights(const RenderDD &rp) { return has_map(rp) || has_light_lists(rp.lights); }

// The candidate filter of a surface trace: the renderer's cut plane and
// (stochastic = true) the stochastic opacity of cutout triangles, drawn
// from the ray's seed at the start of the trace.
BN_HD HitFilter hit_filter(const RenderDD &rp, vec3f org, vec3f dir, uint64_t seed, bool stochastic = true) {
  HitFilter f;
  f.o = org;
  f.d = dir;
  f.cut = rp.cut;
  f.cutOn = rp.cutOn != 0;
  f.seed = seed;
  f.stochastic = stochastic;
  return f;
}

enum BsdfType : int { BSDF_NONE = 0, BSDF_LAMBERT = 1, BSDF_PHASE = 2, BSDF_NVISII = 3, BSDF_GLASS = 4 };

// Packed BSDF (barney/packedBSDF/PackedBSDF.h:20-138): Lambertian albedo,
// phase albedo, NVisii base color + parameters, or Glass ior.
struct Bsdf {
  int type;
  vec3f albedo;     // Lambert/Phase albedo, NVisii base color
  float metallic;   // NVisii
  float roughness;  // NVisii
  float ior;        // NVisii / Glass
};

// One ray of a path: Ray + PathState of barney/render/Ray.h:18-99 (the
// throughput / normal stay fp32 here; barney packs them as halfs).
struct PathRay {
  vec3f org, dir;
  float tMax;
  RNGSeed rng;
  vec3f tp;
  float mis;
  int numDiffuse;
  bool isShadow;
  bool isInMedium;
  vec3f missColor;
  float missAlpha;
};

// What the closest-hit programs leave in the ray (Ray::setHit).
struct ShadeHit {
  bool hit;
  vec3f P;
  vec3f N;  // zero for a volume scatter event
  float t;
  Bsdf bsdf;
};

struct Fragment {
  vec3f rgb;
};

// What a primary ray of row iy shows when it escapes (generateRays.cu:199-210):
// the background color, or (bg alpha < 0) the default vertical gradient.
BN_HD void miss_color(const RenderDD &rp, int iy, vec3f &color, float &alpha) {
#pragma clang fp contract(off)
  const float t = (iy + .5f) / float(rp.height);
  if (rp.bg[3] >= 0.f) {
    color = {rp.bg[0], rp.bg[1], rp.bg[2]};
    alpha = rp.bg[3];
  } else {
    color = (1.0f - t) * vec3f(0.9f, 0.9f, 0.9f) + t * vec3f(0.15f, 0.25f, .8f);
    alpha = (1.0f - t) * 1.f + t * 1.f;
  }
}

// generateRays.cu:73-129, 199-210 for a perspective camera without depth of
// field: pixel jitter 0.5 on the first frame, else from Random(ix+W*a, iy+H*a).
BN_HD PathRay camera_ray(const CameraDD &cam, const RenderDD &rp, int ix, int iy) {
#pragma clang fp contract(off)
  PathRay ray;
  const int a = rp.accumID;
  Random2 rand((unsigned)(ix + rp.width * a), (unsigned)(iy + rp.height * a));
  ray.rng = RNGSeed();
  ray.rng.seed((uint32_t)(ix + a * rp.width), (uint32_t)iy);
  const float pixel_u = (a == 0) ? .5f : rand();
  const float pixel_v = (a == 0) ? .5f : rand();
  const float image_u = ((ix + pixel_u) / float(rp.width));
  const float image_v = ((iy + pixel_v) / float(rp.height));
  const float aspect = rp.width / float(rp.height);
  const vec3f d = cam.dir_00 + (1.f * aspect * (image_u - .5f)) * cam.dir_du + (1.f * (image_v - .5f)) * cam.dir_dv;
  ray.org = cam.org;
  ray.dir = normalize(d);
  ray.tMax = 1e30f;
  if (rp.cutOn) {
    // the cutting plane (generateRays.cu:171-193): clip where the ray
    // leaves the visible half-space, or start it where it enters; a ray
    // that never enters it keeps tMax = 0 and shows the background
    const vec3f N = vec3f(rp.cut.x, rp.cut.y, rp.cut.z);
    const float distOrg = dot(N, ray.org) + rp.cut.w;
    const float denom = dot(N, ray.dir);
    if (distOrg >= 0.f) {
      if (denom < 0.f) ray.tMax = fminf(ray.tMax, -distOrg / denom);
    } else if (denom > 0.f) {
      ray.org = ray.org + (-distOrg / denom) * ray.dir;
    } else {
      ray.tMax = 0.f;
    }
  }
  ray.tp = vec3f(1.f);
  ray.mis = 0.f;
  ray.numDiffuse = 0;
  ray.isShadow = false;
  ray.isInMedium = false;
  miss_color(rp, iy, ray.missColor, ray.missAlpha);
  return ray;
}

// DG.h:184-194
BN_HD vec3f random_direction(Random2 &rng) {
#pragma clang fp contract(off)
  while (true) {
    vec3f v;
    v.x = 1.f - 2.f * rng();
    v.y = 1.f - 2.f * rng();
    v.z = 1.f - 2.f * rng();
    if (dot(v, v) <= 1.f) return normalize(v);
  }
}

struct EvalRes {
  vec3f value;
  float pdf;
};

BN_HD bool is_inf(float f) { return f == BN_INF || f == -BN_INF; }

// ---------------------------------------------------------------------------
// NVisii (Disney-style) BSDF, restating barney/packedBSDF/NVisii.h for the
// parameters AnariPBR::createBSDF sets (AnariPBR.h:68-94): anisotropy,
// sheen, clearcoat, flatness and specular transmission at their defaults
// (0), whose lobes then contribute exactly zero.
// ---------------------------------------------------------------------------
namespace nvisii {
#define BN_NV_MIN_ALPHA .002f
#define BN_NV_SMALL_EPS 2e-10f
BN_HD float sqr(float f) { return f * f; }
BN_HD float lerp_r(float a, float b, float t) { return (1.f - t) * a + t * b; }
BN_HD vec3f lerp_r(vec3f a, vec3f b, vec3f t) { return (vec3f(1.f) - t) * a + t * b; }
BN_HD float saturate(float f) { return fmaxf(0.f, fminf(f, 1.f)); }
BN_HD float schlick_weight(float c) { return bnf::pow5(saturate(1.f - c)); }
BN_HD bool same_hemisphere(vec3f wo, vec3f wi, vec3f n) { return dot(wo, n) * dot(wi, n) > 0.f; }
// NVisii.h:138-145 (no alpha^2 numerator, single power in the denominator)
BN_HD float gtr_2(float cos_theta_h, float alpha) {
  const float a2 = alpha * alpha;
  const float den1 = 1.f + (a2 - 1.f) * cos_theta_h * cos_theta_h;
  const float den2 = fmaxf(den1, BN_NV_SMALL_EPS);
  return (float)M_1_PI * 1.f / den2;
}
BN_HD float smith_ggx(float n_dot_o, float alpha_g) {
  const float a = alpha_g * alpha_g;
  const float b = n_dot_o * n_dot_o;
  return 1.f / (n_dot_o + sqrtf(a + b - a * b));
}
BN_HD float luminance(vec3f c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
BN_HD vec3f reflect(vec3f v, vec3f n) { return v - (2.f * dot(v, n)) * n; }
// render/DG.h:121-128: the scalar eta*d is added to the vector sqrtf(k)*n
// there (vec3f(eta*d) + sqrtf(k)*n); kept as written
BN_HD vec3f refract(vec3f v, vec3f n, float eta) {
#pragma clang fp contract(off)
  const float d = dot(n, v);
  const float k = 1.f - eta * eta * (1.f - d * d);
  if (!(k >= 0.f)) return vec3f(0.f);
  return eta * v - (vec3f(eta * d) + sqrtf(k) * n);
}
// sample_gtr_2_h (NVisii.h:186-194); sin_theta_h = 1 - cos^2 as written
BN_HD vec3f sample_gtr_2_h(vec3f n, vec3f v_x, vec3f v_y, float alpha, float ux, float uy) {
#pragma clang fp contract(off)
  const float phi_h = 2.f * (float)M_PI * ux;
  const float c2 = (1.f - uy) / (1.f + (alpha * alpha - 1.f) * uy);
  const float cos_h = sqrtf(c2);
  const float sin_h = 1.f - c2;
  float sphi, cphi;
  bnf::sincos(phi_h, sphi, cphi);
  const vec3f hd = normalize(vec3f(sin_h * cphi, sin_h * sphi, cos_h));
  return hd.x * v_x + hd.y * v_y + hd.z * n;
}

struct Mat {
  vec3f base_color;
  float metallic, specular, roughness, specular_tint, ior, transmission_roughness;
};

// NVisii::setDefaults (NVisii.h:1040-1060, stored as halfs) with the
// AnariPBR overrides; specular .5 and transmission roughness .04 as halfs
BN_HD Mat unpack(const Bsdf &b) {
  Mat m;
  m.base_color = b.albedo;
  m.metallic = b.metallic;
  m.specular = .5f;
  m.roughness = b.roughness;
  m.specular_tint = 0.f;
  m.ior = b.ior;
  m.transmission_roughness = 0.0400390625f;  // half(.04)
  return m;
}

// disney_brdf (NVisii.h:600-660) for the active lobes
BN_HD vec3f brdf(const Mat &mat, vec3f n, vec3f w_o, vec3f w_i, vec3f w_h) {
#pragma clang fp contract(off)
  const float n_dot_o = fabsf(dot(w_o, n));
  const float n_dot_i = fabsf(dot(w_i, n));
  const float i_dot_h = dot(w_i, w_h);
  // disney_diffuse (no 1/pi, as in NVisii.h:338-352)
  const float fd90 = 0.5f + 1.0f * mat.roughness * i_dot_h * i_dot_h;
  const float fi = schlick_weight(n_dot_i), fo = schlick_weight(n_dot_o);
  const float diffuse = lerp_r(1.f, fd90, fi) * lerp_r(1.f, fd90, fo);
  // disney_microfacet_isotropic (NVisii.h:430-455)
  const float lum = luminance(mat.base_color);
  const vec3f tint = lum > 0.f ? mat.base_color * (1.f / lum) : vec3f(1.f);
  const vec3f spec = lerp_r(mat.specular * 0.08f * lerp_r(vec3f(1.f), tint, vec3f(mat.specular_tint)),
                            mat.base_color, vec3f(mat.metallic));
  const float alpha = fmaxf(BN_NV_MIN_ALPHA, mat.roughness * mat.roughness);
  const float d = gtr_2(fabsf(dot(n, w_h)), alpha);
  const vec3f f = lerp_r(spec, vec3f(1.f),
                         vec3f(schlick_weight(fabsf(dot(w_i, w_h))) * lerp_r(.5f, 1.f, fmaxf(mat.metallic, alpha))));
  const float g = smith_ggx(fabsf(dot(n, w_i)), alpha) * smith_ggx(fabsf(dot(n, w_o)), alpha);
  const vec3f gloss = d * f * g;
  const vec3f flat = diffuse * mat.base_color;  // flatness 0
  return flat * (1.f - mat.metallic) * (1.f - 0.f) + gloss;
}

// disney_pdf (NVisii.h:700-790), IMPORTANCE_SAMPLE_BRDF weights (unnormalized)
BN_HD float pdf(const Mat &mat, vec3f n, vec3f w_o, vec3f w_i, vec3f w_h) {
#pragma clang fp contract(off)
  const float diffuse_weight = (1.f - mat.metallic) * (1.f - 0.f);
  const float glossy_weight = 1.f + mat.metallic + 0.f + mat.specular + mat.roughness;
  const float alpha = fmaxf(BN_NV_MIN_ALPHA, mat.roughness * mat.roughness);
  const float d_i = dot(w_i, n);
  const float diffuse = d_i > 0.f ? d_i * (float)M_1_PI : 0.f;
  float microfacet = 0.f;
  if (same_hemisphere(w_o, w_i, n)) {
    const float c = fabsf(dot(n, w_h));
    microfacet = gtr_2(c, alpha) * c / (4.f * fabsf(dot(w_o, w_h)));
  }
  return diffuse_weight * diffuse + glossy_weight * microfacet;
}

// concentric disk mapping (NVisii.h:62-82)
BN_HD vec3f cos_sample_hemisphere(float ux, float uy) {
#pragma clang fp contract(off)
  const float sx = 2.f * ux - 1.f, sy = 2.f * uy - 1.f;
  float radius = 0.f, theta = 0.f;
  if (!(sx == 0.f && sy == 0.f)) {
    if (fabsf(sx) > fabsf(sy)) {
      radius = sx;
      theta = (float)M_PI / 4.f * (sy / sx);
    } else {
      radius = sy;
      theta = (float)M_PI / 2.f - (float)M_PI / 4.f * (sx / sy);
    }
  }
  float st, ct;
  bnf::sincos(theta, st, ct);
  const float dx = radius * ct, dy = radius * st;
  return {dx, dy, sqrtf(fmaxf(0.f, 1.f - dx * dx - dy * dy))};
}
}  // namespace nvisii

// RobustDielectric::sample of barney/packedBSDF/Glass.h (ospray-derived)
struct GlassSample {
  vec3f wi;
  bool transmit;
};

BN_HD GlassSample glass_sample(float iorEta, vec3f wo, vec3f Ns, vec3f Ng, Random2 &random) {
#pragma clang fp contract(off)
  const float eta = 1.f / iorEta;
  const vec3f incoming = -wo;
  vec3f sn = Ns, gn = Ng;
  const float glass_epsilon = 1.e-6f;
  const bool canReflect = (eta < 1.f - glass_epsilon || eta > 1.f + glass_epsilon);
  if (dot(incoming, sn) > 0.f) sn = -sn;
  if (dot(incoming, gn) > 0.f) gn = -gn;
  // reflectionDirection
  const vec3f par = dot(incoming, sn) * sn;
  const vec3f dirR = (incoming - par) - par;
  // refractionDirection (the head-on branch can never trigger: its test
  // reads 1-1e-6 < c < 1-1e-6)
  vec3f dirT;
  float fresnelR;
  bool mustReflect = false;
  {
    const float cos1 = dot(incoming, sn);
    float cos1_abs = cos1 < 0.f ? -cos1 : cos1;
    const vec3f dpar = cos1 * sn;
    const vec3f dperp = incoming - dpar;
    const vec3f b = normalize(dperp);
    const float sin1 = dot(dperp, b);
    const float sin2 = sin1 / eta;
    if (sin2 > 1.f) {
      dirT = vec3f(0.f);
      fresnelR = 1.f;
      mustReflect = true;
    } else {
      const float cos2 = sqrtf(fminf(fmaxf(1.f - sin2 * sin2, 0.f), 1.f));
      const vec3f N2 = normalize(dpar);
      dirT = N2 * cos2 + b * sin2;
      const float c1 = cos1_abs, c2 = cos2;
      const float r1 = (eta * c1 - c2) / (eta * c1 + c2);
      const float r2 = (c1 - eta * c2) / (c1 + eta * c2);
      fresnelR = 0.5f * (r1 * r1 + r2 * r2);
    }
    fresnelR = fminf(fmaxf(fresnelR, 0.f), 1.f);
  }
  GlassSample res;
  if ((random() < fresnelR && canReflect) || mustReflect) {
    res.wi = dirR;
    res.transmit = false;
    const float gc = dot(res.wi, gn);
    if (gc >= 0.f) {
      if (gc < glass_epsilon) res.wi = normalize(res.wi + glass_epsilon * gn);
    } else {
      res.wi = res.wi - 2.f * gn * gc;
    }
  } else {
    res.wi = dirT;
    res.transmit = true;
    const float gc = dot(res.wi, gn);
    if (gc <= 0.f) {
      if (gc > -glass_epsilon) res.wi = normalize(res.wi - glass_epsilon * gn);
    } else {
      res.transmit = false;  // counted as a reflection
    }
  }
  return res;
}

// BSDF::eval (dg.Ns = dg.Ng = Ng, the unflipped geometric normal)
BN_HD EvalRes bsdf_eval(const Bsdf &b, vec3f Ns, vec3f wi, vec3f wo) {
#pragma clang fp contract(off)
  EvalRes r;
  if (b.type == BSDF_LAMBERT) {
    const float cosThetaI = fmaxf(dot(wi, Ns), 0.f);
    r.pdf = cosThetaI * BN_ONE_OVER_PI;
    r.value = b.albedo * BN_ONE_OVER_PI * cosThetaI;
  } else if (b.type == BSDF_NVISII) {
    const nvisii::Mat m = nvisii::unpack(b);
    const vec3f w_h = normalize(wi + wo);
    r.value = nvisii::brdf(m, Ns, wo, wi, w_h);
    r.pdf = nvisii::pdf(m, Ns, wo, wi, w_h);
  } else if (b.type == BSDF_GLASS) {
    r.value = vec3f(0.f);
    r.pdf = 0.f;
  } else {
    const float density = BN_ONE_OVER_FOUR_PI;
    r.value = density * b.albedo;
    r.pdf = density;
  }
  return r;
}

BN_HD float bsdf_pdf(const Bsdf &b, vec3f Ng, vec3f wi, vec3f wo) {
#pragma clang fp contract(off)
  if (b.type == BSDF_LAMBERT) {
    vec3f N = Ng;
    if (dot(wi, N) < 0.f) N = -N;
    const float cosThetaI = fmaxf(dot(wi, N), 0.f);
    return cosThetaI * BN_ONE_OVER_PI;
  }
  if (b.type == BSDF_NVISII) {
    const nvisii::Mat m = nvisii::unpack(b);
    return nvisii::pdf(m, Ng, wo, wi, normalize(wi + wo));
  }
  if (b.type == BSDF_GLASS) return 0.f;
  return BN_ONE_OVER_FOUR_PI;
}

enum ScatterType : int { SCATTER_DIFFUSE = 0, SCATTER_GLOSSY = 1, SCATTER_SPECULAR = 2, SCATTER_VOLUME = 3 };

struct ScatterRes {
  vec3f f_r, dir;
  float pdf;
  int type;
  float offsetDirection;
  bool changedMedium;
};

BN_HD ScatterRes bsdf_scatter(const Bsdf &b, vec3f Ng, vec3f wo, bool insideMedium, Random2 &rng) {
#pragma clang fp contract(off)
  ScatterRes s;
  s.offsetDirection = 1.f;
  s.changedMedium = false;
  if (b.type == BSDF_LAMBERT) {
    vec3f N = Ng;
    if (dot(wo, N) < 0.f) N = -N;
    const float sx = rng();
    const float sy = rng();
    // DG.h:130-136 cosineSampleHemisphere + DG.h:107-121 cartesian
    const float phi = BN_TWO_PI * sx;
    const float cosTheta = sqrtf(sy);
    const float sinTheta = sqrtf(1.0f - sy);
    float sinPhi, cosPhi;
    bnf::sincos(phi, sinPhi, cosPhi);
    const vec3f local = {cosPhi * sinTheta, sinPhi * sinTheta, cosTheta};
    s.dir = xfmVector(make_frame(N), local);
    s.pdf = local.z * BN_ONE_OVER_PI;
    s.f_r = b.albedo;
    s.type = SCATTER_DIFFUSE;
  } else if (b.type == BSDF_NVISII) {
    // NVisii::scatter + sample_disney_brdf (NVisii.h:1092-1163, 820-975)
    const nvisii::Mat m = nvisii::unpack(b);
    const vec3f n = Ng;
    const vec3f v_y = normalize(cross(Ng, wo));
    const vec3f v_x = normalize(cross(Ng, v_y));
    float dw = (1.f - m.metallic) * (1.f - 0.f);
    float gw = 1.f + m.metallic + 0.f + m.specular + m.roughness;
    const float sum = dw + gw;
    const float scale = 1.f / sum;
    dw *= scale;
    gw *= scale;
    // lobe selection as written (:853-870); with clear coat weight 0 the
    // clear-coat branch cannot be taken, and the transmission branch only
    // when the normalized weights round to a sum below type_rng
    const float type_rng = rng();
    const int lobe = type_rng < dw ? 0 : ((type_rng - dw) < gw ? 1 : 3);
    const float ux = rng();
    const float uy = rng();
    vec3f w_i;
    s.type = lobe == 0 ? SCATTER_DIFFUSE : SCATTER_GLOSSY;
    if (lobe == 0) {
      const vec3f h = normalize(nvisii::cos_sample_hemisphere(ux, uy));
      w_i = h.x * v_x + h.y * v_y + h.z * n;
    } else if (lobe == 1) {
      const float alpha = fmaxf(BN_NV_MIN_ALPHA, m.roughness * m.roughness);
      const vec3f w_h = nvisii::sample_gtr_2_h(n, v_x, v_y, alpha, ux, uy);
      w_i = nvisii::reflect(-wo, w_h);
      if (!nvisii::same_hemisphere(wo, w_i, n)) {
        s.pdf = 0.f;
        s.f_r = vec3f(0.f);
        s.dir = vec3f(0.f);
        return s;
      }
    } else {
      // DISNEY_TRANSMISSION_BRDF sampling (:955-973)
      const float alpha = fmaxf(BN_NV_MIN_ALPHA, m.transmission_roughness * m.transmission_roughness);
      const vec3f w_h = nvisii::sample_gtr_2_h(n, v_x, v_y, alpha, ux, uy);
      const bool entering = dot(wo, w_h) > 0.f;
      const float eta_i = entering ? 1.f : m.ior, eta_o = entering ? m.ior : 1.f;
      const vec3f hn = entering ? w_h : -w_h;
      w_i = nvisii::refract(-wo, hn, eta_i / eta_o);
      if (w_i.x == 0.f && w_i.y == 0.f && w_i.z == 0.f) {
        w_i = nvisii::reflect(-wo, hn);
        s.pdf = 1.f;
        s.f_r = vec3f(1.f);
        s.dir = normalize(w_i);
        return s;
      }
    }
    const vec3f w_h = normalize(w_i + wo);
    s.pdf = nvisii::pdf(m, n, wo, w_i, w_h);
    s.f_r = nvisii::brdf(m, n, wo, w_i, w_h);
    s.dir = normalize(w_i);
  } else if (b.type == BSDF_GLASS) {
    // Glass::scatter (Glass.h:219-241)
    const float eta = (!insideMedium) ? 1.f / b.ior : b.ior;
    const GlassSample g = glass_sample(eta, wo, Ng, Ng, rng);
    s.f_r = vec3f(1.f);
    s.dir = g.wi;
    s.pdf = BN_INF;
    s.offsetDirection = g.transmit ? -1.f : 1.f;
    s.changedMedium = g.transmit;
    s.type = SCATTER_SPECULAR;
  } else {
    const float r1 = rng();
    const float r2 = rng();
    float sphi, cphi;
    bnf::sincos(BN_TWO_PI * r1, sphi, cphi);
    const float x = cphi * sqrtf(r2 * (1.f - r2));
    const float y = sphi * sqrtf(r2 * (1.f - r2));
    const float z = (1.f - 2.f * r2);
    s.pdf = BN_ONE_OVER_FOUR_PI;
    s.f_r = b.albedo * BN_ONE_OVER_FOUR_PI;
    s.dir = {x, y, z};
    s.type = SCATTER_VOLUME;
  }
  return s;
}

// shadeRays.cu:476-878 with sampleLights (:295-404) over the environment
// (HDRI map or uniform ambient radiance) and the quad, directional and
// point lights. Returns the fragment this ray adds to its pixel. On return
// ray.tMax > 0 means the path continues with the updated ray; shadow.tMax > 0
// means a shadow ray was spawned.
// kLights: rp carries an HDRI map or light lists (has_lights); kernels are
// instantiated for both cases so that the samplers cost the ambient-light
// kernels no registers.
template <bool kLights>
BN_HD vec3f bounce(const RenderDD &rp, PathRay &ray, const ShadeHit &h, PathRay &shadow, int pathDepth) {
#pragma clang fp contract(off)
  const float EPS = 1e-4f;
  shadow.tMax = -1.f;
  vec3f fragment = vec3f(0.f);
  const vec3f incomingTp = ray.tp;
  const bool map = kLights && has_map(rp);
  if (ray.isShadow) {
    if (!h.hit) fragment = ray.mis * ray.tp;
    ray.tMax = -1.f;
    return fragment;
  }
  vec3f Ng = h.hit ? h.N : vec3f(0.f);
  const bool isVolumeHit = (Ng == vec3f(0.f));
  if (!isVolumeHit) Ng = normalize(Ng);
  const bool hitWasOnFront = dot(ray.dir, Ng) < 0.f;
  const vec3f Ngff = hitWasOnFront ? Ng : -Ng;
  if (!h.hit) {
    // primaryRayMissColor / radianceFromEnv (shadeRays.cu:437-470, 556-575)
    if (pathDepth == 0)
      fragment = map ? env_radiance(rp.env, ray.dir) : ray.missColor;
    else
      fragment = ray.tp * (map ? env_radiance(rp.env, ray.dir) : vec3f(rp.ambient)) * ray.mis;
    ray.tMax = -1.f;
    return fragment;
  }
  Random2 random(ray.rng, (uint64_t)float_as_uint(h.t));
  const Bsdf &bsdf = h.bsdf;
  const vec3f P = h.P;
  const vec3f wo = -normalize(ray.dir);
  const float offsetEps = fmaxf(EPS, 1e-6f * reduce_max(vabs(P)));
  const vec3f ffOffset = isVolumeHit ? wo : Ngff;

  // ---- sampleEnvLight (shadeRays.cu:267-293): the map's importance
  // sample, or randomDirection into the normal's hemisphere ----
  vec3f lsDir;
  vec3f lsRad;
  float lsPdf;
  float lsDist = BN_INF;
  if (map) {
    env_sample(rp.env, random, lsDir, lsRad, lsPdf);
  } else {
    lsDir = random_direction(random);
    if (!(Ngff == vec3f(0.f))) {
      lsRad = vec3f(BN_TWO_PI * rp.ambient);
      if (dot(lsDir, Ngff) < 0.f) lsDir = -lsDir;
      lsPdf = BN_ONE_OVER_TWO_PI;
    } else {
      lsRad = vec3f(BN_FOUR_PI * rp.ambient);
      lsPdf = BN_ONE_OVER_FOUR_PI;
    }
  }
  // ---- sampleLights (shadeRays.cu:295-404): the environment's sample, then
  // the area, directional and point lights' (in that draw order); one is
  // picked by its share of the summed weights. A light whose sampler found
  // nothing has weight 0: picking it (a draw at a boundary, or past the
  // rounded sum) gives no light, as the default Light::Sample's pdf 0 does.
  bool haveLight = false;
  bool lightNeedsMIS = false;
  {
    const float elsWeight = reduce_max(lsRad);
    LightSample als, dls, pls;
    float alsWeight = 0.f, dlsWeight = 0.f, plsWeight = 0.f;
    bool alsOk = false, dlsOk = false, plsOk = false;
    if (kLights) {
      alsOk = sample_quad_lights(rp.lights, P, Ngff, random, als);
      if (alsOk) alsWeight = reduce_max(als.rad) / als.pdf;
      dlsOk = sample_dir_lights(rp.lights, Ngff, random, dls);
      if (dlsOk) dlsWeight = reduce_max(dls.rad);
      plsOk = sample_point_lights(rp.lights, P, Ngff, random, pls);
      if (plsOk) plsWeight = reduce_max(pls.rad);
    }
    const float sumWeights = ((alsWeight + dlsWeight) + elsWeight) + plsWeight;
    if (sumWeights != 0.f) {
      const float alsW = alsWeight * (1.f / sumWeights);
      const float elsW = elsWeight * (1.f / sumWeights);
      const float dlsW = dlsWeight * (1.f / sumWeights);
      const float plsW = plsWeight * (1.f / sumWeights);
      const float r = random();
      if (r <= alsW) {
        if (kLights && alsOk) {
          lsDir = als.dir;
          lsRad = als.rad * (1.f / alsW);
          lsDist = als.dist;
          lsPdf = als.pdf;
          haveLight = !(lsPdf <= 0.f) && !(lsPdf != lsPdf);
        }
      } else if (r <= alsW + elsW) {
        lsRad *= (1.f / elsW);
        lightNeedsMIS = true;
        haveLight = !(lsPdf <= 0.f) && !(lsPdf != lsPdf);
      } else if (kLights) {
        const bool point = r <= (alsW + elsW) + plsW;
        const LightSample &ls = point ? pls : dls;
        if (point ? plsOk : dlsOk) {
          lsDir = ls.dir;
          lsRad = ls.rad * (1.f / (point ? plsW : dlsW));
          lsDist = ls.dist;
          lsPdf = ls.pdf;  // inf
          haveLight = !(lsPdf <= 0.f) && !(lsPdf != lsPdf);
        }
      }
    }
  }
  if (haveLight) {
    EvalRes f_r = bsdf_eval(bsdf, Ng, lsDir, wo);
    if (!(f_r.pdf > 0.f) || reduce_max(f_r.value) < 1e-4f) {
      shadow.tMax = -1.f;
    } else {
      f_r.value = vmin(f_r.value, 13.f);
      const vec3f tp_sr = incomingTp * (is_inf(lsPdf) ? 1.f : 1.f / lsPdf) * f_r.value * lsRad *
                          (isVolumeHit ? 1.f : (BN_ONE_OVER_PI * fabsf(dot(Ng, lsDir))));
      shadow.isShadow = true;
      shadow.dir = lsDir;
      shadow.org = P + offsetEps * ffOffset;
      shadow.tMax = lsDist * (1.f - 2.f * offsetEps);
      shadow.tp = tp_sr;
      shadow.rng = ray.rng;
      ray.rng.next_hash((uint64_t)float_as_uint(h.t));
      shadow.mis = 1.f;
      shadow.numDiffuse = ray.numDiffuse;
      shadow.isInMedium = ray.isInMedium;
      if (lightNeedsMIS) {
        // EnvMapLight::DD::pdf (1/4pi without a texture)
        const float pl = map ? env_pdf(rp.env, lsDir) : BN_ONE_OVER_FOUR_PI;
        const float pb = bsdf_pdf(bsdf, Ng, lsDir, wo);
        shadow.mis = pl / (pl + pb + 1e-10f);
        if (shadow.mis < 1e-5f) shadow.tMax = -1.f;
      }
    }
  }

  if (pathDepth >= rp.maxPathDepth) {
    ray.tMax = -1.f;
    return fragment;
  }
  ray.tMax = BN_INF;
  const ScatterRes sr = bsdf_scatter(bsdf, Ng, wo, ray.isInMedium, random);
  if (!(sr.pdf > 0.f) || sr.pdf <= 1e-6f) return fragment;  // shadeRays.cu:782-783
  if (sr.type == SCATTER_VOLUME) {  // SCI_VIS_MODE: volumes shadow, no bounce
    ray.tMax = -1.f;
    return fragment;
  }
  if (sr.type == SCATTER_DIFFUSE) {
    if (ray.numDiffuse >= rp.maxDiffuseBounces) {
      ray.tMax = -1.f;
      return fragment;
    }
    ray.numDiffuse = ray.numDiffuse + 1;
  }
  ray.org = P + sr.offsetDirection * offsetEps * ffOffset;
  ray.dir = normalize(sr.dir);
  vec3f scatterFactor = sr.f_r * (isVolumeHit ? 1.f : fabsf(dot(Ng, normalize(ray.dir)))) /
                        (is_inf(sr.pdf) ? 1.f : (sr.pdf + 1e-10f));
  scatterFactor = vmin(scatterFactor, 2.5f);
  ray.tp = ray.tp * scatterFactor;
  if (sr.changedMedium) ray.isInMedium = !ray.isInMedium;
  if (lightNeedsMIS && !is_inf(sr.pdf)) {
    const float pb = bsdf_pdf(bsdf, Ng, ray.dir, wo);
    const float pl = map ? env_pdf(rp.env, ray.dir) : BN_ONE_OVER_FOUR_PI;
    ray.mis = pb / (pb + pl);
  } else {
    ray.mis = 1.f;
  }
  if (rp.rrDepth > 0 && pathDepth + 1 >= rp.rrDepth) {
    // Russian roulette (not in barney; opt-in): survive with p = max(tp)
    const float p = fminf(reduce_max(ray.tp), 0.95f);
    if (random() >= p) {
      ray.tMax = -1.f;
      return fragment;
    }
    ray.tp = ray.tp * (1.f / p);
  }
  return fragment;
}

// DeviceMaterial::createBSDF (material/DeviceMaterial.h:47-62): matte ->
// Lambertian; PBR -> Glass when ior != 1 and transmission >= 1e-3
// (AnariPBR.h:75-80), else NVisii with roughness clamped to [.001, .9]. The
// NVisii parameters are halfs in barney (NVisii.h:1062-1083); the material
// table holds them already rounded to half.
BN_HD Bsdf material_bsdf(const MaterialDev &m, vec3f color) {
#pragma clang fp contract(off)
  Bsdf b;
  if (m.type == BNAMD_MAT_PBR) {
    if (m.ior != 1.f && m.transmission >= 1e-3f) {
      b.type = BSDF_GLASS;
      b.albedo = vec3f(1.f);
      b.ior = m.ior;
      b.metallic = b.roughness = 0.f;
      return b;
    }
    b.type = BSDF_NVISII;
    b.albedo = {__half2float(__float2half(color.x)), __half2float(__float2half(color.y)),
                __half2float(__float2half(color.z))};
    b.metallic = m.metallic;
    b.roughness = m.roughness;
    b.ior = m.ior;
    return b;
  }
  b.type = BSDF_LAMBERT;
  b.albedo = .85f * color * BN_ONE_OVER_PI;
  b.metallic = b.roughness = b.ior = 0.f;
  return b;
}

// AnariMatte::createBSDF: albedo = 0.85 * color / pi (AnariMatte.h:56-62)
BN_HD Bsdf matte_bsdf(vec3f color) {
#pragma clang fp contract(off)
  Bsdf b;
  b.type = BSDF_LAMBERT;
  b.albedo = .85f * color * BN_ONE_OVER_PI;
  b.metallic = b.roughness = b.ior = 0.f;
  return b;
}

// Phase(color) with avg_reflectance 0.7 (Phase.h:137-141)
BN_HD Bsdf phase_bsdf(vec3f color) {
#pragma clang fp contract(off)
  Bsdf b;
  b.type = BSDF_PHASE;
  b.albedo = .7f * color;
  b.metallic = b.roughness = b.ior = 0.f;
  return b;
}

}  // namespace bnamd
