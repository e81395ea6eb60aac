// Camera rays and the shade/bounce step of barney's path tracer with the
// environment (ambient) light and single scattering:
//   camera           barney/Camera.cu:81-100
//   ray generation   barney/kernels/generateRays.cu:73-129, 199-210
//   bounce()         barney/kernels/shadeRays.cu:476-878
//   env light        shadeRays.cu:267-293 (sampleEnvLight), 295-404 (sampleLights)
//   BSDFs            packedBSDF/Lambertian.h:55-87, Phase.h:120-172
//   materials        material/AnariMatte.h:51-63
//   accumulation     shadeRays.cu:970-1038
#pragma once
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
};

inline RenderDD make_render_dd(const bnamd_render_params &p) {
  RenderDD r;
  r.width = p.width;
  r.height = p.height;
  r.accumID = p.accumID;
  r.ambient = p.ambientRadiance;
  for (int k = 0; k < 4; ++k) r.bg[k] = p.bgColor[k];
  r.maxPathDepth = p.maxPathDepth > 0 ? p.maxPathDepth : 10;
  r.maxDiffuseBounces = p.maxDiffuseBounces >= 0 ? p.maxDiffuseBounces : 1;
  r.rrDepth = p.russianRoulette;
  return r;
}

enum BsdfType : int { BSDF_NONE = 0, BSDF_LAMBERT = 1, BSDF_PHASE = 2 };

struct Bsdf {
  int type;
  vec3f albedo;
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
  ray.tp = vec3f(1.f);
  ray.mis = 0.f;
  ray.numDiffuse = 0;
  ray.isShadow = false;
  const float t = (iy + .5f) / float(rp.height);
  if (rp.bg[3] >= 0.f) {
    ray.missColor = {rp.bg[0], rp.bg[1], rp.bg[2]};
    ray.missAlpha = rp.bg[3];
  } else {
    ray.missColor = (1.0f - t) * vec3f(0.9f, 0.9f, 0.9f) + t * vec3f(0.15f, 0.25f, .8f);
    ray.missAlpha = (1.0f - t) * 1.f + t * 1.f;
  }
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

// Lambertian::eval / Phase::eval (single-scatter phase is isotropic 1/4pi)
BN_HD EvalRes bsdf_eval(const Bsdf &b, vec3f Ns, vec3f wi) {
#pragma clang fp contract(off)
  EvalRes r;
  if (b.type == BSDF_LAMBERT) {
    const float cosThetaI = fmaxf(dot(wi, Ns), 0.f);
    r.pdf = cosThetaI * BN_ONE_OVER_PI;
    r.value = b.albedo * BN_ONE_OVER_PI * cosThetaI;
  } else {
    const float density = BN_ONE_OVER_FOUR_PI;
    r.value = density * b.albedo;
    r.pdf = density;
  }
  return r;
}

BN_HD float bsdf_pdf(const Bsdf &b, vec3f Ng, vec3f wi) {
#pragma clang fp contract(off)
  if (b.type == BSDF_LAMBERT) {
    vec3f N = Ng;
    if (dot(wi, N) < 0.f) N = -N;
    const float cosThetaI = fmaxf(dot(wi, N), 0.f);
    return cosThetaI * BN_ONE_OVER_PI;
  }
  return BN_ONE_OVER_FOUR_PI;
}

struct ScatterRes {
  vec3f f_r, dir;
  float pdf;
  bool diffuse;  // DIFFUSE vs VOLUME
};

BN_HD ScatterRes bsdf_scatter(const Bsdf &b, vec3f Ng, vec3f wo, Random2 &rng) {
#pragma clang fp contract(off)
  ScatterRes s;
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
    sincosf(phi, &sinPhi, &cosPhi);
    const vec3f local = {cosPhi * sinTheta, sinPhi * sinTheta, cosTheta};
    s.dir = xfmVector(make_frame(N), local);
    s.pdf = local.z * BN_ONE_OVER_PI;
    s.f_r = b.albedo;
    s.diffuse = true;
  } else {
    const float r1 = rng();
    const float r2 = rng();
    const float x = cosf(BN_TWO_PI * r1) * sqrtf(r2 * (1.f - r2));
    const float y = sinf(BN_TWO_PI * r1) * sqrtf(r2 * (1.f - r2));
    const float z = (1.f - 2.f * r2);
    s.pdf = BN_ONE_OVER_FOUR_PI;
    s.f_r = b.albedo * BN_ONE_OVER_FOUR_PI;
    s.dir = {x, y, z};
    s.diffuse = false;
  }
  return s;
}

BN_HD bool is_inf(float f) { return f == BN_INF || f == -BN_INF; }

// shadeRays.cu:476-878 for env/ambient lighting. Returns the fragment this
// ray adds to its pixel. On return ray.tMax > 0 means the path continues with
// the updated ray; shadow.tMax > 0 means a shadow ray was spawned.
BN_HD vec3f bounce(const RenderDD &rp, PathRay &ray, const ShadeHit &h, PathRay &shadow, int pathDepth) {
#pragma clang fp contract(off)
  const float EPS = 1e-4f;
  shadow.tMax = -1.f;
  vec3f fragment = vec3f(0.f);
  const vec3f incomingTp = ray.tp;
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
    if (pathDepth == 0)
      fragment = ray.missColor;
    else
      fragment = ray.tp * vec3f(rp.ambient) * ray.mis;
    ray.tMax = -1.f;
    return fragment;
  }
  Random2 random(ray.rng, (uint64_t)float_as_uint(h.t));
  const Bsdf &bsdf = h.bsdf;
  const vec3f P = h.P;
  const vec3f wo = -normalize(ray.dir);
  const float offsetEps = fmaxf(EPS, 1e-6f * reduce_max(vabs(P)));
  const vec3f ffOffset = isVolumeHit ? wo : Ngff;

  // ---- sampleLights(): only the environment light is present ----
  vec3f lsDir = random_direction(random);
  vec3f lsRad;
  float lsPdf;
  if (!(Ngff == vec3f(0.f))) {
    lsRad = vec3f(BN_TWO_PI * rp.ambient);
    if (dot(lsDir, Ngff) < 0.f) lsDir = -lsDir;
    lsPdf = BN_ONE_OVER_TWO_PI;
  } else {
    lsRad = vec3f(BN_FOUR_PI * rp.ambient);
    lsPdf = BN_ONE_OVER_FOUR_PI;
  }
  bool haveLight = false;
  bool lightNeedsMIS = false;
  {
    float elsWeight = reduce_max(lsRad);
    const float sumWeights = elsWeight;
    if (sumWeights != 0.f) {
      elsWeight *= 1.f / sumWeights;
      const float r = random();
      (void)r;  // r <= 0 selects the (absent) area light in barney; env otherwise
      lsRad *= (1.f / elsWeight);
      lightNeedsMIS = true;
      haveLight = !(lsPdf <= 0.f) && !(lsPdf != lsPdf);
    }
  }
  if (haveLight) {
    EvalRes f_r = bsdf_eval(bsdf, Ng, lsDir);
    if (!(f_r.pdf > 0.f) || reduce_max(f_r.value) < 1e-4f) {
      shadow.tMax = -1.f;
    } else {
      f_r.value = vmin(f_r.value, 13.f);
      const vec3f tp_sr = incomingTp * (is_inf(lsPdf) ? 1.f : 1.f / lsPdf) * f_r.value * lsRad *
                          (isVolumeHit ? 1.f : (BN_ONE_OVER_PI * fabsf(dot(Ng, lsDir))));
      shadow.isShadow = true;
      shadow.dir = lsDir;
      shadow.org = P + offsetEps * ffOffset;
      shadow.tMax = BN_INF * (1.f - 2.f * offsetEps);
      shadow.tp = tp_sr;
      shadow.rng = ray.rng;
      ray.rng.next_hash((uint64_t)float_as_uint(h.t));
      shadow.mis = 1.f;
      shadow.numDiffuse = ray.numDiffuse;
      if (lightNeedsMIS) {
        const float pl = BN_ONE_OVER_FOUR_PI;  // EnvMapLight::DD::pdf without a texture
        const float pb = bsdf_pdf(bsdf, Ng, lsDir);
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
  const ScatterRes sr = bsdf_scatter(bsdf, Ng, wo, random);
  if (!(sr.pdf > 0.f) || sr.pdf <= 1e-6f) return fragment;  // shadeRays.cu:782-783
  if (!sr.diffuse) {  // VOLUME scatter in SCI_VIS_MODE: volumes shadow, no bounce
    ray.tMax = -1.f;
    return fragment;
  }
  if (ray.numDiffuse >= rp.maxDiffuseBounces) {
    ray.tMax = -1.f;
    return fragment;
  }
  ray.numDiffuse = ray.numDiffuse + 1;
  ray.org = P + 1.f * offsetEps * ffOffset;
  ray.dir = normalize(sr.dir);
  vec3f scatterFactor = sr.f_r * (isVolumeHit ? 1.f : fabsf(dot(Ng, normalize(ray.dir)))) /
                        (is_inf(sr.pdf) ? 1.f : (sr.pdf + 1e-10f));
  scatterFactor = vmin(scatterFactor, 2.5f);
  ray.tp = ray.tp * scatterFactor;
  if (lightNeedsMIS && !is_inf(sr.pdf)) {
    const float pb = bsdf_pdf(bsdf, Ng, ray.dir);
    const float pl = BN_ONE_OVER_FOUR_PI;
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

// AnariMatte::createBSDF: albedo = 0.85 * color / pi (AnariMatte.h:56-62)
BN_HD Bsdf matte_bsdf(vec3f color) {
#pragma clang fp contract(off)
  Bsdf b;
  b.type = BSDF_LAMBERT;
  b.albedo = .85f * color * BN_ONE_OVER_PI;
  return b;
}

// Phase(color) with avg_reflectance 0.7 (Phase.h:137-141)
BN_HD Bsdf phase_bsdf(vec3f color) {
#pragma clang fp contract(off)
  Bsdf b;
  b.type = BSDF_PHASE;
  b.albedo = .7f * color;
  return b;
}

}  // namespace bnamd
