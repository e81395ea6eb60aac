// Light lists of the C ABI: World::set(quadLights / dirLights / pointLights)
// (barney/render/World.cpp:80-120) with the getDD of each light
// (light/QuadLight.cu:9-18, DirLight.cu:10-20, PointLight.cu:9-17).
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/bnamd.h"
#include "lights.h"
#include "runtime.h"

using namespace bnamd;

namespace {

struct LightsImpl : bnamd_lights_s {
  DevBuf<QuadLightDev> quad;
  DevBuf<DirLightDev> dir;
  DevBuf<PointLightDev> point;
};

vec3f load3(const float *p) { return {p[0], p[1], p[2]}; }
bool finite3(vec3f v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

template <class T>
void upload(DevBuf<T> &buf, const std::vector<T> &h) {
  if (h.empty()) return;
  buf.alloc(h.size());
  BN_HIP(hipMemcpy(buf.ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

}  // namespace

extern "C" int bnamd_lights_create(const bnamd_quad_light *quads, int numQuads, const bnamd_dir_light *dirs,
                                   int numDirs, const bnamd_point_light *points, int numPoints,
                                   bnamd_lights **out) {
  return guarded([&] {
#pragma clang fp contract(off)
    BN_REQUIRE(out, "bnamd_lights_create: NULL argument");
    *out = nullptr;
    BN_REQUIRE(numQuads >= 0 && numDirs >= 0 && numPoints >= 0, "bnamd_lights_create: negative light count");
    BN_REQUIRE((quads || !numQuads) && (dirs || !numDirs) && (points || !numPoints),
               "bnamd_lights_create: NULL light array with a non-zero count");
    std::vector<QuadLightDev> q(numQuads);
    for (int i = 0; i < numQuads; ++i) {
      QuadLightDev &d = q[i];
      d.corner = load3(quads[i].corner);
      d.edge0 = load3(quads[i].edge0);
      d.edge1 = load3(quads[i].edge1);
      d.emission = load3(quads[i].emission);
      const vec3f c = cross(d.edge0, d.edge1);
      d.area = length(c);
      BN_REQUIRE(finite3(d.corner) && finite3(d.edge0) && finite3(d.edge1) && d.area > 0.f,
                 "bnamd_lights_create: a quad light needs finite corner and edges spanning a non-zero area");
      d.normal = normalize(c);
    }
    std::vector<DirLightDev> dl(numDirs);
    for (int i = 0; i < numDirs; ++i) {
      const vec3f dir = load3(dirs[i].direction);
      BN_REQUIRE(finite3(dir) && dot(dir, dir) > 0.f, "bnamd_lights_create: a directional light needs a direction");
      dl[i].direction = normalize(dir);
      dl[i].radiance = load3(dirs[i].color) * dirs[i].radiance;
    }
    std::vector<PointLightDev> pl(numPoints);
    for (int i = 0; i < numPoints; ++i) {
      pl[i].position = load3(points[i].position);
      const vec3f color = load3(points[i].color);
      // PointLight::DD::radianceTowards (PointLight.h:14-20)
      pl[i].radiance = std::isnan(points[i].intensity) ? color * points[i].power * BN_ONE_OVER_FOUR_PI
                                                       : color * points[i].intensity;
    }
    std::unique_ptr<LightsImpl> L(new LightsImpl);
    upload(L->quad, q);
    upload(L->dir, dl);
    upload(L->point, pl);
    L->dd.quad = L->quad.ptr;
    L->dd.dir = L->dir.ptr;
    L->dd.point = L->point.ptr;
    L->dd.numQuad = numQuads;
    L->dd.numDir = numDirs;
    L->dd.numPoint = numPoints;
    L->dd.pad = 0;
    *out = L.release();
  });
}

extern "C" int bnamd_lights_destroy(bnamd_lights *lights) {
  return guarded([&] { delete static_cast<LightsImpl *>(lights); });
}
