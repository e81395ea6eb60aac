// World = flattened instance group with its BVH8, geometry table and
// materials (the AccelHandle behind rtc::InstanceGroup, rtcore/cuda/Group.h:59-85).
#pragma once
#include "bvh8.h"
#include "runtime.h"
#include "../../include/bnamd.h"

namespace bnamd {

// Internal geometry kind for point-location BVHs over tetrahedra:
// vertices = float4 (x, y, z, scalar), indices = int4 per tet.
constexpr int GEOM_TETS = 3;

// Device-side geometry table entry (one per instanced geom).
struct GeomDev {
  int kind;
  int numPrims;
  uint32_t primOffset;  // first flat index of this geom
  int geomID;           // index within its group
  int instIndex;        // flattened instance index
  int instID;           // user instance ID
  int materialID;
  int userID;
  const float *vertices;
  const int *indices;
  const float *radii;
  const float *colors;
  float defaultRadius;
  int pad;
};

struct MaterialDev {
  int type;
  float color[3];
  float opacity, metallic, roughness, transmission, ior, specular;
};

struct WorldDev {
  const BNode8 *nodes;
  const PrimRec *prims;
  const GeomDev *geoms;
  const MaterialDev *materials;
  int numGeoms;
  int depth;
  uint32_t numPrims;
  int pad;
};

struct World {
  DevBuf<BNode8> nodes;
  DevBuf<PrimRec> prims;
  DevBuf<GeomDev> geoms;
  DevBuf<MaterialDev> materials;
  std::vector<GeomDev> hgeoms;
  uint32_t numPrims = 0, numNodes = 0;
  int depth = 1;
  float buildMs = 0.f;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  WorldDev dd() const {
    WorldDev w;
    w.nodes = nodes.ptr;
    w.prims = prims.ptr;
    w.geoms = geoms.ptr;
    w.materials = materials.ptr;
    w.numGeoms = (int)hgeoms.size();
    w.depth = depth;
    w.numPrims = numPrims;
    w.pad = 0;
    return w;
  }
};

// Builds the BVH8 over the given prim table (implemented in bvh_build.hip).
void build_world_bvh(World &w, hipStream_t stream);

// Maps a flat prim index back to its geom (binary search over primOffset).
BN_D int find_geom(const GeomDev *__restrict geoms, int numGeoms, uint32_t flat) {
  int lo = 0, hi = numGeoms - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (geoms[mid].primOffset <= flat) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Closest-hit epilogue output (what barney's anyHit/closestHit programs
// derive for Ray::setHit): world-space point and geometric normal.
struct SurfaceHit {
  vec3f P;
  vec3f Ng;  // normalized
  int geom;  // index into the world geom table
  int primID;
};

// Rebuilds hit position and normal for the accepted primitive, following
// Triangles.dev.cu:57-74 (geometric normal, interpolated P),
// Spheres.dev.cu:57-90 (normal P - center; the object-space point is the
// shifted-frame hit) and Cylinders.dev.cu:208-216 (normal from intersect).
BN_D SurfaceHit surface_epilogue(const WorldDev &w, vec3f org, vec3f dir, float tmin, float tmax,
                                 const HitRec &h) {
#pragma clang fp contract(off)
  SurfaceHit s;
  const int g = find_geom(w.geoms, w.numGeoms, h.flat);
  const GeomDev &gd = w.geoms[g];
  const int primID = (int)(h.flat - gd.primOffset);
  s.geom = g;
  s.primID = primID;
  if (gd.kind == BNAMD_GEOM_TRIANGLES) {
    const int *ix = gd.indices + 3 * primID;
    const vec3f v0 = load3(gd.vertices + 3 * ix[0]);
    const vec3f v1 = load3(gd.vertices + 3 * ix[1]);
    const vec3f v2 = load3(gd.vertices + 3 * ix[2]);
    s.Ng = normalize(cross(v1 - v0, v2 - v0));
    s.P = (1.f - h.u - h.v) * v0 + h.u * v1 + h.v * v2;
  } else if (gd.kind == BNAMD_GEOM_SPHERES) {
    const vec3f c = load3(gd.vertices + 3 * primID);
    const vec3f o = org + h.v * dir;  // h.v = t_move
    const vec3f P = o + h.u * dir;    // h.u = shifted-frame hit
    s.P = P;
    const vec3f N = (P == c) ? vec3f(1.f, 0.f, 0.f) : (P - c);
    s.Ng = normalize(N);
  } else {
    const int *ix = gd.indices + 2 * primID;
    const vec3f v0 = load3(gd.vertices + 3 * ix[0]);
    const vec3f v1 = load3(gd.vertices + 3 * ix[1]);
    const float r = gd.radii ? gd.radii[primID] : gd.defaultRadius;
    float t;
    vec3f N = {0.f, 0.f, 0.f};
    cylinder_hit(org, dir, tmin, tmax, v0, v1, r, t, &N);
    vec3f P = org + t * dir;
    P = P + 1e-4f * normalize(N);
    s.P = P;
    s.Ng = normalize(N);
  }
  return s;
}

}  // namespace bnamd
