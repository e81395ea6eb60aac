// World = flattened instance group with its BVH8, geometry table and
// materials (the AccelHandle behind rtc::InstanceGroup, rtcore/cuda/Group.h:59-85).
#pragma once
#include "bvh8.h"
#include "runtime.h"
#include "../../include/bnamd.h"
#include <memory>

namespace bnamd {

// Internal geometry kind for point-location BVHs over tetrahedra:
// vertices = float4 (x, y, z, scalar), indices = int4 per tet.
constexpr int GEOM_TETS = 3;
// Internal geometry kind for the top level of a two-level world:
// vertices = float4 pairs (world box lo, hi) per instance.
constexpr int GEOM_INSTANCES = 4;
// element boxes of a general unstructured mesh ('vertices': lo, hi float4 pairs)
constexpr int GEOM_ELEM_BOXES = 5;

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
  const float *normals;  // triangles: per-vertex normals or nullptr
  float defaultRadius;
  int pad;
  const void *programs;  // program record (programmable traversal) or nullptr
};

struct MaterialDev {
  int type;
  float color[3];
  float opacity, metallic, roughness, transmission, ior, specular;
  // the BSDF's alpha that barney's triangle anyHit tests (getOpacity):
  // Lambertian baseColor.w * opacity (AnariMatte.h:63), NVisii
  // (1 - transmission) * baseColor.w * opacity stored as a half
  // (AnariPBR.h:88-91, NVisii.h:1080); Glass is handled by
  // PRIM_SHADOW_TRANSPARENT instead (1 here)
  float alpha;
};

// One instance of a two-level world: affine transforms as 12 floats
// (columns vx, vy, vz, then p; object to world and its inverse, as
// InstanceGroup keeps xfms and rcp(xfms), rtcore/embree/Group.cpp:196-199)
// and the bottom-level BVH of its group.
struct InstDev {
  float o2w[12];
  float w2o[12];
  const BNode8 *nodes;
  const PrimRec *prims;
  const GeomDev *geoms;
  int numGeoms;
  int instID;
};

struct WorldDev {
  const BNode8 *nodes;     // flattened world, or the instance level
  const PrimRec *prims;
  const GeomDev *geoms;    // flattened world only
  const MaterialDev *materials;
  int numGeoms;
  int depth;               // stack levels - 1 (both levels together)
  uint32_t numPrims;
  int pad;
  // two-level worlds
  const InstDev *insts;
  int tlasDepth;           // stack levels of the instance level
  int localBits;           // tie key = (instance << localBits) | group-local flat index
};

BN_HD vec3f xfm_point(const float *m, vec3f p) {
#pragma clang fp contract(off)
  return {m[0] * p.x + m[3] * p.y + m[6] * p.z + m[9], m[1] * p.x + m[4] * p.y + m[7] * p.z + m[10],
          m[2] * p.x + m[5] * p.y + m[8] * p.z + m[11]};
}
BN_HD vec3f xfm_vector(const float *m, vec3f v) {
#pragma clang fp contract(off)
  return {m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
          m[2] * v.x + m[5] * v.y + m[8] * v.z};
}

// Launch configuration of the batched renderers for one world, chosen by
// bnamd_batch_autotune (defaults until it runs).
struct BatchTune {
  int traceBlocksPerCU = 0;  // wavefront trace/shadow workgroups per CU; 0 = as many as are resident
  int chunk = 0;             // rays per dynamic work fetch; 0 = the kernels' default
  int sppBatch = 0;          // paths per pixel per wavefront batch (queue size); 0 = all in one batch
  int megaLanes = 0;         // megakernel lanes per pixel (tile = 256 / lanes pixels); 0 = from spp
  int mode = 1;              // the faster renderer (BNAMD_BATCH_MEGAKERNEL 0 / WAVEFRONT 1)
  bool tuned = false;
};

struct World {
  DevBuf<BNode8> nodes;
  DevBuf<PrimRec> prims;
  DevBuf<GeomDev> geoms;
  DevBuf<MaterialDev> materials;
  std::vector<MaterialDev> hmats;  // bottom levels: the table they were built with
  std::vector<GeomDev> hgeoms;
  uint32_t numPrims = 0, numNodes = 0;
  int depth = 1;
  float buildMs = 0.f;
  float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  // two-level worlds: one bottom level per group, instance records, and the
  // instance level in nodes/prims above
  bool twoLevel = false;
  std::vector<std::unique_ptr<World>> blas;  // bottom levels this world owns
  std::vector<const World *> instBlas;       // each instance's bottom level (owned or the caller's)
  std::vector<int> instIDs;
  int numBlas = 0;                           // distinct bottom levels referenced
  bool hasCutouts = false;                   // some triangle's material has alpha < 1
  bool hasUser = false;                      // user geometry (programmable traversal only)
  mutable DevBuf<WorldDev> devRecord;        // dd() in device memory (bnamd_world_device)
  mutable DevBuf<unsigned long long> counters;  // the surface render's counters
  mutable BatchTune tune;                        // batched renderers' launch configuration
  DevBuf<InstDev> insts;
  DevBuf<float4> instBoxes;
  int tlasDepth = 0, localBits = 0;
  uint64_t totalPrims = 0, totalNodes = 0;  // over all instances / levels
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
    w.insts = insts.ptr;
    w.tlasDepth = tlasDepth;
    w.localBits = localBits;
    return w;
  }
};

// Builds the BVH8 over the given prim table (implemented in bvh_build.hip).
// BVH8 of w's primitives. The binary tree under the 8-wide collapse is
// PLOC's (surfaces: fewer node visits per ray) or, with mortonOrder, the
// radix tree of the Morton codes, whose leaf order follows the space-filling
// curve (tet meshes: point location tries the leaves next to the last hit
// first); BNAMD_BVH_BUILD=ploc|radix overrides the choice.
void build_world_bvh(World &w, hipStream_t stream, bool mortonOrder = false);
// Host-side structural check of a built BVH8 (throws on a violation).
void check_world_bvh(const World &w);

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
  vec3f Ng;   // normalized (flattened worlds); transformed by the instance
              // for two-level worlds (bounce() normalizes)
  const GeomDev *g;
  int primID;
  int instIndex;  // flattened instance index
  int instID;     // user instance ID
};

// Rebuilds hit position and normal in the space of the geometry table
// 'geoms' for the accepted primitive with group-local flat index 'flat',
// following Triangles.dev.cu:57-75 (geometric or interpolated vertex
// normal, interpolated P),
// Spheres.dev.cu:57-90 (normal P - center; the object-space point is the
// shifted-frame hit) and Cylinders.dev.cu:208-216 (normal from intersect).
BN_D SurfaceHit object_epilogue(const GeomDev *__restrict geoms, int numGeoms, vec3f org, vec3f dir, float tmin,
                                float tmax, const HitRec &h, uint32_t flat) {
#pragma clang fp contract(off)
  SurfaceHit s;
  const int g = find_geom(geoms, numGeoms, flat);
  const GeomDev &gd = geoms[g];
  const int primID = (int)(flat - gd.primOffset);
  s.g = &gd;
  s.primID = primID;
  s.instIndex = gd.instIndex;
  s.instID = gd.instID;
  if (gd.kind == BNAMD_GEOM_TRIANGLES) {
    const int *ix = gd.indices + 3 * primID;
    const vec3f v0 = load3(gd.vertices + 3 * ix[0]);
    const vec3f v1 = load3(gd.vertices + 3 * ix[1]);
    const vec3f v2 = load3(gd.vertices + 3 * ix[2]);
    // n = normalize(cross), or the normalized interpolated vertex normal;
    // the object normal is normalize(n) (Triangles.dev.cu:65-75)
    vec3f n = normalize(cross(v1 - v0, v2 - v0));
    if (gd.normals)
      n = normalize((1.f - h.u - h.v) * load3(gd.normals + 3 * ix[0]) + h.u * load3(gd.normals + 3 * ix[1]) +
                    h.v * load3(gd.normals + 3 * ix[2]));
    s.Ng = normalize(n);
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

// World-space hit of a closest-hit result. Two-level worlds: the ray is
// taken to object space as the traversal did, the object-space point goes
// back through the instance transform (transformPointFromObjectToWorldSpace)
// and the normal through its linear part, as the embree backend's
// transformNormal/VectorFromObjectToWorldSpace do
// (rtcore/embree/Device.cpp:26-39).
template <bool TL>
BN_D SurfaceHit surface_epilogue(const WorldDev &w, vec3f org, vec3f dir, float tmin, float tmax, const HitRec &h) {
  if (!TL) return object_epilogue(w.geoms, w.numGeoms, org, dir, tmin, tmax, h, h.flat);
  const uint32_t inst = h.flat >> w.localBits;
  const uint32_t local = h.flat & ((1u << w.localBits) - 1u);
  const InstDev &I = w.insts[inst];
  const vec3f o = xfm_point(I.w2o, org);
  const vec3f d = fix_dir(xfm_vector(I.w2o, dir));
  SurfaceHit s = object_epilogue(I.geoms, I.numGeoms, o, d, tmin, tmax, h, local);
  s.P = xfm_point(I.o2w, s.P);
  s.Ng = xfm_vector(I.o2w, s.Ng);
  s.instIndex = (int)inst;
  s.instID = I.instID;
  return s;
}

// The ID-buffer record of a primary hit, as the anyHit / intersect programs
// write it for the first path of the first frame (Triangles.dev.cu:45-60;
// Spheres.dev.cu:178-198, which writes 13 + rank as the instance ID;
// Cylinders.dev.cu:244-255): {primID, instID, objID (the geometry's user
// ID), depth as float bits} of the closest hit; a miss keeps
// {-1, -1, -1, inf}.
template <bool TL>
BN_D int4 hit_ids(const WorldDev &w, vec3f org, vec3f dir, float tmax, const HitRec &h) {
  if (h.flat == 0xffffffffu) return make_int4(-1, -1, -1, __float_as_int(BN_INF));
  const SurfaceHit s = surface_epilogue<TL>(w, org, dir, 0.f, tmax, h);
  const GeomDev &g = *s.g;
  return make_int4(s.primID, g.kind == BNAMD_GEOM_SPHERES ? 13 : s.instID, g.userID, __float_as_int(h.t));
}

// Geometry record of a hit (material lookups).
template <bool TL>
BN_D const GeomDev &hit_geom(const WorldDev &w, uint32_t flat) {
  if (!TL) return w.geoms[find_geom(w.geoms, w.numGeoms, flat)];
  const InstDev &I = w.insts[flat >> w.localBits];
  const uint32_t local = flat & ((1u << w.localBits) - 1u);
  return I.geoms[find_geom(I.geoms, I.numGeoms, local)];
}

// Closest hit (kAny = false) or occlusion (kAny = true) against the world.
// Two-level: the instance level is walked with the world ray; at an instance
// the ray goes to object space (W2O; the direction is not renormalized, so
// t is the same parameter in both spaces) and its group's BVH continues
// from the best hit so far, with tie keys prefixed by the instance index.
template <bool TL, bool kAny, class Filt = NoFilt>
BN_D bool world_traverse(const WorldDev &w, vec3f org, vec3f dir, float tmin, float tmax, HitRec &best,
                         const TraverseCtx &ctx, TraceCount *tc = nullptr, const Filt &filt = Filt()) {
  if (!TL) return traverse<kAny>(w.nodes, w.prims, org, dir, tmin, tmax, best, ctx, tc, filt);
  best.t = tmax;
  best.flat = 0xffffffffu;
  best.u = best.v = 0.f;
  const TraverseCtx top{ctx.stack, ctx.stride, w.tlasDepth + 1};
  const TraverseCtx sub{ctx.stack + (w.tlasDepth + 1) * ctx.stride, ctx.stride, ctx.maxDepth - (w.tlasDepth + 1)};
  const vec3f inv = rcp(dir);
  const uint32_t oct = ray_octant(dir);
  uint32_t grpBase = 0;
  uint32_t grpBits = (1u << oct) | (1u << 8);
  uint32_t primBase = 0, primMask = 0;
  int sp = 0;
  while (true) {
    if ((grpBits & 0xffu) != 0) {
      const uint32_t k = __builtin_ctz(grpBits);
      grpBits &= ~(1u << k);
      const uint32_t slot = k ^ oct;
      const uint32_t imask = grpBits >> 8;
      const uint32_t nodeIdx = grpBase + __builtin_popcount(imask & ((1u << slot) - 1u));
      if ((grpBits & 0xffu) != 0) {
        if (sp >= top.maxDepth) return kAny;
        top.stack[sp * top.stride] = make_uint2(grpBase, grpBits);
        ++sp;
      }
      bvh8_visit(w.nodes, nodeIdx, org, inv, oct, tmin, best.t, grpBase, grpBits, primBase, primMask);
      if (tc) ++tc->nodes;
    }
    while (primMask) {
      const uint32_t i = __builtin_ctz(primMask);
      primMask &= primMask - 1u;
      const uint32_t inst = __float_as_uint(w.prims[primBase + i].a.w);
      const InstDev &I = w.insts[inst];
      const vec3f o = xfm_point(I.w2o, org);
      const vec3f d = fix_dir(xfm_vector(I.w2o, dir));
      if (kAny) {
        HitRec tmp;
        tmp.t = tmax;
        tmp.flat = 0xffffffffu;
        if (traverse_from<true>(I.nodes, I.prims, o, d, tmin, tmax, tmp, sub, 0u, tc, filt)) return true;
      } else {
        traverse_from<false>(I.nodes, I.prims, o, d, tmin, tmax, best, sub, inst << w.localBits, tc, filt);
      }
    }
    if ((grpBits & 0xffu) == 0) {
      if (sp == 0) break;
      --sp;
      const uint2 e = top.stack[sp * top.stride];
      grpBase = e.x;
      grpBits = e.y;
    }
  }
  return !kAny;
}

}  // namespace bnamd
