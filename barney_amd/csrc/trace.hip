// World construction (rtc::Device::createInstanceGroup + buildAccel) and the
// batch closest-hit trace entry point (rtc::TraceInterface::traceRay).
#include "world.h"
#include <hip/hip_fp16.h>
#include <cmath>
#include <algorithm>

namespace bnamd {

namespace {

__global__ void __launch_bounds__(256) k_trace(WorldDev w, const bnamd_ray *__restrict rays,
                                               bnamd_hit *__restrict hits, int n) {
  extern __shared__ uint2 lds_stack[];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bnamd_ray r = rays[i];
  const vec3f org = {r.org[0], r.org[1], r.org[2]};
  const vec3f dir = fix_dir({r.dir[0], r.dir[1], r.dir[2]});
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, w.depth + 1};
  HitRec h;
  bnamd_hit out;
  if (!traverse_closest(w.nodes, w.prims, org, dir, r.tmin, r.tmax, h, ctx)) {
    h.flat = 0xffffffffu;
    h.t = __builtin_nanf("");
  }
  out.t = h.t;
  out.u = h.u;
  out.v = h.v;
  if (h.flat == 0xffffffffu) {
    out.primID = -1;
    out.geomID = -1;
    out.instID = -1;
    out.instIndex = -1;
    out.N[0] = out.N[1] = out.N[2] = 0.f;
  } else {
    const SurfaceHit s = surface_epilogue(w, org, dir, r.tmin, r.tmax, h);
    const GeomDev &g = w.geoms[s.geom];
    out.primID = s.primID;
    out.geomID = g.geomID;
    out.instID = g.instID;
    out.instIndex = g.instIndex;
    out.N[0] = s.Ng.x;
    out.N[1] = s.Ng.y;
    out.N[2] = s.Ng.z;
  }
  hits[i] = out;
}

float round_half(float f) { return __half2float(__float2half(f)); }

bool is_identity(const float *x) {
  static const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  for (int k = 0; k < 12; ++k)
    if (x[k] != id[k]) return false;
  return true;
}

}  // namespace

size_t trace_lds_bytes(int depth, int block) { return (size_t)(depth + 1) * block * sizeof(uint2); }

}  // namespace bnamd

using namespace bnamd;

struct bnamd_world_s {
  World w;
};

extern "C" int bnamd_world_build(const bnamd_group *groups, int numGroups, const bnamd_instance *instances,
                                 int numInstances, const bnamd_material *materials, int numMaterials,
                                 void *stream, bnamd_world **out) {
  return guarded([&] {
    BN_REQUIRE(out, "bnamd_world_build: out is NULL");
    BN_REQUIRE(numGroups >= 0 && numInstances >= 0, "bnamd_world_build: negative counts");
    auto *wo = new bnamd_world_s;
    try {
      World &w = wo->w;
      uint64_t offset = 0;
      for (int ii = 0; ii < numInstances; ++ii) {
        const bnamd_instance &inst = instances[ii];
        BN_REQUIRE(inst.group >= 0 && inst.group < numGroups, "bnamd_world_build: bad group index");
        // Instances are flattened into one BVH8; this is exact only for
        // identity transforms (object space == world space).
        BN_REQUIRE(is_identity(inst.xfm),
                   "bnamd_world_build: only identity instance transforms are supported");
        const bnamd_group &grp = groups[inst.group];
        for (int gi = 0; gi < grp.numGeoms; ++gi) {
          const bnamd_geom &src = grp.geoms[gi];
          BN_REQUIRE(src.kind >= 0 && src.kind <= 2, "bnamd_world_build: unknown geom kind");
          BN_REQUIRE(src.numPrims >= 0, "bnamd_world_build: negative prim count");
          BN_REQUIRE(src.numPrims == 0 || src.vertices, "bnamd_world_build: geom without vertices");
          BN_REQUIRE(src.kind == BNAMD_GEOM_SPHERES || src.numPrims == 0 || src.indices,
                     "bnamd_world_build: geom without indices");
          BN_REQUIRE(src.kind != BNAMD_GEOM_CYLINDERS || src.numPrims == 0 || src.radii,
                     "bnamd_world_build: cylinders need per-primitive radii");
          BN_REQUIRE(src.materialID >= 0 && src.materialID < std::max(numMaterials, 1),
                     "bnamd_world_build: materialID out of range");
          if (src.numPrims == 0) continue;
          GeomDev g{};
          g.kind = src.kind;
          g.numPrims = src.numPrims;
          g.primOffset = (uint32_t)offset;
          g.geomID = gi;
          g.instIndex = ii;
          g.instID = inst.instID;
          g.materialID = src.materialID;
          g.userID = src.userID;
          g.vertices = src.vertices;
          g.indices = src.indices;
          g.radii = src.radii;
          g.colors = src.colors;
          g.defaultRadius = src.defaultRadius;
          w.hgeoms.push_back(g);
          offset += (uint64_t)src.numPrims;
        }
      }
      BN_REQUIRE(offset < 0xfffffff0ull, "bnamd_world_build: too many primitives");
      w.numPrims = (uint32_t)offset;
      std::vector<MaterialDev> mats(std::max(numMaterials, 1));
      for (int m = 0; m < (int)mats.size(); ++m) {
        MaterialDev md{};
        if (m < numMaterials) {
          const bnamd_material &s = materials[m];
          md.type = s.type;
          md.color[0] = s.color[0]; md.color[1] = s.color[1]; md.color[2] = s.color[2];
          md.opacity = s.opacity; md.metallic = s.metallic; md.roughness = s.roughness;
          md.transmission = s.transmission; md.ior = s.ior; md.specular = s.specular;
          if (s.type == BNAMD_MAT_PBR && !(s.ior != 1.f && s.transmission >= 1e-3f)) {
            // NVisii stores its parameters as halfs (NVisii.h:1062-1083);
            // roughness is clamped to [.001, .9] (AnariPBR.h:86)
            md.metallic = round_half(s.metallic);
            md.roughness = round_half(std::min(std::max(s.roughness, .001f), .9f));
            md.ior = round_half(s.ior);
          }
        } else {
          md.type = BNAMD_MAT_MATTE;
          md.color[0] = md.color[1] = md.color[2] = 0.8f;
          md.opacity = 1.f;
        }
        mats[m] = md;
      }
      hipStream_t s = (hipStream_t)stream;
      w.materials.upload(mats.data(), mats.size(), s);
      build_world_bvh(w, s);
      BN_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete wo;
      throw;
    }
    *out = wo;
  });
}

extern "C" int bnamd_world_destroy(bnamd_world *world) {
  return guarded([&] { delete world; });
}

extern "C" int bnamd_world_info_get(const bnamd_world *world, bnamd_world_info *info) {
  return guarded([&] {
    BN_REQUIRE(world && info, "bnamd_world_info_get: NULL argument");
    info->numPrims = world->w.numPrims;
    info->numNodes = world->w.numNodes;
    info->depth = world->w.depth;
    info->buildMs = world->w.buildMs;
    for (int k = 0; k < 3; ++k) {
      info->sceneLower[k] = world->w.lo[k];
      info->sceneUpper[k] = world->w.hi[k];
    }
  });
}

extern "C" int bnamd_trace(const bnamd_world *world, const bnamd_ray *rays, bnamd_hit *hits, int numRays,
                           void *stream) {
  return guarded([&] {
    BN_REQUIRE(world, "bnamd_trace: world is NULL");
    BN_REQUIRE(numRays >= 0, "bnamd_trace: negative ray count");
    if (numRays == 0) return;
    BN_REQUIRE(rays && hits, "bnamd_trace: NULL ray or hit buffer");
    const int block = 256;
    const size_t lds = trace_lds_bytes(world->w.depth, block);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_trace: BVH too deep for the LDS stack");
    hipStream_t s = (hipStream_t)stream;
    k_trace<<<div_up(numRays, block), block, lds, s>>>(world->w.dd(), rays, hits, numRays);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

// Used by other translation units (world accessors).
namespace bnamd {
const World &world_of(const bnamd_world *w) { return w->w; }
}
