// World construction (rtc::Device::createInstanceGroup + buildAccel) and the
// batch closest-hit trace entry point (rtc::TraceInterface::traceRay).
#include "world.h"
#include <hip/hip_fp16.h>
#include <cmath>
#include <algorithm>
#include <memory>

namespace bnamd {

namespace {

template <bool TL>
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
  if (!world_traverse<TL, false>(w, org, dir, r.tmin, r.tmax, h, ctx)) {
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
    const SurfaceHit s = surface_epilogue<TL>(w, org, dir, r.tmin, r.tmax, h);
    out.primID = s.primID;
    out.geomID = s.g->geomID;
    out.instID = s.instID;
    out.instIndex = s.instIndex;
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

namespace bnamd {
namespace {

void check_geom(const bnamd_geom &src, int numMaterials) {
  BN_REQUIRE((src.kind >= 0 && src.kind <= 2) || src.kind == BNAMD_GEOM_USER, "bnamd_world_build: unknown geom kind");
  BN_REQUIRE(src.numPrims >= 0, "bnamd_world_build: negative prim count");
  BN_REQUIRE(src.numPrims == 0 || src.vertices, "bnamd_world_build: geom without vertices");
  BN_REQUIRE(src.kind == BNAMD_GEOM_SPHERES || src.kind == BNAMD_GEOM_USER || src.numPrims == 0 || src.indices,
             "bnamd_world_build: geom without indices");
  BN_REQUIRE(src.kind != BNAMD_GEOM_USER || src.programs,
             "bnamd_world_build: a user geometry needs its program record (intersect program)");
  BN_REQUIRE(src.kind != BNAMD_GEOM_CYLINDERS || src.numPrims == 0 || src.radii,
             "bnamd_world_build: cylinders need per-primitive radii");
  BN_REQUIRE(src.materialID >= 0 && src.materialID < std::max(numMaterials, 1),
             "bnamd_world_build: materialID out of range");
}

GeomDev make_geom(const bnamd_geom &src, int geomID, int instIndex, int instID, uint64_t offset) {
  GeomDev g{};
  g.kind = src.kind;
  g.numPrims = src.numPrims;
  g.primOffset = (uint32_t)offset;
  g.geomID = geomID;
  g.instIndex = instIndex;
  g.instID = instID;
  g.materialID = src.materialID;
  g.userID = src.userID;
  g.vertices = src.vertices;
  g.indices = src.indices;
  g.radii = src.radii;
  g.colors = src.colors;
  g.normals = src.kind == BNAMD_GEOM_TRIANGLES ? src.normals : nullptr;
  g.defaultRadius = src.defaultRadius;
  g.programs = src.programs;
  return g;
}

bool has_user(const std::vector<GeomDev> &geoms) {
  for (const GeomDev &g : geoms)
    if (g.kind == BNAMD_GEOM_USER) return true;
  return false;
}

std::vector<MaterialDev> make_materials(const bnamd_material *materials, int numMaterials) {
  std::vector<MaterialDev> mats(std::max(numMaterials, 1));
  for (int m = 0; m < (int)mats.size(); ++m) {
    MaterialDev md{};
    if (m < numMaterials) {
      const bnamd_material &s = materials[m];
      md.type = s.type;
      md.color[0] = s.color[0]; md.color[1] = s.color[1]; md.color[2] = s.color[2];
      md.opacity = s.opacity; md.metallic = s.metallic; md.roughness = s.roughness;
      md.transmission = s.transmission; md.ior = s.ior; md.specular = s.specular;
      md.alpha = 1.f * s.opacity;
      if (s.type == BNAMD_MAT_PBR && !(s.ior != 1.f && s.transmission >= 1e-3f)) {
        // NVisii stores its parameters as halfs (NVisii.h:1062-1083);
        // roughness is clamped to [.001, .9] (AnariPBR.h:86)
        md.metallic = round_half(s.metallic);
        md.roughness = round_half(std::min(std::max(s.roughness, .001f), .9f));
        md.ior = round_half(s.ior);
        md.alpha = round_half(((1.f - s.transmission) * 1.f) * s.opacity);
      } else if (s.type == BNAMD_MAT_PBR) {
        md.alpha = 1.f;  // Glass
      }
    } else {
      md.type = BNAMD_MAT_MATTE;
      md.color[0] = md.color[1] = md.color[2] = 0.8f;
      md.opacity = 1.f;
      md.alpha = 1.f;
    }
    mats[m] = md;
  }
  return mats;
}

// rcp(affine3f) in double, rounded once to float: linear part inverted by
// the adjugate, p' = -(L^-1 p). m = columns vx, vy, vz, then p.
void invert_affine(const float *m, float *out) {
  const double a = m[0], b = m[3], c = m[6];  // row 0: vx.x vy.x vz.x
  const double d = m[1], e = m[4], f = m[7];
  const double g = m[2], h = m[5], i = m[8];
  const double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
  const double det = a * A + b * B + c * Cc;
  BN_REQUIRE(det != 0.0 && std::isfinite(det), "bnamd_world_build: singular instance transform");
  const double r = 1.0 / det;
  // inverse rows
  const double inv[3][3] = {{A * r, -(b * i - c * h) * r, (b * f - c * e) * r},
                            {B * r, (a * i - c * g) * r, -(a * f - c * d) * r},
                            {Cc * r, -(a * h - b * g) * r, (a * e - b * d) * r}};
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) out[3 * col + row] = (float)inv[row][col];
  for (int row = 0; row < 3; ++row)
    out[9 + row] = (float)-(inv[row][0] * m[9] + inv[row][1] * m[10] + inv[row][2] * m[11]);
}

// world box of an object box under o2w, rounded outward
void xfm_box(const float *m, const float *lo, const float *hi, float4 &blo, float4 &bhi) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int c = 0; c < 8; ++c) {
    const double p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
    for (int k = 0; k < 3; ++k) {
      const double v = m[k] * p[0] + m[3 + k] * p[1] + m[6 + k] * p[2] + m[9 + k];
      mn[k] = std::min(mn[k], v);
      mx[k] = std::max(mx[k], v);
    }
  }
  float l[3], u[3];
  for (int k = 0; k < 3; ++k) {
    l[k] = (float)mn[k];
    if ((double)l[k] > mn[k]) l[k] = std::nextafter(l[k], -INFINITY);
    u[k] = (float)mx[k];
    if ((double)u[k] < mx[k]) u[k] = std::nextafter(u[k], INFINITY);
  }
  blo = make_float4(l[0], l[1], l[2], 0.f);
  bhi = make_float4(u[0], u[1], u[2], 0.f);
}

// Two levels: one BVH8 per referenced group in object space, and a BVH8
// over the instances' world boxes (InstanceGroup::buildAccel,
// rtcore/cuda/cudaGroup.cu:144-258).
// One group's bottom level in object space (Group::buildAccel of a
// triangles or user-geometry group, rtcore/cuda/cudaGroup.cu:261-562). The
// material table is part of the build: a triangle primitive records whether
// its material makes it transparent to shadow rays (Glass), and the BVH
// build reads the table for that.
bool shadow_transparent(const MaterialDev &m) {
  return m.type == BNAMD_MAT_PBR && m.ior != 1.f && m.transmission >= 1e-3f;
}

// Whether some triangle's material makes its hits stochastic (alpha < 1,
// not glass): the BVH records carry PRIM_CUTOUT then.
bool has_cutouts(const std::vector<GeomDev> &geoms, const std::vector<MaterialDev> &mats) {
  for (const GeomDev &g : geoms)
    if (g.kind == BNAMD_GEOM_TRIANGLES && !shadow_transparent(mats[g.materialID]) && mats[g.materialID].alpha < 1.f)
      return true;
  return false;
}

std::unique_ptr<World> build_blas(const bnamd_group &grp, const std::vector<MaterialDev> &mats, int numMaterials,
                                  hipStream_t s) {
  auto b = std::make_unique<World>();
  uint64_t offset = 0;
  for (int gi = 0; gi < grp.numGeoms; ++gi) {
    const bnamd_geom &src = grp.geoms[gi];
    check_geom(src, numMaterials);
    if (src.numPrims == 0) continue;
    b->hgeoms.push_back(make_geom(src, gi, -1, 0, offset));
    offset += (uint64_t)src.numPrims;
  }
  BN_REQUIRE(offset < 0xfffffff0ull, "bnamd_world_build: too many primitives in a group");
  b->numPrims = (uint32_t)offset;
  b->materials.upload(mats.data(), mats.size(), s);
  build_world_bvh(*b, s);
  b->hmats = mats;
  b->hasCutouts = has_cutouts(b->hgeoms, mats);
  b->hasUser = has_user(b->hgeoms);
  return b;
}

// A bottom level built with one material table used in a world with
// another: the IDs must be in range, and every triangle geom must be as
// transparent to shadow rays under the world's table as its primitives
// recorded at the build.
void check_materials(const World &b, const std::vector<MaterialDev> &mats, int numMaterials) {
  for (const GeomDev &g : b.hgeoms) {
    BN_REQUIRE(g.materialID >= 0 && g.materialID < std::max(numMaterials, 1),
               "bnamd_world_build: materialID out of range");
    if (g.kind == BNAMD_GEOM_TRIANGLES) {
      BN_REQUIRE(shadow_transparent(b.hmats[g.materialID]) == shadow_transparent(mats[g.materialID]),
                 "bnamd_world_build_instances: a triangle geom's material is glass in one material table and "
                 "not in the other (its BLAS was built with a different table)");
      BN_REQUIRE(b.hmats[g.materialID].alpha == mats[g.materialID].alpha,
                 "bnamd_world_build_instances: a triangle geom's material opacity differs from the table its "
                 "BLAS was built with");
    }
  }
}

// The instance level over w.instBlas with the given transforms: instance
// records, their world boxes, and a BVH8 over those boxes
// (InstanceGroup::buildAccel, rtcore/cuda/cudaGroup.cu:144-258).
void assemble_tlas(World &w, const float *xfms, hipStream_t s) {
  const int numInstances = (int)w.instBlas.size();
  uint32_t maxPrims = 1;
  int maxDepth = 1;
  uint64_t total = 0, nodes = 0;
  std::vector<const World *> distinct;
  w.hasCutouts = false;
  w.hasUser = false;
  for (const World *b : w.instBlas) {
    w.hasCutouts = w.hasCutouts || b->hasCutouts;
    w.hasUser = w.hasUser || b->hasUser;
    maxPrims = std::max(maxPrims, b->numPrims);
    maxDepth = std::max(maxDepth, b->depth);
    total += b->numPrims;
    if (std::find(distinct.begin(), distinct.end(), b) == distinct.end()) {
      distinct.push_back(b);
      nodes += b->numNodes;
    }
  }
  int bits = 1;
  while (bits < 32 && (1ull << bits) < (uint64_t)maxPrims) ++bits;
  BN_REQUIRE(bits < 32 && (uint64_t)numInstances < (1ull << (32 - bits)),
             "bnamd_world_build: instances x primitives per group exceed the 32-bit hit key");
  w.localBits = bits;
  std::vector<InstDev> insts(numInstances);
  std::vector<float4> boxes(2ull * numInstances);
  for (int ii = 0; ii < numInstances; ++ii) {
    const World &b = *w.instBlas[ii];
    const float *xfm = xfms + 12 * (size_t)ii;
    InstDev &I = insts[ii];
    for (int k = 0; k < 12; ++k) I.o2w[k] = xfm[k];
    invert_affine(xfm, I.w2o);
    I.nodes = b.nodes.ptr;
    I.prims = b.prims.ptr;
    I.geoms = b.geoms.ptr;
    I.numGeoms = (int)b.hgeoms.size();
    I.instID = w.instIDs[ii];
    const float zero[3] = {0.f, 0.f, 0.f};
    xfm_box(xfm, b.numPrims ? b.lo : zero, b.numPrims ? b.hi : zero, boxes[2 * ii], boxes[2 * ii + 1]);
  }
  w.insts.upload(insts.data(), insts.size(), s);
  w.instBoxes.upload(boxes.data(), boxes.size(), s);
  GeomDev top{};
  top.kind = GEOM_INSTANCES;
  top.numPrims = numInstances;
  top.primOffset = 0;
  top.vertices = (const float *)w.instBoxes.ptr;
  w.hgeoms.assign(1, top);
  w.numPrims = (uint32_t)numInstances;
  w.twoLevel = true;
  build_world_bvh(w, s);
  w.tlasDepth = w.depth;
  w.depth = w.tlasDepth + maxDepth + 1;
  w.numBlas = (int)distinct.size();
  w.totalPrims = total;
  w.totalNodes = nodes + w.numNodes;
}

std::vector<float> instance_xfms(const bnamd_instance *instances, int numInstances) {
  std::vector<float> x(12 * (size_t)numInstances);
  for (int ii = 0; ii < numInstances; ++ii)
    for (int k = 0; k < 12; ++k) x[12 * (size_t)ii + k] = instances[ii].xfm[k];
  return x;
}

// Two levels: one BVH8 per referenced group in object space, and a BVH8
// over the instances' world boxes.
void build_two_level(World &w, const bnamd_group *groups, int numGroups, const bnamd_instance *instances,
                     int numInstances, const std::vector<MaterialDev> &mats, int numMaterials, hipStream_t s) {
  std::vector<int> blasOf(numGroups, -1);
  w.instBlas.clear();
  w.instIDs.clear();
  for (int ii = 0; ii < numInstances; ++ii) {
    const int gidx = instances[ii].group;
    if (blasOf[gidx] < 0) {
      blasOf[gidx] = (int)w.blas.size();
      w.blas.push_back(build_blas(groups[gidx], mats, numMaterials, s));
    }
    w.instBlas.push_back(w.blas[blasOf[gidx]].get());
    w.instIDs.push_back(instances[ii].instID);
  }
  w.materials.upload(mats.data(), mats.size(), s);
  assemble_tlas(w, instance_xfms(instances, numInstances).data(), s);
}

}  // namespace
}  // namespace bnamd

extern "C" int bnamd_world_build(const bnamd_group *groups, int numGroups, const bnamd_instance *instances,
                                 int numInstances, const bnamd_material *materials, int numMaterials,
                                 void *stream, bnamd_world **out) {
  return guarded([&] {
    BN_REQUIRE(out, "bnamd_world_build: out is NULL");
    BN_REQUIRE(numGroups >= 0 && numInstances >= 0, "bnamd_world_build: negative counts");
    BN_REQUIRE(numInstances == 0 || (groups && instances), "bnamd_world_build: NULL group or instance array");
    auto *wo = new bnamd_world_s;
    try {
      World &w = wo->w;
      hipStream_t s = (hipStream_t)stream;
      bool allIdentity = true;
      for (int ii = 0; ii < numInstances; ++ii) {
        BN_REQUIRE(instances[ii].group >= 0 && instances[ii].group < numGroups, "bnamd_world_build: bad group index");
        allIdentity = allIdentity && is_identity(instances[ii].xfm);
      }
      const std::vector<MaterialDev> mats = make_materials(materials, numMaterials);
      if (!allIdentity) {
        build_two_level(w, groups, numGroups, instances, numInstances, mats, numMaterials, s);
      } else {
        // identity instances: one flattened BVH8 over all primitives
        uint64_t offset = 0;
        for (int ii = 0; ii < numInstances; ++ii) {
          const bnamd_instance &inst = instances[ii];
          const bnamd_group &grp = groups[inst.group];
          for (int gi = 0; gi < grp.numGeoms; ++gi) {
            const bnamd_geom &src = grp.geoms[gi];
            check_geom(src, numMaterials);
            if (src.numPrims == 0) continue;
            w.hgeoms.push_back(make_geom(src, gi, ii, inst.instID, offset));
            offset += (uint64_t)src.numPrims;
          }
        }
        BN_REQUIRE(offset < 0xfffffff0ull, "bnamd_world_build: too many primitives");
        w.numPrims = (uint32_t)offset;
        w.materials.upload(mats.data(), mats.size(), s);
        build_world_bvh(w, s);
        w.hasCutouts = has_cutouts(w.hgeoms, mats);
        w.hasUser = has_user(w.hgeoms);
        w.totalPrims = w.numPrims;
        w.totalNodes = w.numNodes;
      }
      BN_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete wo;
      throw;
    }
    *out = wo;
  });
}

extern "C" int bnamd_world_device(const bnamd_world *world, const void **deviceRecord) {
  return guarded([&] {
    BN_REQUIRE(world && deviceRecord, "bnamd_world_device: NULL argument");
    const World &w = world->w;
    const WorldDev d = w.dd();
    w.devRecord.upload(&d, 1);
    BN_HIP(hipDeviceSynchronize());
    *deviceRecord = w.devRecord.ptr;
  });
}

extern "C" int bnamd_world_validate(const bnamd_world *world) {
  return guarded([&] {
    BN_REQUIRE(world, "bnamd_world_validate: NULL world");
    const World &w = world->w;
    BN_HIP(hipDeviceSynchronize());
    check_world_bvh(w);
    std::vector<const World *> seen;
    for (const World *b : w.instBlas)
      if (std::find(seen.begin(), seen.end(), b) == seen.end()) {
        seen.push_back(b);
        check_world_bvh(*b);
      }
  });
}

extern "C" int bnamd_world_destroy(bnamd_world *world) {
  return guarded([&] { delete world; });
}

struct bnamd_blas_s {
  std::unique_ptr<World> b;
};

extern "C" int bnamd_blas_build(const bnamd_group *group, const bnamd_material *materials, int numMaterials,
                                void *stream, bnamd_blas **out) {
  return guarded([&] {
    BN_REQUIRE(group && out, "bnamd_blas_build: NULL argument");
    BN_REQUIRE(group->numGeoms >= 0 && (group->numGeoms == 0 || group->geoms), "bnamd_blas_build: bad group");
    BN_REQUIRE(numMaterials >= 0 && (numMaterials == 0 || materials), "bnamd_blas_build: bad material table");
    hipStream_t s = (hipStream_t)stream;
    auto *bo = new bnamd_blas_s;
    try {
      bo->b = build_blas(*group, make_materials(materials, numMaterials), numMaterials, s);
      BN_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete bo;
      throw;
    }
    *out = bo;
  });
}

extern "C" int bnamd_blas_destroy(bnamd_blas *blas) {
  return guarded([&] { delete blas; });
}

extern "C" int bnamd_world_build_instances(const bnamd_blas *const *blas, int numBlas,
                                           const bnamd_instance *instances, int numInstances,
                                           const bnamd_material *materials, int numMaterials, void *stream,
                                           bnamd_world **out) {
  return guarded([&] {
    BN_REQUIRE(out, "bnamd_world_build_instances: out is NULL");
    BN_REQUIRE(numBlas >= 0 && numInstances >= 0, "bnamd_world_build_instances: negative counts");
    BN_REQUIRE(numInstances == 0 || (blas && instances), "bnamd_world_build_instances: NULL array");
    auto *wo = new bnamd_world_s;
    try {
      World &w = wo->w;
      hipStream_t s = (hipStream_t)stream;
      const std::vector<MaterialDev> mats = make_materials(materials, numMaterials);
      for (int ii = 0; ii < numInstances; ++ii) {
        const int g = instances[ii].group;
        BN_REQUIRE(g >= 0 && g < numBlas && blas[g], "bnamd_world_build_instances: bad group index");
        check_materials(*blas[g]->b, mats, numMaterials);
        w.instBlas.push_back(blas[g]->b.get());
        w.instIDs.push_back(instances[ii].instID);
      }
      w.materials.upload(mats.data(), mats.size(), s);
      assemble_tlas(w, instance_xfms(instances, numInstances).data(), s);
      BN_HIP(hipStreamSynchronize(s));
    } catch (...) {
      delete wo;
      throw;
    }
    *out = wo;
  });
}

extern "C" int bnamd_world_set_transforms(bnamd_world *world, const float *xfms, int numInstances, void *stream) {
  return guarded([&] {
    BN_REQUIRE(world && (xfms || numInstances == 0), "bnamd_world_set_transforms: NULL argument");
    World &w = world->w;
    BN_REQUIRE(w.twoLevel, "bnamd_world_set_transforms: the world has no instance level (built flattened)");
    BN_REQUIRE(numInstances == (int)w.instBlas.size(), "bnamd_world_set_transforms: one transform per instance");
    hipStream_t s = (hipStream_t)stream;
    assemble_tlas(w, xfms, s);
    BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_world_info_get(const bnamd_world *world, bnamd_world_info *info) {
  return guarded([&] {
    BN_REQUIRE(world && info, "bnamd_world_info_get: NULL argument");
    info->numPrims = (long long)world->w.totalPrims;
    info->numNodes = (long long)world->w.totalNodes;
    info->depth = world->w.depth;
    info->buildMs = world->w.buildMs;
    info->numBlas = world->w.twoLevel ? world->w.numBlas : 0;
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
    BN_REQUIRE(!world->w.hasUser, "bnamd_trace: the world holds user geometry (BNAMD_GEOM_USER): only the "
                                  "programmable traversal (rtc::amd::TraceInterface::traceRay) can trace it");
    const int block = 256;
    const size_t lds = trace_lds_bytes(world->w.depth, block);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_trace: BVH too deep for the LDS stack");
    hipStream_t s = (hipStream_t)stream;
    if (world->w.twoLevel)
      k_trace<true><<<div_up(numRays, block), block, lds, s>>>(world->w.dd(), rays, hits, numRays);
    else
      k_trace<false><<<div_up(numRays, block), block, lds, s>>>(world->w.dd(), rays, hits, numRays);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

// Used by other translation units (world accessors).
namespace bnamd {
// The world of a fused renderer: its programs are the native ones, so a
// world with user geometry (whose hits only its intersect program finds)
// is traced by rtc::amd::TraceInterface instead.
const World &world_of(const bnamd_world *w) {
  BN_REQUIRE(!w->w.hasUser, "the world holds user geometry (BNAMD_GEOM_USER): only the programmable traversal "
                            "(rtc::amd::TraceInterface::traceRay) can trace it");
  return w->w;
}
}
