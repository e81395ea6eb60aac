// rtc::amd -- device side of barney's rtcore backend for MI355X (gfx950),
// for barney's own device programs and kernels compiled with hipcc.
//
// It restates the device interface barney programs against:
//   ComputeInterface, fatomicMin / fatomicMax   rtcore/cudaCommon/ComputeInterface.h:15-60
//   RTC_EXPORT_COMPUTE1D/2D/3D, RTC_IMPORT_*    ComputeInterface.h:180-264, ComputeKernel.h:35-44
//   TraceInterface (the accessors programs     rtcore/cuda/ProgramInterface.h:20-120
//     call, traceRay)                          rtcore/cuda/TraceInterface.h:102-230
//   RTC_EXPORT_TRACE2D, RTC_IMPORT_TRACE2D     TraceInterface.h:509-536, TraceKernel.h:33-43
//   RTC_EXPORT_USER_GEOM,
//   RTC_EXPORT_TRIANGLES_GEOM                  ProgramInterface.h:150-260
// with the same names, argument meanings and call semantics, over this
// backend's BVH8s. traceRay walks the world's 8-wide BVHs (instance level,
// then the instance's bottom level in object space) with the native slab
// test; a candidate is a triangle (this backend's triangle test,
// rtcore/cuda/TraceInterface.h:32-98), a native sphere or cylinder, or a
// user primitive whose intersect program reports a closer t; its anyHit
// program may ignore it; the accepted candidate shrinks the ray; at the end
// the closestHit program of the accepted hit runs (TraceInterface.h:135-230).
// Program addresses travel in each geometry's ProgramRecord (rtcore_amd.h),
// read by a one-thread kernel of the module that defines the programs, so
// the traversal (inlined into the calling kernel) and the programs share
// that module's code object, as the reference's do.
//
// Vector types are rtc::amd's (rtcore_amd.h); a build that has barney's own
// vector types makes them aliases of those (same layouts).
#pragma once
#include <hip/hip_runtime.h>

#include "rtcore_amd.h"
#include "../csrc/world.h"
#include "../csrc/envmap.h"
#include "../csrc/volume.h"

namespace rtc {
namespace amd {

struct vec2f {
  float x, y;
};

// ------------------------------------------------------------------ textures
// tex2D / tex3D on a TextureObject (the device address of a
// TextureObjectDD, Texture::getDD()), sampled as the embree backend's
// TextureSampler does (rtcore/embree/Texture.cpp): tex2D with normalized
// coordinates and the texture's filter and per-axis addressing (the
// native renderers' sampler, envmap.h); tex3D with unnormalized
// coordinates, texel centres at +0.5 and clamping (lerpCoords_notNormalized
// and the trilinear blend of Texture.cpp:324-390, as the structured-volume
// sampler of volume.h), FLOAT texels.
__device__ inline float4 tex2d_sample(TextureObject t, float u, float v) {
  const TextureObjectDD &o = *(const TextureObjectDD *)t;
  bnamd::Tex2DDD d;
  d.texels = (const float4 *)o.texels;
  d.w = o.w;
  d.h = o.h;
  d.filter = o.filter;
  d.addr[0] = o.addr[0];
  d.addr[1] = o.addr[1];
  d.pad = 0;
  return bnamd::tex2d(d, u, v);
}

__device__ inline float tex3d_sample(TextureObject t, float x, float y, float z) {
#pragma clang fp contract(off)
  const TextureObjectDD &o = *(const TextureObjectDD *)t;
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  bnamd::lerp_coords(o.w, x, x0, x1, fx);
  bnamd::lerp_coords(o.h, y, y0, y1, fy);
  bnamd::lerp_coords(o.d, z, z0, z1, fz);
  const float *s = (const float *)o.texels;
  const size_t sy = (size_t)o.w, sz = (size_t)o.w * o.h;
  auto at = [&](int i, int j, int k) { return s[(size_t)i + sy * j + sz * k]; };
  const float v00 = bnamd::lerp_l(fx, at(x0, y0, z0), at(x1, y0, z0));
  const float v01 = bnamd::lerp_l(fx, at(x0, y1, z0), at(x1, y1, z0));
  const float v10 = bnamd::lerp_l(fx, at(x0, y0, z1), at(x1, y0, z1));
  const float v11 = bnamd::lerp_l(fx, at(x0, y1, z1), at(x1, y1, z1));
  const float v0 = bnamd::lerp_l(fy, v00, v01);
  const float v1 = bnamd::lerp_l(fy, v10, v11);
  return bnamd::lerp_l(fz, v0, v1);
}

template <class T>
__device__ T tex2D(TextureObject t, float u, float v);
template <>
__device__ inline vec4f tex2D<vec4f>(TextureObject t, float u, float v) {
  const float4 c = tex2d_sample(t, u, v);
  return {c.x, c.y, c.z, c.w};
}
template <>
__device__ inline float tex2D<float>(TextureObject t, float u, float v) {
  return tex2d_sample(t, u, v).x;
}

template <class T>
__device__ T tex3D(TextureObject t, float x, float y, float z);
template <>
__device__ inline float tex3D<float>(TextureObject t, float x, float y, float z) {
  return tex3d_sample(t, x, y, z);
}
template <>
__device__ inline vec4f tex3D<vec4f>(TextureObject t, float x, float y, float z) {
  const float v = tex3d_sample(t, x, y, z);  // FLOAT texels read as (v, v, v, v)
  return {v, v, v, v};
}

// ------------------------------------------------------------------ compute
struct ComputeInterface {
  __device__ vec3ui launchIndex() const {
    return {threadIdx.x + blockIdx.x * blockDim.x, threadIdx.y + blockIdx.y * blockDim.y,
            threadIdx.z + blockIdx.z * blockDim.z};
  }
  __device__ vec3ui getThreadIdx() const { return {threadIdx.x, threadIdx.y, threadIdx.z}; }
  __device__ vec3ui getBlockDim() const { return {blockDim.x, blockDim.y, blockDim.z}; }
  __device__ vec3ui getBlockIdx() const { return {blockIdx.x, blockIdx.y, blockIdx.z}; }
  __device__ int atomicAdd(int *ptr, int inc) const { return ::atomicAdd(ptr, inc); }
  __device__ float atomicAdd(float *ptr, float inc) const { return ::atomicAdd(ptr, inc); }
};

// float min / max by compare-and-swap on the bits (order-independent)
__device__ inline void fatomicMin(float *addr, float value) {
  float old = *(volatile float *)addr;
  while (!(old <= value)) {
    const int found = atomicCAS((int *)addr, __float_as_int(old), __float_as_int(value));
    if (found == __float_as_int(old)) return;
    old = __int_as_float(found);
  }
}
__device__ inline void fatomicMax(float *addr, float value) {
  float old = *(volatile float *)addr;
  while (!(old >= value)) {
    const int found = atomicCAS((int *)addr, __float_as_int(old), __float_as_int(value));
    if (found == __float_as_int(old)) return;
    old = __int_as_float(found);
  }
}

// -------------------------------------------------------------------- trace
struct TraceInterface;
typedef void (*ProgramFct)(TraceInterface &);

struct TraceInterface {
  // the interface pipeline programs use (ProgramInterface.h:20-120)
  __device__ void ignoreIntersection() { rejectThisHit = true; }
  __device__ void reportIntersection(float t, int) { current.tMax = t; }
  __device__ void *getPRD() const { return prd; }
  __device__ const void *getProgramData() const { return geomData; }
  __device__ const void *getLPData() const { return lpData; }
  __device__ vec3i getLaunchDims() const {
    return {(int)(blockDim.x * gridDim.x), (int)(blockDim.y * gridDim.y), (int)(blockDim.z * gridDim.z)};
  }
  __device__ vec3i getLaunchIndex() const {
    return {(int)(threadIdx.x + blockIdx.x * blockDim.x), (int)(threadIdx.y + blockIdx.y * blockDim.y),
            (int)(threadIdx.z + blockIdx.z * blockDim.z)};
  }
  __device__ vec2f getTriangleBarycentrics() const { return current.bary; }
  __device__ int getGeometryIndex() const { return current.geomID; }
  __device__ int getPrimitiveIndex() const { return current.primID; }
  // the user ID of the current instance, and its index in the world
  __device__ int getInstanceID() const { return currentInstanceID; }
  __device__ int getRTCInstanceIndex() const { return current.instID; }
  __device__ float getRayTmax() const { return current.tMax; }
  __device__ float getRayTmin() const { return tMin; }
  __device__ vec3f getObjectRayDirection() const { return object.dir; }
  __device__ vec3f getObjectRayOrigin() const { return object.org; }
  __device__ vec3f getWorldRayDirection() const { return world.dir; }
  __device__ vec3f getWorldRayOrigin() const { return world.org; }
  // the current instance's transforms (identity in a flattened world); as
  // the reference, normals go through the linear part of the transform
  __device__ vec3f transformNormalFromObjectToWorldSpace(vec3f v) const { return xvec(o2w, v); }
  __device__ vec3f transformPointFromObjectToWorldSpace(vec3f v) const { return xpt(o2w, v); }
  __device__ vec3f transformVectorFromObjectToWorldSpace(vec3f v) const { return xvec(o2w, v); }
  __device__ vec3f transformNormalFromWorldToObjectSpace(vec3f v) const { return xvec(w2o, v); }
  __device__ vec3f transformPointFromWorldToObjectSpace(vec3f v) const { return xpt(w2o, v); }
  __device__ vec3f transformVectorFromWorldToObjectSpace(vec3f v) const { return xvec(w2o, v); }

  // closest hit of the ray [t0, t1] in the world 'accel'
  // (getAccelHandle), with the programs of the geometries it meets; prd is
  // what the programs' getPRD() returns
  __device__ void traceRay(AccelHandle accel, vec3f org, vec3f dir, float t0, float t1, void *prdPtr);

  // launch parameters
  const void *lpData = nullptr;
  // ray and traversal state
  void *prd = nullptr;
  const void *geomData = nullptr;
  float tMin = 0.f;
  struct Hit {
    vec2f bary;
    int primID, geomID, instID;
    float tMax;
  } current{}, accepted{};
  struct Ray {
    vec3f org, dir;
  } world{}, object{};
  const float *o2w = nullptr, *w2o = nullptr;  // current instance (null: identity)
  int currentInstanceID = 0;
  bool rejectThisHit = false;
  const ProgramRecord *acceptedSBT = nullptr;
  const float *acceptedO2W = nullptr, *acceptedW2O = nullptr;
  int acceptedInstanceID = 0;

 private:
  static __device__ vec3f xpt(const float *m, vec3f v) {
    if (!m) return v;
    const bnamd::vec3f r = bnamd::xfm_point(m, {v.x, v.y, v.z});
    return {r.x, r.y, r.z};
  }
  static __device__ vec3f xvec(const float *m, vec3f v) {
    if (!m) return v;
    const bnamd::vec3f r = bnamd::xfm_vector(m, {v.x, v.y, v.z});
    return {r.x, r.y, r.z};
  }
  // one candidate primitive of a bottom level
  __device__ void candidate(const bnamd::PrimRec *__restrict prims, uint32_t idx, const bnamd::GeomDev *geoms,
                            int numGeoms, int instIndex, int instID);
  // the closest-hit walk of one bottom level with the object-space ray
  __device__ void walkBlas(const bnamd::BNode8 *nodes, const bnamd::PrimRec *prims, const bnamd::GeomDev *geoms,
                           int numGeoms, int instIndex, int instID);
};

// BVH8 levels a walk can stack (the builder's collapse allows 64)
constexpr int kTraceStack = 64;

__device__ inline void TraceInterface::candidate(const bnamd::PrimRec *__restrict prims, uint32_t idx,
                                                 const bnamd::GeomDev *geoms, int numGeoms, int instIndex,
                                                 int instID) {
  const uint32_t bits = __float_as_uint(prims[idx].c.w);
  const uint32_t type = bits & 0xffu;
  const uint32_t flat = __float_as_uint(prims[idx].a.w);
  const int g = bnamd::find_geom(geoms, numGeoms, flat);
  const bnamd::GeomDev &gd = geoms[g];
  current.primID = (int)(flat - gd.primOffset);
  current.geomID = gd.geomID;
  current.instID = instIndex >= 0 ? instIndex : gd.instIndex;
  currentInstanceID = instIndex >= 0 ? instID : gd.instID;
  current.tMax = accepted.tMax;
  const ProgramRecord *rec = (const ProgramRecord *)gd.programs;
  geomData = rec ? (const void *)(rec + 1) : nullptr;
  if (type == bnamd::PRIM_USER) {
    if (!rec || !rec->intersect) return;
    reinterpret_cast<ProgramFct>(rec->intersect)(*this);
    if (!(current.tMax < accepted.tMax)) return;
  } else {
    // native primitive tests in object space, strictly closer than the
    // accepted hit (ties keep the earlier candidate, as the reference)
    bnamd::HitRec h;
    h.t = accepted.tMax;
    h.flat = 0xffffffffu;
    h.u = h.v = 0.f;
    bnamd::test_prim(prims, idx, {object.org.x, object.org.y, object.org.z},
                     {object.dir.x, object.dir.y, object.dir.z}, tMin, accepted.tMax, h);
    if (h.flat == 0xffffffffu) return;
    current.tMax = h.t;
    current.bary = type == bnamd::PRIM_TRIANGLE ? vec2f{h.u, h.v} : vec2f{0.f, 0.f};
  }
  rejectThisHit = false;
  if (rec && rec->anyHit) reinterpret_cast<ProgramFct>(rec->anyHit)(*this);
  if (!rejectThisHit) {
    accepted = current;
    acceptedSBT = rec;
    acceptedO2W = o2w;
    acceptedW2O = w2o;
    acceptedInstanceID = currentInstanceID;
  }
}

__device__ inline void TraceInterface::walkBlas(const bnamd::BNode8 *nodes, const bnamd::PrimRec *prims,
                                                const bnamd::GeomDev *geoms, int numGeoms, int instIndex,
                                                int instID) {
  const bnamd::vec3f O = {object.org.x, object.org.y, object.org.z};
  const bnamd::vec3f D = {object.dir.x, object.dir.y, object.dir.z};
  const bnamd::vec3f inv = bnamd::rcp(D);
  const uint32_t oct = bnamd::ray_octant(D);
  uint2 stack[kTraceStack];
  int sp = 0;
  uint32_t grpBase = 0, grpBits = (1u << oct) | (1u << 8), primBase = 0, primMask = 0;
  while (true) {
    if ((grpBits & 0xffu) != 0) {
      const uint32_t k = __builtin_ctz(grpBits);
      grpBits &= ~(1u << k);
      const uint32_t slot = k ^ oct;
      const uint32_t imask = grpBits >> 8;
      const uint32_t nodeIdx = grpBase + __builtin_popcount(imask & ((1u << slot) - 1u));
      if ((grpBits & 0xffu) != 0 && sp < kTraceStack) stack[sp++] = make_uint2(grpBase, grpBits);
      bnamd::bvh8_visit(nodes, nodeIdx, O, inv, oct, tMin, accepted.tMax, grpBase, grpBits, primBase, primMask);
    }
    while (primMask) {
      const uint32_t i = __builtin_ctz(primMask);
      primMask &= primMask - 1u;
      candidate(prims, primBase + i, geoms, numGeoms, instIndex, instID);
    }
    if ((grpBits & 0xffu) == 0) {
      if (sp == 0) break;
      --sp;
      grpBase = stack[sp].x;
      grpBits = stack[sp].y;
    }
  }
}

__device__ inline void TraceInterface::traceRay(AccelHandle accel, vec3f org, vec3f dir, float t0, float t1,
                                                void *prdPtr) {
  // as the reference: no direction component below 1e-6 in magnitude
  if (fabsf(dir.x) < 1e-6f) dir.x = 1e-6f;
  if (fabsf(dir.y) < 1e-6f) dir.y = 1e-6f;
  if (fabsf(dir.z) < 1e-6f) dir.z = 1e-6f;
  prd = prdPtr;
  world.org = org;
  world.dir = dir;
  tMin = t0;
  accepted = Hit{};
  accepted.tMax = t1;
  accepted.primID = -1;
  accepted.geomID = -1;
  accepted.instID = -1;
  acceptedSBT = nullptr;
  acceptedO2W = acceptedW2O = nullptr;
  const bnamd::WorldDev &W = *(const bnamd::WorldDev *)accel;
  if (!W.insts) {
    // a flattened world: object space is world space
    o2w = w2o = nullptr;
    object = world;
    walkBlas(W.nodes, W.prims, W.geoms, W.numGeoms, -1, 0);
  } else {
    // the instance level with the world ray; each instance met walks its
    // bottom level with the ray in its object space
    const bnamd::vec3f O = {org.x, org.y, org.z}, D = {dir.x, dir.y, dir.z};
    const bnamd::vec3f inv = bnamd::rcp(D);
    const uint32_t oct = bnamd::ray_octant(D);
    uint2 stack[kTraceStack];
    int sp = 0;
    uint32_t grpBase = 0, grpBits = (1u << oct) | (1u << 8), primBase = 0, primMask = 0;
    while (true) {
      if ((grpBits & 0xffu) != 0) {
        const uint32_t k = __builtin_ctz(grpBits);
        grpBits &= ~(1u << k);
        const uint32_t slot = k ^ oct;
        const uint32_t imask = grpBits >> 8;
        const uint32_t nodeIdx = grpBase + __builtin_popcount(imask & ((1u << slot) - 1u));
        if ((grpBits & 0xffu) != 0 && sp < kTraceStack) stack[sp++] = make_uint2(grpBase, grpBits);
        bnamd::bvh8_visit(W.nodes, nodeIdx, O, inv, oct, tMin, accepted.tMax, grpBase, grpBits, primBase, primMask);
      }
      while (primMask) {
        const uint32_t i = __builtin_ctz(primMask);
        primMask &= primMask - 1u;
        const uint32_t inst = __float_as_uint(W.prims[primBase + i].a.w);
        const bnamd::InstDev &I = W.insts[inst];
        o2w = I.o2w;
        w2o = I.w2o;
        const bnamd::vec3f oo = bnamd::xfm_point(I.w2o, O);
        const bnamd::vec3f od = bnamd::fix_dir(bnamd::xfm_vector(I.w2o, D));
        object.org = {oo.x, oo.y, oo.z};
        object.dir = {od.x, od.y, od.z};
        walkBlas(I.nodes, I.prims, I.geoms, I.numGeoms, (int)inst, I.instID);
      }
      if ((grpBits & 0xffu) == 0) {
        if (sp == 0) break;
        --sp;
        grpBase = stack[sp].x;
        grpBits = stack[sp].y;
      }
    }
  }
  if (acceptedSBT && acceptedSBT->closestHit) {
    current = accepted;
    geomData = acceptedSBT + 1;
    o2w = acceptedO2W;
    w2o = acceptedW2O;
    currentInstanceID = acceptedInstanceID;
    if (o2w) {  // the accepted instance's object-space ray
      const bnamd::vec3f oo = bnamd::xfm_point(w2o, {world.org.x, world.org.y, world.org.z});
      const bnamd::vec3f od = bnamd::fix_dir(bnamd::xfm_vector(w2o, {world.dir.x, world.dir.y, world.dir.z}));
      object.org = {oo.x, oo.y, oo.z};
      object.dir = {od.x, od.y, od.z};
    } else {
      object = world;
    }
    reinterpret_cast<ProgramFct>(acceptedSBT->closestHit)(*this);
  }
}

}  // namespace amd
}  // namespace rtc

// ------------------------------------------------------------------ exports
// Compute kernels: RTC_EXPORT_COMPUTE<N>D(name, className) defines the
// kernel running className::run(ComputeInterface&) on the kernel data and
// createCompute_<name>(Device*); RTC_IMPORT_COMPUTE<N>D declares the latter.
#define RTC_AMD_COMPUTE_KERNEL(name, className)                  \
  __global__ void rtc_amd_compute_##name(className dd) {         \
    ::rtc::amd::ComputeInterface ci;                             \
    dd.run(ci);                                                  \
  }

#define RTC_EXPORT_COMPUTE1D(name, className)                                                   \
  RTC_AMD_COMPUTE_KERNEL(name, className)                                                       \
  struct rtc_amd_compute1D_##name : public ::rtc::amd::ComputeKernel1D {                        \
    void launch(unsigned int nb, unsigned int bs, const void *dd) override {                    \
      if (nb > 0) rtc_amd_compute_##name<<<nb, bs>>>(*(const className *)dd);                   \
    }                                                                                           \
  };                                                                                            \
  ::rtc::amd::ComputeKernel1D *createCompute_##name(::rtc::amd::Device *) {                     \
    return new rtc_amd_compute1D_##name;                                                        \
  }

#define RTC_EXPORT_COMPUTE2D(name, className)                                                   \
  RTC_AMD_COMPUTE_KERNEL(name, className)                                                       \
  struct rtc_amd_compute2D_##name : public ::rtc::amd::ComputeKernel2D {                        \
    void launch(::rtc::amd::vec2ui nb, ::rtc::amd::vec2ui bs, const void *dd) override {        \
      if (nb.x > 0 && nb.y > 0)                                                                 \
        rtc_amd_compute_##name<<<dim3(nb.x, nb.y), dim3(bs.x, bs.y)>>>(*(const className *)dd);  \
    }                                                                                           \
  };                                                                                            \
  ::rtc::amd::ComputeKernel2D *createCompute_##name(::rtc::amd::Device *) {                     \
    return new rtc_amd_compute2D_##name;                                                        \
  }

#define RTC_EXPORT_COMPUTE3D(name, className)                                                       \
  RTC_AMD_COMPUTE_KERNEL(name, className)                                                           \
  struct rtc_amd_compute3D_##name : public ::rtc::amd::ComputeKernel3D {                            \
    void launch(::rtc::amd::vec3ui nb, ::rtc::amd::vec3ui bs, const void *dd) override {            \
      if (nb.x > 0 && nb.y > 0 && nb.z > 0)                                                         \
        rtc_amd_compute_##name<<<dim3(nb.x, nb.y, nb.z), dim3(bs.x, bs.y, bs.z)>>>(                 \
            *(const className *)dd);                                                                \
    }                                                                                               \
  };                                                                                                \
  ::rtc::amd::ComputeKernel3D *createCompute_##name(::rtc::amd::Device *) {                         \
    return new rtc_amd_compute3D_##name;                                                            \
  }

#define RTC_IMPORT_COMPUTE1D(name) ::rtc::amd::ComputeKernel1D *createCompute_##name(::rtc::amd::Device *dev);
#define RTC_IMPORT_COMPUTE2D(name) ::rtc::amd::ComputeKernel2D *createCompute_##name(::rtc::amd::Device *dev);
#define RTC_IMPORT_COMPUTE3D(name) ::rtc::amd::ComputeKernel3D *createCompute_##name(::rtc::amd::Device *dev);

// Trace kernels: RTC_EXPORT_TRACE2D(name, Class) defines the kernel running
// Class::run(TraceInterface&) over 16x16-thread blocks covering the launch
// dims, with the launch parameters at getLPData(); RTC_IMPORT_TRACE2D
// defines createTrace_<name>(Device*) over it.
#define RTC_EXPORT_TRACE2D(name, Class)                                                          \
  __global__ void rtc_amd_trace_##name(::rtc::amd::TraceInterface ti) { Class::run(ti); }       \
  void rtc_amd_launch_##name(::rtc::amd::Device *, ::rtc::amd::vec2i dims, const void *lpData) { \
    ::rtc::amd::TraceInterface ti;                                                               \
    ti.lpData = lpData;                                                                          \
    const dim3 bs(16, 16), nb((dims.x + 15) / 16, (dims.y + 15) / 16);                          \
    rtc_amd_trace_##name<<<nb, bs>>>(ti);                                                        \
  }

#define RTC_IMPORT_TRACE2D(fileNameBase, name, sizeOfLP)                                               \
  void rtc_amd_launch_##name(::rtc::amd::Device *device, ::rtc::amd::vec2i dims, const void *lpData); \
  ::rtc::amd::TraceKernel2D *createTrace_##name(::rtc::amd::Device *device) {                          \
    return new ::rtc::amd::TraceKernel2D(device, sizeOfLP, rtc_amd_launch_##name);                     \
  }

// Geometry types: the program addresses are read on the device by a
// one-thread kernel of this module, then handed to the GeomType; user
// geometry also exports the launcher of its bounds program.
#define RTC_AMD_PROGRAM_ADDRESSES(name, Programs, has_ah, has_ch, has_is)                  \
  __global__ void rtc_amd_programs_##name(::rtc::amd::ProgramRecord *r) {                  \
    r->anyHit = (has_ah) ? (uint64_t)(&Programs::anyHit) : 0ull;                           \
    r->closestHit = (has_ch) ? (uint64_t)(&Programs::closestHit) : 0ull;                   \
    r->intersect = rtc_amd_intersect_##name();                                            \
    r->reserved = 0ull;                                                                    \
  }                                                                                        \
  static ::rtc::amd::ProgramRecord rtc_amd_read_programs_##name() {                        \
    ::rtc::amd::ProgramRecord *d = nullptr, h{};                                           \
    ::rtc::amd::check(bnamd_device_malloc((void **)&d, sizeof(h)), #name);                 \
    rtc_amd_programs_##name<<<1, 1>>>(d);                                                  \
    ::rtc::amd::check(bnamd_memcpy_d2h(&h, d, sizeof(h)), #name);                          \
    ::rtc::amd::check(bnamd_device_free(d), #name);                                        \
    return h;                                                                              \
  }

#define RTC_EXPORT_USER_GEOM(name, DD, Programs, has_ah, has_ch)                                      \
  __device__ __forceinline__ uint64_t rtc_amd_intersect_##name() {                                   \
    return (uint64_t)(&Programs::intersect);                                                          \
  }                                                                                                   \
  RTC_AMD_PROGRAM_ADDRESSES(name, Programs, has_ah, has_ch, true)                                     \
  __global__ void rtc_amd_bounds_##name(const void *geom, float4 *boxes, int numPrims) {              \
    const int primID = (int)(blockIdx.x * blockDim.x + threadIdx.x);                                  \
    if (primID >= numPrims) return;                                                                   \
    ::rtc::amd::TraceInterface ti;                                                                    \
    ::rtc::amd::box3f b;                                                                              \
    Programs::bounds(ti, geom, b, primID);                                                            \
    boxes[2 * primID] = make_float4(b.lower.x, b.lower.y, b.lower.z, 0.f);                            \
    boxes[2 * primID + 1] = make_float4(b.upper.x, b.upper.y, b.upper.z, 0.f);                        \
  }                                                                                                   \
  static void rtc_amd_write_bounds_##name(const void *geom, void *boxes, int numPrims) {              \
    if (numPrims > 0) rtc_amd_bounds_##name<<<(numPrims + 255) / 256, 256>>>(geom, (float4 *)boxes, numPrims); \
  }                                                                                                   \
  ::rtc::amd::GeomType *createGeomType_##name(::rtc::amd::Device *device) {                           \
    return new ::rtc::amd::GeomType(device, ::rtc::amd::GeomKind::User, #name, sizeof(DD),          \
                                    rtc_amd_read_programs_##name(), rtc_amd_write_bounds_##name);    \
  }

#define RTC_EXPORT_TRIANGLES_GEOM(name, DD, Programs, has_ah, has_ch)                                 \
  __device__ __forceinline__ uint64_t rtc_amd_intersect_##name() { return 0ull; }                    \
  RTC_AMD_PROGRAM_ADDRESSES(name, Programs, has_ah, has_ch, false)                                    \
  ::rtc::amd::GeomType *createGeomType_##name(::rtc::amd::Device *device) {                           \
    return new ::rtc::amd::GeomType(device, ::rtc::amd::GeomKind::Triangles, #name, sizeof(DD),      \
                                    rtc_amd_read_programs_##name(), nullptr);                         \
  }

#define RTC_IMPORT_USER_GEOM(moduleName, typeName, DD, has_ah, has_ch) \
  ::rtc::amd::GeomType *createGeomType_##typeName(::rtc::amd::Device *);
#define RTC_IMPORT_TRIANGLES_GEOM(moduleName, typeName, DD, has_ah, has_ch) \
  ::rtc::amd::GeomType *createGeomType_##typeName(::rtc::amd::Device *);

#define RTC_DECLARE_GLOBALS(ignore) /* ignore */
