// rtc::amd -- host side of barney's rtcore backend for MI355X (gfx950).
//
// The classes mirror the backend interface barney's core programs against
// (rtcore/cuda/Device.h:36-70, rtcore/cudaCommon/Device.h:34-64,
// rtcore/cuda/Buffer.h, Geom.h, GeomType.h, Group.h): same class and method
// names, same argument meaning, errors thrown as std::runtime_error (as
// rtcore/cudaCommon/cuda-helper.h:41-60 does). Every call goes through the
// C ABI of libbnamd (include/bnamd.h); no HIP or torch type appears here, so
// this layer builds with any C++17 host compiler.
//
// Geometry programs are native in libbnamd: the geometry types barney
// registers with RTC_EXPORT_USER_GEOM / RTC_EXPORT_TRIANGLES_GEOM
// (Triangles, Spheres, Cylinders) come from createGeomType_<Name>() below,
// and their device data is given through the typed setters of Geom rather
// than as raw DD bytes.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bnamd.h"

namespace rtc {
namespace amd {

struct vec3f {
  float x, y, z;
};
struct vec3i {
  int x, y, z;
};
// affine3f as rtcore uses it: linear part as three columns (vx, vy, vz),
// then the translation p (object to world)
struct affine3f {
  vec3f vx{1, 0, 0}, vy{0, 1, 0}, vz{0, 0, 1}, p{0, 0, 0};
};

// throws std::runtime_error with libbnamd's message when status != 0
void check(int status, const char *what);

enum class DataType { FLOAT };

struct Device;
struct Geom;
struct Group;

struct Buffer {
  Buffer(Device *device, size_t numBytes, const void *initValues);
  virtual ~Buffer();
  void *getDD() const { return d_data; }
  void upload(const void *data, size_t numBytes, size_t offset = 0);
  void resize(size_t newNumBytes);
  void *d_data = nullptr;
  size_t numBytes = 0;
  Device *const device;
};

// 3D scalar texture data (structured volumes); LINEAR/CLAMP filtering is
// done by the sampler of the volume renderer, so this is device memory.
struct TextureData {
  TextureData(Device *device, vec3i dims, DataType format, const void *texels);
  ~TextureData();
  Device *const device;
  vec3i dims;
  float *d_texels = nullptr;
};

enum class GeomKind { Triangles = BNAMD_GEOM_TRIANGLES, Spheres = BNAMD_GEOM_SPHERES,
                      Cylinders = BNAMD_GEOM_CYLINDERS };

struct GeomType {
  GeomType(Device *device, GeomKind kind, const char *name);
  virtual ~GeomType() = default;
  Geom *createGeom();
  Device *const device;
  const GeomKind kind;
  const std::string name;
};

struct Geom {
  explicit Geom(GeomType *gt);
  virtual ~Geom() = default;
  // Triangles (TrianglesGeom::setVertices / setIndices): float3 / int3
  // device buffers; numIndices = number of triangles
  void setVertices(Buffer *vertices, int numVertices);
  void setIndices(Buffer *indices, int numIndices);
  // user geometry (UserGeom::setPrimCount)
  void setPrimCount(int primCount);
  // Spheres::DD / Cylinders::DD contents (barney/geometry/Spheres.h,
  // Cylinders.h): per-prim radii (may be null: defaultRadius), per-prim
  // colors (may be null), cylinder vertex pairs through setIndices (int2)
  void setRadii(Buffer *radii, float defaultRadius);
  void setColors(Buffer *colors);
  void setMaterial(int materialID, int userID);
  bnamd_geom desc() const;
  GeomType *const gt;
  Buffer *vertices = nullptr, *indices = nullptr, *radii = nullptr, *colors = nullptr;
  int numVertices = 0, numIndices = 0, primCount = 0;
  float defaultRadius = 1.f;
  int materialID = 0, userID = 0;
};

struct Group {
  explicit Group(Device *device) : device(device) {}
  virtual ~Group() = default;
  virtual void buildAccel() = 0;
  virtual void refitAccel() { buildAccel(); }
  Device *const device;
};

struct GeomGroup : public Group {
  GeomGroup(Device *device, const std::vector<Geom *> &geoms) : Group(device), geoms(geoms) {}
  void buildAccel() override {}  // built as part of the instance group
  const std::vector<Geom *> geoms;
};
struct TrianglesGeomGroup : public GeomGroup {
  using GeomGroup::GeomGroup;
};
struct UserGeomGroup : public GeomGroup {
  using GeomGroup::GeomGroup;
};

struct InstanceGroup : public Group {
  InstanceGroup(Device *device, const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                const std::vector<affine3f> &xfms);
  ~InstanceGroup() override;
  void buildAccel() override;
  void setTransforms(const std::vector<affine3f> &xfms);
  // closest hit per ray (TraceInterface::traceRay over this accel);
  // rays/hits are device arrays
  void traceRays(const bnamd_ray *d_rays, bnamd_hit *d_hits, int numRays);
  bnamd_world *getDD() const { return world; }
  std::vector<bnamd_material> materials;  // the world's material table
  const std::vector<Group *> groups;
  const std::vector<int> instIDs;
  std::vector<affine3f> xfms;
  bnamd_world *world = nullptr;
};

struct Denoiser;

struct Device {
  explicit Device(int physicalGPU);
  virtual ~Device();
  std::string toString() const { return "rtc::amd::Device(physical=" + std::to_string(physicalID) + ")"; }
  void destroy() {}

  void copyAsync(void *dst, const void *src, size_t numBytes);
  void copy(void *dst, const void *src, size_t numBytes) {
    copyAsync(dst, src, numBytes);
    sync();
  }
  void *allocHost(size_t numBytes);
  void freeHost(void *mem);
  void memsetAsync(void *mem, int value, size_t size);
  void *allocMem(size_t numBytes);
  void freeMem(void *mem);
  void sync();
  int setActive() const;
  void restoreActive(int oldActive) const;

  TextureData *createTextureData(vec3i dims, DataType format, const void *texels);
  void freeTextureData(TextureData *);

  Buffer *createBuffer(size_t numBytes, const void *initValues = nullptr);
  void freeBuffer(Buffer *);

  void freeGeomType(GeomType *);
  void freeGeom(Geom *);
  Group *createTrianglesGroup(const std::vector<Geom *> &geoms);
  Group *createUserGeomsGroup(const std::vector<Geom *> &geoms);
  Group *createInstanceGroup(const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                             const std::vector<affine3f> &xfms);
  void freeGroup(Group *);
  void buildPipeline() {}  // programs are native: nothing to link
  void buildSBT() {}
  Denoiser *createDenoiser() { return nullptr; }

  const int physicalID;
};

// the geometry types barney imports (RTC_IMPORT_*_GEOM)
GeomType *createGeomType_Triangles(Device *device);
GeomType *createGeomType_Spheres(Device *device);
GeomType *createGeomType_Cylinders(Device *device);

}  // namespace amd
}  // namespace rtc
