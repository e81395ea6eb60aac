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
// Two ways to use it. Native: the geometry types barney imports
// (Triangles, Spheres, Cylinders) come from createGeomType_<Name>() below,
// with programs native in libbnamd, and barney's frame loop calls the fused
// renderers; their device data arrives as barney writes it: Geom::setDD(&dd)
// with the bytes of barney's Triangles::DD / Spheres::DD / Cylinders::DD,
// whose layout barney_dd below restates (or through the typed setters).
// Programmable: barney's own device programs and kernels, compiled with
// hipcc against rtcore_amd_device.h, register through its RTC_EXPORT_*
// macros (geometry types with anyHit / closestHit / intersect / bounds
// programs, trace and compute kernels); TraceKernel2D and ComputeKernel*D
// below launch the kernels, and TraceInterface::traceRay walks the BVH8s
// built here and calls the programs.
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
struct vec2i {
  int x, y;
};
struct vec2ui {
  unsigned x, y;
};
struct vec3ui {
  unsigned x, y, z;
};
struct box3f {
  vec3f lower, upper;
};
// affine3f as rtcore uses it: linear part as three columns (vx, vy, vz),
// then the translation p (object to world)
struct affine3f {
  vec3f vx{1, 0, 0}, vy{0, 1, 0}, vz{0, 0, 1}, p{0, 0, 0};
};

// throws std::runtime_error with libbnamd's message when status != 0
void check(int status, const char *what);

// rtc::DataType, AddressMode, FilterMode, ColorSpace, TextureDesc and the
// opaque handle types of rtcore/common/rtcore-common.h, same values.
enum DataType {
  UCHAR = 0, UCHAR4, INT = 10, INT2, INT3, INT4, LONG = 20, LONG2, LONG3, LONG4,
  FLOAT = 30, FLOAT2, FLOAT3, FLOAT4, USHORT = 40
};
enum AddressMode { WRAP, CLAMP, BORDER, MIRROR };
enum FilterMode { FILTER_MODE_POINT, FILTER_MODE_LINEAR };
enum ColorSpace { COLOR_SPACE_LINEAR, COLOR_SPACE_SRGB };
struct vec4f {
  float x, y, z, w;
};
struct TextureDesc {
  FilterMode filterMode = FILTER_MODE_LINEAR;
  AddressMode addressMode[3] = {CLAMP, CLAMP, CLAMP};
  vec4f borderColor = {0.f, 0.f, 0.f, 0.f};
  bool normalizedCoords = true;
  ColorSpace colorSpace = COLOR_SPACE_LINEAR;
};
typedef struct _TextureObject *TextureObject;
typedef struct _AccelHandle *AccelHandle;

// The device data barney hands to Geom::setDD: Geometry::DD
// (barney/geometry/Geometry.h:34-45, with GeometryAttributes::DD of
// barney/render/GeometryAttributes.h:15-60 and numAttributes = 4 from
// barney/render/HitAttributes.h:12) followed by each geometry's fields
// (Triangles.h:28-34, Spheres.h:14-20, Cylinders.h:19-25). The vec4f in
// GeometryAttribute::DD is 16-byte aligned, so each attribute record is 48
// bytes; the derived structs' fields follow at kGeometryDDSize, the base's
// data size (its vec4f has constructors, so the base is not POD for layout
// and its tail padding is reused). barney built against this backend can
// check the assumption: offsetof(Spheres::DD, origins) == kGeometryDDSize,
// and likewise for Triangles::DD::indices and Cylinders::DD::vertices.
namespace barney_dd {
struct alignas(16) Vec4 {
  float x, y, z, w;
};
struct AttributeArray {
  const void *ptr;
  int type;  // BNDataType
  int size;
};
struct GeometryAttribute {
  Vec4 value;
  AttributeArray fromArray;
  int scope;  // INVALID, CONSTANT, PER_PRIM, PER_VERTEX, FACE_VARYING
};
enum { SCOPE_INVALID = 0, SCOPE_CONSTANT, SCOPE_PER_PRIM, SCOPE_PER_VERTEX, SCOPE_FACE_VARYING };
enum { BN_FLOAT32_VEC3 = 182 };  // barney/include/barney.h:147-150
constexpr int kNumAttributes = 4;
struct GeometryAttributes {
  GeometryAttribute attribute[kNumAttributes];
  GeometryAttribute colorAttribute;
  GeometryAttribute normalAttribute;
};
struct Geometry {
  GeometryAttributes attributes;
  int userID;
  int materialID;
};
constexpr size_t kGeometryDDSize = sizeof(GeometryAttributes) + 2 * sizeof(int);
static_assert(sizeof(GeometryAttribute) == 48, "GeometryAttribute::DD is 48 bytes");
static_assert(kGeometryDDSize == 296, "Geometry::DD holds 296 bytes of data");
struct TrianglesFields {
  const int *indices;       // vec3i
  const float *vertices;    // vec3f
  const float *normals;     // vec3f, may be null
  const float *texcoords;   // vec2f, may be null
};
struct SpheresFields {
  const float *origins;  // vec3f
  const float *radii;    // may be null: defaultRadius
  const float *colors;   // vec3f, may be null
  float defaultRadius;
};
struct CylindersFields {
  const float *vertices;  // vec3f
  const int *indices;     // vec2i
  const float *radii;
};
}  // namespace barney_dd

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

struct Texture;

// Texel data (rtcore/cudaCommon/TextureData.h:15-31): 3D scalar fields
// (dims.z > 0, structured volumes) and 2D images (dims.z == 0). A 3D
// field's samplers read the device texels directly; a 2D image keeps a host
// copy from which each Texture builds its sampled bnamd_texture2d.
struct TextureData {
  TextureData(Device *device, vec3i dims, DataType format, const void *texels);
  ~TextureData();
  Texture *createTexture(const TextureDesc &desc);
  Device *const device;
  const vec3i dims;
  const DataType format;
  void *d_texels = nullptr;
  size_t numBytes = 0;
  std::vector<unsigned char> hostTexels;  // 2D images only
};

// What a TextureObject points to on the device: the sampling description
// and the texels, so that device code samples it with tex2D / tex3D
// (rtcore_amd_device.h) as barney's programs do.
struct TextureObjectDD {
  int dim;               // 2: image (normalized coordinates), 3: scalar field
  int w, h, d;
  int filter;            // FilterMode
  int addr[3];           // AddressMode per axis
  const void *texels;    // 2D: w*h float4 (widened as tex2D reads them); 3D: w*h*d float, x fastest
};

// A view of texel data with its sampling description
// (rtcore/cudaCommon/Texture.h:12-24). getDD() is the TextureObject the
// device side of barney stores: the device address of a TextureObjectDD.
// The native entry points take the texels themselves (texels3D() for a
// structured field) or the bnamd_texture2d handle (tex2d, for an env map).
struct Texture {
  Texture(TextureData *data, const TextureDesc &desc);
  ~Texture();
  TextureObject getDD() const { return (TextureObject)d_obj; }
  const float *texels3D() const { return data->dims.z > 0 ? (const float *)data->d_texels : nullptr; }
  Device *const device;
  TextureData *const data;
  const TextureDesc desc;
  bnamd_texture2d *tex2d = nullptr;
  void *d_obj = nullptr;  // the TextureObjectDD
};

// barney's EnvMapLight (barney/light/EnvMap.cu:59-230) over a 2D texture:
// commit() becomes the constructor (frame from direction / up, importance
// CDFs); handle goes into bnamd_render_params.envMap for the renders.
struct EnvMapLight {
  EnvMapLight(Texture *texture, const float direction[3], const float up[3], float scale);
  ~EnvMapLight();
  Texture *const texture;
  bnamd_envmap *handle = nullptr;
};

enum class GeomKind { Triangles = BNAMD_GEOM_TRIANGLES, Spheres = BNAMD_GEOM_SPHERES,
                      Cylinders = BNAMD_GEOM_CYLINDERS, User = BNAMD_GEOM_USER };

// A geometry's program record in device memory, as TraceInterface::traceRay
// reads it (the SBT record of rtcore/cuda/Geom.h:16-69): the device addresses
// of its anyHit, closestHit and intersect programs (0: none), then the
// geometry's DD bytes, which getProgramData() returns.
struct ProgramRecord {
  uint64_t anyHit, closestHit, intersect, reserved;
};
static_assert(sizeof(ProgramRecord) == 32, "program record header is 32 bytes");

// Launches a user geometry's bounds program over its primitives: the boxes
// of numPrims primitives of the geometry whose DD is at deviceDD, written to
// deviceBoxes as float4 pairs (lo, hi); RTC_EXPORT_USER_GEOM defines it.
typedef void (*BoundsFct)(const void *deviceDD, void *deviceBoxes, int numPrims);

struct GeomType {
  // a native geometry type (programs in libbnamd)
  GeomType(Device *device, GeomKind kind, const char *name);
  // a programmable one (RTC_EXPORT_TRIANGLES_GEOM / RTC_EXPORT_USER_GEOM):
  // the size of its DD, its programs' addresses and, for user geometry, the
  // launcher of its bounds program
  GeomType(Device *device, GeomKind kind, const char *name, size_t sizeOfDD, const ProgramRecord &programs,
           BoundsFct bounds);
  virtual ~GeomType() = default;
  Geom *createGeom();
  Device *const device;
  const GeomKind kind;
  const std::string name;
  const bool programmable = false;
  const size_t sizeOfDD = 0;
  const ProgramRecord programs{};
  const BoundsFct bounds = nullptr;
};

struct Geom {
  explicit Geom(GeomType *gt);
  virtual ~Geom();
  // the bytes of barney's Triangles::DD / Spheres::DD / Cylinders::DD
  // (rtcore/cuda/Geom.h:16-69: Geom::setDD copies sizeof(DD) bytes): the
  // geometry's buffers, userID, materialID, and per-primitive float3 colors
  // from the color attribute (other attribute scopes are not read)
  void setDD(const void *dd);
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
  // programmable geometry: writes its program record (programs + DD) to
  // device memory and, for user geometry, runs its bounds program
  void prepare();
  bnamd_geom desc() const;
  GeomType *const gt;
  Buffer *vertices = nullptr, *indices = nullptr, *radii = nullptr, *colors = nullptr;
  int numVertices = 0, numIndices = 0, primCount = 0;
  float defaultRadius = 1.f;
  int materialID = 0, userID = 0;
  // device pointers given through setDD (take precedence over the Buffers)
  const void *ddVertices = nullptr, *ddIndices = nullptr, *ddRadii = nullptr, *ddColors = nullptr;
  const void *ddNormals = nullptr;
  bool haveDD = false;
  // programmable geometry: the DD bytes given to setDD, the device program
  // record and the primitive boxes (user geometry)
  std::vector<uint8_t> ddBytes;
  void *d_record = nullptr;
  void *d_boxes = nullptr;
};

struct Group {
  explicit Group(Device *device) : device(device) {}
  virtual ~Group() = default;
  virtual void buildAccel() = 0;
  virtual void refitAccel() { buildAccel(); }
  Device *const device;
};

// A triangles or user-geometry group: buildAccel builds its bottom-level
// BVH8 (rtcore/cuda/cudaGroup.cu:261-562), which every instance of the
// group shares.
struct GeomGroup : public Group {
  GeomGroup(Device *device, const std::vector<Geom *> &geoms) : Group(device), geoms(geoms) {}
  ~GeomGroup() override;
  void buildAccel() override;
  const std::vector<Geom *> geoms;
  bnamd_blas *blas = nullptr;
};
struct TrianglesGeomGroup : public GeomGroup {
  using GeomGroup::GeomGroup;
};
struct UserGeomGroup : public GeomGroup {
  using GeomGroup::GeomGroup;
};

// Instances of geometry groups (rtcore/cuda/cudaGroup.cu:144-258): one
// bottom level per distinct group (built on demand), instances over them.
// refitAccel after setTransforms rebuilds the instance level only.
struct InstanceGroup : public Group {
  InstanceGroup(Device *device, const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                const std::vector<affine3f> &xfms);
  ~InstanceGroup() override;
  void buildAccel() override;
  void refitAccel() override;
  void setTransforms(const std::vector<affine3f> &xfms);
  // distinct bottom levels the built world references
  int numBottomLevels() const;
  // closest hit per ray (TraceInterface::traceRay over this accel);
  // rays/hits are device arrays
  void traceRays(const bnamd_ray *d_rays, bnamd_hit *d_hits, int numRays);
  bnamd_world *getDD() const { return world; }
  // the world's material table; empty: the device's
  std::vector<bnamd_material> materials;
  const std::vector<Group *> groups;
  const std::vector<int> instIDs;
  std::vector<affine3f> xfms;
  bnamd_world *world = nullptr;
};

// Trace kernels (rtcore/cuda/TraceKernel.h:16-28): launch(dims, lp) copies
// the sizeOfLP bytes of the launch parameters to device memory and runs the
// kernel RTC_EXPORT_TRACE2D defined (16x16 threads per block over dims).
typedef void (*TraceLaunchFct)(Device *device, vec2i dims, const void *d_lpData);
struct TraceKernel2D {
  TraceKernel2D(Device *device, size_t sizeOfLP, TraceLaunchFct traceLaunchFct);
  ~TraceKernel2D();
  void launch(vec2i launchDims, const void *kernelData);
  Device *const device;
  const TraceLaunchFct traceLaunchFct;
  const size_t sizeOfLP;
  void *d_lpData = nullptr;
};

// Compute kernels (rtcore/cudaCommon/ComputeKernel.h:16-32): the kernel
// RTC_EXPORT_COMPUTE*D defined runs dd.run(ComputeInterface) per thread
// over nb blocks of bs threads, the kernel data passed by value.
struct ComputeKernel1D {
  virtual ~ComputeKernel1D() = default;
  virtual void launch(unsigned int nb, unsigned int bs, const void *pKernelData) = 0;
};
struct ComputeKernel2D {
  virtual ~ComputeKernel2D() = default;
  virtual void launch(vec2ui nb, vec2ui bs, const void *pKernelData) = 0;
};
struct ComputeKernel3D {
  virtual ~ComputeKernel3D() = default;
  virtual void launch(vec3ui nb, vec3ui bs, const void *pKernelData) = 0;
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
  void freeTexture(Texture *);
  EnvMapLight *createEnvMapLight(Texture *texture, const float direction[3], const float up[3], float scale);
  void freeEnvMapLight(EnvMapLight *);

  Buffer *createBuffer(size_t numBytes, const void *initValues = nullptr);
  void freeBuffer(Buffer *);

  void freeGeomType(GeomType *);
  void freeGeom(Geom *);
  Group *createTrianglesGroup(const std::vector<Geom *> &geoms);
  Group *createUserGeomsGroup(const std::vector<Geom *> &geoms);
  Group *createInstanceGroup(const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                             const std::vector<affine3f> &xfms);
  void freeGroup(Group *);
  void buildPipeline() {}  // nothing to link: program addresses are read when a geometry type is made
  void buildSBT() {}
  Denoiser *createDenoiser() { return nullptr; }

  const int physicalID;
  // the material table of this device's groups and worlds (the
  // bnamd_material view of barney's per-device material registry); bottom
  // levels are built with it (it marks glass triangles for shadow rays)
  std::vector<bnamd_material> materials;
};

// The acceleration handle barney's launch parameters carry
// (rtcore/cuda/Device.h:88): the device address of the instance group's
// world record, which TraceInterface::traceRay reads.
AccelHandle getAccelHandle(Group *instanceGroup);

// the geometry types barney imports (RTC_IMPORT_*_GEOM)
GeomType *createGeomType_Triangles(Device *device);
GeomType *createGeomType_Spheres(Device *device);
GeomType *createGeomType_Cylinders(Device *device);

}  // namespace amd
}  // namespace rtc
