// rtc::amd host layer over the libbnamd C ABI (see rtcore_amd.h).
#include "rtcore_amd.h"

#include <cstring>

namespace rtc {
namespace amd {

void check(int status, const char *what) {
  if (status != BNAMD_OK) throw std::runtime_error(std::string(what) + ": " + bnamd_last_error());
}

// ---------------------------------------------------------------- Buffer
Buffer::Buffer(Device *device, size_t numBytes, const void *initValues) : device(device) {
  resize(numBytes);
  if (initValues && numBytes) upload(initValues, numBytes);
}

Buffer::~Buffer() {
  if (d_data) (void)bnamd_device_free(d_data);
}

void Buffer::upload(const void *data, size_t bytes, size_t offset) {
  if (offset + bytes > numBytes) throw std::runtime_error("rtc::amd::Buffer::upload: out of range");
  check(bnamd_memcpy_h2d((uint8_t *)d_data + offset, data, bytes), "rtc::amd::Buffer::upload");
}

void Buffer::resize(size_t newNumBytes) {
  if (d_data) check(bnamd_device_free(d_data), "rtc::amd::Buffer::resize");
  d_data = nullptr;
  numBytes = newNumBytes;
  if (newNumBytes) check(bnamd_device_malloc(&d_data, newNumBytes), "rtc::amd::Buffer::resize");
}

// ----------------------------------------------------------- TextureData
namespace {
size_t texel_bytes(DataType f) {
  switch (f) {
    case UCHAR: return 1;
    case UCHAR4: return 4;
    case USHORT: return 2;
    case FLOAT: return 4;
    case FLOAT2: return 8;
    case FLOAT3: return 12;
    case FLOAT4: return 16;
    default: throw std::runtime_error("rtc::amd::TextureData: unsupported texel format");
  }
}
}  // namespace

TextureData::TextureData(Device *device, vec3i dims, DataType format, const void *texels)
    : device(device), dims(dims), format(format) {
  if (dims.x <= 0 || dims.y <= 0 || dims.z < 0) throw std::runtime_error("rtc::amd::TextureData: bad dims");
  numBytes = texel_bytes(format) * (size_t)dims.x * dims.y * (dims.z > 0 ? dims.z : 1);
  check(bnamd_device_malloc(&d_texels, numBytes), "rtc::amd::TextureData");
  if (texels) check(bnamd_memcpy_h2d(d_texels, texels, numBytes), "rtc::amd::TextureData");
  if (dims.z == 0 && texels) hostTexels.assign((const unsigned char *)texels, (const unsigned char *)texels + numBytes);
}

TextureData::~TextureData() {
  if (d_texels) (void)bnamd_device_free(d_texels);
}

Texture *TextureData::createTexture(const TextureDesc &desc) { return new Texture(this, desc); }

Texture::Texture(TextureData *data, const TextureDesc &desc) : device(data->device), data(data), desc(desc) {
  // the structured-volume sampler filters trilinearly with clamping on
  // texel centres (StructuredData.cu:154-173 asks for exactly that)
  if (data->dims.z > 0 && (desc.addressMode[0] != CLAMP || desc.addressMode[1] != CLAMP ||
                           desc.addressMode[2] != CLAMP || desc.normalizedCoords))
    throw std::runtime_error("rtc::amd::Texture: 3D textures are sampled with CLAMP, unnormalized coordinates");
  if (data->dims.z == 0) {
    // 2D images: barney creates them with normalized coordinates
    // (barney/common/Texture.cpp:98-101), which is what tex2D samples
    if (!desc.normalizedCoords)
      throw std::runtime_error("rtc::amd::Texture: 2D textures are sampled with normalized coordinates");
    if (data->hostTexels.empty()) throw std::runtime_error("rtc::amd::Texture: 2D texture without texels");
    bnamd_texture2d_desc td{};
    td.width = data->dims.x;
    td.height = data->dims.y;
    td.format = (int)data->format;
    td.texels = data->hostTexels.data();
    td.filterMode = (int)desc.filterMode;
    td.addressMode[0] = (int)desc.addressMode[0];
    td.addressMode[1] = (int)desc.addressMode[1];
    check(bnamd_texture2d_create(&td, nullptr, &tex2d), "rtc::amd::Texture");
  }
  // the device descriptor getDD() points to
  TextureObjectDD o{};
  o.filter = (int)desc.filterMode;
  for (int k = 0; k < 3; ++k) o.addr[k] = (int)desc.addressMode[k];
  if (tex2d) {
    o.dim = 2;
    check(bnamd_texture2d_device_view(tex2d, &o.texels, &o.w, &o.h), "rtc::amd::Texture");
    o.d = 1;
  } else {
    if (data->format != FLOAT) throw std::runtime_error("rtc::amd::Texture: 3D textures hold FLOAT texels");
    o.dim = 3;
    o.w = data->dims.x;
    o.h = data->dims.y;
    o.d = data->dims.z;
    o.texels = data->d_texels;
  }
  check(bnamd_device_malloc(&d_obj, sizeof o), "rtc::amd::Texture");
  check(bnamd_memcpy_h2d(d_obj, &o, sizeof o), "rtc::amd::Texture");
}

Texture::~Texture() {
  if (tex2d) (void)bnamd_texture2d_destroy(tex2d);
  if (d_obj) (void)bnamd_device_free(d_obj);
}

EnvMapLight::EnvMapLight(Texture *texture, const float direction[3], const float up[3], float scale)
    : texture(texture) {
  if (!texture || !texture->tex2d) throw std::runtime_error("rtc::amd::EnvMapLight: needs a 2D texture");
  bnamd_envmap_desc d{};
  for (int k = 0; k < 3; ++k) {
    d.direction[k] = direction[k];
    d.up[k] = up[k];
  }
  d.scale = scale;
  check(bnamd_envmap_create(texture->tex2d, &d, nullptr, &handle), "rtc::amd::EnvMapLight");
}

EnvMapLight::~EnvMapLight() {
  if (handle) (void)bnamd_envmap_destroy(handle);
}

// -------------------------------------------------------- GeomType / Geom
GeomType::GeomType(Device *device, GeomKind kind, const char *name) : device(device), kind(kind), name(name) {}

GeomType::GeomType(Device *device, GeomKind kind, const char *name, size_t sizeOfDD, const ProgramRecord &programs,
                   BoundsFct bounds)
    : device(device), kind(kind), name(name), programmable(true), sizeOfDD(sizeOfDD), programs(programs),
      bounds(bounds) {
  if (kind != GeomKind::Triangles && kind != GeomKind::User)
    throw std::runtime_error("rtc::amd::GeomType: programmable types are triangles or user geometry");
  if (kind == GeomKind::User && (!bounds || !programs.intersect))
    throw std::runtime_error("rtc::amd::GeomType: user geometry needs bounds and intersect programs");
}

Geom *GeomType::createGeom() { return new Geom(this); }

Geom::Geom(GeomType *gt) : gt(gt) {}

Geom::~Geom() {
  if (d_record) (void)bnamd_device_free(d_record);
  if (d_boxes) (void)bnamd_device_free(d_boxes);
}

void Geom::setDD(const void *dd) {
  if (!dd) throw std::runtime_error("rtc::amd::Geom::setDD: NULL");
  if (gt->programmable) {
    ddBytes.assign((const uint8_t *)dd, (const uint8_t *)dd + gt->sizeOfDD);
    if (gt->kind == GeomKind::User) return;  // its programs read the DD; the boxes come from bounds
  }
  barney_dd::Geometry base;
  std::memcpy(&base, dd, barney_dd::kGeometryDDSize);
  const uint8_t *fields = (const uint8_t *)dd + barney_dd::kGeometryDDSize;
  userID = base.userID;
  materialID = base.materialID;
  const barney_dd::GeometryAttribute &col = base.attributes.colorAttribute;
  ddColors = (col.scope == barney_dd::SCOPE_PER_PRIM && col.fromArray.type == barney_dd::BN_FLOAT32_VEC3)
                 ? col.fromArray.ptr : nullptr;
  ddNormals = nullptr;
  switch (gt->kind) {
    case GeomKind::Triangles: {
      barney_dd::TrianglesFields f;
      std::memcpy(&f, fields, sizeof(f));
      ddIndices = f.indices;
      ddVertices = f.vertices;
      ddNormals = f.normals;
      ddRadii = nullptr;
      break;
    }
    case GeomKind::Spheres: {
      barney_dd::SpheresFields f;
      std::memcpy(&f, fields, sizeof(f));
      ddVertices = f.origins;
      ddRadii = f.radii;
      if (f.colors) ddColors = f.colors;
      ddIndices = nullptr;
      defaultRadius = f.defaultRadius;
      break;
    }
    case GeomKind::Cylinders: {
      barney_dd::CylindersFields f;
      std::memcpy(&f, fields, sizeof(f));
      ddVertices = f.vertices;
      ddIndices = f.indices;
      ddRadii = f.radii;
      break;
    }
    case GeomKind::User:
      break;
  }
  haveDD = true;
}

void Geom::prepare() {
  if (!gt->programmable) return;
  const size_t bytes = sizeof(ProgramRecord) + gt->sizeOfDD;
  std::vector<uint8_t> rec(bytes, 0);
  std::memcpy(rec.data(), &gt->programs, sizeof(ProgramRecord));
  if (!ddBytes.empty()) std::memcpy(rec.data() + sizeof(ProgramRecord), ddBytes.data(), ddBytes.size());
  if (d_record) check(bnamd_device_free(d_record), "rtc::amd::Geom::prepare");
  d_record = nullptr;
  check(bnamd_device_malloc(&d_record, bytes), "rtc::amd::Geom::prepare");
  check(bnamd_memcpy_h2d(d_record, rec.data(), bytes), "rtc::amd::Geom::prepare");
  if (gt->kind == GeomKind::User) {
    if (d_boxes) check(bnamd_device_free(d_boxes), "rtc::amd::Geom::prepare");
    d_boxes = nullptr;
    if (primCount > 0) {
      check(bnamd_device_malloc(&d_boxes, 32 * (size_t)primCount), "rtc::amd::Geom::prepare");
      gt->bounds((const uint8_t *)d_record + sizeof(ProgramRecord), d_boxes, primCount);
      check(bnamd_synchronize(), "rtc::amd::Geom::prepare (bounds program)");
    }
  }
}

void Geom::setVertices(Buffer *v, int n) {
  vertices = v;
  numVertices = n;
  if (gt->kind == GeomKind::Spheres) primCount = n;
}

void Geom::setIndices(Buffer *i, int n) {
  if (gt->kind == GeomKind::Spheres) throw std::runtime_error("rtc::amd::Geom::setIndices: spheres take no indices");
  indices = i;
  numIndices = n;
  primCount = n;
}

void Geom::setPrimCount(int n) {
  if (gt->kind == GeomKind::Triangles)
    throw std::runtime_error("rtc::amd::Geom::setPrimCount: triangle geoms count their indices");
  primCount = n;
}

void Geom::setRadii(Buffer *r, float defRadius) {
  radii = r;
  defaultRadius = defRadius;
}

void Geom::setColors(Buffer *c) { colors = c; }

void Geom::setMaterial(int mat, int user) {
  materialID = mat;
  userID = user;
}

bnamd_geom Geom::desc() const {
  bnamd_geom g{};
  g.kind = (int)gt->kind;
  g.numPrims = primCount;
  g.numVertices = numVertices;
  g.materialID = materialID;
  g.programs = d_record;
  if (gt->kind == GeomKind::User) {
    g.vertices = (const float *)d_boxes;
    g.userID = userID;
    return g;
  }
  if (haveDD) {
    g.vertices = (const float *)ddVertices;
    g.indices = (const int *)ddIndices;
    g.radii = (const float *)ddRadii;
    g.colors = (const float *)ddColors;
  } else {
    g.vertices = vertices ? (const float *)vertices->getDD() : nullptr;
    g.indices = indices ? (const int *)indices->getDD() : nullptr;
    g.radii = radii ? (const float *)radii->getDD() : nullptr;
    g.colors = colors ? (const float *)colors->getDD() : nullptr;
  }
  g.defaultRadius = defaultRadius;
  g.userID = userID;
  return g;
}

// -------------------------------------------------------------- GeomGroup
GeomGroup::~GeomGroup() {
  if (blas) (void)bnamd_blas_destroy(blas);
}

void GeomGroup::buildAccel() {
  std::vector<bnamd_geom> gs;
  for (Geom *g : geoms) {
    g->prepare();
    gs.push_back(g->desc());
  }
  bnamd_group grp{gs.data(), (int)gs.size()};
  bnamd_blas *b = nullptr;
  const std::vector<bnamd_material> &mats = device->materials;
  check(bnamd_blas_build(&grp, mats.empty() ? nullptr : mats.data(), (int)mats.size(), nullptr, &b),
        "rtc::amd::GeomGroup::buildAccel");
  if (blas) (void)bnamd_blas_destroy(blas);
  blas = b;
}

// ---------------------------------------------------------- InstanceGroup
InstanceGroup::InstanceGroup(Device *device, const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                             const std::vector<affine3f> &xfms)
    : Group(device), groups(groups), instIDs(instIDs), xfms(xfms) {
  if (!instIDs.empty() && instIDs.size() != groups.size())
    throw std::runtime_error("rtc::amd::InstanceGroup: one instance ID per group expected");
  if (xfms.size() != groups.size())
    throw std::runtime_error("rtc::amd::InstanceGroup: one transform per group expected");
}

InstanceGroup::~InstanceGroup() {
  if (world) (void)bnamd_world_destroy(world);
}

void InstanceGroup::setTransforms(const std::vector<affine3f> &newXfms) {
  if (newXfms.size() != groups.size())
    throw std::runtime_error("rtc::amd::InstanceGroup::setTransforms: one transform per instance expected");
  xfms = newXfms;
}

namespace {
std::vector<float> flat_xfms(const std::vector<affine3f> &xfms) {
  std::vector<float> out;
  for (const affine3f &x : xfms) {
    const float m[12] = {x.vx.x, x.vx.y, x.vx.z, x.vy.x, x.vy.y, x.vy.z, x.vz.x, x.vz.y, x.vz.z, x.p.x, x.p.y, x.p.z};
    out.insert(out.end(), m, m + 12);
  }
  return out;
}
}  // namespace

void InstanceGroup::buildAccel() {
  // one bottom level per distinct group (built here if its group has not
  // built it yet), referenced by every instance of that group
  std::vector<GeomGroup *> distinct;
  std::vector<bnamd_instance> inst(groups.size());
  const std::vector<float> m = flat_xfms(xfms);
  for (size_t i = 0; i < groups.size(); ++i) {
    auto *gg = dynamic_cast<GeomGroup *>(groups[i]);
    if (!gg) throw std::runtime_error("rtc::amd::InstanceGroup::buildAccel: nested instance groups");
    size_t k = 0;
    while (k < distinct.size() && distinct[k] != gg) ++k;
    if (k == distinct.size()) {
      if (!gg->blas) gg->buildAccel();
      distinct.push_back(gg);
    }
    inst[i].group = (int)k;
    inst[i].instID = instIDs.empty() ? (int)i : instIDs[i];
    std::memcpy(inst[i].xfm, &m[12 * i], 12 * sizeof(float));
  }
  std::vector<const bnamd_blas *> blas;
  for (GeomGroup *gg : distinct) blas.push_back(gg->blas);
  const std::vector<bnamd_material> &mats = materials.empty() ? device->materials : materials;
  bnamd_world *w = nullptr;
  check(bnamd_world_build_instances(blas.data(), (int)blas.size(), inst.data(), (int)inst.size(),
                                    mats.empty() ? nullptr : mats.data(), (int)mats.size(), nullptr, &w),
        "rtc::amd::InstanceGroup::buildAccel");
  if (world) (void)bnamd_world_destroy(world);
  world = w;
}

void InstanceGroup::refitAccel() {
  if (!world) {
    buildAccel();
    return;
  }
  const std::vector<float> m = flat_xfms(xfms);
  check(bnamd_world_set_transforms(world, m.data(), (int)xfms.size(), nullptr), "rtc::amd::InstanceGroup::refitAccel");
}

int InstanceGroup::numBottomLevels() const {
  if (!world) return 0;
  bnamd_world_info info{};
  check(bnamd_world_info_get(world, &info), "rtc::amd::InstanceGroup::numBottomLevels");
  return info.numBlas;
}

void InstanceGroup::traceRays(const bnamd_ray *d_rays, bnamd_hit *d_hits, int numRays) {
  if (!world) throw std::runtime_error("rtc::amd::InstanceGroup::traceRays: buildAccel() first");
  check(bnamd_trace(world, d_rays, d_hits, numRays, nullptr), "rtc::amd::InstanceGroup::traceRays");
}

// ------------------------------------------------------------------ Device
Device::Device(int physicalGPU) : physicalID(physicalGPU) { check(bnamd_init(physicalGPU), "rtc::amd::Device"); }

Device::~Device() = default;

void Device::copyAsync(void *dst, const void *src, size_t n) {
  check(bnamd_memcpy_async(dst, src, n, nullptr), "rtc::amd::Device::copyAsync");
}

void *Device::allocHost(size_t n) {
  void *p = nullptr;
  check(bnamd_host_malloc(&p, n), "rtc::amd::Device::allocHost");
  return p;
}

void Device::freeHost(void *mem) { check(bnamd_host_free(mem), "rtc::amd::Device::freeHost"); }

void Device::memsetAsync(void *mem, int value, size_t size) {
  check(bnamd_memset_async(mem, value, size, nullptr), "rtc::amd::Device::memsetAsync");
}

void *Device::allocMem(size_t n) {
  void *p = nullptr;
  check(bnamd_device_malloc(&p, n), "rtc::amd::Device::allocMem");
  return p;
}

void Device::freeMem(void *mem) { check(bnamd_device_free(mem), "rtc::amd::Device::freeMem"); }

void Device::sync() { check(bnamd_synchronize(), "rtc::amd::Device::sync"); }

int Device::setActive() const {
  int old = 0;
  check(bnamd_get_device(&old), "rtc::amd::Device::setActive");
  check(bnamd_set_device(physicalID), "rtc::amd::Device::setActive");
  return old;
}

void Device::restoreActive(int oldActive) const {
  check(bnamd_set_device(oldActive), "rtc::amd::Device::restoreActive");
}

TextureData *Device::createTextureData(vec3i dims, DataType format, const void *texels) {
  return new TextureData(this, dims, format, texels);
}

void Device::freeTextureData(TextureData *t) { delete t; }

void Device::freeTexture(Texture *t) { delete t; }

EnvMapLight *Device::createEnvMapLight(Texture *texture, const float direction[3], const float up[3], float scale) {
  return new EnvMapLight(texture, direction, up, scale);
}

void Device::freeEnvMapLight(EnvMapLight *l) { delete l; }

AccelHandle getAccelHandle(Group *ig) {
  auto *g = dynamic_cast<InstanceGroup *>(ig);
  if (!g || !g->world) throw std::runtime_error("rtc::amd::getAccelHandle: a built InstanceGroup expected");
  const void *rec = nullptr;
  check(bnamd_world_device(g->world, &rec), "rtc::amd::getAccelHandle");
  return (AccelHandle)rec;
}

// ------------------------------------------------------------ TraceKernel2D
TraceKernel2D::TraceKernel2D(Device *device, size_t sizeOfLP, TraceLaunchFct traceLaunchFct)
    : device(device), traceLaunchFct(traceLaunchFct), sizeOfLP(sizeOfLP) {
  if (sizeOfLP) check(bnamd_device_malloc(&d_lpData, sizeOfLP), "rtc::amd::TraceKernel2D");
}

TraceKernel2D::~TraceKernel2D() {
  if (d_lpData) (void)bnamd_device_free(d_lpData);
}

void TraceKernel2D::launch(vec2i launchDims, const void *kernelData) {
  // the launch parameters (host or device memory) are copied before the
  // call returns, so the caller may reuse its buffer at once; the kernel
  // follows on the same (default) stream
  if (sizeOfLP) {
    check(bnamd_memcpy_async(d_lpData, kernelData, sizeOfLP, nullptr), "rtc::amd::TraceKernel2D::launch");
    check(bnamd_synchronize(), "rtc::amd::TraceKernel2D::launch");
  }
  if (launchDims.x > 0 && launchDims.y > 0) traceLaunchFct(device, launchDims, d_lpData);
}

Buffer *Device::createBuffer(size_t numBytes, const void *initValues) { return new Buffer(this, numBytes, initValues); }

void Device::freeBuffer(Buffer *b) { delete b; }

void Device::freeGeomType(GeomType *gt) { delete gt; }

void Device::freeGeom(Geom *g) { delete g; }

Group *Device::createTrianglesGroup(const std::vector<Geom *> &geoms) {
  for (Geom *g : geoms)
    if (g->gt->kind != GeomKind::Triangles)
      throw std::runtime_error("rtc::amd::Device::createTrianglesGroup: non-triangle geom");
  return new TrianglesGeomGroup(this, geoms);
}

Group *Device::createUserGeomsGroup(const std::vector<Geom *> &geoms) {
  for (Geom *g : geoms)
    if (g->gt->kind == GeomKind::Triangles)
      throw std::runtime_error("rtc::amd::Device::createUserGeomsGroup: triangle geom in a user group");
  return new UserGeomGroup(this, geoms);
}

Group *Device::createInstanceGroup(const std::vector<Group *> &groups, const std::vector<int> &instIDs,
                                   const std::vector<affine3f> &xfms) {
  return new InstanceGroup(this, groups, instIDs, xfms);
}

void Device::freeGroup(Group *g) { delete g; }

GeomType *createGeomType_Triangles(Device *device) { return new GeomType(device, GeomKind::Triangles, "Triangles"); }
GeomType *createGeomType_Spheres(Device *device) { return new GeomType(device, GeomKind::Spheres, "Spheres"); }
GeomType *createGeomType_Cylinders(Device *device) { return new GeomType(device, GeomKind::Cylinders, "Cylinders"); }

}  // namespace amd
}  // namespace rtc
