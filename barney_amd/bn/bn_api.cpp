// barney's bn* host API (barney/include/barney.h, barney/api/barney.cu) over
// the bnamd C ABI: objects hold staged parameters (Object::set*,
// barney/Object.cpp), bnBuild turns the model's instances of groups of
// geometries into one bnamd world (ModelSlot::build,
// barney/ModelSlot.cpp:210-216) and/or builds its one volume (structured,
// unstructured, block-structured AMR or NanoVDB field; next to geometry, a
// structured one), and its groups'
// lights (ModelSlot::updateWorldLightsFromInstances, ModelSlot.cpp:90-125);
// bnRender runs the fused path tracer into the frame buffer's accumulation
// buffer (Context::render, barney/Context.cpp:142-185) and
// bnFrameBufferRead resolves it (FrameBuffer::read,
// barney/fb/FrameBuffer.cu:188-450).
#include "bn_api.h"

#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bnamd.h"

namespace {

void check(int rc, const char *what) {
  if (rc != BNAMD_OK) throw std::runtime_error(std::string(what) + ": " + bnamd_last_error());
}

struct DevMem {
  void *ptr = nullptr;
  DevMem() = default;
  explicit DevMem(size_t bytes) { check(bnamd_device_malloc(&ptr, bytes ? bytes : 1), "device allocation"); }
  DevMem(const void *host, size_t bytes) : DevMem(bytes) {
    if (bytes) check(bnamd_memcpy_h2d(ptr, host, bytes), "upload");
  }
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() {
    if (ptr) (void)bnamd_device_free(ptr);
  }
};

struct Context;

struct Object : _BNObject {
  explicit Object(Context *c, std::string t) : ctx(c), type(std::move(t)) {}
  virtual ~Object() = default;
  Context *ctx;
  std::string type;
  int appRefs = 1;
  std::map<std::string, std::array<float, 4>> f;
  std::map<std::string, int> i;
  std::map<std::string, std::array<int, 4>> iv;  // bnSet2i / 3i / 4i
  std::map<std::string, std::array<float, 16>> mat;  // bnSet4x3fv / 4x4fv
  std::map<std::string, std::string> str;
  std::map<std::string, Object *> o;
  float getf(const std::string &n, float def) const {
    auto it = f.find(n);
    return it == f.end() ? def : it->second[0];
  }
  std::array<float, 4> get4(const std::string &n, std::array<float, 4> def) const {
    auto it = f.find(n);
    return it == f.end() ? def : it->second;
  }
  int geti(const std::string &n, int def) const {
    auto it = i.find(n);
    return it == i.end() ? def : it->second;
  }
  Object *geto(const std::string &n) const {
    auto it = o.find(n);
    return it == o.end() ? nullptr : it->second;
  }
};

struct Data : Object {
  Data(Context *c, BNDataType t, size_t n, const void *items) : Object(c, "data"), dataType(t), count(n) {
    const size_t sz = elem_size(t);
    if (!sz) throw std::runtime_error("bnDataCreate: unsupported data type");
    if (n && !items) throw std::runtime_error("bnDataCreate: NULL items");
    if (t == BN_OBJECT)
      objs.assign((Object *const *)items, (Object *const *)items + n);
    else
      bytes.assign((const uint8_t *)items, (const uint8_t *)items + sz * n);
  }
  static size_t elem_size(BNDataType t) {
    switch (t) {
      case BN_OBJECT: return sizeof(void *);
      case BN_UINT8: return 1;
      case BN_UINT8_VEC4: case BN_INT32: case BN_UINT32: case BN_FLOAT32: return 4;
      case BN_INT32_VEC2: case BN_FLOAT32_VEC2: case BN_INT64: case BN_UINT64: return 8;
      case BN_INT32_VEC3: case BN_FLOAT32_VEC3: return 12;
      case BN_INT32_VEC4: case BN_FLOAT32_VEC4: return 16;
      default: return 0;
    }
  }
  // integer items widened to 64 bits (offsets may come as any integer type)
  std::vector<unsigned long long> as_u64() const {
    std::vector<unsigned long long> r(count);
    for (size_t k = 0; k < count; ++k) {
      switch (dataType) {
        case BN_INT32: r[k] = (unsigned long long)((const int32_t *)bytes.data())[k]; break;
        case BN_UINT32: r[k] = ((const uint32_t *)bytes.data())[k]; break;
        case BN_INT64: r[k] = (unsigned long long)((const int64_t *)bytes.data())[k]; break;
        case BN_UINT64: r[k] = ((const uint64_t *)bytes.data())[k]; break;
        default: throw std::runtime_error("bnBuild: integer data expected");
      }
    }
    return r;
  }
  BNDataType dataType;
  size_t count;
  std::vector<uint8_t> bytes;
  std::vector<Object *> objs;  // BN_OBJECT items
};

// bnTexture2DCreate (barney/common/Texture.cpp:85-135): host texels and the
// sampling description; the device texture is made when a model uses it
struct Texture2D : Object {
  Texture2D(Context *c, BNDataType fmt, int w, int h, const void *texels, BNTextureFilterMode filter,
            BNTextureAddressMode ax, BNTextureAddressMode ay)
      : Object(c, "texture2D") {
    desc.width = w;
    desc.height = h;
    size_t bpp;
    switch (fmt) {  // rtc::DataType codes of the ABI
      case BN_FLOAT32_VEC4: desc.format = 33; bpp = 16; break;
      case BN_FLOAT32: desc.format = 30; bpp = 4; break;
      case BN_UFIXED8_RGBA: case BN_UINT8_VEC4: desc.format = 1; bpp = 4; break;
      case BN_UFIXED8: case BN_UINT8: desc.format = 0; bpp = 1; break;
      default: throw std::runtime_error("bnTexture2DCreate: unsupported texel format");
    }
    if (w <= 0 || h <= 0 || !texels) throw std::runtime_error("bnTexture2DCreate: bad size or NULL texels");
    bytes.assign((const uint8_t *)texels, (const uint8_t *)texels + bpp * (size_t)w * h);
    desc.filterMode = filter == BN_TEXTURE_NEAREST ? 0 : 1;
    desc.addressMode[0] = (int)ax;  // BNTextureAddressMode = rtc::AddressMode order
    desc.addressMode[1] = (int)ay;
  }
  ~Texture2D() override {
    if (tex) (void)bnamd_texture2d_destroy(tex);
  }
  bnamd_texture2d *device() {
    if (!tex) {
      desc.texels = bytes.data();
      check(bnamd_texture2d_create(&desc, nullptr, &tex), "bnTexture2DCreate");
    }
    return tex;
  }
  bnamd_texture2d_desc desc{};
  std::vector<uint8_t> bytes;
  bnamd_texture2d *tex = nullptr;
};

struct TexData : Object {
  TexData(Context *c, int w, int h, int d, const void *items) : Object(c, "textureData"), dims{w, h, d} {
    texels.assign((const float *)items, (const float *)items + (size_t)w * h * d);
  }
  int dims[3];
  std::vector<float> texels;
};
struct VolumeObj : Object {
  VolumeObj(Context *c, Object *f) : Object(c, "volume"), field(f) {}
  Object *field;
  float domain[2] = {0.f, 1.f};
  std::vector<float> xf = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 1.f};
  float densityAt1 = 1.f;
};
struct Group : Object {
  using Object::Object;
  std::vector<Object *> geoms;
  std::vector<VolumeObj *> volumes;
};
struct Model : Object {
  using Object::Object;
  std::vector<Group *> groups;
  std::vector<BNTransform> xfms;
  bnamd_world *world = nullptr;
  bnamd_volume *volume = nullptr;
  bnamd_umesh *umesh = nullptr;
  bnamd_amr *amr = nullptr;
  bnamd_nanovdb *nvdb = nullptr;
  bnamd_envmap *env = nullptr;
  bnamd_lights *lights = nullptr;
  std::vector<std::unique_ptr<DevMem>> mem;
  void release() {
    if (world) (void)bnamd_world_destroy(world);
    if (volume) (void)bnamd_volume_destroy(volume);
    if (umesh) (void)bnamd_umesh_destroy(umesh);
    if (amr) (void)bnamd_amr_destroy(amr);
    if (nvdb) (void)bnamd_nanovdb_destroy(nvdb);
    if (env) (void)bnamd_envmap_destroy(env);
    if (lights) (void)bnamd_lights_destroy(lights);
    world = nullptr;
    volume = nullptr;
    umesh = nullptr;
    amr = nullptr;
    nvdb = nullptr;
    env = nullptr;
    lights = nullptr;
    mem.clear();
  }
  ~Model() override { release(); }
};
struct FrameBuffer : Object {
  using Object::Object;
  int w = 0, h = 0;
  uint32_t channels = BN_FB_COLOR;
  int paths = 0;  // paths accumulated per pixel
  std::unique_ptr<DevMem> accum, ids;
};

struct Context {
  std::vector<std::unique_ptr<Object>> objects;
  template <class T>
  T *add(T *p) {
    objects.emplace_back(p);
    return p;
  }
};

Object *obj(BNObject h) {
  if (!h) throw std::runtime_error("bn*: NULL object");
  return static_cast<Object *>(h);
}

void slot0(int s) {
  if (s != 0) throw std::runtime_error("bn*: this backend has one data slot (0)");
}

// AnariMatte / AnariPBR parameters (material/AnariMatte.cpp:29-90,
// AnariPBR.cpp:36-80) to the ABI's material record
bnamd_material make_material(const Object *m) {
  bnamd_material r{};
  r.opacity = 1.f;
  r.ior = 1.45f;
  r.specular = 0.5f;
  if (!m || m->type == "matte" || m->type == "AnariMatte") {
    const auto c = m ? m->get4("color", {.8f, .8f, .8f, 1.f}) : std::array<float, 4>{.8f, .8f, .8f, 1.f};
    r.type = BNAMD_MAT_MATTE;
    r.color[0] = c[0]; r.color[1] = c[1]; r.color[2] = c[2];
    if (m) r.opacity = m->getf("opacity", 1.f);
    return r;
  }
  if (m->type != "physicallyBased" && m->type != "AnariPBR") throw std::runtime_error("bnMaterialCreate: unknown type " + m->type);
  const auto c = m->get4("baseColor", {1.f, 1.f, 1.f, 1.f});
  r.type = BNAMD_MAT_PBR;
  r.color[0] = c[0]; r.color[1] = c[1]; r.color[2] = c[2];
  r.metallic = m->getf("metallic", 1.f);
  r.roughness = m->getf("roughness", 1.f);
  r.transmission = m->getf("transmission", 0.f);
  r.ior = m->getf("ior", 1.45f);  // AnariPBR.h:57
  r.opacity = m->getf("opacity", 1.f);
  return r;
}

const Data *data_param(const Object *g, const char *name, BNDataType want, bool required) {
  const Object *d = g->geto(name);
  if (!d) {
    if (required) throw std::runtime_error("bnBuild: " + g->type + " without '" + name + "'");
    return nullptr;
  }
  const Data *dd = dynamic_cast<const Data *>(d);
  if (!dd || dd->dataType != want) throw std::runtime_error("bnBuild: '" + std::string(name) + "' has the wrong type");
  return dd;
}

}  // namespace

BN_API BNContext bnContextCreate(const int *, int, const int *gpuIDs, int numGPUs) {
  check(bnamd_init((gpuIDs && numGPUs > 0) ? gpuIDs[0] : 0), "bnContextCreate");
  return reinterpret_cast<BNContext>(new Context);
}

BN_API void bnContextDestroy(BNContext context) { delete reinterpret_cast<Context *>(context); }

BN_API void bnRelease(BNObject object) { --obj(object)->appRefs; }
BN_API void bnAddReference(BNObject object) { ++obj(object)->appRefs; }

BN_API BNCamera bnCameraCreate(BNContext context, const char *type) {
  const std::string t = type ? type : "";
  if (t != "perspective" && t != "pinhole") throw std::runtime_error("bnCameraCreate: unknown type " + t);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNCamera>(static_cast<_BNObject *>(c->add(new Object(c, "perspective"))));
}

BN_API BNFrameBuffer bnFrameBufferCreate(BNContext context, int) {
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNFrameBuffer>(static_cast<_BNObject *>(c->add(new FrameBuffer(c, "frameBuffer"))));
}

BN_API BNModel bnModelCreate(BNContext ctx) {
  auto *c = reinterpret_cast<Context *>(ctx);
  return static_cast<BNModel>(static_cast<_BNObject *>(c->add(new Model(c, "model"))));
}

BN_API BNRenderer bnRendererCreate(BNContext ctx, const char *type) {
  const std::string t = type ? type : "default";
  if (t != "default" && t != "pathTracer") throw std::runtime_error("bnRendererCreate: unknown type " + t);
  auto *c = reinterpret_cast<Context *>(ctx);
  return static_cast<BNRenderer>(static_cast<_BNObject *>(c->add(new Object(c, "renderer"))));
}

BN_API void bnCommit(BNObject target) { (void)obj(target); }  // parameters are read at build / render time

BN_API void bnSetData(BNObject target, const char *paramName, BNData value) {
  obj(target)->o[paramName] = value ? obj(value) : nullptr;
}
BN_API void bnSetObject(BNObject target, const char *paramName, const BNObject value) {
  obj(target)->o[paramName] = value ? obj(value) : nullptr;
}
BN_API void bnSetString(BNObject target, const char *paramName, const char *value) {
  obj(target)->str[paramName] = value ? value : "";
}
BN_API void bnSetLight(BNObject target, const char *paramName, BNLight value) {
  obj(target)->o[paramName] = value ? obj(value) : nullptr;
}
BN_API void bnSet1i(BNObject target, const char *paramName, int value) { obj(target)->i[paramName] = value; }
BN_API void bnSet2i(BNObject target, const char *paramName, int x, int y) {
  obj(target)->iv[paramName] = {x, y, 0, 0};
}
BN_API void bnSet3i(BNObject target, const char *paramName, int x, int y, int z) {
  obj(target)->iv[paramName] = {x, y, z, 0};
}
BN_API void bnSet4i(BNObject target, const char *paramName, int x, int y, int z, int w) {
  obj(target)->iv[paramName] = {x, y, z, w};
}
BN_API void bnSet1f(BNObject target, const char *paramName, float value) {
  obj(target)->f[paramName] = {value, 0.f, 0.f, 0.f};
}
BN_API void bnSet2f(BNObject target, const char *paramName, float x, float y) {
  obj(target)->f[paramName] = {x, y, 0.f, 0.f};
}
BN_API void bnSet4x3fv(BNObject target, const char *paramName, const BNTransform *x) {
  if (!x) throw std::runtime_error("bnSet4x3fv: NULL matrix");
  obj(target)->mat[paramName] = {x->l.vx.x, x->l.vx.y, x->l.vx.z, x->l.vy.x, x->l.vy.y, x->l.vy.z,
                                 x->l.vz.x, x->l.vz.y, x->l.vz.z, x->p.x,    x->p.y,    x->p.z,
                                 0.f,       0.f,       0.f,       1.f};
}
BN_API void bnSet4x4fv(BNObject target, const char *paramName, const bn_float4 *x) {
  if (!x) throw std::runtime_error("bnSet4x4fv: NULL matrix");
  std::array<float, 16> m{};
  std::memcpy(m.data(), x, sizeof m);
  obj(target)->mat[paramName] = m;
}
BN_API void bnSet3f(BNObject target, const char *paramName, float x, float y, float z) {
  obj(target)->f[paramName] = {x, y, z, 0.f};
}
BN_API void bnSet4f(BNObject target, const char *paramName, float x, float y, float z, float w) {
  obj(target)->f[paramName] = {x, y, z, w};
}

BN_API BNData bnDataCreate(BNContext context, int whichSlot, BNDataType dataType, size_t numItems, const void *items) {
  slot0(whichSlot);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNData>(static_cast<_BNObject *>(c->add(new Data(c, dataType, numItems, items))));
}

BN_API BNGeom bnGeometryCreate(BNContext context, int whichSlot, const char *type) {
  slot0(whichSlot);
  const std::string t = type ? type : "";
  if (t != "triangles" && t != "spheres" && t != "cylinders") throw std::runtime_error("bnGeometryCreate: unknown type " + t);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNGeom>(static_cast<_BNObject *>(c->add(new Object(c, t))));
}

// Light::create (barney/light/Light.cpp:22-35)
BN_API BNLight bnLightCreate(BNContext context, int whichSlot, const char *type) {
  slot0(whichSlot);
  const std::string t = type ? type : "";
  if (t != "envmap" && t != "directional" && t != "point" && t != "quad")
    throw std::runtime_error("bnLightCreate: unknown type " + t);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNLight>(static_cast<_BNObject *>(c->add(new Object(c, t))));
}

BN_API BNTexture2D bnTexture2DCreate(BNContext context, int whichSlot, BNDataType texelFormat, uint32_t size_x,
                                     uint32_t size_y, const void *texels, BNTextureFilterMode filterMode,
                                     BNTextureAddressMode addressMode_x, BNTextureAddressMode addressMode_y,
                                     BNTextureColorSpace colorSpace) {
  slot0(whichSlot);
  if (colorSpace != BN_COLOR_SPACE_LINEAR)
    throw std::runtime_error("bnTexture2DCreate: only the linear color space");
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNTexture2D>(static_cast<_BNObject *>(
      c->add(new Texture2D(c, texelFormat, (int)size_x, (int)size_y, texels, filterMode, addressMode_x,
                           addressMode_y))));
}

BN_API BNMaterial bnMaterialCreate(BNContext context, int whichSlot, const char *type) {
  slot0(whichSlot);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNMaterial>(static_cast<_BNObject *>(c->add(new Object(c, type ? type : "matte"))));
}

BN_API BNTextureData bnTextureData3DCreate(BNContext context, int whichSlot, BNDataType texelFormat, int width,
                                           int height, int depth, const void *items) {
  slot0(whichSlot);
  if (texelFormat != BN_FLOAT32) throw std::runtime_error("bnTextureData3DCreate: only BN_FLOAT32 texels");
  if (width < 2 || height < 2 || depth < 2 || !items) throw std::runtime_error("bnTextureData3DCreate: bad size");
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNTextureData>(static_cast<_BNObject *>(c->add(new TexData(c, width, height, depth, items))));
}

BN_API BNScalarField bnScalarFieldCreate(BNContext context, int whichSlot, const char *type) {
  slot0(whichSlot);
  const std::string t = type ? type : "";
  if (t != "structured" && t != "unstructured" && t != "BlockStructuredAMR" && t != "NanoVDB")
    throw std::runtime_error("bnScalarFieldCreate: unknown type " + t);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNScalarField>(static_cast<_BNObject *>(c->add(new Object(c, t))));
}

BN_API BNVolume bnVolumeCreate(BNContext context, int whichSlot, BNScalarField sf) {
  slot0(whichSlot);
  auto *c = reinterpret_cast<Context *>(context);
  return static_cast<BNVolume>(static_cast<_BNObject *>(c->add(new VolumeObj(c, obj(sf)))));
}

// Volume::setXF (barney/volume/Volume.cpp, TransferFunction.cpp)
BN_API void bnVolumeSetXF(BNVolume volume, bn_float2 domain, const bn_float4 *colorMap, int numColorMapValues,
                          float densityAt1) {
  auto *v = static_cast<VolumeObj *>(obj(volume));
  if (!colorMap || numColorMapValues < 2) throw std::runtime_error("bnVolumeSetXF: need at least 2 values");
  v->domain[0] = domain.x;
  v->domain[1] = domain.y;
  v->xf.assign((const float *)colorMap, (const float *)colorMap + 4 * (size_t)numColorMapValues);
  v->densityAt1 = densityAt1;
}

BN_API BNGroup bnGroupCreate(BNContext context, int whichSlot, BNGeom *geoms, int numGeoms, BNVolume *volumes,
                             int numVolumes) {
  slot0(whichSlot);
  auto *c = reinterpret_cast<Context *>(context);
  auto *g = c->add(new Group(c, "group"));
  for (int k = 0; k < numGeoms; ++k) g->geoms.push_back(obj(geoms[k]));
  for (int k = 0; k < numVolumes; ++k) g->volumes.push_back(static_cast<VolumeObj *>(obj(volumes[k])));
  return static_cast<BNGroup>(static_cast<_BNObject *>(g));
}

BN_API void bnGroupBuild(BNGroup group) { (void)obj(group); }  // built with the model (one world per model)

BN_API void bnSetInstances(BNModel model, int whichSlot, BNGroup *groups, BNTransform *xfms, int numInstances) {
  slot0(whichSlot);
  auto *m = static_cast<Model *>(obj(model));
  m->groups.clear();
  m->xfms.clear();
  for (int k = 0; k < numInstances; ++k) {
    m->groups.push_back(static_cast<Group *>(obj(groups[k])));
    m->xfms.push_back(xfms[k]);
  }
}

namespace {

// the structured, unstructured, block-structured or NanoVDB field of a
// volume model
void build_volume(Model *m, const VolumeObj *vol) {
  const Object *f = vol->field;
  auto upload = [&](const void *p, size_t bytes) {
    m->mem.emplace_back(new DevMem(p, bytes));
    return m->mem.back()->ptr;
  };
  if (f->type == "structured") {
    // StructuredData::commit and buildMCs with barney's 8 cells per
    // macrocell (StructuredData.cu:25-95)
    const TexData *td = dynamic_cast<const TexData *>(f->geto("textureData"));
    if (!td) throw std::runtime_error("bnBuild: structured field without 'textureData'");
    bnamd_structured_desc d{};
    d.scalars = (const float *)upload(td->texels.data(), td->texels.size() * 4);
    const auto o = f->get4("gridOrigin", {0.f, 0.f, 0.f, 0.f});
    const auto sp = f->get4("gridSpacing", {1.f, 1.f, 1.f, 0.f});
    for (int k = 0; k < 3; ++k) {
      d.dims[k] = td->dims[k];
      d.gridOrigin[k] = o[k];
      d.gridSpacing[k] = sp[k];
    }
    d.cellsPerMC = 8;
    check(bnamd_volume_create(&d, &m->volume), "bnBuild");
    check(bnamd_volume_set_tf(m->volume, vol->xf.data(), (int)vol->xf.size() / 4, vol->domain[0], vol->domain[1],
                              vol->densityAt1),
          "bnBuild");
  } else if (f->type == "unstructured") {
    // UMeshField::setData / commit (UMeshField.cu:204-236)
    if (f->geto("cell.data")) throw std::runtime_error("bnBuild: per-cell scalars ('cell.data') are not supported");
    const Data *v = data_param(f, "vertex.position", BN_FLOAT32_VEC3, true);
    const Data *sc = data_param(f, "vertex.data", BN_FLOAT32, true);
    const Data *ix = data_param(f, "index", BN_INT32, true);
    const Data *co = data_param(f, "cell.index", BN_INT32, true);
    const Data *ct = data_param(f, "cell.type", BN_UINT8, true);
    if (sc->count != v->count || ct->count != co->count)
      throw std::runtime_error("bnBuild: unstructured field arrays of mismatched sizes");
    bnamd_umesh_desc d{};
    d.vertices = (const float *)upload(v->bytes.data(), v->bytes.size());
    d.scalars = (const float *)upload(sc->bytes.data(), sc->bytes.size());
    d.numVertices = (int)v->count;
    d.indices = (const int *)upload(ix->bytes.data(), ix->bytes.size());
    d.numIndices = (int)ix->count;
    d.cellOffsets = (const int *)upload(co->bytes.data(), co->bytes.size());
    d.cellTypes = (const unsigned char *)upload(ct->bytes.data(), ct->bytes.size());
    d.numCells = (int)co->count;
    check(bnamd_umesh_create(&d, nullptr, &m->umesh), "bnBuild");
    check(bnamd_umesh_set_tf(m->umesh, vol->xf.data(), (int)vol->xf.size() / 4, vol->domain[0], vol->domain[1],
                             vol->densityAt1),
          "bnBuild");
  } else if (f->type == "NanoVDB") {
    // NanoVDBData::setData("data") / commit: the grid buffer as BN_UINT8
    // data (NanoVDB.cu:190-235; anari/SpatialField.cpp hands it so)
    const Data *g = data_param(f, "data", BN_UINT8, true);
    bnamd_nanovdb_desc d{};
    d.data = g->bytes.data();
    d.bytes = g->bytes.size();
    check(bnamd_nanovdb_create(&d, nullptr, &m->nvdb), "bnBuild");
    check(bnamd_nanovdb_set_tf(m->nvdb, vol->xf.data(), (int)vol->xf.size() / 4, vol->domain[0], vol->domain[1],
                               vol->densityAt1),
          "bnBuild");
  } else {
    // BlockStructuredField::setData / commit (BlockStructuredField.cu:225-300)
    const Data *orig = data_param(f, "grid.origins", BN_INT32_VEC3, true);
    const Data *dims = data_param(f, "grid.dims", BN_INT32_VEC3, true);
    const Data *lev = data_param(f, "grid.levels", BN_INT32, true);
    const Data *sc = data_param(f, "scalars", BN_FLOAT32, true);
    const Data *ref = data_param(f, "level.refinements", BN_INT32, true);
    const Object *offo = f->geto("grid.offsets");
    const Data *off = dynamic_cast<const Data *>(offo);
    if (!off) throw std::runtime_error("bnBuild: BlockStructuredAMR field without 'grid.offsets'");
    const std::vector<unsigned long long> offsets = off->as_u64();
    const size_t nb = orig->count;
    if (dims->count != nb || lev->count != nb || offsets.size() != nb)
      throw std::runtime_error("bnBuild: BlockStructuredAMR arrays of mismatched sizes");
    bnamd_amr_desc d{};
    d.numBlocks = (int)nb;
    d.origins = (const int *)orig->bytes.data();
    d.dims = (const int *)dims->bytes.data();
    d.levels = (const int *)lev->bytes.data();
    d.offsets = offsets.data();
    d.numLevels = (int)ref->count;
    d.refinements = (const int *)ref->bytes.data();
    d.scalars = (const float *)upload(sc->bytes.data(), sc->bytes.size());
    d.numScalars = sc->count;
    check(bnamd_amr_create(&d, nullptr, &m->amr), "bnBuild");
    check(bnamd_amr_set_tf(m->amr, vol->xf.data(), (int)vol->xf.size() / 4, vol->domain[0], vol->domain[1],
                           vol->densityAt1),
          "bnBuild");
  }
}

// affine3f applied to a point / a vector (owl's xfmPoint / xfmVector)
std::array<float, 3> xfm_vector(const BNTransform &x, const std::array<float, 4> &v) {
  return {x.l.vx.x * v[0] + x.l.vy.x * v[1] + x.l.vz.x * v[2], x.l.vx.y * v[0] + x.l.vy.y * v[1] + x.l.vz.y * v[2],
          x.l.vx.z * v[0] + x.l.vy.z * v[1] + x.l.vz.z * v[2]};
}
std::array<float, 3> xfm_point(const BNTransform &x, const std::array<float, 4> &v) {
  const auto r = xfm_vector(x, v);
  return {r[0] + x.p.x, r[1] + x.p.y, r[2] + x.p.z};
}

// ModelSlot::updateWorldLightsFromInstances (ModelSlot.cpp:90-125): each
// instance's group lights, quad / directional / point lights through their
// getDD with the instance transform (QuadLight.cu:9-18, DirLight.cu:10-20,
// PointLight.cu:9-17), and the env-map light (the last one met);
// EnvMapLight::commit with its "texture", "direction", "up", "scale"
// (EnvMap.cu:157-171, 280-330)
void build_lights(Model *m) {
#pragma clang fp contract(off)
  const Object *env = nullptr;
  std::vector<bnamd_quad_light> quads;
  std::vector<bnamd_dir_light> dirs;
  std::vector<bnamd_point_light> points;
  const float nan = std::numeric_limits<float>::quiet_NaN();
  for (size_t k = 0; k < m->groups.size(); ++k) {
    const Group *g = m->groups[k];
    const BNTransform &x = m->xfms[k];
    const Data *lights = dynamic_cast<const Data *>(g->geto("lights"));
    if (!lights) continue;
    if (lights->dataType != BN_OBJECT) throw std::runtime_error("bnBuild: 'lights' must be BN_OBJECT data");
    for (const Object *l : lights->objs) {
      if (!l) continue;
      const auto color = l->get4("color", {1.f, 1.f, 1.f, 0.f});  // Light::color
      if (l->type == "envmap") {
        env = l;
      } else if (l->type == "quad") {  // QuadLight::DD defaults (QuadLight.h:15-22)
        bnamd_quad_light q{};
        const auto c = xfm_point(x, l->get4("corner", {0.f, 0.f, 0.f, 0.f}));
        const auto e0 = xfm_vector(x, l->get4("edge0", {1.f, 0.f, 0.f, 0.f}));
        const auto e1 = xfm_vector(x, l->get4("edge1", {0.f, 1.f, 0.f, 0.f}));
        const auto em = l->get4("emission", {1.f, 1.f, 1.f, 0.f});
        for (int j = 0; j < 3; ++j) {
          q.corner[j] = c[j];
          q.edge0[j] = e0[j];
          q.edge1[j] = e1[j];
          q.emission[j] = em[j];
        }
        quads.push_back(q);
      } else if (l->type == "directional") {  // DirLight.h defaults: (0,0,-1), radiance 1
        bnamd_dir_light d{};
        const auto v = xfm_vector(x, l->get4("direction", {0.f, 0.f, -1.f, 0.f}));
        for (int j = 0; j < 3; ++j) {
          d.direction[j] = v[j];
          d.color[j] = color[j];
        }
        const float irr = l->getf("irradiance", nan);
        d.radiance = std::isnan(irr) ? l->getf("radiance", 1.f) : irr;
        dirs.push_back(d);
      } else if (l->type == "point") {  // PointLight.h defaults: origin, power 1
        bnamd_point_light pt{};
        const auto v = xfm_point(x, l->get4("position", {0.f, 0.f, 0.f, 0.f}));
        for (int j = 0; j < 3; ++j) {
          pt.position[j] = v[j];
          pt.color[j] = color[j];
        }
        pt.power = l->getf("power", 1.f);
        pt.intensity = l->getf("intensity", nan);
        points.push_back(pt);
      } else {
        throw std::runtime_error("bnBuild: un-handled type of light '" + l->type + "'");
      }
    }
  }
  if (!quads.empty() || !dirs.empty() || !points.empty())
    check(bnamd_lights_create(quads.data(), (int)quads.size(), dirs.data(), (int)dirs.size(), points.data(),
                              (int)points.size(), &m->lights),
          "bnBuild");
  if (!env) return;
  auto *tex = dynamic_cast<Texture2D *>(env->geto("texture"));
  if (!tex) throw std::runtime_error("bnBuild: envmap light without a 2D 'texture'");
  bnamd_envmap_desc d{};
  const auto dir = env->get4("direction", {1.f, 0.f, 0.f, 0.f});
  const auto up = env->get4("up", {0.f, 0.f, 1.f, 0.f});
  for (int k = 0; k < 3; ++k) {
    d.direction[k] = dir[k];
    d.up[k] = up[k];
  }
  d.scale = env->getf("scale", 1.f);
  check(bnamd_envmap_create(tex->device(), &d, nullptr, &m->env), "bnBuild");
}

}  // namespace

BN_API void bnBuild(BNModel model, int whichDataSlot) {
  slot0(whichDataSlot);
  auto *m = static_cast<Model *>(obj(model));
  m->release();
  build_lights(m);
  size_t nvol = 0, ngeom = 0;
  const VolumeObj *vol = nullptr;
  for (const Group *g : m->groups) {
    nvol += g->volumes.size();
    ngeom += g->geoms.size();
    if (!g->volumes.empty()) vol = g->volumes[0];
  }
  if (nvol > 0 && ngeom == 0) {
    if (nvol != 1 || m->groups.size() != 1)
      throw std::runtime_error("bnBuild: a volume model holds one volume in one instance in this backend");
    build_volume(m, vol);
    return;
  }
  if (nvol > 0) {
    // geometry and a volume in one model: the surface world plus one
    // structured volume, rendered by bnamd_render_mixed
    if (nvol != 1 || vol->field->type != "structured")
      throw std::runtime_error("bnBuild: a model mixing geometry and volumes takes one structured volume in this "
                               "backend");
    build_volume(m, vol);
  }
  std::vector<Group *> groups;  // distinct groups, in first-use order
  std::vector<bnamd_instance> insts;
  for (size_t k = 0; k < m->groups.size(); ++k) {
    if (m->groups[k]->geoms.empty()) continue;  // the volume's own instance
    size_t gi = 0;
    while (gi < groups.size() && groups[gi] != m->groups[k]) ++gi;
    if (gi == groups.size()) groups.push_back(m->groups[k]);
    bnamd_instance in{};
    in.group = (int)gi;
    in.instID = (int)k;
    const BNTransform &x = m->xfms[k];
    const float v[12] = {x.l.vx.x, x.l.vx.y, x.l.vx.z, x.l.vy.x, x.l.vy.y, x.l.vy.z,
                         x.l.vz.x, x.l.vz.y, x.l.vz.z, x.p.x,    x.p.y,    x.p.z};
    std::memcpy(in.xfm, v, sizeof v);
    insts.push_back(in);
  }
  std::vector<bnamd_material> mats;
  std::vector<const Object *> matObjs;
  auto matIndex = [&](const Object *mo) {
    for (size_t k = 0; k < matObjs.size(); ++k)
      if (matObjs[k] == mo) return (int)k;
    matObjs.push_back(mo);
    mats.push_back(make_material(mo));
    return (int)mats.size() - 1;
  };
  auto upload = [&](const Data *d) -> const void * {
    if (!d) return nullptr;
    m->mem.emplace_back(new DevMem(d->bytes.data(), d->bytes.size()));
    return m->mem.back()->ptr;
  };
  std::vector<std::vector<bnamd_geom>> gg(groups.size());
  std::vector<bnamd_group> bg(groups.size());
  for (size_t k = 0; k < groups.size(); ++k) {
    for (const Object *g : groups[k]->geoms) {
      bnamd_geom e{};
      e.materialID = matIndex(g->geto("material"));
      e.userID = g->geti("userID", 0);
      if (g->type == "triangles") {  // Triangles.cpp:29-40
        const Data *v = data_param(g, "vertices", BN_FLOAT32_VEC3, true);
        const Data *ix = data_param(g, "indices", BN_INT32_VEC3, true);
        const Data *nr = data_param(g, "normals", BN_FLOAT32_VEC3, false);  // Triangles.cpp:37-40
        if (nr && nr->count != v->count) throw std::runtime_error("bnBuild: 'normals' needs one per vertex");
        e.kind = BNAMD_GEOM_TRIANGLES;
        e.numPrims = (int)ix->count;
        e.numVertices = (int)v->count;
        e.vertices = (const float *)upload(v);
        e.indices = (const int *)upload(ix);
        e.normals = (const float *)upload(nr);
      } else if (g->type == "spheres") {  // Spheres.cpp:48-70, default radius .1
        const Data *v = data_param(g, "origins", BN_FLOAT32_VEC3, true);
        const Data *r = data_param(g, "radii", BN_FLOAT32, false);
        const Data *col = data_param(g, "colors", BN_FLOAT32_VEC3, false);  // Spheres.cpp:59-62
        if (col && col->count != v->count) throw std::runtime_error("bnBuild: 'colors' needs one per sphere");
        e.kind = BNAMD_GEOM_SPHERES;
        e.numPrims = e.numVertices = (int)v->count;
        e.vertices = (const float *)upload(v);
        e.radii = (const float *)upload(r);
        e.colors = (const float *)upload(col);
        e.defaultRadius = g->getf("radius", .1f);
      } else {  // cylinders: Cylinders.cpp:60-71
        const Data *v = data_param(g, "vertices", BN_FLOAT32_VEC3, true);
        const Data *ix = data_param(g, "indices", BN_INT32_VEC2, true);
        const Data *r = data_param(g, "radii", BN_FLOAT32, false);
        e.kind = BNAMD_GEOM_CYLINDERS;
        e.numPrims = (int)ix->count;
        e.numVertices = (int)v->count;
        e.vertices = (const float *)upload(v);
        e.indices = (const int *)upload(ix);
        e.radii = (const float *)upload(r);
        e.defaultRadius = g->getf("radius", .1f);
      }
      gg[k].push_back(e);
    }
    bg[k].geoms = gg[k].data();
    bg[k].numGeoms = (int)gg[k].size();
  }
  if (mats.empty()) mats.push_back(make_material(nullptr));
  check(bnamd_world_build(bg.data(), (int)bg.size(), insts.data(), (int)insts.size(), mats.data(), (int)mats.size(),
                          nullptr, &m->world),
        "bnBuild");
}

BN_API void bnAccumReset(BNFrameBuffer fb) { static_cast<FrameBuffer *>(obj(fb))->paths = 0; }

BN_API void bnFrameBufferResize(BNFrameBuffer fbh, BNDataType, int sizeX, int sizeY, uint32_t requiredChannels) {
  auto *fb = static_cast<FrameBuffer *>(obj(fbh));
  if (sizeX <= 0 || sizeY <= 0) throw std::runtime_error("bnFrameBufferResize: bad size");
  if (requiredChannels & BN_FB_NORMAL) throw std::runtime_error("bnFrameBufferResize: no normal channel in this backend");
  fb->w = sizeX;
  fb->h = sizeY;
  fb->channels = requiredChannels | BN_FB_COLOR;
  fb->paths = 0;
  fb->accum.reset(new DevMem((size_t)sizeX * sizeY * 16));
  const bool ids = requiredChannels & (BN_FB_DEPTH | BN_FB_PRIMID | BN_FB_INSTID | BN_FB_OBJID);
  fb->ids.reset(ids ? new DevMem((size_t)sizeX * sizeY * 16) : nullptr);
}

BN_API void bnFrameBufferGetSize(BNFrameBuffer fbh, int *sizeX, int *sizeY) {
  auto *fb = static_cast<FrameBuffer *>(obj(fbh));
  if (sizeX) *sizeX = fb->w;
  if (sizeY) *sizeY = fb->h;
}

BN_API void bnRender(BNRenderer renderer, BNModel model, BNCamera camera, BNFrameBuffer fbh) {
  const Object *r = obj(renderer), *cam = obj(camera);
  auto *m = static_cast<Model *>(obj(model));
  auto *fb = static_cast<FrameBuffer *>(obj(fbh));
  if (!m->world && !m->volume && !m->umesh && !m->amr && !m->nvdb)
    throw std::runtime_error("bnRender: model not built (bnBuild)");
  if (!fb->accum) throw std::runtime_error("bnRender: frame buffer not sized (bnFrameBufferResize)");
  // PerspectiveCamera parameters (Camera.cu:30-80)
  bnamd_camera c{};
  const auto pos = cam->get4("position", {0.f, 0.f, 0.f, 0.f});
  const auto dir = cam->get4("direction", {0.f, 0.f, 1.f, 0.f});
  const auto up = cam->get4("up", {0.f, 1.f, 0.f, 0.f});
  for (int k = 0; k < 3; ++k) {
    c.position[k] = pos[k];
    c.direction[k] = dir[k];
    c.up[k] = up[k];
  }
  c.fovy = cam->getf("fovy", 60.f);
  // Renderer parameters (render/Renderer.cpp:26-95; Renderer.h:55-57)
  bnamd_render_params p{};
  p.width = fb->w;
  p.height = fb->h;
  p.accumID = fb->paths;
  p.pathsPerPixel = r->geti("pathsPerPixel", 1);
  p.ambientRadiance = r->getf("ambientRadiance", 1.f);
  const auto bg = r->get4("bgColor", {0.f, 0.f, 0.f, 1.f});
  for (int k = 0; k < 4; ++k) p.bgColor[k] = bg[k];
  p.maxPathDepth = 10;
  p.maxDiffuseBounces = 1;
  p.russianRoulette = 0;
  p.envMap = m->env;
  p.lights = m->lights;
  const auto cut = r->get4("cutPlane", {0.f, 0.f, 0.f, -1e30f});  // Renderer.cpp:95-98
  for (int k = 0; k < 4; ++k) p.cutPlane[k] = cut[k];
  int *ids = (fb->ids && fb->paths == 0) ? (int *)fb->ids->ptr : nullptr;
  if (!m->world) {
    // the volume renderer writes no ID channels: a volume model reads back
    // as all misses (primID/instID/objID -1, depth +inf)
    if (ids) {
      const size_t n = (size_t)fb->w * fb->h;
      std::vector<int32_t> miss(4 * n, -1);
      const float inf = std::numeric_limits<float>::infinity();
      for (size_t q = 0; q < n; ++q) std::memcpy(&miss[4 * q + 3], &inf, 4);
      check(bnamd_memcpy_h2d(ids, miss.data(), miss.size() * 4), "bnRender");
    }
    if (m->volume)
      check(bnamd_volume_render(m->volume, &c, &p, (float *)fb->accum->ptr, nullptr, nullptr), "bnRender");
    else if (m->amr)
      check(bnamd_amr_render(m->amr, &c, &p, (float *)fb->accum->ptr, nullptr, nullptr), "bnRender");
    else if (m->nvdb)
      check(bnamd_nanovdb_render(m->nvdb, &c, &p, (float *)fb->accum->ptr, nullptr, nullptr), "bnRender");
    else {
      bnamd_umesh_render_opts o{};
      o.volume = 1;
      check(bnamd_umesh_render(m->umesh, &c, &p, &o, (float *)fb->accum->ptr, nullptr, nullptr), "bnRender");
    }
  } else if (m->volume)
    check(bnamd_render_mixed(m->world, m->volume, &c, &p, (float *)fb->accum->ptr, ids, nullptr, nullptr, nullptr),
          "bnRender");
  else
    check(bnamd_render_surface(m->world, &c, &p, (float *)fb->accum->ptr, ids, nullptr, nullptr), "bnRender");
  fb->paths += p.pathsPerPixel;
}

BN_API void bnFrameBufferRead(BNFrameBuffer fbh, BNFrameBufferChannel channel, void *out, BNDataType format) {
  auto *fb = static_cast<FrameBuffer *>(obj(fbh));
  if (!fb->accum || fb->paths == 0) throw std::runtime_error("bnFrameBufferRead: nothing rendered");
  const size_t n = (size_t)fb->w * fb->h;
  if (channel == BN_FB_COLOR) {
    int fmt;
    size_t bpp;
    switch (format) {
      case BN_FLOAT32_VEC4: fmt = BNAMD_FB_FLOAT4; bpp = 16; break;
      case BN_UFIXED8_RGBA: fmt = BNAMD_FB_UFIXED8_RGBA; bpp = 4; break;
      case BN_UFIXED8_RGBA_SRGB: fmt = BNAMD_FB_UFIXED8_RGBA_SRGB; bpp = 4; break;
      default: throw std::runtime_error("bnFrameBufferRead: unsupported color format");
    }
    DevMem tmp(n * bpp);
    check(bnamd_fb_resolve((const float *)fb->accum->ptr, fb->w, fb->h, 1.f / (float)fb->paths, fmt, tmp.ptr, nullptr),
          "bnFrameBufferRead");
    check(bnamd_synchronize(), "bnFrameBufferRead");
    check(bnamd_memcpy_d2h(out, tmp.ptr, n * bpp), "bnFrameBufferRead");
    return;
  }
  if (!fb->ids || !(fb->channels & channel)) throw std::runtime_error("bnFrameBufferRead: channel not requested at resize");
  std::vector<int32_t> rec(4 * n);
  check(bnamd_memcpy_d2h(rec.data(), fb->ids->ptr, rec.size() * 4), "bnFrameBufferRead");
  const int k = channel == BN_FB_DEPTH ? 3 : channel == BN_FB_PRIMID ? 0 : channel == BN_FB_INSTID ? 1 : 2;
  if (channel == BN_FB_DEPTH && format != BN_FLOAT32) throw std::runtime_error("bnFrameBufferRead: depth is BN_FLOAT32");
  if (channel != BN_FB_DEPTH && format != BN_INT32) throw std::runtime_error("bnFrameBufferRead: IDs are BN_INT32");
  for (size_t q = 0; q < n; ++q) std::memcpy((char *)out + 4 * q, &rec[4 * q + k], 4);
}
