// Conformance check of the rtc::amd host layer (built with g++ against
// librtcore_amd.so / libbnamd.so): device memory and pinned host memory,
// the active-device guard, Triangles / Spheres / Cylinders geometry types,
// groups, an instance group, closest-hit tracing with analytic answers,
// setDD with barney's DD bytes, one bottom level shared by 1,000 instances
// and a top-level refit, getAccelHandle, texture data and textures, and
// error reporting as std::runtime_error.
// Exit codes: 0 = all checks passed, 1 = a check failed, 3 = no usable GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../barney_amd/rtcore_amd/rtcore_amd.h"

using namespace rtc::amd;

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("CHECK FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static bool near(float a, float b, float rel) { return std::fabs(a - b) <= rel * std::fmax(1.f, std::fabs(b)); }

int main() {
  Device *dev = nullptr;
  try {
    dev = new Device(0);
  } catch (const std::runtime_error &e) {
    std::printf("NO_GPU: %s\n", e.what());
    return 3;
  }
  std::printf("%s\n", dev->toString().c_str());

  // memory: pinned host <-> device round trip, memset, active device
  {
    const size_t n = 1024;
    int *h = (int *)dev->allocHost(n * sizeof(int));
    for (size_t i = 0; i < n; ++i) h[i] = (int)(i * 7);
    void *d = dev->allocMem(n * sizeof(int));
    dev->copy(d, h, n * sizeof(int));
    std::memset(h, 0, n * sizeof(int));
    dev->copy(h, d, n * sizeof(int));
    bool ok = true;
    for (size_t i = 0; i < n; ++i) ok = ok && h[i] == (int)(i * 7);
    CHECK(ok);
    dev->memsetAsync(d, 0, n * sizeof(int));
    dev->sync();
    dev->copy(h, d, n * sizeof(int));
    ok = true;
    for (size_t i = 0; i < n; ++i) ok = ok && h[i] == 0;
    CHECK(ok);
    dev->freeMem(d);
    dev->freeHost(h);
    const int old = dev->setActive();
    dev->restoreActive(old);
  }

  // geometry: a quad (2 triangles) in z = 5 over [-1,1]^2, a sphere of
  // radius 1 at (3,0,5), a cylinder of radius 0.5 from (-3,-1,5) to (-3,1,5)
  const float quadV[] = {-1, -1, 5, 1, -1, 5, 1, 1, 5, -1, 1, 5};
  const int quadI[] = {0, 1, 2, 0, 2, 3};
  const float sphV[] = {3, 0, 5};
  const float sphR[] = {1.f};
  const float cylV[] = {-3, -1, 5, -3, 1, 5};
  const int cylI[] = {0, 1};
  const float cylR[] = {0.5f};
  GeomType *triT = createGeomType_Triangles(dev);
  GeomType *sphT = createGeomType_Spheres(dev);
  GeomType *cylT = createGeomType_Cylinders(dev);
  Buffer *qv = dev->createBuffer(sizeof(quadV), quadV), *qi = dev->createBuffer(sizeof(quadI), quadI);
  Buffer *sv = dev->createBuffer(sizeof(sphV), sphV), *sr = dev->createBuffer(sizeof(sphR), sphR);
  Buffer *cv = dev->createBuffer(sizeof(cylV), cylV), *ci = dev->createBuffer(sizeof(cylI), cylI);
  Buffer *cr = dev->createBuffer(sizeof(cylR), cylR);
  Geom *quad = triT->createGeom();
  quad->setVertices(qv, 4);
  quad->setIndices(qi, 2);
  Geom *sph = sphT->createGeom();
  sph->setVertices(sv, 1);
  sph->setRadii(sr, 1.f);
  Geom *cyl = cylT->createGeom();
  cyl->setVertices(cv, 2);
  cyl->setIndices(ci, 1);
  cyl->setRadii(cr, 0.5f);
  Group *tg = dev->createTrianglesGroup({quad});
  Group *ug = dev->createUserGeomsGroup({sph, cyl});
  // mixing the kinds the reference keeps apart is rejected
  bool threw = false;
  try {
    delete dev->createTrianglesGroup({sph});
  } catch (const std::runtime_error &) {
    threw = true;
  }
  CHECK(threw);
  auto *ig = (InstanceGroup *)dev->createInstanceGroup({tg, ug}, {7, 8}, {affine3f(), affine3f()});
  ig->buildAccel();

  // rays from the origin; expected closest hits
  struct Want {
    float dir[3];
    float t;
    int instID, geomID, primID;
  };
  const float d2s = std::sqrt(3.f * 3.f + 5.f * 5.f);
  const float d2c = std::sqrt(3.f * 3.f + 5.f * 5.f);
  std::vector<Want> want = {
      {{0.f, 0.f, 1.f}, 5.f, 7, 0, -2},                        // quad (either triangle of the diagonal)
      {{0.5f, 0.25f, 5.f}, std::sqrt(0.25f + 0.0625f + 25.f), 7, 0, 0},
      {{3.f / d2s, 0.f, 5.f / d2s}, d2s - 1.f, 8, 0, 0},      // sphere, front point
      {{-3.f / d2c, 0.f, 5.f / d2c}, d2c - 0.5f, 8, 1, 0},    // cylinder side
      {{0.f, 0.f, -1.f}, INFINITY, -1, -1, -1},                // miss
  };
  std::vector<bnamd_ray> rays(want.size());
  for (size_t i = 0; i < want.size(); ++i) {
    bnamd_ray &r = rays[i];
    r.org[0] = r.org[1] = r.org[2] = 0.f;
    const float *d = want[i].dir;
    const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) r.dir[k] = d[k] / l;
    r.tmin = 0.f;
    r.tmax = 1e30f;
  }
  void *dRays = dev->allocMem(rays.size() * sizeof(bnamd_ray));
  void *dHits = dev->allocMem(rays.size() * sizeof(bnamd_hit));
  dev->copy(dRays, rays.data(), rays.size() * sizeof(bnamd_ray));
  ig->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, (int)rays.size());
  std::vector<bnamd_hit> hits(rays.size());
  dev->copy(hits.data(), dHits, hits.size() * sizeof(bnamd_hit));
  for (size_t i = 0; i < want.size(); ++i) {
    const bnamd_hit &h = hits[i];
    const Want &w = want[i];
    std::printf("ray %zu: t=%g inst=%d geom=%d prim=%d\n", i, h.t, h.instID, h.geomID, h.primID);
    CHECK(h.instID == w.instID && h.geomID == w.geomID);
    if (w.primID != -2) CHECK(h.primID == w.primID);
    if (std::isinf(w.t)) CHECK(h.primID == -1);
    else CHECK(near(h.t, w.t, 1e-5f));
  }
  // instance transforms: the quad group moved by +1 in z and scaled by 2 in
  // x/y (two-level world); the head-on ray now meets it at t = 6
  affine3f moved;
  moved.vx = {2.f, 0.f, 0.f};
  moved.vy = {0.f, 2.f, 0.f};
  moved.p = {0.f, 0.f, 1.f};
  ig->setTransforms({moved, affine3f()});
  ig->buildAccel();
  ig->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, (int)rays.size());
  dev->copy(hits.data(), dHits, hits.size() * sizeof(bnamd_hit));
  std::printf("moved: t=%g inst=%d prim=%d N=(%g %g %g)\n", hits[0].t, hits[0].instID, hits[0].primID,
              hits[0].N[0], hits[0].N[1], hits[0].N[2]);
  CHECK(hits[0].instID == 7 && near(hits[0].t, 6.f, 1e-6f));
  CHECK(std::fabs(hits[0].N[2]) > 0.99f);
  // the sphere behind it is unaffected
  CHECK(hits[2].instID == 8 && near(hits[2].t, d2s - 1.f, 1e-5f));
  // a singular transform is rejected with a message
  affine3f flat;
  flat.vz = {0.f, 0.f, 0.f};
  ig->setTransforms({flat, affine3f()});
  threw = false;
  try {
    ig->buildAccel();
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "singular") != nullptr;
  }
  CHECK(threw);

  // setDD: the bytes barney's Spheres::commit / Cylinders::commit /
  // Triangles::commit write (Geometry::DD head, then the geometry's fields)
  // give the same geometry as the typed setters
  {
    using namespace barney_dd;
    auto ddBytes = [](int userID, int materialID, const void *fields, size_t n) {
      std::vector<uint8_t> b(kGeometryDDSize + n, 0);
      Geometry g;
      std::memset(&g, 0, sizeof(g));
      g.userID = userID;
      g.materialID = materialID;
      std::memcpy(b.data(), &g, kGeometryDDSize);
      std::memcpy(b.data() + kGeometryDDSize, fields, n);
      return b;
    };
    SpheresFields sf{(const float *)sv->getDD(), (const float *)sr->getDD(), nullptr, 1.f};
    CylindersFields cf{(const float *)cv->getDD(), (const int *)ci->getDD(), (const float *)cr->getDD()};
    TrianglesFields tf{(const int *)qi->getDD(), (const float *)qv->getDD(), nullptr, nullptr};
    Geom *sph2 = sphT->createGeom();
    sph2->setPrimCount(1);
    const auto sb = ddBytes(41, 0, &sf, sizeof(sf));
    sph2->setDD(sb.data());
    Geom *cyl2 = cylT->createGeom();
    cyl2->setPrimCount(1);
    const auto cb = ddBytes(42, 0, &cf, sizeof(cf));
    cyl2->setDD(cb.data());
    Geom *quad2 = triT->createGeom();
    quad2->setVertices(qv, 4);
    quad2->setIndices(qi, 2);
    const auto tb = ddBytes(43, 0, &tf, sizeof(tf));
    quad2->setDD(tb.data());
    const bnamd_geom d1 = sph2->desc(), d2 = cyl2->desc(), d3 = quad2->desc();
    CHECK(d1.userID == 41 && d1.vertices == sf.origins && d1.radii == sf.radii && d1.numPrims == 1);
    CHECK(d2.userID == 42 && d2.vertices == cf.vertices && d2.indices == cf.indices && d2.radii == cf.radii);
    CHECK(d3.userID == 43 && d3.vertices == tf.vertices && d3.indices == tf.indices && d3.numPrims == 2);
    Group *tg2 = dev->createTrianglesGroup({quad2});
    Group *ug2 = dev->createUserGeomsGroup({sph2, cyl2});
    auto *ig2 = (InstanceGroup *)dev->createInstanceGroup({tg2, ug2}, {7, 8}, {affine3f(), affine3f()});
    ig2->buildAccel();
    ig->setTransforms({affine3f(), affine3f()});
    ig->buildAccel();
    std::vector<bnamd_hit> h1(rays.size()), h2(rays.size());
    ig->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, (int)rays.size());
    dev->copy(h1.data(), dHits, h1.size() * sizeof(bnamd_hit));
    ig2->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, (int)rays.size());
    dev->copy(h2.data(), dHits, h2.size() * sizeof(bnamd_hit));
    CHECK(std::memcmp(h1.data(), h2.data(), h1.size() * sizeof(bnamd_hit)) == 0);
    {  // the handle is the world's device record
      const void *rec = nullptr;
      CHECK(bnamd_world_device(ig2->world, &rec) == BNAMD_OK);
      CHECK((const void *)getAccelHandle(ig2) == rec && rec != nullptr);
    }
    dev->freeGroup(ig2);
    dev->freeGroup(tg2);
    dev->freeGroup(ug2);
    for (Geom *g : {sph2, cyl2, quad2}) dev->freeGeom(g);
  }
  // 1,000 instances of one group share one bottom level; moving them
  // (setTransforms + refitAccel) rebuilds only the instance level
  {
    const int n = 1000;
    std::vector<Group *> gs(n, ug);
    std::vector<int> ids(n);
    std::vector<affine3f> xs(n);
    for (int i = 0; i < n; ++i) {
      ids[i] = 1000 + i;
      xs[i].p = {0.f, 0.f, 10.f * (float)i};  // instance i's sphere at (3, 0, 5 + 10 i)
    }
    auto *many = (InstanceGroup *)dev->createInstanceGroup(gs, ids, xs);
    many->buildAccel();
    std::printf("1000 instances: %d bottom level(s)\n", many->numBottomLevels());
    CHECK(many->numBottomLevels() == 1);
    // a ray along +z through the sphere column at x = 3 meets instance 0
    bnamd_ray r{};
    r.org[0] = 3.f;
    r.org[2] = -10.f;
    r.dir[2] = 1.f;
    r.tmax = 1e30f;
    dev->copy(dRays, &r, sizeof(r));
    bnamd_hit h{};
    many->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, 1);
    dev->copy(&h, dHits, sizeof(h));
    CHECK(h.instID == 1000 && near(h.t, 14.f, 1e-5f));
    // drop the first 500 instances far away: the ray now meets instance 500
    for (int i = 0; i < 500; ++i) xs[i].p.y = 1000.f;
    many->setTransforms(xs);
    many->refitAccel();
    many->traceRays((const bnamd_ray *)dRays, (bnamd_hit *)dHits, 1);
    dev->copy(&h, dHits, sizeof(h));
    std::printf("after refit: inst=%d t=%g\n", h.instID, h.t);
    CHECK(h.instID == 1500 && near(h.t, 5014.f, 1e-5f));
    CHECK(many->numBottomLevels() == 1);
    dev->freeGroup(many);
  }
  // texture data and textures: a 3D scalar field sampled with CLAMP and
  // unnormalized coordinates (StructuredData's description)
  {
    const vec3i dims = {4, 3, 2};
    std::vector<float> texels(24);
    for (int i = 0; i < 24; ++i) texels[i] = 0.5f * (float)i;
    TextureData *td = dev->createTextureData(dims, FLOAT, texels.data());
    TextureDesc desc;
    desc.normalizedCoords = false;
    Texture *tex = td->createTexture(desc);
    CHECK(tex->texels3D() == (const float *)td->d_texels && td->numBytes == 24 * sizeof(float));
    std::vector<float> back(24, -1.f);
    dev->copy(back.data(), tex->texels3D(), 24 * sizeof(float));
    CHECK(back == texels);
    // getDD(): the device descriptor device code samples with tex3D
    TextureObjectDD o{};
    dev->copy(&o, (const void *)tex->getDD(), sizeof o);
    CHECK(o.dim == 3 && o.w == 4 && o.h == 3 && o.d == 2 && o.texels == td->d_texels);
    threw = false;
    try {
      delete td->createTexture(TextureDesc());  // normalized coordinates on a 3D field
    } catch (const std::runtime_error &) {
      threw = true;
    }
    CHECK(threw);
    dev->freeTexture(tex);
    dev->freeTextureData(td);
  }
  // a 2D image (normalized coordinates, WRAP x CLAMP) and the environment
  // light over it: the light's importance CDFs end at 1
  {
    const vec3i dims = {8, 4, 0};
    std::vector<float> texels(8 * 4 * 4);
    for (size_t i = 0; i < texels.size(); ++i) texels[i] = 0.25f + 0.01f * (float)(i % 37);
    TextureData *td = dev->createTextureData(dims, FLOAT4, texels.data());
    TextureDesc desc;
    desc.addressMode[0] = WRAP;
    Texture *tex = td->createTexture(desc);
    CHECK(tex->tex2d != nullptr);
    TextureObjectDD o{};
    dev->copy(&o, (const void *)tex->getDD(), sizeof o);
    CHECK(o.dim == 2 && o.w == 8 && o.h == 4 && o.addr[0] == WRAP && o.addr[1] == CLAMP && o.texels != nullptr);
    const float dir[3] = {1.f, 0.f, 0.f}, up[3] = {0.f, 0.f, 1.f};
    EnvMapLight *env = dev->createEnvMapLight(tex, dir, up, 2.f);
    std::vector<float> cy(4), cx(32);
    CHECK(bnamd_envmap_get_cdfs(env->handle, cy.data(), cx.data()) == 0);
    CHECK(cy[3] == 1.f && cx[7] == 1.f && cx[31] == 1.f && cy[0] > 0.f && cy[0] < cy[1]);
    std::printf("env map: cdf_y %g %g %g %g\n", cy[0], cy[1], cy[2], cy[3]);
    dev->freeEnvMapLight(env);
    dev->freeTexture(tex);
    dev->freeTextureData(td);
  }

  dev->freeMem(dRays);
  dev->freeMem(dHits);
  dev->freeGroup(ig);
  dev->freeGroup(tg);
  dev->freeGroup(ug);
  for (Geom *g : {quad, sph, cyl}) dev->freeGeom(g);
  for (Buffer *b : {qv, qi, sv, sr, cv, ci, cr}) dev->freeBuffer(b);
  for (GeomType *t : {triT, sphT, cylT}) dev->freeGeomType(t);
  delete dev;
  std::printf(failures ? "FAILED %d\n" : "OK\n", failures);
  return failures ? 1 : 0;
}
