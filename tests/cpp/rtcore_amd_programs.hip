// Conformance program for the programmable side of rtc::amd: device
// programs and kernels written against rtcore_amd_device.h the way barney
// writes its own (RTC_EXPORT_TRIANGLES_GEOM with anyHit / closestHit,
// RTC_EXPORT_USER_GEOM with bounds / intersect / closestHit,
// RTC_EXPORT_TRACE2D, RTC_EXPORT_COMPUTE1D), registered and launched through
// the rtc::amd host classes, and traced with TraceInterface::traceRay over a
// two-level world (a triangle mesh instanced twice, user-geometry spheres).
//
// Checks, per ray: the closest hit equals the closer of (a) the native
// closest hit of bnamd_trace over the same instances of the triangles whose
// anyHit program does not ignore them and (b) a float64 brute force over the
// spheres' intersect program arithmetic; t bit for bit, primitive, geometry
// and instance IDs, and the instance transforms the closestHit program sees.
// Exit 0 and "OK" on success, 1 on a mismatch, 3 without a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../barney_amd/rtcore_amd/rtcore_amd_device.h"

using namespace rtc::amd;

namespace {

struct PRD {
  float t;
  int primID, geomID, instIndex, instID;
  float u, v;
  float wx;  // the closest hit's world-space x through transformPointFromObjectToWorldSpace
};

// barney's Triangles::DD layout: Geometry::DD's 296 bytes, then the fields
struct TestTriDD {
  unsigned char bytes[barney_dd::kGeometryDDSize + sizeof(barney_dd::TrianglesFields)];
};
struct SphereDD {
  const float4 *spheres;  // center, radius
};

__device__ void write_hit(TraceInterface &ti) {
  PRD &prd = *(PRD *)ti.getPRD();
  prd.t = ti.getRayTmax();
  prd.primID = ti.getPrimitiveIndex();
  prd.geomID = ti.getGeometryIndex();
  prd.instIndex = ti.getRTCInstanceIndex();
  prd.instID = ti.getInstanceID();
  const vec2f b = ti.getTriangleBarycentrics();
  prd.u = b.x;
  prd.v = b.y;
  const vec3f o = ti.getObjectRayOrigin(), d = ti.getObjectRayDirection();
  const float t = ti.getRayTmax();
  prd.wx = ti.transformPointFromObjectToWorldSpace({o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}).x;
}

struct TriPrograms {
  // every fifth triangle (primID % 5 == 2) is a cutout its anyHit ignores
  static __device__ void anyHit(TraceInterface &ti) {
    if (ti.getPrimitiveIndex() % 5 == 2) ti.ignoreIntersection();
  }
  static __device__ void closestHit(TraceInterface &ti) { write_hit(ti); }
};

struct SpherePrograms {
  static __device__ void bounds(const TraceInterface &, const void *geomData, box3f &b, int primID) {
    const float4 s = ((const SphereDD *)geomData)->spheres[primID];
    b.lower = {s.x - s.w, s.y - s.w, s.z - s.w};
    b.upper = {s.x + s.w, s.y + s.w, s.z + s.w};
  }
  // the nearer root of |o + t d - c| = r in float64, reported when it lies
  // in (tmin, tmax)
  static __device__ void intersect(TraceInterface &ti) {
    const float4 s = ((const SphereDD *)ti.getProgramData())->spheres[ti.getPrimitiveIndex()];
    const vec3f o = ti.getObjectRayOrigin(), d = ti.getObjectRayDirection();
    const double ox = (double)o.x - s.x, oy = (double)o.y - s.y, oz = (double)o.z - s.z;
    const double a = (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z;
    const double bb = ox * d.x + oy * d.y + oz * d.z;
    const double c = ox * ox + oy * oy + oz * oz - (double)s.w * s.w;
    const double disc = bb * bb - a * c;
    if (disc < 0.0) return;
    const double sq = sqrt(disc);
    double t = (-bb - sq) / a;
    if (!(t > ti.getRayTmin())) t = (-bb + sq) / a;
    const float tf = (float)t;
    if (tf > ti.getRayTmin() && tf < ti.getRayTmax()) ti.reportIntersection(tf, 0);
  }
  static __device__ void anyHit(TraceInterface &) {}
  static __device__ void closestHit(TraceInterface &ti) { write_hit(ti); }
};

struct LP {
  AccelHandle world;
  const bnamd_ray *rays;
  PRD *out;
  int width, n;
};

struct RayGen {
  static __device__ void run(TraceInterface &ti) {
    const LP &lp = *(const LP *)ti.getLPData();
    const vec3i ix = ti.getLaunchIndex();
    const int i = ix.x + lp.width * ix.y;
    if (ix.x >= lp.width || i >= lp.n) return;
    const bnamd_ray r = lp.rays[i];
    PRD prd;
    prd.t = INFINITY;
    prd.primID = prd.geomID = prd.instIndex = prd.instID = -1;
    prd.u = prd.v = prd.wx = 0.f;
    ti.traceRay(lp.world, {r.org[0], r.org[1], r.org[2]}, {r.dir[0], r.dir[1], r.dir[2]}, r.tmin, r.tmax, &prd);
    lp.out[i] = prd;
  }
};

struct ClearOut {
  PRD *out;
  int n;
  __device__ void run(const ComputeInterface &ci) {
    const int i = (int)ci.launchIndex().x;
    if (i >= n) return;
    out[i].t = -1.f;
    out[i].primID = -7;
  }
};

// samples a 2D image and a 3D field through their TextureObjects, as
// barney's material and structured-volume programs do
struct TexSample {
  TextureObject img, field;
  const float *uvw;  // n x 3: (u, v) for the image, (x, y, z) for the field
  float *out;        // n x 5: tex2D<vec4f> then tex3D<float>
  int n;
  __device__ void run(const ComputeInterface &ci) const {
    const int i = (int)ci.launchIndex().x;
    if (i >= n) return;
    const vec4f c = tex2D<vec4f>(img, uvw[3 * i], uvw[3 * i + 1]);
    out[5 * i] = c.x;
    out[5 * i + 1] = c.y;
    out[5 * i + 2] = c.z;
    out[5 * i + 3] = c.w;
    out[5 * i + 4] = tex3D<float>(field, uvw[3 * i] * 7.f, uvw[3 * i + 1] * 6.f, uvw[3 * i + 2] * 5.f);
  }
};

}  // namespace

RTC_EXPORT_COMPUTE1D(texSample, TexSample)
RTC_EXPORT_TRIANGLES_GEOM(TestTriangles, TestTriDD, TriPrograms, true, true)
RTC_EXPORT_USER_GEOM(TestSpheres, SphereDD, SpherePrograms, false, true)
RTC_EXPORT_TRACE2D(testTrace, RayGen)
RTC_IMPORT_TRACE2D(programs, testTrace, sizeof(LP))
RTC_EXPORT_COMPUTE1D(clearOut, ClearOut)

static int fail(const char *what, int i = -1) {
  std::printf("FAIL: %s (ray %d)\n", what, i);
  return 1;
}

// the library's world-to-object map of a pure translation, as the device applies it
static void to_object(const float *p, float tx, float ty, float tz, float *q) {
  q[0] = p[0] * 1.f + p[1] * 0.f + p[2] * 0.f + -tx;
  q[1] = p[0] * 0.f + p[1] * 1.f + p[2] * 0.f + -ty;
  q[2] = p[0] * 0.f + p[1] * 0.f + p[2] * 1.f + -tz;
}

int main() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    std::printf("NO_GPU\n");
    return 3;
  }
  Device dev(0);
  // ---- a triangle mesh: a 24x24 wavy sheet over [0,24]^2 at y ~ 0
  const int N = 24;
  std::vector<float> verts;
  std::vector<int> tris;
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i <= N; ++i) {
      verts.push_back((float)i + 0.13f * std::sin(1.7f * j));
      verts.push_back(0.6f * std::sin(0.9f * i) * std::cos(0.7f * j));
      verts.push_back((float)j + 0.11f * std::cos(2.3f * i));
    }
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) {
      const int a = i + (N + 1) * j, b = a + 1, c = a + N + 1, d = c + 1;
      tris.insert(tris.end(), {a, b, d, a, d, c});
    }
  const int numTris = (int)tris.size() / 3;
  Buffer *vb = dev.createBuffer(verts.size() * 4, verts.data());
  Buffer *ib = dev.createBuffer(tris.size() * 4, tris.data());
  GeomType *triType = createGeomType_TestTriangles(&dev);
  Geom *tg = triType->createGeom();
  tg->setVertices(vb, (int)verts.size() / 3);
  tg->setIndices(ib, numTris);
  {
    TestTriDD dd{};
    barney_dd::Geometry head{};
    head.userID = 5;
    head.materialID = 0;
    std::memcpy(dd.bytes, &head, barney_dd::kGeometryDDSize);
    barney_dd::TrianglesFields f{};
    f.indices = (const int *)ib->getDD();
    f.vertices = (const float *)vb->getDD();
    std::memcpy(dd.bytes + barney_dd::kGeometryDDSize, &f, sizeof(f));
    tg->setDD(&dd);
  }
  // ---- user geometry: 40 spheres above the sheet
  std::vector<float4> spheres;
  for (int k = 0; k < 40; ++k)
    spheres.push_back(make_float4(1.f + 0.57f * k, 1.2f + 0.9f * std::sin(0.8f * k), 2.f + 0.5f * k,
                                  0.35f + 0.2f * std::fabs(std::cos(1.3f * k))));
  Buffer *sb = dev.createBuffer(spheres.size() * sizeof(float4), spheres.data());
  GeomType *sphType = createGeomType_TestSpheres(&dev);
  Geom *sg = sphType->createGeom();
  sg->setPrimCount((int)spheres.size());
  SphereDD sdd{(const float4 *)sb->getDD()};
  sg->setDD(&sdd);
  // ---- world: the sheet twice (identity and a translation), the spheres once
  Group *triGroup = dev.createTrianglesGroup({tg});
  Group *sphGroup = dev.createUserGeomsGroup({sg});
  affine3f id, moved;
  moved.p = {0.f, 1.5f, 30.f};
  auto *ig = (InstanceGroup *)dev.createInstanceGroup({triGroup, sphGroup, triGroup}, {100, 200, 300},
                                                      {id, id, moved});
  ig->buildAccel();
  AccelHandle accel = getAccelHandle(ig);
  // ---- rays from above, toward the sheets and spheres
  const int W = 96, H = 64, n = W * H;
  std::vector<bnamd_ray> rays(n);
  uint32_t seed = 12345u;
  auto rnd = [&] {
    seed = seed * 1664525u + 1013904223u;
    return (float)(seed >> 8) * (1.f / 16777216.f);
  };
  for (int i = 0; i < n; ++i) {
    bnamd_ray &r = rays[i];
    r.org[0] = -5.f + 34.f * rnd();
    r.org[1] = 9.f + 3.f * rnd();
    r.org[2] = -5.f + 70.f * rnd();
    const float tx = 2.f + 20.f * rnd(), ty = 0.f, tz = 2.f + 52.f * rnd();
    float d[3] = {tx - r.org[0], ty - r.org[1], tz - r.org[2]};
    const float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) r.dir[k] = d[k] / l;
    r.tmin = 0.f;
    r.tmax = 1e30f;
  }
  void *dRays = nullptr, *dOut = nullptr;
  rtc::amd::check(bnamd_device_malloc(&dRays, n * sizeof(bnamd_ray)), "malloc");
  rtc::amd::check(bnamd_device_malloc(&dOut, n * sizeof(PRD)), "malloc");
  rtc::amd::check(bnamd_memcpy_h2d(dRays, rays.data(), n * sizeof(bnamd_ray)), "copy");
  ComputeKernel1D *clear = createCompute_clearOut(&dev);
  ClearOut co{(PRD *)dOut, n};
  clear->launch((n + 127) / 128, 128, &co);
  TraceKernel2D *trace = createTrace_testTrace(&dev);
  LP lp{accel, (const bnamd_ray *)dRays, (PRD *)dOut, W, n};
  trace->launch({W, H}, &lp);
  dev.sync();
  std::vector<PRD> got(n);
  rtc::amd::check(bnamd_memcpy_d2h(got.data(), dOut, n * sizeof(PRD)), "copy back");

  // ---- expected: native triangles (without the ignored ones) ...
  std::vector<int> kept, keptPrim;
  for (int p = 0; p < numTris; ++p)
    if (p % 5 != 2) {
      kept.insert(kept.end(), {tris[3 * p], tris[3 * p + 1], tris[3 * p + 2]});
      keptPrim.push_back(p);
    }
  void *dKept = nullptr;
  rtc::amd::check(bnamd_device_malloc(&dKept, kept.size() * 4), "malloc");
  rtc::amd::check(bnamd_memcpy_h2d(dKept, kept.data(), kept.size() * 4), "copy");
  bnamd_geom g{};
  g.kind = BNAMD_GEOM_TRIANGLES;
  g.numPrims = (int)keptPrim.size();
  g.numVertices = (int)verts.size() / 3;
  g.vertices = (const float *)vb->getDD();
  g.indices = (const int *)dKept;
  bnamd_group grp{&g, 1};
  bnamd_instance inst[2] = {};
  inst[0].group = 0;
  inst[0].instID = 100;
  inst[1].group = 0;
  inst[1].instID = 300;
  const float I3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  std::memcpy(inst[0].xfm, I3, sizeof(I3));
  std::memcpy(inst[1].xfm, I3, sizeof(I3));
  inst[1].xfm[10] = 1.5f;
  inst[1].xfm[11] = 30.f;
  bnamd_world *native = nullptr;
  rtc::amd::check(bnamd_world_build(&grp, 1, inst, 2, nullptr, 0, nullptr, &native), "native world");
  void *dHits = nullptr;
  rtc::amd::check(bnamd_device_malloc(&dHits, n * sizeof(bnamd_hit)), "malloc");
  rtc::amd::check(bnamd_trace(native, (const bnamd_ray *)dRays, (bnamd_hit *)dHits, n, nullptr), "native trace");
  std::vector<bnamd_hit> nh(n);
  rtc::amd::check(bnamd_memcpy_d2h(nh.data(), dHits, n * sizeof(bnamd_hit)), "copy back");
  // ... and the spheres by brute force (the intersect program's arithmetic)
  int triHits = 0, sphHits = 0;
  for (int i = 0; i < n; ++i) {
    const bnamd_ray &r = rays[i];
    float d[3] = {r.dir[0], r.dir[1], r.dir[2]};
    for (int k = 0; k < 3; ++k)
      if (std::fabs(d[k]) < 1e-6f) d[k] = 1e-6f;
    float bestT = r.tmax;
    int bestPrim = -1, bestInst = -1;
    for (int s = 0; s < (int)spheres.size(); ++s) {
      const float4 sp = spheres[s];
      float o[3];
      to_object(r.org, 0.f, 0.f, 0.f, o);
      const double ox = (double)o[0] - sp.x, oy = (double)o[1] - sp.y, oz = (double)o[2] - sp.z;
      const double a = (double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2];
      const double bb = ox * d[0] + oy * d[1] + oz * d[2];
      const double c = ox * ox + oy * oy + oz * oz - (double)sp.w * sp.w;
      const double disc = bb * bb - a * c;
      if (disc < 0.0) continue;
      const double sq = std::sqrt(disc);
      double t = (-bb - sq) / a;
      if (!(t > r.tmin)) t = (-bb + sq) / a;
      const float tf = (float)t;
      if (tf > r.tmin && tf < bestT) {
        bestT = tf;
        bestPrim = s;
        bestInst = 1;
      }
    }
    const PRD &p = got[i];
    const bool tri = nh[i].primID >= 0 && nh[i].t < bestT;
    if (tri) {
      ++triHits;
      const int inst = nh[i].instIndex == 0 ? 0 : 2;
      if (p.primID != keptPrim[nh[i].primID] || p.geomID != 0 || p.instIndex != inst ||
          p.instID != (inst == 0 ? 100 : 300))
        return fail("triangle hit IDs", i);
      if (std::memcmp(&p.t, &nh[i].t, 4) != 0) return fail("triangle hit t", i);
      if (p.u != nh[i].u || p.v != nh[i].v) return fail("triangle barycentrics", i);
      if (inst == 2 && std::fabs(p.wx - (r.org[0] + p.t * r.dir[0])) > 1e-3f) return fail("instance transform", i);
    } else if (bestPrim >= 0) {
      ++sphHits;
      if (p.primID != bestPrim || p.geomID != 0 || p.instIndex != bestInst || p.instID != 200)
        return fail("sphere hit IDs", i);
      if (std::memcmp(&p.t, &bestT, 4) != 0) return fail("sphere hit t", i);
    } else {
      if (p.primID != -1 || p.t != INFINITY) return fail("expected a miss", i);
    }
  }
  std::printf("rays %d: triangle hits %d, sphere hits %d, misses %d\n", n, triHits, sphHits, n - triHits - sphHits);
  if (triHits < n / 4 || sphHits < 50) return fail("too few hits to test");
  // the fused renderers refuse a world with user geometry
  bnamd_hit dummy;
  if (bnamd_trace((bnamd_world *)ig->getDD(), (const bnamd_ray *)dRays, (bnamd_hit *)dHits, 1, nullptr) == BNAMD_OK)
    return fail("bnamd_trace accepted user geometry");
  (void)dummy;
  delete trace;
  delete clear;
  bnamd_world_destroy(native);
  for (void *b : {dRays, dOut, dKept, dHits}) bnamd_device_free(b);
  dev.freeGroup(ig);
  dev.freeGroup(triGroup);
  dev.freeGroup(sphGroup);
  dev.freeGeom(tg);
  dev.freeGeom(sg);
  dev.freeGeomType(triType);
  dev.freeGeomType(sphType);
  dev.freeBuffer(vb);
  dev.freeBuffer(ib);
  dev.freeBuffer(sb);
  // tex2D / tex3D through TextureObjects: the image equals the native 2D
  // sampler (bnamd_texture2d_sample) bit for bit, the field equals the
  // embree backend's unnormalized trilinear restated on the host
  {
    const int TW = 9, TH = 5;
    std::vector<float> img(4 * TW * TH);
    for (size_t i = 0; i < img.size(); ++i) img[i] = 0.1f + 0.37f * (float)((i * 7) % 23);
    TextureData *itd = dev.createTextureData({TW, TH, 0}, FLOAT4, img.data());
    TextureDesc idesc;
    idesc.addressMode[0] = WRAP;
    idesc.addressMode[1] = MIRROR;
    Texture *itex = itd->createTexture(idesc);
    const int FX = 6, FY = 5, FZ = 4;
    std::vector<float> field(FX * FY * FZ);
    for (size_t i = 0; i < field.size(); ++i) field[i] = std::sin(0.3f * (float)i) + 0.01f * (float)i;
    TextureData *ftd = dev.createTextureData({FX, FY, FZ}, FLOAT, field.data());
    TextureDesc fdesc;
    fdesc.normalizedCoords = false;
    Texture *ftex = ftd->createTexture(fdesc);
    const int m = 4096;
    std::vector<float> uvw(3 * m);
    uint32_t st = 12345u;
    for (float &x : uvw) {
      st = st * 1664525u + 1013904223u;
      x = -0.3f + 1.6f * (float)(st >> 8) * (1.f / 16777216.f);
    }
    void *dUvw = nullptr, *dTo = nullptr, *dUv = nullptr, *dNat = nullptr;
    rtc::amd::check(bnamd_device_malloc(&dUvw, uvw.size() * 4), "malloc");
    rtc::amd::check(bnamd_device_malloc(&dTo, 5 * m * 4), "malloc");
    rtc::amd::check(bnamd_device_malloc(&dUv, 2 * m * 4), "malloc");
    rtc::amd::check(bnamd_device_malloc(&dNat, 4 * m * 4), "malloc");
    rtc::amd::check(bnamd_memcpy_h2d(dUvw, uvw.data(), uvw.size() * 4), "copy");
    std::vector<float> uv(2 * m);
    for (int i = 0; i < m; ++i) {
      uv[2 * i] = uvw[3 * i];
      uv[2 * i + 1] = uvw[3 * i + 1];
    }
    rtc::amd::check(bnamd_memcpy_h2d(dUv, uv.data(), uv.size() * 4), "copy");
    ComputeKernel1D *ts = createCompute_texSample(&dev);
    TexSample tsd{itex->getDD(), ftex->getDD(), (const float *)dUvw, (float *)dTo, m};
    ts->launch((m + 127) / 128, 128, &tsd);
    dev.sync();
    rtc::amd::check(bnamd_texture2d_sample(itex->tex2d, (const float *)dUv, (float *)dNat, m, nullptr), "sample");
    std::vector<float> got(5 * m), nat(4 * m);
    rtc::amd::check(bnamd_memcpy_d2h(got.data(), dTo, got.size() * 4), "copy");
    rtc::amd::check(bnamd_memcpy_d2h(nat.data(), dNat, nat.size() * 4), "copy");
    auto coords = [](int N, float tc, int &i0, int &i1, float &f) {
      if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
      if (tc - .5f >= N - 1) { i0 = i1 = N - 1; f = 0.f; return; }
      f = (tc - .5f) - int(tc - .5f);
      i0 = int(tc - .5f);
      i1 = i0 + 1;
    };
    auto lerp = [](float f, float a, float b) { return (1.f - f) * a + f * b; };
    for (int i = 0; i < m; ++i) {
      for (int k = 0; k < 4; ++k)
        if (std::memcmp(&got[5 * i + k], &nat[4 * i + k], 4) != 0) return fail("tex2D differs from the native sampler", i);
      int x0, x1, y0, y1, z0, z1;
      float fx, fy, fz;
      coords(FX, uvw[3 * i] * 7.f, x0, x1, fx);
      coords(FY, uvw[3 * i + 1] * 6.f, y0, y1, fy);
      coords(FZ, uvw[3 * i + 2] * 5.f, z0, z1, fz);
      auto at = [&](int a, int b, int c) { return field[a + FX * (b + FY * c)]; };
      const float v0 = lerp(fy, lerp(fx, at(x0, y0, z0), at(x1, y0, z0)), lerp(fx, at(x0, y1, z0), at(x1, y1, z0)));
      const float v1 = lerp(fy, lerp(fx, at(x0, y0, z1), at(x1, y0, z1)), lerp(fx, at(x0, y1, z1), at(x1, y1, z1)));
      const float e = lerp(fz, v0, v1);
      if (std::memcmp(&got[5 * i + 4], &e, 4) != 0) return fail("tex3D differs from the trilinear restatement", i);
    }
    std::printf("tex2D / tex3D through TextureObjects: %d samples equal\n", m);
    delete ts;
    for (void *b : {dUvw, dTo, dUv, dNat}) bnamd_device_free(b);
    dev.freeTexture(itex);
    dev.freeTexture(ftex);
    dev.freeTextureData(itd);
    dev.freeTextureData(ftd);
  }
  std::printf("OK\n");
  return 0;
}
