// Renders one scene twice -- through the bn* API (libbn_api.so) and through
// the bnamd C ABI directly -- and checks that colors, depth and IDs agree
// bit for bit, over two progressive frames. Exit 3 without a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "bn_api.h"
#include "bnamd.h"

static int fail(const char *what) {
  std::printf("FAIL: %s\n", what);
  return 1;
}

int main(int argc, char **argv) {
  // optional: a NanoVDB grid file for the "NanoVDB" scalar field case
  const char *nvdbPath = argc > 1 ? argv[1] : nullptr;
  BNContext ctx;
  try {
    ctx = bnContextCreate();
  } catch (const std::exception &e) {
    std::printf("NO_GPU: %s\n", e.what());
    return 3;
  }
  const int W = 64, H = 48, SPP = 2;
  // scene: two triangles (a floor) and three spheres of two materials
  const float tv[] = {-4, 0, -4, 4, 0, -4, 4, 0, 4, -4, 0, 4};
  const int ti[] = {0, 1, 2, 0, 2, 3};
  const float so[] = {0, 1, 0, 1.5f, 0.6f, 0.5f, -1.2f, 0.5f, -0.3f};
  const float sr[] = {1.f, 0.6f, 0.5f};
  const float pos[3] = {0, 2.5f, -6}, dir[3] = {0, -0.35f, 1}, up[3] = {0, 1, 0};

  BNMaterial floorMat = bnMaterialCreate(ctx, 0, "matte");
  bnSet3f(floorMat, "color", .7f, .6f, .5f);
  BNMaterial ball = bnMaterialCreate(ctx, 0, "physicallyBased");
  bnSet3f(ball, "baseColor", .3f, .5f, .9f);
  bnSet1f(ball, "metallic", 0.f);
  bnSet1f(ball, "roughness", .4f);
  BNGeom tri = bnGeometryCreate(ctx, 0, "triangles");
  bnSetData(tri, "vertices", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 4, tv));
  bnSetData(tri, "indices", bnDataCreate(ctx, 0, BN_INT32_VEC3, 2, ti));
  bnSetObject(tri, "material", floorMat);
  bnSet1i(tri, "userID", 7);
  bnCommit(tri);
  BNGeom sph = bnGeometryCreate(ctx, 0, "spheres");
  bnSetData(sph, "origins", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 3, so));
  bnSetData(sph, "radii", bnDataCreate(ctx, 0, BN_FLOAT32, 3, sr));
  bnSetObject(sph, "material", ball);
  bnSet1i(sph, "userID", 9);
  bnCommit(sph);
  BNGeom geoms[2] = {tri, sph};
  BNGroup group = bnGroupCreate(ctx, 0, geoms, 2, nullptr, 0);
  bnGroupBuild(group);
  BNModel model = bnModelCreate(ctx);
  BNTransform xfm = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {0, 0, 0}};
  bnSetInstances(model, 0, &group, &xfm, 1);
  bnBuild(model, 0);
  BNCamera cam = bnCameraCreate(ctx, "perspective");
  bnSet3f(cam, "position", pos[0], pos[1], pos[2]);
  bnSet3f(cam, "direction", dir[0], dir[1], dir[2]);
  bnSet3f(cam, "up", up[0], up[1], up[2]);
  bnSet1f(cam, "fovy", 50.f);
  bnCommit(cam);
  BNRenderer rnd = bnRendererCreate(ctx, "default");
  bnSet1i(rnd, "pathsPerPixel", SPP);
  bnSet1f(rnd, "ambientRadiance", 0.9f);
  bnSet4f(rnd, "bgColor", 0.1f, 0.2f, 0.3f, 1.f);
  bnCommit(rnd);
  BNFrameBuffer fb = bnFrameBufferCreate(ctx);
  bnFrameBufferResize(fb, BN_FLOAT4, W, H, BN_FB_COLOR | BN_FB_DEPTH | BN_FB_PRIMID | BN_FB_INSTID | BN_FB_OBJID);
  int sx = 0, sy = 0;
  bnFrameBufferGetSize(fb, &sx, &sy);
  if (sx != W || sy != H) return fail("frame buffer size");
  bnRender(rnd, model, cam, fb);
  bnRender(rnd, model, cam, fb);  // second progressive frame
  std::vector<float> color(4 * W * H), depth(W * H);
  std::vector<int> prim(W * H), inst(W * H), objid(W * H);
  bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
  bnFrameBufferRead(fb, BN_FB_DEPTH, depth.data(), BN_FLOAT32);
  bnFrameBufferRead(fb, BN_FB_PRIMID, prim.data(), BN_INT32);
  bnFrameBufferRead(fb, BN_FB_INSTID, inst.data(), BN_INT32);
  bnFrameBufferRead(fb, BN_FB_OBJID, objid.data(), BN_INT32);

  // the same through the C ABI
  void *dv, *di, *dso, *dsr, *acc, *ids, *out;
  bnamd_device_malloc(&dv, sizeof tv);
  bnamd_device_malloc(&di, sizeof ti);
  bnamd_device_malloc(&dso, sizeof so);
  bnamd_device_malloc(&dsr, sizeof sr);
  bnamd_memcpy_h2d(dv, tv, sizeof tv);
  bnamd_memcpy_h2d(di, ti, sizeof ti);
  bnamd_memcpy_h2d(dso, so, sizeof so);
  bnamd_memcpy_h2d(dsr, sr, sizeof sr);
  bnamd_geom g[2] = {};
  g[0].kind = BNAMD_GEOM_TRIANGLES; g[0].numPrims = 2; g[0].numVertices = 4; g[0].materialID = 0;
  g[0].vertices = (const float *)dv; g[0].indices = (const int *)di; g[0].userID = 7;
  g[1].kind = BNAMD_GEOM_SPHERES; g[1].numPrims = 3; g[1].numVertices = 3; g[1].materialID = 1;
  g[1].vertices = (const float *)dso; g[1].radii = (const float *)dsr; g[1].defaultRadius = .1f; g[1].userID = 9;
  bnamd_group grp = {g, 2};
  bnamd_instance in = {0, 0, {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0}};
  bnamd_material m[2] = {};
  m[0].type = BNAMD_MAT_MATTE; m[0].color[0] = .7f; m[0].color[1] = .6f; m[0].color[2] = .5f; m[0].opacity = 1;
  m[0].ior = 1.45f; m[0].specular = .5f;
  m[1].type = BNAMD_MAT_PBR; m[1].color[0] = .3f; m[1].color[1] = .5f; m[1].color[2] = .9f; m[1].opacity = 1;
  m[1].metallic = 0; m[1].roughness = .4f; m[1].transmission = 0; m[1].ior = 1.5f; m[1].specular = .5f;
  bnamd_world *world = nullptr;
  if (bnamd_world_build(&grp, 1, &in, 1, m, 2, nullptr, &world) != BNAMD_OK) return fail(bnamd_last_error());
  bnamd_camera c = {{pos[0], pos[1], pos[2]}, {dir[0], dir[1], dir[2]}, {up[0], up[1], up[2]}, 50.f};
  bnamd_render_params p = {W, H, 0, SPP, 0.9f, {0.1f, 0.2f, 0.3f, 1.f}, 10, 1, 0};
  bnamd_device_malloc(&acc, 16 * W * H);
  bnamd_device_malloc(&ids, 16 * W * H);
  bnamd_device_malloc(&out, 16 * W * H);
  if (bnamd_render_surface(world, &c, &p, (float *)acc, (int *)ids, nullptr, nullptr) != BNAMD_OK) return fail("render");
  p.accumID = SPP;
  if (bnamd_render_surface(world, &c, &p, (float *)acc, nullptr, nullptr, nullptr) != BNAMD_OK) return fail("render 2");
  bnamd_fb_resolve((const float *)acc, W, H, 1.f / (2 * SPP), BNAMD_FB_FLOAT4, out, nullptr);
  bnamd_synchronize();
  std::vector<float> rc(4 * W * H);
  std::vector<int> rid(4 * W * H);
  bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
  bnamd_memcpy_d2h(rid.data(), ids, rid.size() * 4);
  if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("color differs from the C ABI render");
  int hits = 0;
  for (int q = 0; q < W * H; ++q) {
    float d;
    std::memcpy(&d, &rid[4 * q + 3], 4);
    if (prim[q] != rid[4 * q] || inst[q] != rid[4 * q + 1] || objid[q] != rid[4 * q + 2] ||
        std::memcmp(&d, &depth[q], 4) != 0)
      return fail("ID channels differ");
    hits += prim[q] >= 0;
  }
  bool sawTri = false, sawSph = false;
  for (int q = 0; q < W * H; ++q) {
    sawTri |= objid[q] == 7;
    sawSph |= objid[q] == 9;
  }
  if (!sawTri || !sawSph || hits < W * H / 4) return fail("scene not visible");
  // bnAccumReset restarts accumulation: one frame equals the C ABI's first
  bnAccumReset(fb);
  bnRender(rnd, model, cam, fb);
  bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
  p.accumID = 0;
  bnamd_render_surface(world, &c, &p, (float *)acc, nullptr, nullptr, nullptr);
  bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
  bnamd_synchronize();
  bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
  if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("color after bnAccumReset");
  // a physicallyBased material that sets only baseColor and transmission
  // takes AnariPBR's default ior 1.45 (AnariPBR.h:57), which makes it Glass
  {
    BNMaterial glass = bnMaterialCreate(ctx, 0, "physicallyBased");
    bnSet3f(glass, "baseColor", .9f, .9f, .9f);
    bnSet1f(glass, "transmission", 1.f);
    BNGeom gs = bnGeometryCreate(ctx, 0, "spheres");
    bnSetData(gs, "origins", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 3, so));
    bnSetData(gs, "radii", bnDataCreate(ctx, 0, BN_FLOAT32, 3, sr));
    bnSetObject(gs, "material", glass);
    bnSet1i(gs, "userID", 9);
    bnCommit(gs);
    BNGeom gg[2] = {tri, gs};
    BNGroup ggrp = bnGroupCreate(ctx, 0, gg, 2, nullptr, 0);
    bnGroupBuild(ggrp);
    BNModel gm = bnModelCreate(ctx);
    bnSetInstances(gm, 0, &ggrp, &xfm, 1);
    bnBuild(gm, 0);
    bnAccumReset(fb);
    bnRender(rnd, gm, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    bnamd_material mg[2] = {m[0], m[1]};
    mg[1].color[0] = mg[1].color[1] = mg[1].color[2] = .9f;
    mg[1].metallic = 1.f; mg[1].roughness = 1.f; mg[1].transmission = 1.f; mg[1].ior = 1.45f;
    bnamd_world *gw = nullptr;
    if (bnamd_world_build(&grp, 1, &in, 1, mg, 2, nullptr, &gw) != BNAMD_OK) return fail(bnamd_last_error());
    p.accumID = 0;
    bnamd_render_surface(gw, &c, &p, (float *)acc, nullptr, nullptr, nullptr);
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("PBR default ior differs from 1.45");
    // ior 1.5 renders a different image: the default is what decided it
    mg[1].ior = 1.5f;
    bnamd_world *gw2 = nullptr;
    if (bnamd_world_build(&grp, 1, &in, 1, mg, 2, nullptr, &gw2) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_render_surface(gw2, &c, &p, (float *)acc, nullptr, nullptr, nullptr);
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) == 0) return fail("ior 1.5 and 1.45 render alike");
    bnamd_world_destroy(gw);
    bnamd_world_destroy(gw2);
  }
  // a structured volume through bn* equals bnamd_volume_* with the same
  // field, transfer function and cellsPerMC 8
  {
    const int N = 17;
    std::vector<float> field(N * N * N);
    for (int z = 0; z < N; ++z)
      for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
          const float dx = x - 8.f, dy = y - 7.f, dz = z - 9.f;
          field[x + N * (y + N * z)] = std::exp(-(dx * dx + dy * dy + dz * dz) / 40.f);
        }
    const bn_float4 xf[3] = {{0.f, 0.f, 1.f, 0.f}, {1.f, .5f, 0.f, .5f}, {1.f, 1.f, 1.f, 1.f}};
    BNTextureData td = bnTextureData3DCreate(ctx, 0, BN_FLOAT32, N, N, N, field.data());
    BNScalarField sf = bnScalarFieldCreate(ctx, 0, "structured");
    bnSetObject(sf, "textureData", td);
    bnSet3f(sf, "gridOrigin", -2.f, -1.f, -2.f);
    bnSet3f(sf, "gridSpacing", .25f, .25f, .25f);
    bnCommit(sf);
    BNVolume vol = bnVolumeCreate(ctx, 0, sf);
    bnVolumeSetXF(vol, {0.f, 1.f}, xf, 3, 4.f);
    BNGroup vg = bnGroupCreate(ctx, 0, nullptr, 0, &vol, 1);
    BNModel vm = bnModelCreate(ctx);
    bnSetInstances(vm, 0, &vg, &xfm, 1);
    bnBuild(vm, 0);
    bnAccumReset(fb);
    bnRender(rnd, vm, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    void *dfield;
    bnamd_device_malloc(&dfield, field.size() * 4);
    bnamd_memcpy_h2d(dfield, field.data(), field.size() * 4);
    bnamd_structured_desc d = {(const float *)dfield, {N, N, N}, {-2.f, -1.f, -2.f}, {.25f, .25f, .25f}, 8};
    bnamd_volume *v = nullptr;
    if (bnamd_volume_create(&d, &v) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_volume_set_tf(v, (const float *)xf, 3, 0.f, 1.f, 4.f);
    p.accumID = 0;
    bnamd_volume_stats vs;
    if (bnamd_volume_render(v, &c, &p, (float *)acc, nullptr, &vs) != BNAMD_OK) return fail("volume render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("volume color differs");
    if (vs.samples <= 0) return fail("volume not sampled");
    // a volume model has no surface hits: its ID channels read as misses
    bnFrameBufferRead(fb, BN_FB_PRIMID, prim.data(), BN_INT32);
    bnFrameBufferRead(fb, BN_FB_DEPTH, depth.data(), BN_FLOAT32);
    for (int q = 0; q < W * H; ++q)
      if (prim[q] != -1 || !std::isinf(depth[q])) return fail("volume model ID channels not cleared");
    std::printf("volume samples %lld\n", vs.samples);
    // geometry and the volume in one model (two instances) through bn*
    // equal bnamd_render_mixed with the same world and volume
    {
      BNGroup both[2] = {group, vg};
      BNTransform bx[2] = {xfm, xfm};
      BNModel mm = bnModelCreate(ctx);
      bnSetInstances(mm, 0, both, bx, 2);
      bnBuild(mm, 0);
      bnAccumReset(fb);
      bnRender(rnd, mm, cam, fb);
      bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
      long long ms = 0;
      if (bnamd_render_mixed(world, v, &c, &p, (float *)acc, nullptr, nullptr, nullptr, &ms) != BNAMD_OK)
        return fail(bnamd_last_error());
      bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
      bnamd_synchronize();
      bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
      if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("mixed model differs");
      if (ms <= 0) return fail("mixed model: volume not sampled");
      std::printf("geometry + volume model: %lld volume samples, bn* = C ABI\n", ms);
    }
    bnamd_volume_destroy(v);
    bnamd_device_free(dfield);
  }
  // environment-map light over a 2D texture, the renderer's cut plane,
  // triangle vertex normals, per-sphere colors and a cutout (opacity .5)
  // material through bn* equal the same through the C ABI
  {
    const int TW = 16, TH = 8;
    std::vector<float> texels(4 * TW * TH);
    for (int i = 0; i < TW * TH; ++i) {
      texels[4 * i] = 0.3f + 0.05f * (i % 7);
      texels[4 * i + 1] = 0.2f + 0.1f * ((i / TW) % 3);
      texels[4 * i + 2] = (i == 37) ? 30.f : 0.6f;
      texels[4 * i + 3] = 1.f;
    }
    const float nrm[12] = {0.1f, 1, 0, 0, 1, 0.1f, -0.1f, 1, 0, 0, 1, -0.1f};
    const float scol[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    BNTexture2D tex = bnTexture2DCreate(ctx, 0, BN_FLOAT32_VEC4, TW, TH, texels.data(), BN_TEXTURE_LINEAR,
                                        BN_TEXTURE_WRAP, BN_TEXTURE_CLAMP);
    BNLight env = bnLightCreate(ctx, 0, "envmap");
    bnSetObject(env, "texture", tex);
    bnSet3f(env, "direction", 0.f, 0.f, 1.f);
    bnSet3f(env, "up", 0.f, 1.f, 0.f);
    bnSet1f(env, "scale", 1.5f);
    bnCommit(env);
    BNMaterial cut = bnMaterialCreate(ctx, 0, "matte");
    bnSet3f(cut, "color", .7f, .6f, .5f);
    bnSet1f(cut, "opacity", .5f);
    BNGeom t2 = bnGeometryCreate(ctx, 0, "triangles");
    bnSetData(t2, "vertices", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 4, tv));
    bnSetData(t2, "indices", bnDataCreate(ctx, 0, BN_INT32_VEC3, 2, ti));
    bnSetData(t2, "normals", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 4, nrm));
    bnSetObject(t2, "material", cut);
    bnSet1i(t2, "userID", 7);
    BNGeom s2 = bnGeometryCreate(ctx, 0, "spheres");
    bnSetData(s2, "origins", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 3, so));
    bnSetData(s2, "radii", bnDataCreate(ctx, 0, BN_FLOAT32, 3, sr));
    bnSetData(s2, "colors", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 3, scol));
    bnSetObject(s2, "material", floorMat);
    bnSet1i(s2, "userID", 9);
    BNGeom g2[2] = {t2, s2};
    BNGroup lg = bnGroupCreate(ctx, 0, g2, 2, nullptr, 0);
    BNObject lights[1] = {env};
    bnSetData(lg, "lights", bnDataCreate(ctx, 0, BN_OBJECT, 1, lights));
    BNModel lm = bnModelCreate(ctx);
    bnSetInstances(lm, 0, &lg, &xfm, 1);
    bnBuild(lm, 0);
    BNRenderer r2 = bnRendererCreate(ctx, "default");
    bnSet1i(r2, "pathsPerPixel", SPP);
    bnSet4f(r2, "bgColor", 0.1f, 0.2f, 0.3f, 1.f);
    bnSet4f(r2, "cutPlane", 0.2f, 0.f, 1.f, 0.5f);
    // setters this backend stores without using: callable, no effect
    bnSet2i(r2, "crosshairs", 1, 2);
    bnSet3i(r2, "unused3i", 1, 2, 3);
    bnSet4i(r2, "unused4i", 1, 2, 3, 4);
    bnSet2f(r2, "unused2f", 1.f, 2.f);
    bnSetString(r2, "name", "renderer");
    bnSet4x3fv(r2, "unusedXfm", &xfm);
    bnAccumReset(fb);
    bnRender(r2, lm, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    bnFrameBufferRead(fb, BN_FB_PRIMID, prim.data(), BN_INT32);
    // the C ABI: the same texture, light, geometry and parameters
    bnamd_texture2d_desc tdsc{};
    tdsc.width = TW; tdsc.height = TH; tdsc.format = 33; tdsc.texels = texels.data();
    tdsc.filterMode = 1; tdsc.addressMode[0] = 0; tdsc.addressMode[1] = 1;
    bnamd_texture2d *t2d = nullptr;
    if (bnamd_texture2d_create(&tdsc, nullptr, &t2d) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_envmap_desc edsc = {{0.f, 0.f, 1.f}, {0.f, 1.f, 0.f}, 1.5f};
    bnamd_envmap *em = nullptr;
    if (bnamd_envmap_create(t2d, &edsc, nullptr, &em) != BNAMD_OK) return fail(bnamd_last_error());
    void *dn, *dc;
    bnamd_device_malloc(&dn, sizeof nrm);
    bnamd_device_malloc(&dc, sizeof scol);
    bnamd_memcpy_h2d(dn, nrm, sizeof nrm);
    bnamd_memcpy_h2d(dc, scol, sizeof scol);
    bnamd_geom gl[2] = {g[0], g[1]};
    gl[0].normals = (const float *)dn;
    gl[1].colors = (const float *)dc;
    gl[1].materialID = 0;
    bnamd_group lgrp = {gl, 2};
    bnamd_material ml[1] = {m[0]};
    ml[0].opacity = .5f;
    bnamd_world *lw = nullptr;
    if (bnamd_world_build(&lgrp, 1, &in, 1, ml, 1, nullptr, &lw) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_render_params pl = p;
    pl.accumID = 0;
    pl.ambientRadiance = 1.f;
    pl.envMap = em;
    pl.cutPlane[0] = 0.2f; pl.cutPlane[1] = 0.f; pl.cutPlane[2] = 1.f; pl.cutPlane[3] = 0.5f;
    if (bnamd_render_surface(lw, &c, &pl, (float *)acc, (int *)ids, nullptr, nullptr) != BNAMD_OK)
      return fail("env render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("env-map render differs");
    // the cut plane removes part of the scene: fewer hits than without it
    int cutHits = 0;
    for (int q = 0; q < W * H; ++q) cutHits += prim[q] >= 0;
    std::printf("env-map + cut plane: %d hits (without the plane %d)\n", cutHits, hits);
    if (cutHits <= 0 || cutHits >= hits) return fail("cut plane has no effect");
    bnamd_world_destroy(lw);
    bnamd_envmap_destroy(em);
    bnamd_texture2d_destroy(t2d);
    // quad, directional and point lights of a translated instance through
    // bn* equal the same lights, moved by the translation, through the C ABI
    BNLight ql = bnLightCreate(ctx, 0, "quad");
    bnSet3f(ql, "corner", -1.f, 3.f, -1.f);
    bnSet3f(ql, "edge0", 2.f, 0.f, 0.f);
    bnSet3f(ql, "edge1", 0.f, 0.f, 2.f);  // edge0 x edge1 = -y: lights the floor
    bnSet3f(ql, "emission", 4.f, 3.f, 2.f);
    bnCommit(ql);
    BNLight dl = bnLightCreate(ctx, 0, "directional");
    bnSet3f(dl, "direction", 0.3f, -1.f, 0.2f);
    bnSet3f(dl, "color", 1.f, .9f, .8f);
    bnSet1f(dl, "irradiance", 2.f);
    bnCommit(dl);
    BNLight ptl = bnLightCreate(ctx, 0, "point");
    bnSet3f(ptl, "position", 0.5f, 2.f, 0.5f);
    bnSet1f(ptl, "power", 30.f);
    bnCommit(ptl);
    BNObject ls3[3] = {ql, dl, ptl};
    bnSetData(lg, "lights", bnDataCreate(ctx, 0, BN_OBJECT, 3, ls3));
    BNTransform xt = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, {0.5f, 0.f, -0.25f}};
    BNModel lm2 = bnModelCreate(ctx);
    bnSetInstances(lm2, 0, &lg, &xt, 1);
    bnBuild(lm2, 0);
    bnAccumReset(fb);
    bnRender(r2, lm2, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    bnamd_quad_light cq = {{-1.f + 0.5f, 3.f, -1.f - 0.25f}, {2.f, 0.f, 0.f}, {0.f, 0.f, 2.f}, {4.f, 3.f, 2.f}};
    bnamd_dir_light cd = {{0.3f, -1.f, 0.2f}, {1.f, .9f, .8f}, 2.f};
    bnamd_point_light cp = {{0.5f + 0.5f, 2.f, 0.5f - 0.25f}, {1.f, 1.f, 1.f}, 30.f, std::nanf("")};
    bnamd_lights *cl = nullptr;
    if (bnamd_lights_create(&cq, 1, &cd, 1, &cp, 1, &cl) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_instance it = {0, 0, {1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5f, 0.f, -0.25f}};
    bnamd_world *tw = nullptr;
    if (bnamd_world_build(&lgrp, 1, &it, 1, ml, 1, nullptr, &tw) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_render_params pq = pl;
    pq.envMap = nullptr;
    pq.lights = cl;
    if (bnamd_render_surface(tw, &c, &pq, (float *)acc, nullptr, nullptr, nullptr) != BNAMD_OK)
      return fail("lights render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("quad/dir/point light render differs");
    std::printf("quad, directional and point lights of a translated instance: bn* = C ABI\n");
    bnamd_world_destroy(tw);
    bnamd_lights_destroy(cl);
    bnamd_device_free(dn);
    bnamd_device_free(dc);
    // a light type barney does not have is rejected
    bool thr = false;
    try {
      (void)bnLightCreate(ctx, 0, "spot");
    } catch (const std::runtime_error &) {
      thr = true;
    }
    if (!thr) return fail("unknown light type accepted");
  }
  // an unstructured field (two hexahedra and a tetrahedron) and a
  // block-structured AMR field through bn* equal bnamd_umesh / bnamd_amr
  {
    const float uv[] = {-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1,
                        3, -1, -1, 3, 1, -1, 3, -1, 1, 3, 1, 1, 1, 3, 0};
    const float us[] = {0.1f, 0.5f, 0.9f, 0.3f, 0.2f, 0.8f, 1.f, 0.4f, 0.6f, 0.7f, 0.3f, 0.95f, 0.5f};
    const int ui[] = {0, 1, 2, 3, 4, 5, 6, 7, 1, 8, 9, 2, 5, 10, 11, 6, 2, 3, 6, 12};
    const int uo[] = {0, 8, 16};
    const uint8_t ut[] = {BN_UNSTRUCTURED_HEX, BN_UNSTRUCTURED_HEX, BN_UNSTRUCTURED_TET};
    const bn_float4 xf[3] = {{0.f, 0.f, 1.f, 0.f}, {1.f, .5f, 0.f, .5f}, {1.f, 1.f, 1.f, 1.f}};
    BNScalarField uf = bnScalarFieldCreate(ctx, 0, "unstructured");
    bnSetData(uf, "vertex.position", bnDataCreate(ctx, 0, BN_FLOAT32_VEC3, 13, uv));
    bnSetData(uf, "vertex.data", bnDataCreate(ctx, 0, BN_FLOAT32, 13, us));
    bnSetData(uf, "index", bnDataCreate(ctx, 0, BN_INT32, 20, ui));
    bnSetData(uf, "cell.index", bnDataCreate(ctx, 0, BN_INT32, 3, uo));
    bnSetData(uf, "cell.type", bnDataCreate(ctx, 0, BN_UINT8, 3, ut));
    BNVolume uvol = bnVolumeCreate(ctx, 0, uf);
    bnVolumeSetXF(uvol, {0.f, 1.f}, xf, 3, 2.f);
    BNGroup ug = bnGroupCreate(ctx, 0, nullptr, 0, &uvol, 1);
    BNModel um = bnModelCreate(ctx);
    bnSetInstances(um, 0, &ug, &xfm, 1);
    bnBuild(um, 0);
    bnAccumReset(fb);
    bnRender(rnd, um, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    void *b0, *b1, *b2, *b3, *b4;
    bnamd_device_malloc(&b0, sizeof uv);
    bnamd_device_malloc(&b1, sizeof us);
    bnamd_device_malloc(&b2, sizeof ui);
    bnamd_device_malloc(&b3, sizeof uo);
    bnamd_device_malloc(&b4, sizeof ut);
    bnamd_memcpy_h2d(b0, uv, sizeof uv);
    bnamd_memcpy_h2d(b1, us, sizeof us);
    bnamd_memcpy_h2d(b2, ui, sizeof ui);
    bnamd_memcpy_h2d(b3, uo, sizeof uo);
    bnamd_memcpy_h2d(b4, ut, sizeof ut);
    bnamd_umesh_desc ud{};
    ud.vertices = (const float *)b0; ud.scalars = (const float *)b1; ud.numVertices = 13;
    ud.indices = (const int *)b2; ud.numIndices = 20; ud.cellOffsets = (const int *)b3;
    ud.cellTypes = (const unsigned char *)b4; ud.numCells = 3;
    bnamd_umesh *mesh = nullptr;
    if (bnamd_umesh_create(&ud, nullptr, &mesh) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_umesh_set_tf(mesh, (const float *)xf, 3, 0.f, 1.f, 2.f);
    bnamd_umesh_render_opts uo2{};
    uo2.volume = 1;
    p.accumID = 0;
    bnamd_volume_stats us2;
    if (bnamd_umesh_render(mesh, &c, &p, &uo2, (float *)acc, nullptr, &us2) != BNAMD_OK) return fail("umesh render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("unstructured color differs");
    if (us2.samples <= 0) return fail("unstructured field not sampled");
    std::printf("unstructured samples %lld\n", us2.samples);
    bnamd_umesh_destroy(mesh);
    for (void *b : {b0, b1, b2, b3, b4}) bnamd_device_free(b);
    // AMR: a 4^3 level-0 block and a 4^3 level-1 block inside it
    const int ao[] = {0, 0, 0, 2, 2, 2};
    const int ad[] = {4, 4, 4, 4, 4, 4};
    const int al[] = {0, 1};
    const uint64_t aoff[] = {0, 64};
    const int aref[] = {2, 2};
    std::vector<float> asc(128);
    for (int i = 0; i < 128; ++i) asc[i] = 0.5f + 0.45f * std::sin(0.37f * (float)i);
    BNScalarField af = bnScalarFieldCreate(ctx, 0, "BlockStructuredAMR");
    bnSetData(af, "grid.origins", bnDataCreate(ctx, 0, BN_INT32_VEC3, 2, ao));
    bnSetData(af, "grid.dims", bnDataCreate(ctx, 0, BN_INT32_VEC3, 2, ad));
    bnSetData(af, "grid.levels", bnDataCreate(ctx, 0, BN_INT32, 2, al));
    bnSetData(af, "grid.offsets", bnDataCreate(ctx, 0, BN_UINT64, 2, aoff));
    bnSetData(af, "scalars", bnDataCreate(ctx, 0, BN_FLOAT32, 128, asc.data()));
    bnSetData(af, "level.refinements", bnDataCreate(ctx, 0, BN_INT32, 2, aref));
    BNVolume avol = bnVolumeCreate(ctx, 0, af);
    bnVolumeSetXF(avol, {0.f, 1.f}, xf, 3, 2.f);
    BNGroup ag = bnGroupCreate(ctx, 0, nullptr, 0, &avol, 1);
    BNModel am = bnModelCreate(ctx);
    bnSetInstances(am, 0, &ag, &xfm, 1);
    bnBuild(am, 0);
    bnAccumReset(fb);
    bnRender(rnd, am, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    void *dsc;
    bnamd_device_malloc(&dsc, asc.size() * 4);
    bnamd_memcpy_h2d(dsc, asc.data(), asc.size() * 4);
    const unsigned long long aoffu[] = {0, 64};
    bnamd_amr_desc ad2 = {2, ao, ad, al, aoffu, 2, aref, (const float *)dsc, 128};
    bnamd_amr *amr = nullptr;
    if (bnamd_amr_create(&ad2, nullptr, &amr) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_amr_set_tf(amr, (const float *)xf, 3, 0.f, 1.f, 2.f);
    bnamd_volume_stats as2;
    if (bnamd_amr_render(amr, &c, &p, (float *)acc, nullptr, &as2) != BNAMD_OK) return fail("amr render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("AMR color differs");
    std::printf("amr samples %lld\n", as2.samples);
    bnamd_amr_destroy(amr);
    bnamd_device_free(dsc);
  }
  // a NanoVDB field (the grid buffer from the file named on the command
  // line) through bn* equals bnamd_nanovdb
  if (nvdbPath) {
    std::vector<uint8_t> grid;
    if (FILE *fg = std::fopen(nvdbPath, "rb")) {
      std::fseek(fg, 0, SEEK_END);
      grid.resize((size_t)std::ftell(fg));
      std::fseek(fg, 0, SEEK_SET);
      if (std::fread(grid.data(), 1, grid.size(), fg) != grid.size()) grid.clear();
      std::fclose(fg);
    }
    if (grid.empty()) return fail("NanoVDB grid file unreadable");
    const bn_float4 xf[3] = {{0.f, 0.f, 1.f, 0.f}, {1.f, .5f, 0.f, .5f}, {1.f, 1.f, 1.f, 1.f}};
    BNScalarField nf = bnScalarFieldCreate(ctx, 0, "NanoVDB");
    bnSetData(nf, "data", bnDataCreate(ctx, 0, BN_UINT8, grid.size(), grid.data()));
    BNVolume nvol = bnVolumeCreate(ctx, 0, nf);
    bnVolumeSetXF(nvol, {0.f, 1.f}, xf, 3, 2.f);
    BNGroup ng = bnGroupCreate(ctx, 0, nullptr, 0, &nvol, 1);
    BNModel nm = bnModelCreate(ctx);
    bnSetInstances(nm, 0, &ng, &xfm, 1);
    bnBuild(nm, 0);
    bnAccumReset(fb);
    bnRender(rnd, nm, cam, fb);
    bnFrameBufferRead(fb, BN_FB_COLOR, color.data(), BN_FLOAT4);
    bnamd_nanovdb_desc nd{};
    nd.data = grid.data();
    nd.bytes = grid.size();
    bnamd_nanovdb *g = nullptr;
    if (bnamd_nanovdb_create(&nd, nullptr, &g) != BNAMD_OK) return fail(bnamd_last_error());
    bnamd_nanovdb_set_tf(g, (const float *)xf, 3, 0.f, 1.f, 2.f);
    p.accumID = 0;
    bnamd_volume_stats ns;
    if (bnamd_nanovdb_render(g, &c, &p, (float *)acc, nullptr, &ns) != BNAMD_OK) return fail("nanovdb render");
    bnamd_fb_resolve((const float *)acc, W, H, 1.f / SPP, BNAMD_FB_FLOAT4, out, nullptr);
    bnamd_synchronize();
    bnamd_memcpy_d2h(rc.data(), out, rc.size() * 4);
    if (std::memcmp(rc.data(), color.data(), rc.size() * 4) != 0) return fail("NanoVDB color differs");
    if (ns.samples <= 0) return fail("NanoVDB field not sampled");
    std::printf("nanovdb samples %lld\n", ns.samples);
    bnamd_nanovdb_destroy(g);
  }
  // errors surface as exceptions
  bool threw = false;
  try {
    bnGeometryCreate(ctx, 0, "capsules");
  } catch (const std::runtime_error &) {
    threw = true;
  }
  if (!threw) return fail("unknown geometry type accepted");
  bnamd_world_destroy(world);
  bnContextDestroy(ctx);
  std::printf("hits %d of %d\nOK\n", hits, W * H);
  return 0;
}
