// Structured volume object: macrocell ranges (StructuredData_computeMCs,
// barney/volume/StructuredData.cu:30-61), majorants (mapMCs,
// barney/volume/MCGrid.cu:94-116) and the delta-tracking renderer with the
// SCI_VIS_MODE shade path (barney/kernels/shadeRays.cu:24-26, 787-790).
#include "shade.h"
#include "volume.h"
#include "volume_render.h"
#include "amr.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace bnamd {

struct Volume {
  VolumeDD dd;
  DevBuf<float> bricks;
  mutable DevBuf<unsigned long long> counters;  // the render's counters
  int cellsPerMC = 8;
  DevBuf<float2> ranges;
  DevBuf<float> majorants;
  DevBuf<float4> tf;
  bool haveTF = false;
};

namespace {

__global__ void k_mc_ranges(VolumeDD v, int cpm, float2 *__restrict ranges) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = v.mx * v.my * v.mz;
  if (tid >= n) return;
  const int mx = tid % v.mx, my = (tid / v.mx) % v.my, mz = tid / (v.mx * v.my);
  float lo = BN_INF, hi = -BN_INF;
  for (int iz = 0; iz <= cpm; ++iz) {
    const int z = mz * cpm + iz;
    if (z >= v.nz) continue;
    for (int iy = 0; iy <= cpm; ++iy) {
      const int y = my * cpm + iy;
      if (y >= v.ny) continue;
      const float *row = v.scalars + ((size_t)z * v.ny + y) * v.nx;
      for (int ix = 0; ix <= cpm; ++ix) {
        const int x = mx * cpm + ix;
        if (x >= v.nx) continue;
        const float f = row[x];
        lo = fminf(lo, f);
        hi = fmaxf(hi, f);
      }
    }
  }
  ranges[tid] = make_float2(lo, hi);
}

// The bricked copy of the field: one thread per voxel of the caller's array
// (coalesced reads), written to its place in the brick layout. Voxels of the
// padded brick grid past the field's end stay unwritten: sampling clamps its
// corner coordinates into the field and never reads them.
__global__ void k_brick_field(VolumeDD v, float *__restrict out) {
  const size_t total = (size_t)v.nx * v.ny * v.nz;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % v.nx), y = (int)((i / v.nx) % v.ny), z = (int)(i / ((size_t)v.nx * v.ny));
    out[brick_x(x) + brick_y(y, v.brickStrideY) + brick_z(z, v.brickStrideZ)] = v.scalars[i];
  }
}


__global__ void k_volume_rays(VolumeDD v, const bnamd_ray *__restrict rays, const uint64_t *__restrict seeds,
                              float4 *__restrict out, int n, int exactLog) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bnamd_ray r = rays[i];
  RNGSeed seed;
  seed.state = seeds[i];
  VolCounters cnt{0, 0, 0};
  const vec3f dir = fix_dir({r.dir[0], r.dir[1], r.dir[2]});
  const VolHit h = volume_intersect(v, {r.org[0], r.org[1], r.org[2]}, dir, r.tmin, r.tmax, r.tmax, seed,
                                    exactLog != 0, cnt);
  out[i] = h.hit ? make_float4(h.t, h.rgb.x, h.rgb.y, h.rgb.z) : make_float4(BN_INF, 0.f, 0.f, 0.f);
}

// The macrocell value ranges of a structured volume in the form
// iso_intersect walks (MCIsoSurfaceAccel over StructuredData).
struct StructuredIsoGrid {
  vec3f worldLo, worldHi, mcOrigin, mcSpacing;
  int mx, my, mz;
  const float2 *ranges;
};

__global__ void __launch_bounds__(256) k_volume_iso_rays(VolumeDD v, StructuredIsoGrid g,
                                                         const bnamd_ray *__restrict rays, float iso,
                                                         float4 *__restrict out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bnamd_ray r = rays[i];
  RNGSeed seed;
  const IsoHit h = iso_intersect(
      g, [&](vec3f P) { return sample_field(v, P); }, iso, {r.org[0], r.org[1], r.org[2]},
      fix_dir({r.dir[0], r.dir[1], r.dir[2]}), r.tmin, r.tmax, r.tmax, seed, bn_hash3(0u, 0u, 0u));
  out[i] = h.hit ? make_float4(h.t, h.N.x, h.N.y, h.N.z) : make_float4(BN_INF, 0.f, 0.f, 0.f);
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

struct bnamd_volume_s {
  Volume v;
};

namespace bnamd {
// The device view of a structured volume for the other renderers (a model
// holding geometry and a volume renders through the surface path tracer).
const VolumeDD &volume_dd(const bnamd_volume *vol, const char *who) {
  BN_REQUIRE(vol, std::string(who) + ": NULL volume");
  BN_REQUIRE(vol->v.haveTF, std::string(who) + ": no transfer function set");
  return vol->v.dd;
}
}  // namespace bnamd

extern "C" int bnamd_volume_create(const bnamd_structured_desc *desc, bnamd_volume **out) {
  return guarded([&] {
    BN_REQUIRE(desc && out, "bnamd_volume_create: NULL argument");
    BN_REQUIRE(desc->scalars, "bnamd_volume_create: no scalars");
    BN_REQUIRE(desc->dims[0] >= 2 && desc->dims[1] >= 2 && desc->dims[2] >= 2,
               "bnamd_volume_create: dims must be >= 2 per axis");
    BN_REQUIRE(desc->cellsPerMC >= 1, "bnamd_volume_create: cellsPerMC must be >= 1");
    auto *vo = new bnamd_volume_s;
    try {
      Volume &V = vo->v;
      VolumeDD &d = V.dd;
      d.scalars = desc->scalars;
      d.nx = desc->dims[0]; d.ny = desc->dims[1]; d.nz = desc->dims[2];
      d.origin = {desc->gridOrigin[0], desc->gridOrigin[1], desc->gridOrigin[2]};
      d.spacing = {desc->gridSpacing[0], desc->gridSpacing[1], desc->gridSpacing[2]};
      // StructuredData::commit (StructuredData.cu:176-180)
      const vec3f cells = {(float)(d.nx - 1), (float)(d.ny - 1), (float)(d.nz - 1)};
      d.worldLo = d.origin;
      d.worldHi = d.origin + d.spacing * cells;
      // StructuredData::buildMCs (StructuredData.cu:70-95)
      const int cpm = desc->cellsPerMC;
      V.cellsPerMC = cpm;
      d.mx = div_up(d.nx - 1, cpm); d.my = div_up(d.ny - 1, cpm); d.mz = div_up(d.nz - 1, cpm);
      BN_REQUIRE((long long)d.my * d.mz < (1ll << 24), "bnamd_volume_create: over 2^24 macrocells in y*z");
      d.mcOrigin = d.worldLo;
      d.mcSpacing = vec3f((float)cpm) * d.spacing;
      const long long nmc = (long long)d.mx * d.my * d.mz;
      BN_REQUIRE(nmc < (1ll << 31), "bnamd_volume_create: too many macrocells");
      V.ranges.alloc(nmc);
      V.majorants.alloc(nmc);
      BN_HIP(hipMemset(V.majorants.ptr, 0, V.majorants.bytes()));
      d.majorants = V.majorants.ptr;
      d.tf = TFDD{nullptr, 0, 0.f, 1.f, 1.f};
      k_mc_ranges<<<div_up(nmc, 128), 128>>>(d, cpm, V.ranges.ptr);
      BN_HIP(hipGetLastError());
      // the bricked copy the renderers sample (32-bit voxel offsets)
      const uint32_t bx = (uint32_t)div_up(d.nx, kBrick), by = (uint32_t)div_up(d.ny, kBrick),
                     bz = (uint32_t)div_up(d.nz, kBrick);
      const size_t total = (size_t)bx * by * bz * kBrickVoxels;
      BN_REQUIRE(total <= 0xffffffffull, "bnamd_volume_create: field over 2^32 voxels (bricked)");
      V.bricks.alloc(total);
      d.brickStrideY = bx * kBrickVoxels;
      d.brickStrideZ = bx * by * kBrickVoxels;
      const size_t nvox = (size_t)d.nx * d.ny * d.nz;
      k_brick_field<<<(unsigned)std::min<size_t>((nvox + 255) / 256, 65536), 256>>>(d, V.bricks.ptr);
      BN_HIP(hipGetLastError());
      d.bricks = V.bricks.ptr;
      BN_HIP(hipDeviceSynchronize());
    } catch (...) {
      delete vo;
      throw;
    }
    *out = vo;
  });
}

extern "C" int bnamd_volume_destroy(bnamd_volume *vol) {
  return guarded([&] { delete vol; });
}

extern "C" int bnamd_volume_set_tf(bnamd_volume *vol, const float *values, int numValues, float domainLo,
                                   float domainHi, float baseDensity) {
  return guarded([&] {
    BN_REQUIRE(vol && values, "bnamd_volume_set_tf: NULL argument");
    BN_REQUIRE(numValues >= 2, "bnamd_volume_set_tf: need at least 2 values");
    BN_REQUIRE(domainHi > domainLo, "bnamd_volume_set_tf: empty domain");
    Volume &V = vol->v;
    V.tf.upload((const float4 *)values, numValues);
    V.dd.tf = TFDD{V.tf.ptr, numValues, domainLo, domainHi, baseDensity};
    V.haveTF = true;
    const int nmc = V.dd.mx * V.dd.my * V.dd.mz;
    k_majorants<<<div_up(nmc, 256), 256>>>(V.dd, V.ranges.ptr, V.majorants.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipDeviceSynchronize());
  });
}

extern "C" int bnamd_volume_mc_dims(const bnamd_volume *vol, int dims[3]) {
  return guarded([&] {
    BN_REQUIRE(vol && dims, "bnamd_volume_mc_dims: NULL argument");
    dims[0] = vol->v.dd.mx; dims[1] = vol->v.dd.my; dims[2] = vol->v.dd.mz;
  });
}

extern "C" int bnamd_volume_get_mcs(const bnamd_volume *vol, float *ranges, float *majorants) {
  return guarded([&] {
    BN_REQUIRE(vol, "bnamd_volume_get_mcs: NULL volume");
    const Volume &V = vol->v;
    if (ranges) BN_HIP(hipMemcpy(ranges, V.ranges.ptr, V.ranges.bytes(), hipMemcpyDeviceToHost));
    if (majorants) BN_HIP(hipMemcpy(majorants, V.majorants.ptr, V.majorants.bytes(), hipMemcpyDeviceToHost));
  });
}

extern "C" int bnamd_volume_render(const bnamd_volume *vol, const bnamd_camera *camera,
                                   const bnamd_render_params *params, float *accum, void *stream,
                                   bnamd_volume_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(vol && camera && params && accum, "bnamd_volume_render: NULL argument");
    BN_REQUIRE(vol->v.haveTF, "bnamd_volume_render: no transfer function set");
    BN_REQUIRE(params->width > 0 && params->height > 0 && params->pathsPerPixel >= 1,
               "bnamd_volume_render: bad frame size or pathsPerPixel");
    hipStream_t s = (hipStream_t)stream;
    const CameraDD cam = make_camera(*camera);
    const RenderDD rp = make_render_dd(*params);
    unsigned long long h[4] = {0, 0, 0, 0};
    const float ms = render_on_stream(s, vol->v.counters, 4, stats != nullptr, h, [&] {
      render_volume_dispatch(vol->v.dd, cam, rp, params->pathsPerPixel, (float4 *)accum, vol->v.counters.ptr, s);
    });
    if (stats) {
      *stats = bnamd_volume_stats{};
      stats->samples = (long long)h[0];
      stats->ddaSteps = (long long)h[1];
      stats->raysTraced = (long long)h[2];
      stats->ms = ms;
      stats->zeroMajorantSteps = (long long)h[3];
    }
  });
}

extern "C" int bnamd_volume_sample_rays(const bnamd_volume *vol, const bnamd_ray *rays, const uint64_t *seeds,
                                        float *out, int numRays, int exactLog, void *stream) {
  return guarded([&] {
    BN_REQUIRE(vol && rays && seeds && out, "bnamd_volume_sample_rays: NULL argument");
    BN_REQUIRE(vol->v.haveTF, "bnamd_volume_sample_rays: no transfer function set");
    if (numRays <= 0) return;
    hipStream_t s = (hipStream_t)stream;
    k_volume_rays<<<div_up(numRays, 256), 256, 0, s>>>(vol->v.dd, rays, seeds, (float4 *)out, numRays,
                                                       exactLog);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_volume_iso_rays(const bnamd_volume *vol, const bnamd_ray *rays, float isoValue, float *out,
                                     int numRays, void *stream) {
  return guarded([&] {
    BN_REQUIRE(vol && rays && out, "bnamd_volume_iso_rays: NULL argument");
    if (numRays <= 0) return;
    const VolumeDD &d = vol->v.dd;
    const StructuredIsoGrid g{d.worldLo, d.worldHi, d.mcOrigin, d.mcSpacing, d.mx, d.my, d.mz, vol->v.ranges.ptr};
    hipStream_t s = (hipStream_t)stream;
    k_volume_iso_rays<<<div_up(numRays, 256), 256, 0, s>>>(d, g, rays, isoValue, (float4 *)out, numRays);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

// ===========================================================================
// Block-structured AMR (barney/amr/BlockStructuredField.{h,cu},
// BlockStructuredCuBQLSampler.h): blocks of cells at levels; a level-L cell
// is (int)powf(refinement[L], L) units wide.
// ===========================================================================
namespace bnamd {

struct Amr {
  AmrDD dd{};
  mutable DevBuf<unsigned long long> counters;  // the render's counters
  int numBlocks = 0;
  int lookup[3] = {0, 0, 0};
  DevBuf<float4> dom;
  DevBuf<int4> dims, origins;
  DevBuf<unsigned long long> offsets;
  DevBuf<uint32_t> cellStart, cellBlocks;
  DevBuf<float2> ranges;
  DevBuf<float> majorants;
  DevBuf<float4> tf;
  bool haveTF = false;
  float buildMs = 0.f;
};

namespace {

__device__ __forceinline__ uint32_t amr_ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_amr_clear(uint2 *__restrict r, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    r[i] = make_uint2(amr_ord(BN_INF), amr_ord(-BN_INF));
}

// BSField_rasterGrids (BlockStructuredField.cu:33-57): every cell's basis
// support, cellBounds = [(origin + cell - 1/2) cs, + 2 cs], with its scalar,
// through rasterBox (MCGrid.cuh:67-91). One workgroup per block.
__global__ void k_amr_raster(AmrDD a, const int4 *__restrict origins, uint2 *__restrict ranges, int numBlocks,
                             vec3f rcpSpacing) {
#pragma clang fp contract(off)
  const int b = blockIdx.x;
  if (b >= numBlocks) return;
  const float cs = a.dom[2 * b].w;
  const int4 o = origins[b], d = a.dims[b];
  const float *sc = a.scalars + a.offsets[b];
  const int n = d.x * d.y * d.z;
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int x = c % d.x, y = (c / d.x) % d.y, z = c / (d.x * d.y);
    const uint32_t s = amr_ord(sc[c]);
    const vec3f lo = {((float)(o.x + x) - .5f) * cs, ((float)(o.y + y) - .5f) * cs, ((float)(o.z + z) - .5f) * cs};
    const vec3f hi = {lo.x + 2.f * cs, lo.y + 2.f * cs, lo.z + 2.f * cs};
    const int x0 = min(max((int)((lo.x - a.mcOrigin.x) * rcpSpacing.x), 0), a.mx - 1);
    const int y0 = min(max((int)((lo.y - a.mcOrigin.y) * rcpSpacing.y), 0), a.my - 1);
    const int z0 = min(max((int)((lo.z - a.mcOrigin.z) * rcpSpacing.z), 0), a.mz - 1);
    const int x1 = min(max((int)((hi.x - a.mcOrigin.x) * rcpSpacing.x), 0), a.mx - 1);
    const int y1 = min(max((int)((hi.y - a.mcOrigin.y) * rcpSpacing.y), 0), a.my - 1);
    const int z1 = min(max((int)((hi.z - a.mcOrigin.z) * rcpSpacing.z), 0), a.mz - 1);
    for (int iz = z0; iz <= z1; ++iz)
      for (int iy = y0; iy <= y1; ++iy)
        for (int ix = x0; ix <= x1; ++ix) {
          uint2 *cell = ranges + ix + (size_t)iy * a.mx + (size_t)iz * a.mx * a.my;
          atomicMin(&cell->x, s);
          atomicMax(&cell->y, s);
        }
  }
}

__global__ void k_amr_finish(const uint2 *__restrict in, float2 *__restrict out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const uint2 u = in[i];
    auto f = [](uint32_t v) { return __uint_as_float((v & 0x80000000u) ? (v & 0x7fffffffu) : ~v); };
    out[i] = make_float2(f(u.x), f(u.y));
  }
}

__global__ void __launch_bounds__(256) k_amr_sample(AmrDD a, const float *__restrict pts, float *__restrict out,
                                                    int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = amr_sample(a, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
}

}  // namespace
}  // namespace bnamd

struct bnamd_amr_s {
  Amr a;
};

extern "C" int bnamd_amr_create(const bnamd_amr_desc *desc, void *stream, bnamd_amr **out) {
  return guarded([&] {
    BN_REQUIRE(desc && out, "bnamd_amr_create: NULL argument");
    BN_REQUIRE(desc->numBlocks > 0 && desc->numLevels > 0, "bnamd_amr_create: no blocks or levels");
    BN_REQUIRE(desc->origins && desc->dims && desc->levels && desc->offsets && desc->refinements && desc->scalars,
               "bnamd_amr_create: NULL array");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = (hipStream_t)stream;
    auto *ua = new bnamd_amr_s;
    try {
      Amr &A = ua->a;
      const int nb = desc->numBlocks;
      A.numBlocks = nb;
      std::vector<float4> dom(2 * (size_t)nb);
      std::vector<int4> dims(nb), origins(nb);
      vec3f wlo = vec3f(BN_INF), whi = vec3f(-BN_INF);
      for (int b = 0; b < nb; ++b) {
        const int *o = desc->origins + 3 * (size_t)b, *d = desc->dims + 3 * (size_t)b;
        const int L = desc->levels[b];
        BN_REQUIRE(L >= 0 && L < desc->numLevels, "bnamd_amr_create: block level out of range");
        BN_REQUIRE(d[0] > 0 && d[1] > 0 && d[2] > 0, "bnamd_amr_create: empty block");
        BN_REQUIRE(desc->offsets[b] + (unsigned long long)d[0] * d[1] * d[2] <= desc->numScalars,
                   "bnamd_amr_create: block scalars out of range");
        // Block::getFrom / getDomain (BlockStructuredField.h:227-248)
        const float cs = (float)(int)(powf((float)desc->refinements[L], (float)L));
        const vec3f lo = {((float)o[0] - .5f) * cs, ((float)o[1] - .5f) * cs, ((float)o[2] - .5f) * cs};
        const vec3f hi = {((float)(o[0] + d[0]) + .5f) * cs, ((float)(o[1] + d[1]) + .5f) * cs,
                          ((float)(o[2] + d[2]) + .5f) * cs};
        dom[2 * b] = make_float4(lo.x, lo.y, lo.z, cs);
        dom[2 * b + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
        dims[b] = make_int4(d[0], d[1], d[2], 0);
        origins[b] = make_int4(o[0], o[1], o[2], 0);
        wlo = vmin(wlo, lo);
        whi = vmax(whi, hi);
      }
      AmrDD &dd = A.dd;
      dd.worldLo = wlo;
      dd.worldHi = whi;
      const vec3f size = whi - wlo;
      const float maxWidth = reduce_max(size);
      // buildMCs (BlockStructuredField.cu:82-105): 256 cells along the widest axis
      {
        const float f = (256 - 1) / maxWidth;
        dd.mx = 1 + (int)(size.x * f);
        dd.my = 1 + (int)(size.y * f);
        dd.mz = 1 + (int)(size.z * f);
        BN_REQUIRE((long long)dd.my * dd.mz < (1ll << 24), "bnamd_amr_create: over 2^24 macrocells in y*z");
        dd.mcOrigin = wlo;
        dd.mcSpacing = size * rcp(vec3f((float)dd.mx, (float)dd.my, (float)dd.mz));
      }
      // block lookup grid (in place of the reference's BVH over block domains)
      {
        const int G = std::min(256, std::max(16, 4 * (int)std::ceil(std::cbrt((double)nb))));
        const float f = (G - 1) / maxWidth;
        dd.gx = 1 + (int)(size.x * f);
        dd.gy = 1 + (int)(size.y * f);
        dd.gz = 1 + (int)(size.z * f);
        dd.gOrigin = wlo;
        const vec3f gsp = size * rcp(vec3f((float)dd.gx, (float)dd.gy, (float)dd.gz));
        dd.gRcp = rcp(gsp);
        A.lookup[0] = dd.gx; A.lookup[1] = dd.gy; A.lookup[2] = dd.gz;
        const size_t ng = (size_t)dd.gx * dd.gy * dd.gz;
        std::vector<uint32_t> start(ng + 1, 0u);
        std::vector<int> r(6 * (size_t)nb);
        auto cell = [](float p, float o, float rc, int n) { return std::min(std::max((int)((p - o) * rc), 0), n - 1); };
        for (int b = 0; b < nb; ++b) {
          const float4 lo = dom[2 * b], hi = dom[2 * b + 1];
          int *q = &r[6 * (size_t)b];
          q[0] = cell(lo.x, wlo.x, dd.gRcp.x, dd.gx); q[3] = cell(hi.x, wlo.x, dd.gRcp.x, dd.gx);
          q[1] = cell(lo.y, wlo.y, dd.gRcp.y, dd.gy); q[4] = cell(hi.y, wlo.y, dd.gRcp.y, dd.gy);
          q[2] = cell(lo.z, wlo.z, dd.gRcp.z, dd.gz); q[5] = cell(hi.z, wlo.z, dd.gRcp.z, dd.gz);
          for (int z = q[2]; z <= q[5]; ++z)
            for (int y = q[1]; y <= q[4]; ++y)
              for (int x = q[0]; x <= q[3]; ++x) ++start[1 + x + (size_t)dd.gx * (y + (size_t)dd.gy * z)];
        }
        for (size_t c = 0; c < ng; ++c) start[c + 1] += start[c];
        std::vector<uint32_t> list(start[ng]), fill(start.begin(), start.end() - 1);
        for (int b = 0; b < nb; ++b) {  // ascending block ids per cell
          const int *q = &r[6 * (size_t)b];
          for (int z = q[2]; z <= q[5]; ++z)
            for (int y = q[1]; y <= q[4]; ++y)
              for (int x = q[0]; x <= q[3]; ++x) list[fill[x + (size_t)dd.gx * (y + (size_t)dd.gy * z)]++] = b;
        }
        A.cellStart.upload(start.data(), start.size(), s);
        A.cellBlocks.upload(list.data(), std::max<size_t>(list.size(), 1), s);
      }
      A.dom.upload(dom.data(), dom.size(), s);
      A.dims.upload(dims.data(), dims.size(), s);
      A.origins.upload(origins.data(), origins.size(), s);
      A.offsets.upload(desc->offsets, nb, s);
      dd.dom = A.dom.ptr;
      dd.dims = A.dims.ptr;
      dd.offsets = A.offsets.ptr;
      dd.scalars = desc->scalars;
      dd.cellStart = A.cellStart.ptr;
      dd.cellBlocks = A.cellBlocks.ptr;
      const size_t nmc = (size_t)dd.mx * dd.my * dd.mz;
      BN_REQUIRE(nmc < (1ull << 31), "bnamd_amr_create: too many macrocells");
      DevBuf<uint2> ord(nmc);
      A.ranges.alloc(nmc);
      A.majorants.alloc(nmc);
      const int gs = (int)std::min<size_t>(div_up(nmc, 256), 8192);
      k_amr_clear<<<gs, 256, 0, s>>>(ord.ptr, nmc);
      k_amr_raster<<<nb, 256, 0, s>>>(dd, A.origins.ptr, ord.ptr, nb, rcp(dd.mcSpacing));
      k_amr_finish<<<gs, 256, 0, s>>>(ord.ptr, A.ranges.ptr, nmc);
      BN_HIP(hipGetLastError());
      BN_HIP(hipMemsetAsync(A.majorants.ptr, 0, A.majorants.bytes(), s));
      BN_HIP(hipStreamSynchronize(s));
      dd.majorants = A.majorants.ptr;
      dd.tf = TFDD{nullptr, 0, 0.f, 1.f, 1.f};
      A.buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } catch (...) {
      delete ua;
      throw;
    }
    *out = ua;
  });
}

extern "C" int bnamd_amr_destroy(bnamd_amr *amr) {
  return guarded([&] { delete amr; });
}

extern "C" int bnamd_amr_info_get(const bnamd_amr *amr, bnamd_amr_info *info) {
  return guarded([&] {
    BN_REQUIRE(amr && info, "bnamd_amr_info_get: NULL argument");
    const Amr &A = amr->a;
    info->numBlocks = A.numBlocks;
    info->mcDims[0] = A.dd.mx; info->mcDims[1] = A.dd.my; info->mcDims[2] = A.dd.mz;
    for (int k = 0; k < 3; ++k) info->lookupDims[k] = A.lookup[k];
    info->worldLower[0] = A.dd.worldLo.x; info->worldLower[1] = A.dd.worldLo.y; info->worldLower[2] = A.dd.worldLo.z;
    info->worldUpper[0] = A.dd.worldHi.x; info->worldUpper[1] = A.dd.worldHi.y; info->worldUpper[2] = A.dd.worldHi.z;
    info->buildMs = A.buildMs;
  });
}

extern "C" int bnamd_amr_set_tf(bnamd_amr *amr, const float *values, int numValues, float domainLo, float domainHi,
                                float baseDensity) {
  return guarded([&] {
    BN_REQUIRE(amr && values, "bnamd_amr_set_tf: NULL argument");
    BN_REQUIRE(numValues >= 2, "bnamd_amr_set_tf: need at least 2 values");
    BN_REQUIRE(domainHi > domainLo, "bnamd_amr_set_tf: empty domain");
    Amr &A = amr->a;
    A.tf.upload((const float4 *)values, numValues);
    A.dd.tf = TFDD{A.tf.ptr, numValues, domainLo, domainHi, baseDensity};
    A.haveTF = true;
    const int nmc = A.dd.mx * A.dd.my * A.dd.mz;
    k_majorants<<<div_up(nmc, 256), 256>>>(A.dd, A.ranges.ptr, A.majorants.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipDeviceSynchronize());
  });
}

extern "C" int bnamd_amr_get_mcs(const bnamd_amr *amr, float *ranges, float *majorants) {
  return guarded([&] {
    BN_REQUIRE(amr, "bnamd_amr_get_mcs: NULL argument");
    const Amr &A = amr->a;
    if (ranges) BN_HIP(hipMemcpy(ranges, A.ranges.ptr, A.ranges.bytes(), hipMemcpyDeviceToHost));
    if (majorants) BN_HIP(hipMemcpy(majorants, A.majorants.ptr, A.majorants.bytes(), hipMemcpyDeviceToHost));
  });
}

extern "C" int bnamd_amr_sample_points(const bnamd_amr *amr, const float *points, float *out, int numPoints,
                                       void *stream) {
  return guarded([&] {
    BN_REQUIRE(amr && points && out, "bnamd_amr_sample_points: NULL argument");
    if (numPoints <= 0) return;
    hipStream_t s = (hipStream_t)stream;
    k_amr_sample<<<div_up(numPoints, 256), 256, 0, s>>>(amr->a.dd, points, out, numPoints);
    BN_HIP(hipGetLastError());
    BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_amr_render(const bnamd_amr *amr, const bnamd_camera *camera, const bnamd_render_params *params,
                                float *accum, void *stream, bnamd_volume_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(amr && camera && params && accum, "bnamd_amr_render: NULL argument");
    BN_REQUIRE(amr->a.haveTF, "bnamd_amr_render: no transfer function set");
    BN_REQUIRE(params->width > 0 && params->height > 0 && params->pathsPerPixel >= 1,
               "bnamd_amr_render: bad frame size or pathsPerPixel");
    hipStream_t s = (hipStream_t)stream;
    const CameraDD cam = make_camera(*camera);
    const RenderDD rp = make_render_dd(*params);
    unsigned long long h[4] = {0, 0, 0, 0};
    const float ms = render_on_stream(s, amr->a.counters, 4, stats != nullptr, h, [&] {
      render_volume_dispatch(amr->a.dd, cam, rp, params->pathsPerPixel, (float4 *)accum, amr->a.counters.ptr, s);
    });
    if (stats) {
      *stats = bnamd_volume_stats{};
      stats->samples = (long long)h[0];
      stats->ddaSteps = (long long)h[1];
      stats->raysTraced = (long long)h[2];
      stats->ms = ms;
      stats->zeroMajorantSteps = (long long)h[3];
    }
  });
}
