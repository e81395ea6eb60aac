// NanoVDB scalar fields (barney/volume/NanoVDB.{h,cu}): the grid buffer is
// checked on the host (NanoVDBData::commit reads its header there too),
// copied once into HBM, its macrocell ranges built by NanoVDBData_computeMCs'
// rule, and it is rendered by the volume path tracer and walked for
// iso-surfaces like a structured field.
#include "nanovdb.h"
#include "volume_render.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace bnamd {

struct Nvdb {
  NvdbDD dd{};
  mutable DevBuf<unsigned long long> counters;  // the render's counters
  DevBuf<uint64_t> grid;                        // the grid buffer (8-byte words)
  unsigned long long bytes = 0;
  DevBuf<float2> ranges;
  DevBuf<float> majorants;
  DevBuf<float4> tf;
  DevBuf<uint32_t> leafTable;
  bool haveTF = false;
  double voxelSize[3] = {0, 0, 0};
  float buildMs = 0.f;
};

namespace {

constexpr size_t kUpperBytes = 270400, kLowerBytes = 33856;

template <class T>
T host_load(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

// Checks that b holds one NanoVDB 32.x grid of a supported value type whose
// nodes, reached from the root, all lie inside the buffer (so that device
// walks stay in bounds).
template <class LeafFn>
void check_grid(const uint8_t *b, size_t bytes, int &type, LeafFn &&onLeaf) {
  BN_REQUIRE(bytes >= 672 + 64 + 64, "bnamd_nanovdb_create: buffer smaller than a grid header");
  const uint64_t magic = host_load<uint64_t>(b);
  BN_REQUIRE((magic & 0x00FFFFFFFFFFFFFFull) == 0x004244566f6e614eull,
             "bnamd_nanovdb_create: not a NanoVDB grid (magic number)");
  const uint32_t version = host_load<uint32_t>(b + 16);
  BN_REQUIRE((version >> 21) == 32u, "bnamd_nanovdb_create: NanoVDB major version is not 32");
  type = (int)host_load<uint32_t>(b + 636);
  BN_REQUIRE(type == kNvFloat || type == kNvFp4 || type == kNvFp8 || type == kNvFp16 || type == kNvFpN,
             "bnamd_nanovdb_create: unsupported grid type (float, Fp4, Fp8, Fp16 and FpN are)");
  BN_REQUIRE(host_load<int64_t>(b + 672 + 24) == 64,
             "bnamd_nanovdb_create: the root node must follow the tree header");
  const uint8_t *root = b + 672 + 64;
  const uint32_t nt = host_load<uint32_t>(root + 24);
  BN_REQUIRE(nt > 0, "bnamd_nanovdb_create: empty grid");
  // a node of n bytes at byte offset 'base + off' lies inside the buffer
  // (checked on the offsets, before any pointer is formed from them)
  auto inside = [&](size_t base, int64_t off, size_t n) {
    if (off < 0 || (uint64_t)off > bytes) return false;
    const size_t at = base + (size_t)off;
    return at <= bytes && n <= bytes - at;
  };
  const size_t rootAt = 672 + 64;
  BN_REQUIRE(nt <= (bytes - rootAt) / 32 && inside(rootAt, 0, 64 + 32 * (size_t)nt),
             "bnamd_nanovdb_create: root tiles outside the buffer");
  for (uint32_t i = 0; i < nt; ++i) {
    const int64_t ch = host_load<int64_t>(root + 64 + 32 * (size_t)i + 8);
    if (!ch) continue;
    BN_REQUIRE(ch > 0 && inside(rootAt, ch, kUpperBytes), "bnamd_nanovdb_create: upper node outside the buffer");
    // the upper node's origin from its tile key (RootData::KeyToCoord)
    const uint64_t key = host_load<uint64_t>(root + 64 + 32 * (size_t)i);
    const int ux = (int)(uint32_t)(((key >> 42) & 0x1FFFFFu) << 12);
    const int uy = (int)(uint32_t)(((key >> 21) & 0x1FFFFFu) << 12);
    const int uz = (int)(uint32_t)((key & 0x1FFFFFu) << 12);
    const size_t upAt = rootAt + (size_t)ch;
    const uint8_t *up = b + upAt;
    for (uint32_t n = 0; n < 32768; ++n) {
      if (!((host_load<uint64_t>(up + 4128 + 8 * (n >> 6)) >> (n & 63)) & 1u)) continue;
      const int64_t lc = host_load<int64_t>(up + 8256 + 8 * (size_t)n);
      BN_REQUIRE(lc > 0 && inside(upAt, lc, kLowerBytes), "bnamd_nanovdb_create: lower node outside the buffer");
      const size_t loAt = upAt + (size_t)lc;
      const uint8_t *lo = b + loAt;
      for (uint32_t m = 0; m < 4096; ++m) {
        if (!((host_load<uint64_t>(lo + 544 + 8 * (m >> 6)) >> (m & 63)) & 1u)) continue;
        const int64_t fc = host_load<int64_t>(lo + 1088 + 8 * (size_t)m);
        BN_REQUIRE(fc > 0 && inside(loAt, fc, 96), "bnamd_nanovdb_create: leaf outside the buffer");
        const size_t leafAt = loAt + (size_t)fc;
        const uint8_t *leaf = b + leafAt;
        size_t size = 0;
        switch (type) {
          case kNvFloat: size = 2144; break;
          case kNvFp4: size = 352; break;
          case kNvFp8: size = 608; break;
          case kNvFp16: size = 1120; break;
          default: {
            const int bw = leaf[15] >> 5;
            BN_REQUIRE(bw <= 4, "bnamd_nanovdb_create: FpN leaf with a bit width over 16");
            size = 96 + ((size_t)1 << bw) * 64;
          }
        }
        BN_REQUIRE(inside(leafAt, 0, size), "bnamd_nanovdb_create: leaf outside the buffer");
        // the leaf's origin: the upper origin plus the two table slots
        onLeaf(leaf, ux + (int)((n >> 10) << 7) + (int)((m >> 8) << 3),
               uy + (int)(((n >> 5) & 31) << 7) + (int)(((m >> 4) & 15) << 3),
               uz + (int)((n & 31) << 7) + (int)((m & 15) << 3));
      }
    }
  }
}

// NanoVDBData_computeMCs (NanoVDB.cu:28-62): the value range of the voxels
// lower + id * gridSize / dims .. lower + (id + 1) * gridSize / dims + 1 of
// each macrocell. One wavefront per macrocell; its lanes share the voxels
// and reduce with fminf / fmaxf (order-independent).
__global__ void __launch_bounds__(256) k_nvdb_mcs(NvdbDD g, float2 *__restrict ranges) {
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= g.mx * g.my * g.mz) return;
  const int id[3] = {wave % g.mx, (wave / g.mx) % g.my, wave / (g.mx * g.my)};
  const int dims[3] = {g.mx, g.my, g.mz};
  int lo[3], n[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int gs = g.ihi[k] - g.ilo[k] + 1;
    lo[k] = g.ilo[k] + (id[k] * gs) / dims[k];
    n[k] = g.ilo[k] + ((id[k] + 1) * gs) / dims[k] + 1 - lo[k] + 1;
  }
  float rlo = BN_INF, rhi = -BN_INF;
  const int total = n[0] * n[1] * n[2];
  for (int i = lane; i < total; i += 64) {
    const int z = i % n[2], y = (i / n[2]) % n[1], x = i / (n[1] * n[2]);
    const float v = nv_value(g, lo[0] + x, lo[1] + y, lo[2] + z);
    rlo = fminf(rlo, v);
    rhi = fmaxf(rhi, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    rlo = fminf(rlo, __shfl_xor(rlo, off));
    rhi = fmaxf(rhi, __shfl_xor(rhi, off));
  }
  if (lane == 0) ranges[wave] = make_float2(rlo, rhi);
}

__global__ void __launch_bounds__(256) k_nvdb_sample(NvdbDD g, const float *__restrict pts, float *__restrict out,
                                                     int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = nvdb_sample(g, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
}

// the macrocell value ranges in the form iso_intersect walks
// (MCIsoSurfaceAccel over NanoVDBDataSampler)
struct NvdbIsoGrid {
  vec3f worldLo, worldHi, mcOrigin, mcSpacing;
  int mx, my, mz;
  const float2 *ranges;
};

__global__ void __launch_bounds__(256) k_nvdb_iso_rays(NvdbDD g, NvdbIsoGrid ig, const bnamd_ray *__restrict rays,
                                                       float iso, float4 *__restrict out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bnamd_ray r = rays[i];
  RNGSeed seed;
  const IsoHit h = iso_intersect(
      ig, [&](vec3f P) { return nvdb_sample(g, P); }, iso, {r.org[0], r.org[1], r.org[2]},
      fix_dir({r.dir[0], r.dir[1], r.dir[2]}), r.tmin, r.tmax, r.tmax, seed, bn_hash3(0u, 0u, 0u));
  out[i] = h.hit ? make_float4(h.t, h.N.x, h.N.y, h.N.z) : make_float4(BN_INF, 0.f, 0.f, 0.f);
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

struct bnamd_nanovdb_s {
  Nvdb g;
};

extern "C" int bnamd_nanovdb_create(const bnamd_nanovdb_desc *desc, void *stream, bnamd_nanovdb **out) {
  return guarded([&] {
    BN_REQUIRE(desc && out && desc->data, "bnamd_nanovdb_create: NULL argument");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)desc->bytes;
    // the host copy the header is read from (NanoVDBData::commit downloads it)
    std::vector<uint64_t> host((bytes + 7) / 8);
    if (desc->dataOnDevice) {
      BN_HIP(hipMemcpyAsync(host.data(), desc->data, bytes, hipMemcpyDeviceToHost, s));
      BN_HIP(hipStreamSynchronize(s));
    } else {
      std::memcpy(host.data(), desc->data, bytes);
    }
    const uint8_t *b = reinterpret_cast<const uint8_t *>(host.data());
    int type = 0;
    struct LeafAt {
      const uint8_t *p;
      int x, y, z;
    };
    std::vector<LeafAt> leaves;
    check_grid(b, bytes, type, [&](const uint8_t *leaf, int x, int y, int z) { leaves.push_back({leaf, x, y, z}); });
    auto *ug = new bnamd_nanovdb_s;
    try {
      Nvdb &G = ug->g;
      G.bytes = bytes;
      G.grid.alloc(host.size());
      BN_HIP(hipMemcpyAsync(G.grid.ptr, host.data(), bytes, hipMemcpyHostToDevice, s));
      NvdbDD &d = G.dd;
      const uint8_t *root = b + 672 + 64;
      d.root = reinterpret_cast<const uint8_t *>(G.grid.ptr) + 672 + 64;
      d.numTiles = host_load<uint32_t>(root + 24);
      d.background = host_load<float>(root + 28);
      d.type = type;
      for (int k = 0; k < 3; ++k) {
        d.ilo[k] = host_load<int32_t>(root + 4 * k);
        d.ihi[k] = host_load<int32_t>(root + 12 + 4 * k);
        BN_REQUIRE(d.ihi[k] >= d.ilo[k], "bnamd_nanovdb_create: empty index box");
      }
      // the leaf table over the index box grown by one voxel (the
      // stencil's far corners), when it stays within 2^24 blocks and every
      // leaf offset is a multiple of 32 bytes (as NanoVDB writers align them)
      {
        long long tcells = 1;
        for (int k = 0; k < 3; ++k) {
          d.tlo[k] = d.ilo[k] >> 3;
          d.tdim[k] = ((d.ihi[k] + 1) >> 3) - d.tlo[k] + 1;
          tcells *= d.tdim[k];
        }
        bool aligned = true;
        for (const LeafAt &l : leaves) aligned = aligned && ((l.p - b) % 32 == 0) && (l.p - b) / 32 < (1ll << 32);
        if (!desc->noLeafTable && tcells <= (1ll << 24) && aligned) {
          std::vector<uint32_t> table((size_t)tcells, 0u);
          for (const LeafAt &l : leaves) {
            const int bx = (l.x >> 3) - d.tlo[0], by = (l.y >> 3) - d.tlo[1], bz = (l.z >> 3) - d.tlo[2];
            if (bx < 0 || by < 0 || bz < 0 || bx >= d.tdim[0] || by >= d.tdim[1] || bz >= d.tdim[2]) continue;
            table[bx + (size_t)d.tdim[0] * (by + (size_t)d.tdim[1] * bz)] = (uint32_t)((l.p - b) / 32);
          }
          G.leafTable.upload(table.data(), table.size(), s);
          d.leafTable = G.leafTable.ptr;
        }
        d.base = reinterpret_cast<const uint8_t *>(G.grid.ptr);
      }
      std::memcpy(d.invMatF, b + 296 + 36, sizeof(d.invMatF));
      std::memcpy(d.vecF, b + 296 + 72, sizeof(d.vecF));
      std::memcpy(G.voxelSize, b + 608, sizeof(G.voxelSize));
      // NanoVDBData::commit: the world box as floats
      double wb[6];
      std::memcpy(wb, b + 560, sizeof(wb));
      d.worldLo = {(float)wb[0], (float)wb[1], (float)wb[2]};
      d.worldHi = {(float)wb[3], (float)wb[4], (float)wb[5]};
      // buildMCs (NanoVDB.cu:82-102): divRoundUp(gridSize, 16) cells over the
      // world box, spacing = size * rcp(dims)
      long long gs[3];
      for (int k = 0; k < 3; ++k) gs[k] = (long long)d.ihi[k] - d.ilo[k] + 1;
      d.mx = (int)((gs[0] + 15) / 16);
      d.my = (int)((gs[1] + 15) / 16);
      d.mz = (int)((gs[2] + 15) / 16);
      // the cell bounds are products in int, as in the reference
      for (int k = 0; k < 3; ++k)
        BN_REQUIRE(gs[k] * (long long)((gs[k] + 15) / 16 + 1) < (1ll << 31),
                   "bnamd_nanovdb_create: index box too large for the macrocell arithmetic");
      BN_REQUIRE((long long)d.my * d.mz < (1ll << 24), "bnamd_nanovdb_create: over 2^24 macrocells in y*z");
      const long long nmc = (long long)d.mx * d.my * d.mz;
      BN_REQUIRE(nmc < (1ll << 25), "bnamd_nanovdb_create: too many macrocells");
      d.mcOrigin = d.worldLo;
      d.mcSpacing = (d.worldHi - d.worldLo) * rcp(vec3f((float)d.mx, (float)d.my, (float)d.mz));
      G.ranges.alloc(nmc);
      G.majorants.alloc(nmc);
      BN_HIP(hipMemsetAsync(G.majorants.ptr, 0, G.majorants.bytes(), s));
      d.majorants = G.majorants.ptr;
      d.tf = TFDD{nullptr, 0, 0.f, 1.f, 1.f};
      k_nvdb_mcs<<<div_up(nmc * 64, 256), 256, 0, s>>>(d, G.ranges.ptr);
      BN_HIP(hipGetLastError());
      BN_HIP(hipStreamSynchronize(s));
      G.buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    } catch (...) {
      delete ug;
      throw;
    }
    *out = ug;
  });
}

extern "C" int bnamd_nanovdb_destroy(bnamd_nanovdb *g) {
  return guarded([&] { delete g; });
}

extern "C" int bnamd_nanovdb_info_get(const bnamd_nanovdb *g, bnamd_nanovdb_info *info) {
  return guarded([&] {
    BN_REQUIRE(g && info, "bnamd_nanovdb_info_get: NULL argument");
    const Nvdb &G = g->g;
    const NvdbDD &d = G.dd;
    *info = bnamd_nanovdb_info{};
    info->gridType = d.type;
    for (int k = 0; k < 3; ++k) {
      info->indexLower[k] = d.ilo[k];
      info->indexUpper[k] = d.ihi[k];
      info->voxelSize[k] = (float)G.voxelSize[k];
    }
    info->mcDims[0] = d.mx; info->mcDims[1] = d.my; info->mcDims[2] = d.mz;
    info->worldLower[0] = d.worldLo.x; info->worldLower[1] = d.worldLo.y; info->worldLower[2] = d.worldLo.z;
    info->worldUpper[0] = d.worldHi.x; info->worldUpper[1] = d.worldHi.y; info->worldUpper[2] = d.worldHi.z;
    info->mcOrigin[0] = d.mcOrigin.x; info->mcOrigin[1] = d.mcOrigin.y; info->mcOrigin[2] = d.mcOrigin.z;
    info->mcSpacing[0] = d.mcSpacing.x; info->mcSpacing[1] = d.mcSpacing.y; info->mcSpacing[2] = d.mcSpacing.z;
    info->gridBytes = G.bytes;
    info->buildMs = G.buildMs;
  });
}

extern "C" int bnamd_nanovdb_set_tf(bnamd_nanovdb *g, const float *values, int numValues, float domainLo,
                                    float domainHi, float baseDensity) {
  return guarded([&] {
    BN_REQUIRE(g && values, "bnamd_nanovdb_set_tf: NULL argument");
    BN_REQUIRE(numValues >= 2, "bnamd_nanovdb_set_tf: need at least 2 values");
    BN_REQUIRE(domainHi > domainLo, "bnamd_nanovdb_set_tf: empty domain");
    Nvdb &G = g->g;
    G.tf.upload((const float4 *)values, numValues);
    G.dd.tf = TFDD{G.tf.ptr, numValues, domainLo, domainHi, baseDensity};
    G.haveTF = true;
    const int nmc = G.dd.mx * G.dd.my * G.dd.mz;
    k_majorants<<<div_up(nmc, 256), 256>>>(G.dd, G.ranges.ptr, G.majorants.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipDeviceSynchronize());
  });
}

extern "C" int bnamd_nanovdb_get_mcs(const bnamd_nanovdb *g, float *ranges, float *majorants) {
  return guarded([&] {
    BN_REQUIRE(g, "bnamd_nanovdb_get_mcs: NULL argument");
    const Nvdb &G = g->g;
    if (ranges) BN_HIP(hipMemcpy(ranges, G.ranges.ptr, G.ranges.bytes(), hipMemcpyDeviceToHost));
    if (majorants) BN_HIP(hipMemcpy(majorants, G.majorants.ptr, G.majorants.bytes(), hipMemcpyDeviceToHost));
  });
}

extern "C" int bnamd_nanovdb_sample_points(const bnamd_nanovdb *g, const float *points, float *out, int numPoints,
                                           void *stream) {
  return guarded([&] {
    BN_REQUIRE(g && points && out, "bnamd_nanovdb_sample_points: NULL argument");
    if (numPoints <= 0) return;
    hipStream_t s = (hipStream_t)stream;
    k_nvdb_sample<<<div_up(numPoints, 256), 256, 0, s>>>(g->g.dd, points, out, numPoints);
    BN_HIP(hipGetLastError());
    BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_nanovdb_render(const bnamd_nanovdb *g, const bnamd_camera *camera,
                                    const bnamd_render_params *params, float *accum, void *stream,
                                    bnamd_volume_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(g && camera && params && accum, "bnamd_nanovdb_render: NULL argument");
    BN_REQUIRE(g->g.haveTF, "bnamd_nanovdb_render: no transfer function set");
    BN_REQUIRE(params->width > 0 && params->height > 0 && params->pathsPerPixel >= 1,
               "bnamd_nanovdb_render: bad frame size or pathsPerPixel");
    hipStream_t s = (hipStream_t)stream;
    const CameraDD cam = make_camera(*camera);
    const RenderDD rp = make_render_dd(*params);
    unsigned long long h[4] = {0, 0, 0, 0};
    const float ms = render_on_stream(s, g->g.counters, 4, stats != nullptr, h, [&] {
      render_volume_dispatch(g->g.dd, cam, rp, params->pathsPerPixel, (float4 *)accum, g->g.counters.ptr, s);
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

extern "C" int bnamd_nanovdb_iso_rays(const bnamd_nanovdb *g, const bnamd_ray *rays, float isoValue, float *out,
                                      int numRays, void *stream) {
  return guarded([&] {
    BN_REQUIRE(g && rays && out, "bnamd_nanovdb_iso_rays: NULL argument");
    if (numRays <= 0) return;
    const NvdbDD &d = g->g.dd;
    const NvdbIsoGrid ig{d.worldLo, d.worldHi, d.mcOrigin, d.mcSpacing, d.mx, d.my, d.mz, g->g.ranges.ptr};
    hipStream_t s = (hipStream_t)stream;
    k_nvdb_iso_rays<<<div_up(numRays, 256), 256, 0, s>>>(d, ig, rays, isoValue, (float4 *)out, numRays);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}
