// Structured volume object: macrocell ranges (StructuredData_computeMCs,
// barney/volume/StructuredData.cu:30-61), majorants (mapMCs,
// barney/volume/MCGrid.cu:94-116) and the delta-tracking renderer with the
// SCI_VIS_MODE shade path (barney/kernels/shadeRays.cu:24-26, 787-790).
#include "shade.h"
#include "volume.h"

namespace bnamd {

struct Volume {
  VolumeDD dd;
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

__global__ void k_majorants(VolumeDD v, const float2 *__restrict ranges, float *__restrict maj) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= v.mx * v.my * v.mz) return;
  const float2 r = ranges[tid];
  maj[tid] = tf_majorant(v.tf, r.x, r.y);
}

__device__ __forceinline__ int xcd_remap_v(int bid, int nb) {
  const int per = (nb + 7) >> 3;
  const int t = (bid & 7) * per + (bid >> 3);
  return t < nb ? t : -1;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

// One thread per (pixel, path): the L = min(64, pow2 >= spp) paths of a pixel
// run in adjacent lanes (so a wavefront marches near-identical rays through
// the same macrocells and voxels), each lane looping over paths sub,
// sub+L, ...; the per-pixel sum is a butterfly over those L lanes.
template <int L>
__global__ void __launch_bounds__(256) k_render_volume(VolumeDD v, CameraDD cam, RenderDD rp, int spp,
                                                       float4 *__restrict accum,
                                                       unsigned long long *__restrict counters) {
  constexpr int PW = 64 / L;            // pixels per wavefront
  constexpr int PB = 4 * PW;            // pixels per workgroup
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  const int tilesX = (rp.width + TW - 1) / TW, tilesY = (rp.height + TH - 1) / TH;
  const int tile = xcd_remap_v(blockIdx.x, tilesX * tilesY);
  if (tile < 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane & (L - 1);
  const int q = wave * PW + lane / L;
  const int ix = (tile % tilesX) * TW + (q % TW);
  const int iy = (tile / tilesX) * TH + (q / TW);
  VolCounters cnt{0, 0};
  unsigned long long nrays = 0;
  float sr = 0.f, sg = 0.f, sb = 0.f, sa = 0.f;
  const bool active = ix < rp.width && iy < rp.height;
  if (active) {
    for (int p = sub; p < spp; p += L) {
      RenderDD rpp = rp;
      rpp.accumID = rp.accumID + p;
      PathRay ray = camera_ray(cam, rpp, ix, iy);
      const float clampMax = 10.f * (1 + rpp.accumID);
      vec3f sum = vec3f(0.f);
      float alpha = 0.f;
      for (int gen = 0; ray.tMax > 0.f; ++gen) {
        const vec3f dir = fix_dir(ray.dir);
        const VolHit vh = volume_intersect(v, ray.org, dir, 0.f, ray.tMax, ray.tMax, ray.rng, false, cnt);
        ++nrays;
        ShadeHit sh;
        sh.hit = vh.hit;
        if (vh.hit) {
          sh.P = ray.org + vh.t * dir;
          sh.N = vec3f(0.f);
          sh.t = vh.t;
          sh.bsdf = phase_bsdf(vh.rgb);
        }
        if (gen == 0) alpha = vh.hit ? 1.f : ray.missAlpha;
        PathRay shadow;
        vec3f frag = vmin(bounce(rpp, ray, sh, shadow, gen), clampMax);
        if (frag.x > 0.f) sum.x += frag.x;
        if (frag.y > 0.f) sum.y += frag.y;
        if (frag.z > 0.f) sum.z += frag.z;
        if (shadow.tMax > 0.f) {
          const vec3f sdir = fix_dir(shadow.dir);
          const VolHit sv =
              volume_intersect(v, shadow.org, sdir, 0.f, shadow.tMax, shadow.tMax, shadow.rng, false, cnt);
          ++nrays;
          ShadeHit shs;
          shs.hit = sv.hit;
          PathRay dummy;
          frag = vmin(bounce(rpp, shadow, shs, dummy, gen + 1), clampMax);
          if (frag.x > 0.f) sum.x += frag.x;
          if (frag.y > 0.f) sum.y += frag.y;
          if (frag.z > 0.f) sum.z += frag.z;
        }
      }
      sr += sum.x;
      sg += sum.y;
      sb += sum.z;
      sa += alpha;
    }
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off);
    sg += __shfl_xor(sg, off);
    sb += __shfl_xor(sb, off);
    sa += __shfl_xor(sa, off);
  }
  if (active && sub == 0) {
    const int pix = ix + iy * rp.width;
    if (rp.accumID == 0) {
      accum[pix] = make_float4(sr, sg, sb, sa);
    } else {
      float4 o = accum[pix];
      o.x += sr; o.y += sg; o.z += sb; o.w += sa;
      accum[pix] = o;
    }
  }
  wave_add_u64(cnt.samples, counters + 0);
  wave_add_u64(cnt.ddaSteps, counters + 1);
  wave_add_u64(nrays, counters + 2);
}

template <int L>
void launch_render_volume(const VolumeDD &v, const CameraDD &cam, const RenderDD &rp, int spp, float4 *accum,
                          unsigned long long *counters, hipStream_t s) {
  constexpr int PB = 4 * (64 / L);
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  const int tiles = div_up(rp.width, TW) * div_up(rp.height, TH);
  // xcd_remap_v hands XCD x the tiles [x*per, (x+1)*per): the grid must
  // hold 8*per workgroups (a multiple of 8) to cover every tile
  const int grid = ((tiles + 7) / 8) * 8;
  k_render_volume<L><<<grid, 256, 0, s>>>(v, cam, rp, spp, accum, counters);
}

void render_volume_dispatch(const VolumeDD &v, const CameraDD &cam, const RenderDD &rp, int spp, float4 *accum,
                            unsigned long long *counters, hipStream_t s) {
  // Lanes per pixel: a few lanes per pixel, each looping over several of
  // its paths, balance coherence (near-identical rays in a wavefront)
  // against per-lane work variance; 4 measured best for 16 spp.
  // BNAMD_VOLUME_LANES overrides (1/2/4/8) for tuning.
  int lanes = 4;
  if (const char *e = getenv("BNAMD_VOLUME_LANES")) lanes = atoi(e);
  if (lanes <= 1) { launch_render_volume<1>(v, cam, rp, spp, accum, counters, s); return; }
  if (lanes == 2) { launch_render_volume<2>(v, cam, rp, spp, accum, counters, s); return; }
  if (lanes == 4) { launch_render_volume<4>(v, cam, rp, spp, accum, counters, s); return; }
  if (lanes == 8) { launch_render_volume<8>(v, cam, rp, spp, accum, counters, s); return; }
  if (spp >= 64) launch_render_volume<64>(v, cam, rp, spp, accum, counters, s);
  else if (spp > 16) launch_render_volume<32>(v, cam, rp, spp, accum, counters, s);
  else if (spp > 8) launch_render_volume<16>(v, cam, rp, spp, accum, counters, s);
  else if (spp > 4) launch_render_volume<8>(v, cam, rp, spp, accum, counters, s);
  else if (spp > 2) launch_render_volume<4>(v, cam, rp, spp, accum, counters, s);
  else if (spp > 1) launch_render_volume<2>(v, cam, rp, spp, accum, counters, s);
  else launch_render_volume<1>(v, cam, rp, spp, accum, counters, s);
}

__global__ void k_volume_rays(VolumeDD v, const bnamd_ray *__restrict rays, const uint64_t *__restrict seeds,
                              float4 *__restrict out, int n, int exactLog) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bnamd_ray r = rays[i];
  RNGSeed seed;
  seed.state = seeds[i];
  VolCounters cnt{0, 0};
  const vec3f dir = fix_dir({r.dir[0], r.dir[1], r.dir[2]});
  const VolHit h = volume_intersect(v, {r.org[0], r.org[1], r.org[2]}, dir, r.tmin, r.tmax, r.tmax, seed,
                                    exactLog != 0, cnt);
  out[i] = h.hit ? make_float4(h.t, h.rgb.x, h.rgb.y, h.rgb.z) : make_float4(BN_INF, 0.f, 0.f, 0.f);
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

struct bnamd_volume_s {
  Volume v;
};

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
    DevBuf<unsigned long long> counters(3);
    BN_HIP(hipMemsetAsync(counters.ptr, 0, counters.bytes(), s));
    const CameraDD cam = make_camera(*camera);
    const RenderDD rp = make_render_dd(*params);
    hipEvent_t e0, e1;
    BN_HIP(hipEventCreate(&e0));
    BN_HIP(hipEventCreate(&e1));
    BN_HIP(hipEventRecord(e0, s));
    render_volume_dispatch(vol->v.dd, cam, rp, params->pathsPerPixel, (float4 *)accum, counters.ptr, s);
    BN_HIP(hipGetLastError());
    BN_HIP(hipEventRecord(e1, s));
    BN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    BN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    unsigned long long h[3];
    BN_HIP(hipMemcpy(h, counters.ptr, sizeof(h), hipMemcpyDeviceToHost));
    if (stats) {
      stats->samples = (long long)h[0];
      stats->ddaSteps = (long long)h[1];
      stats->raysTraced = (long long)h[2];
      stats->ms = ms;
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
