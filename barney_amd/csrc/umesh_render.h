// The unstructured-mesh render kernel (k_render_umesh) and its launch, shared
// by umesh.hip (the plain variants) and umesh_count.hip (the instrumented
// variants that count point location's work), so that the two halves of its
// instantiations compile in parallel.
#pragma once
#include "shade.h"
#include "umesh.h"

namespace bnamd {

__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int per = (nb + 7) >> 3;
  const int t = (bid & 7) * per + (bid >> 3);
  return t < nb ? t : -1;
}

__device__ __forceinline__ void wave_sum_u64(unsigned long long v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

struct UMeshOpts {
  int volume, iso;
  float isoValue;
  vec3f isoColor;
};

// Closest of the volume (geometry 0) and iso-surface (geometry 1) hits for
// one ray; each intersection program advances the ray's seed once.
template <bool kVol, bool kIso, bool kMixed, bool kCount>
BN_D ShadeHit umesh_trace(const UMeshDD &u, const UMeshOpts &o, vec3f org, vec3f dir, float tMax,
                                RNGSeed &seed, const TraverseCtx &ctx, VolCounters &cnt, uint32_t &hint,
                                TravCount &tc) {
  ShadeHit sh;
  sh.hit = false;
  float tClosest = tMax;
  if (kVol) {
    const VolHit vh = volume_intersect_g(
        u, [&](vec3f P) { return umesh_sample_and_map<kMixed, kCount>(u, P, ctx, hint, &tc); }, org, dir, 0.f, tMax,
        tMax, seed, false, cnt, bn_hash3(0u, 0u, 0u));
    if (vh.hit) {
      sh.hit = true;
      sh.P = org + vh.t * dir;
      sh.N = vec3f(0.f);
      sh.t = vh.t;
      sh.bsdf = phase_bsdf(vh.rgb);
      tClosest = vh.t;
    }
  }
  if (kIso) {
    const IsoHit ih = iso_intersect(
        u,
        [&](vec3f P) {
          cnt.samples++;
          return umesh_locate<kMixed, kCount>(u, P, ctx, hint, &tc);
        },
        o.isoValue, org, dir, 0.f, tMax, tClosest, seed, bn_hash3(0u, 1u, 0u));
    if (ih.hit) {
      sh.hit = true;
      sh.P = ih.P;
      sh.N = ih.N;
      sh.t = ih.t;
      sh.bsdf = matte_bsdf(o.isoColor);
    }
  }
  return sh;
}

// L lanes per pixel, each looping over its share of the pixel's paths.
// The kernel waits on memory (point-location loads, 19 % VALU issue), so
// it is built for six wavefronts per SIMD (<= 80 VGPRs, ~300 bytes of
// spill per lane): 1024x1024x8 spp with volume + iso-surface takes 147 ms
// at six wavefronts and 8 lanes per pixel, against 232 ms at the three
// that the unconstrained allocation (168 VGPRs) leaves with 4 lanes; five
// and eight wavefronts measured 156 and 159 ms.
// kCount: the instrumented variant that also counts point location's node
// visits, element tests and their bytes (counters 4, 5, 6); it renders the
// same image with the same paths, so one counted frame gives the
// uncounted kernel's per-frame figures.
template <int L, bool kVol, bool kIso, bool kMixed, bool kLights, bool kCount = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_render_umesh(UMeshDD u, UMeshOpts o, CameraDD cam, RenderDD rp, int spp, float4 *__restrict accum,
               unsigned long long *__restrict counters) {
  extern __shared__ uint2 lds_stack[];
  constexpr int PW = 64 / L;
  constexpr int PB = 4 * PW;
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  const int tilesX = (rp.width + TW - 1) / TW, tilesY = (rp.height + TH - 1) / TH;
  const int tile = xcd_tile(blockIdx.x, tilesX * tilesY);
  if (tile < 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane & (L - 1);
  const int q = wave * PW + lane / L;
  const int ix = (tile % tilesX) * TW + (q % TW);
  const int iy = (tile / tilesX) * TH + (q / TW);
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, u.depth + 1};
  VolCounters cnt{0, 0, 0};
  TravCount tc{0, 0, 0};
  unsigned long long nrays = 0;
  uint32_t hint = ~0u;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const bool active = ix < rp.width && iy < rp.height;
  for (int r = 0; r * L < spp; ++r) {
    const int p = r * L + sub;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active && p < spp) {
      RenderDD rpp = rp;
      rpp.accumID = rp.accumID + p;
      PathRay ray = camera_ray(cam, rpp, ix, iy);
      const float clampMax = 10.f * (1 + rpp.accumID);
      vec3f sum = vec3f(0.f);
      float alpha = 0.f;
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const vec3f dir = fix_dir(ray.dir);
        const ShadeHit sh =
            umesh_trace<kVol, kIso, kMixed, kCount>(u, o, ray.org, dir, ray.tMax, ray.rng, ctx, cnt, hint, tc);
        ++nrays;
        if (gen == 0) alpha = sh.hit ? 1.f : ray.missAlpha;
        PathRay shadow;
        vec3f frag = vmin(bounce<kLights>(rpp, ray, sh, shadow, gen), clampMax);
        if (frag.x > 0.f) sum.x += frag.x;
        if (frag.y > 0.f) sum.y += frag.y;
        if (frag.z > 0.f) sum.z += frag.z;
        if (shadow.tMax > 0.f) {
          const ShadeHit shs =
              umesh_trace<kVol, kIso, kMixed, kCount>(u, o, shadow.org, fix_dir(shadow.dir), shadow.tMax, shadow.rng,
                                                      ctx, cnt, hint, tc);
          ++nrays;
          PathRay dummy;
          frag = vmin(bounce<kLights>(rpp, shadow, shs, dummy, gen + 1), clampMax);
          if (frag.x > 0.f) sum.x += frag.x;
          if (frag.y > 0.f) sum.y += frag.y;
          if (frag.z > 0.f) sum.z += frag.z;
        }
      }
      v = make_float4(sum.x, sum.y, sum.z, alpha);
    }
    add_round_in_path_order<L>(acc, v, min(L, spp - r * L));
  }
  if (active && sub == 0) {
    const int pix = ix + iy * rp.width;
    if (rp.accumID == 0) {
      accum[pix] = acc;
    } else {
      float4 a = accum[pix];
      a.x += acc.x; a.y += acc.y; a.z += acc.z; a.w += acc.w;
      accum[pix] = a;
    }
  }
  wave_sum_u64(cnt.samples, counters + 0);
  wave_sum_u64(cnt.ddaSteps, counters + 1);
  wave_sum_u64(cnt.zeroSteps, counters + 3);
  wave_sum_u64(nrays, counters + 2);
  if (kCount) {
    wave_sum_u64(tc.nodes, counters + 4);
    wave_sum_u64(tc.tests, counters + 5);
    wave_sum_u64(tc.bytes, counters + 6);
  }
}

template <int L, bool kMixed = false, bool kCount = false>
void launch_render_umesh_t(const UMeshDD &u, const UMeshOpts &o, const CameraDD &cam, const RenderDD &rp, int spp,
                           float4 *accum, unsigned long long *counters, size_t lds, hipStream_t s) {
  constexpr int PB = 4 * (64 / L);
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  const int tiles = div_up(rp.width, TW) * div_up(rp.height, TH);
  const int grid = ((tiles + 7) / 8) * 8;
  auto go = [&](auto kernel) { kernel<<<grid, 256, lds, s>>>(u, o, cam, rp, spp, accum, counters); };
  const bool env = has_lights(rp);
  if (o.volume && o.iso)
    env ? go(k_render_umesh<L, true, true, kMixed, true, kCount>) : go(k_render_umesh<L, true, true, kMixed, false, kCount>);
  else if (o.volume)
    env ? go(k_render_umesh<L, true, false, kMixed, true, kCount>)
        : go(k_render_umesh<L, true, false, kMixed, false, kCount>);
  else if (o.iso)
    env ? go(k_render_umesh<L, false, true, kMixed, true, kCount>)
        : go(k_render_umesh<L, false, true, kMixed, false, kCount>);
  else
    env ? go(k_render_umesh<L, false, false, kMixed, true, kCount>)
        : go(k_render_umesh<L, false, false, kMixed, false, kCount>);
}

}  // namespace bnamd
