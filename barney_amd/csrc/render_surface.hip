// Plain per-pixel megakernel: for each pixel and path, generate the camera
// ray, then trace / shade / trace-shadow until the path ends, accumulating
// into the pixel. This fuses barney's generateRays -> {traceRays, shadeRays}*
// generation loop (barney/Context.cpp:142-185) into one thread per pixel:
// each ray of a path only depends on its own RNG state, so per-pixel
// depth-first order computes the same values as the queue-based order.
// With a structured volume (a model holding geometry and a volume), each
// ray's closest surface hit is found first and the volume is delta-tracked
// up to it (MCVolumeAccel::isProg with the surface hit as tMax,
// barney/volume/MCAccelerator.h:569-708): a collision before the surface is
// a volume hit; shadow rays no surface blocks are tracked through the volume.
#include "shade.h"
#include "volume.h"

namespace bnamd {

size_t trace_lds_bytes(int depth, int block);
const VolumeDD &volume_dd(const bnamd_volume *vol, const char *who);

namespace {

// Pixel tiles of 8x8 per wavefront, 4 wavefronts per workgroup (16x16).
// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD
// receives a contiguous band of tiles and keeps its L2 working set local.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int per = (nb + 7) >> 3;
  const int xcd = bid & 7, local = bid >> 3;
  const int t = xcd * per + local;
  return t < nb ? t : -1;
}

template <bool TL>
__device__ __forceinline__ ShadeHit shade_surface(const WorldDev &w, vec3f org, vec3f dir, float tmax,
                                                  const HitRec &h) {
  ShadeHit sh;
  if (h.flat == 0xffffffffu) {
    sh.hit = false;
    return sh;
  }
  const SurfaceHit s = surface_epilogue<TL>(w, org, dir, 0.f, tmax, h);
  const GeomDev &g = *s.g;
  const MaterialDev &m = w.materials[g.materialID];
  vec3f color = {m.color[0], m.color[1], m.color[2]};
  if (g.colors) color = load3(g.colors + 3 * s.primID);
  sh.hit = true;
  sh.P = s.P;
  sh.N = s.Ng;
  sh.t = h.t;
  sh.bsdf = material_bsdf(m, color);
  return sh;
}

// shadeRays.cu:1030-1037 adds only positive channels of a clamped fragment.
__device__ __forceinline__ void add_fragment(vec3f &sum, vec3f frag, float clampMax) {
  frag = vmin(frag, clampMax);
  if (frag.x > 0.f) sum.x += frag.x;
  if (frag.y > 0.f) sum.y += frag.y;
  if (frag.z > 0.f) sum.z += frag.z;
}

// kLights: an HDRI map or light lists light the scene; kHit: a cut plane or cutout
// triangles, so candidate hits pass the anyHit filter (kernels without
// either keep the plain traversal); kVol: 'vol' is in the scene.
template <bool TL, bool kLights, bool kHit, bool kVol>
__global__ void __launch_bounds__(256) k_render_surface(WorldDev w, CameraDD cam, RenderDD rp, int spp,
                                                        float4 *__restrict accum, int4 *__restrict ids,
                                                        unsigned long long *__restrict counters, bool cutouts,
                                                        VolumeDD vol) {
  extern __shared__ uint2 lds_stack[];
  const int tilesX = (rp.width + 15) >> 4, tilesY = (rp.height + 15) >> 4;
  const int tile = xcd_remap(blockIdx.x, tilesX * tilesY);
  if (tile < 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ix = (tile % tilesX) * 16 + (wave & 1) * 8 + (lane & 7);
  const int iy = (tile / tilesX) * 16 + (wave >> 1) * 8 + (lane >> 3);
  const bool active = ix < rp.width && iy < rp.height;
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, w.depth + 1};
  unsigned long long nrays = 0;
  TraceCount tc;
  VolCounters vc{0, 0, 0};
  if (active) {
    const int pix = ix + iy * rp.width;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int accum0 = rp.accumID;
    for (int p = 0; p < spp; ++p) {
      RenderDD rpp = rp;
      rpp.accumID = accum0 + p;
      PathRay ray = camera_ray(cam, rpp, ix, iy);
      const float clampMax = 10.f * (1 + rpp.accumID);
      vec3f sum = vec3f(0.f);
      float alpha = 0.f;
      // a primary ray clipped away by the cut plane (tMax = 0) is still
      // traced and shows the background
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const vec3f dir = fix_dir(ray.dir);
        HitRec h;
        if (kHit)
          world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, h, ctx, &tc,
                                    hit_filter(rpp, ray.org, dir, ray.rng.state));
        else
          world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, h, ctx, &tc);
        ++nrays;
        ShadeHit sh = shade_surface<TL>(w, ray.org, dir, ray.tMax, h);
        if (kVol) {
          const float tS = sh.hit ? sh.t : ray.tMax;
          const VolHit vh = volume_intersect(vol, ray.org, dir, 0.f, tS, tS, ray.rng, false, vc);
          if (vh.hit) {
            sh.hit = true;
            sh.P = ray.org + vh.t * dir;
            sh.N = vec3f(0.f);
            sh.t = vh.t;
            sh.bsdf = phase_bsdf(vh.rgb);
          }
        }
        if (gen == 0) {
          alpha = sh.hit ? 1.f : ray.missAlpha;
          if (ids && p == 0 && accum0 == 0) {
            // the ID channel records the closest candidate whatever its
            // opacity (Triangles.dev.cu:45-60 writes it before the test)
            HitRec hid = h;
            if (kHit && cutouts)
              world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, hid, ctx, nullptr,
                                        hit_filter(rpp, ray.org, dir, ray.rng.state, false));
            ids[pix] = hit_ids<TL>(w, ray.org, dir, ray.tMax, hid);
          }
        }
        PathRay shadow;
        vec3f frag = bounce<kLights>(rpp, ray, sh, shadow, gen);
        add_fragment(sum, frag, clampMax);
        if (shadow.tMax > 0.f) {
          const vec3f sdir = fix_dir(shadow.dir);
          ShadeHit shs;
          HitRec hs;
          if (kHit)
            shs.hit = world_traverse<TL, true>(w, shadow.org, sdir, 0.f, shadow.tMax, hs, ctx, &tc,
                                               hit_filter(rpp, shadow.org, sdir, shadow.rng.state));
          else
            shs.hit = world_traverse<TL, true>(w, shadow.org, sdir, 0.f, shadow.tMax, hs, ctx, &tc);
          if (kVol && !shs.hit)
            shs.hit = volume_intersect(vol, shadow.org, sdir, 0.f, shadow.tMax, shadow.tMax, shadow.rng, false, vc).hit;
          ++nrays;
          PathRay dummy;
          frag = bounce<kLights>(rpp, shadow, shs, dummy, gen + 1);
          add_fragment(sum, frag, clampMax);
        }
      }
      acc.x += sum.x;
      acc.y += sum.y;
      acc.z += sum.z;
      acc.w += alpha;
    }
    if (accum0 == 0) {
      accum[pix] = acc;
    } else {
      float4 o = accum[pix];
      o.x += acc.x; o.y += acc.y; o.z += acc.z; o.w += acc.w;
      accum[pix] = o;
    }
  }
  // one atomic per wavefront per counter
  unsigned long long nv = tc.nodes, pb = tc.primBytes;
  for (int off = 32; off > 0; off >>= 1) {
    nrays += __shfl_xor(nrays, off);
    nv += __shfl_xor(nv, off);
    pb += __shfl_xor(pb, off);
  }
  unsigned long long ns = vc.samples;
  if (kVol)
    for (int off = 32; off > 0; off >>= 1) ns += __shfl_xor(ns, off);
  if (lane == 0) {
    atomicAdd(counters + 0, nrays);
    atomicAdd(counters + 1, nv);
    atomicAdd(counters + 2, pb);
    if (kVol) atomicAdd(counters + 3, ns);
  }
}

}  // namespace

const World &world_of(const bnamd_world *w);

}  // namespace bnamd

using namespace bnamd;

namespace {

void render_surface_impl(const char *who, const bnamd_world *world, const bnamd_volume *volume,
                         const bnamd_camera *camera, const bnamd_render_params *params, float *accum, int *ids,
                         void *stream, bnamd_render_stats *stats, long long *volumeSamples) {
  BN_REQUIRE(world && camera && params && accum, std::string(who) + ": NULL argument");
  BN_REQUIRE(params->width > 0 && params->height > 0, std::string(who) + ": empty frame");
  BN_REQUIRE(params->pathsPerPixel >= 1, std::string(who) + ": pathsPerPixel must be >= 1");
  const World &w = world_of(world);
  const bool withVol = volume != nullptr;
  const VolumeDD vol = withVol ? volume_dd(volume, who) : VolumeDD{};
  const CameraDD cam = make_camera(*camera);
  const RenderDD rp = make_render_dd(*params);
  hipStream_t s = (hipStream_t)stream;
  const int tiles = div_up(params->width, 16) * div_up(params->height, 16);
  // xcd_remap hands XCD x the tiles [x*per, (x+1)*per): the grid must hold
  // 8*per workgroups (a multiple of 8) to cover every tile
  const int grid = ((tiles + 7) / 8) * 8;
  const int block = 256;
  const size_t lds = trace_lds_bytes(w.depth, block);
  BN_REQUIRE(lds <= 64 * 1024, std::string(who) + ": BVH too deep for the LDS stack");
  unsigned long long n[4] = {0, 0, 0, 0};
  const int numCounters = withVol ? 4 : 3;
  const bool wantStats = stats != nullptr || volumeSamples != nullptr;
  const float ms = render_on_stream(s, w.counters, numCounters, wantStats, n, [&] {
    unsigned long long *c = w.counters.ptr;
    auto go = [&](auto kernel) {
      kernel<<<grid, block, lds, s>>>(w.dd(), cam, rp, params->pathsPerPixel, (float4 *)accum, (int4 *)ids, c,
                                      w.hasCutouts, vol);
    };
    const int variant = (w.twoLevel ? 4 : 0) | (has_lights(rp) ? 2 : 0) | ((rp.cutOn || w.hasCutouts) ? 1 : 0);
    if (withVol) {
      switch (variant) {
        case 0: go(k_render_surface<false, false, false, true>); break;
        case 1: go(k_render_surface<false, false, true, true>); break;
        case 2: go(k_render_surface<false, true, false, true>); break;
        case 3: go(k_render_surface<false, true, true, true>); break;
        case 4: go(k_render_surface<true, false, false, true>); break;
        case 5: go(k_render_surface<true, false, true, true>); break;
        case 6: go(k_render_surface<true, true, false, true>); break;
        default: go(k_render_surface<true, true, true, true>); break;
      }
      return;
    }
    switch (variant) {
      case 0: go(k_render_surface<false, false, false, false>); break;
      case 1: go(k_render_surface<false, false, true, false>); break;
      case 2: go(k_render_surface<false, true, false, false>); break;
      case 3: go(k_render_surface<false, true, true, false>); break;
      case 4: go(k_render_surface<true, false, false, false>); break;
      case 5: go(k_render_surface<true, false, true, false>); break;
      case 6: go(k_render_surface<true, true, false, false>); break;
      default: go(k_render_surface<true, true, true, false>); break;
    }
  });
  if (stats) {
    stats->raysTraced = (long long)n[0];
    stats->nodeVisits = (long long)n[1];
    stats->primBytes = (long long)n[2];
    stats->primaryRays = (long long)params->width * params->height * params->pathsPerPixel;
    stats->ms = ms;
  }
  if (volumeSamples) *volumeSamples = (long long)n[3];
}

}  // namespace

extern "C" int bnamd_render_surface(const bnamd_world *world, const bnamd_camera *camera,
                                    const bnamd_render_params *params, float *accum, int *ids, void *stream,
                                    bnamd_render_stats *stats) {
  return guarded([&] {
    render_surface_impl("bnamd_render_surface", world, nullptr, camera, params, accum, ids, stream, stats, nullptr);
  });
}

extern "C" int bnamd_render_mixed(const bnamd_world *world, const bnamd_volume *volume, const bnamd_camera *camera,
                                  const bnamd_render_params *params, float *accum, int *ids, void *stream,
                                  bnamd_render_stats *stats, long long *volumeSamples) {
  return guarded([&] {
    BN_REQUIRE(volume, "bnamd_render_mixed: NULL volume");
    render_surface_impl("bnamd_render_mixed", world, volume, camera, params, accum, ids, stream, stats,
                        volumeSamples);
  });
}
