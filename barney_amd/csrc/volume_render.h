// The volume path tracer over any scalar field with barney's macrocell grid
// and transfer function (MCVolumeAccel::isProg, MCAccelerator.h:569-708, in
// generate / trace / shade with the SCI_VIS_MODE bounce), the macrocell
// majorant kernel (mapMCs, barney/volume/MCGrid.cu:94-116), and their
// launchers. F supplies the fields DeltaTrack reads (worldLo/Hi, majorants,
// mx/my/mz, mcOrigin, mcSpacing, tf) and sample_and_map(F, P).
#pragma once
#include "shade.h"
#include "volume.h"

namespace bnamd {
namespace {

template <class F>
__global__ void k_majorants(F v, const float2 *__restrict ranges, float *__restrict maj) {
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

// Per-lane stage of the volume path tracer (k_render_volume).
enum VolStage : int {
  VS_DEAD = 0,   // the wavefront's pool of paths is used up
  VS_GEN = 1,    // needs a path from the pool (and its camera ray)
  VS_INIT = 2,   // a ray waits in the walk's org/dir/rayTMax for set-up
  VS_MARCH = 3,  // the walk is running (or has ended: see DeltaTrack::mode)
};

// Paths a workgroup owns at a time (one LDS result slot each).
constexpr int kPoolPaths = 2048;
// Weights of the waiting-lane counts when a wavefront picks its next kind of
// work: cheap steps, field sample, camera ray and walk set-up, collision
// shading, escape shading. A cheap-step iteration runs up to kCheapSteps
// steps per lane, so it is worth running for fewer lanes than the others.
constexpr int kPhaseWeight[5] = {1, 3, 3, 3, 3};
constexpr int kCheapSteps = 12;
// Wavefronts per SIMD the kernel is built for (a register budget of 128).
constexpr int kVolumeWaves = 4;

__device__ __forceinline__ void wave_add_u64(unsigned long long v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(dst, v);
}

// Each workgroup owns PB pixels (a tile) and the PB * spp paths of those
// pixels as a pool (in units of at most kPoolPaths paths when spp is large):
// its four wavefronts take consecutive paths from the pool through an LDS
// counter whenever lanes' previous paths have ended, so lanes whose rays leave
// the volume at once keep working while others walk long macrocell paths,
// and no wavefront idles while its neighbours still have paths. A path's sum
// (rgb, alpha) goes to an LDS slot; when the pool is used up, each pixel adds
// its paths in path order, as barney accumulates them.
//
// A path is barney's generate / trace / shade loop for a volume-only scene
// with SCI_VIS_MODE shading: the camera ray's walk either escapes (the pixel
// shows the background) or collides; at a collision bounce() (NEE toward
// the environment light with the phase function) spawns one shadow ray and
// ends the path ray, since scattering in a volume ends a path there
// (shadeRays.cu:787-790; the phase function's pdf, 1/(4 pi), never fails
// bounce()'s pdf test, so that branch is always reached). The path's value is
// then the shadow ray's weight if its walk escapes, else zero. Its work comes
// in small kinds: camera ray with walk set-up (DeltaTrack::init), cheap walk
// steps (macrocell DDA step or collision distance), field samples at
// collisions, shading of a collision, and shading of an escaped ray. Every
// wavefront iteration runs ONE kind, the one most lanes are waiting for
// (weighted by kPhaseWeight), for all of them; cheap steps, most of a path's
// work, run several to an iteration (kCheapSteps).
//
// A lane carries only the walk and {path index, stage, shadow weight}: the
// camera ray is generated again (same arithmetic) where shading needs it
// instead of being kept across the walk, which keeps the register count at
// the four wavefronts per SIMD that this issue-bound kernel needs.
//
// F is the field: VolumeDD (structured data) or AmrDD (block-structured
// AMR); both carry the macrocell grid and transfer function, and
// sample_and_map(F, P) gives the mapped sample.
template <int PB, class F, bool kLights>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kVolumeWaves)))
k_render_volume(F v, CameraDD cam, RenderDD rp, int spp, float4 *__restrict accum,
                unsigned long long *__restrict counters) {
  constexpr int TW = (PB >= 1024) ? 32 : (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  static_assert(PB >= 1 && PB <= kPoolPaths && PB <= 1024, "a pool holds at least one path per pixel");
  constexpr int kUnit = kPoolPaths / PB;  // paths per pixel in one pool
  __shared__ float4 slots[kPoolPaths];
  __shared__ int poolNext;
  const int tilesX = (rp.width + TW - 1) / TW, tilesY = (rp.height + TH - 1) / TH;
  const int tile = xcd_remap_v(blockIdx.x, tilesX * tilesY);
  if (tile < 0) return;
  const int lane = threadIdx.x & 63;
  const int tx0 = (tile % tilesX) * TW, ty0 = (tile / tilesX) * TH;
  const unsigned long long lanesBelow = (1ull << lane) - 1ull;
  VolCounters cnt{0, 0, 0};
  unsigned nrays = 0;
  const auto sampler = [&](vec3f P) { return sample_and_map(v, P); };
  const uint64_t geomHash = bn_hash3(0u, 0u, 0u);
  float4 pixSum = make_float4(0.f, 0.f, 0.f, 0.f);  // thread j's pixel, across pools (PB <= 256)

  for (int s0 = 0; s0 < spp; s0 += kUnit) {
    const int nPaths = PB * min(kUnit, spp - s0);
    if (threadIdx.x == 0) poolNext = 0;
    __syncthreads();
    int k = 0;                      // this lane's path: pixel k % PB, sample s0 + k / PB
    int stage = VS_GEN;
    bool inShadow = false;          // the walk holds the path's shadow ray
    vec3f shadowW = vec3f(0.f);     // the shadow ray's contribution if it escapes
    DeltaTrack tr;
    tr.mode = DeltaTrack::DONE;
    // the path's pixel and camera ray (generateRays), computed where needed
    auto pixelOf = [&](int kk, int &ix, int &iy) {
      const int j = kk % PB;
      ix = tx0 + j % TW;
      iy = ty0 + j / TW;
    };
    auto pathParams = [&](int kk) {
      RenderDD r = rp;
      r.accumID = rp.accumID + s0 + kk / PB;
      return r;
    };

    for (;;) {
      const bool march = stage == VS_MARCH;
      const bool wCheap = march && tr.mode != DeltaTrack::DONE && tr.mode != DeltaTrack::SAMPLE;
      const bool wSample = march && tr.mode == DeltaTrack::SAMPLE;
      const bool wHit = march && tr.mode == DeltaTrack::DONE && tr.hit && !inShadow;
      const bool wMiss = march && tr.mode == DeltaTrack::DONE && !(tr.hit && !inShadow);
      const bool wGen = stage == VS_GEN, wInit = stage == VS_INIT;
      const unsigned long long genMask = __ballot(wGen);
      if (__ballot(stage != VS_DEAD) == 0ull) break;
      // the kind of work most lanes wait for (counts from ballots are
      // wavefront-uniform, so the branch below is too)
      int best = 0, most = (int)__popcll(__ballot(wCheap)) * kPhaseWeight[0];
      const int waiting[4] = {(int)__popcll(__ballot(wSample)) * kPhaseWeight[1],
                              (int)(__popcll(genMask) + __popcll(__ballot(wInit))) * kPhaseWeight[2],
                              (int)__popcll(__ballot(wHit)) * kPhaseWeight[3],
                              (int)__popcll(__ballot(wMiss)) * kPhaseWeight[4]};
#pragma unroll
      for (int w = 0; w < 4; ++w)
        if (waiting[w] > most) {
          most = waiting[w];
          best = w + 1;
        }
      if (best == 0) {
        if (wCheap) {  // steps until the walk needs the field or ends
          int i = 0;
          do {
            tr.advance<false>(v, cnt);
          } while (++i < kCheapSteps && tr.mode != DeltaTrack::DONE && tr.mode != DeltaTrack::SAMPLE);
        }
      } else if (best == 1) {
        if (wSample) tr.sample<false>(sampler, cnt);
      } else if (best == 2) {
        // the workgroup's pool hands out consecutive paths to this wavefront
        int base = 0;
        if (genMask) {
          const int leader = __ffsll((long long)genMask) - 1;
          if (lane == leader) base = atomicAdd(&poolNext, (int)__popcll(genMask));
          base = __shfl(base, leader);
        }
        RNGSeed seed;
        if (wGen) {  // take the next path of the pool; generateRays for it
          k = base + (int)__popcll(genMask & lanesBelow);
          int ix, iy;
          pixelOf(k, ix, iy);
          if (k >= nPaths) {
            stage = VS_DEAD;
          } else if (ix >= rp.width || iy >= rp.height) {
            slots[k] = make_float4(0.f, 0.f, 0.f, 0.f);  // outside the frame: take another path
          } else {
            const PathRay ray = camera_ray(cam, pathParams(k), ix, iy);
            seed = ray.rng;
            inShadow = false;
            tr.org = ray.org;
            tr.dir = fix_dir(ray.dir);
            tr.rayTMax = ray.tMax;
            stage = VS_INIT;
          }
        }
        // the walks of the new camera rays and of the shadow rays shading
        // left (whose seed waits in tr.rng)
        if (stage == VS_INIT) {
          if (inShadow) seed.state = tr.rng.state;
          tr.init(v, tr.org, tr.dir, 0.f, tr.rayTMax, tr.rayTMax, seed, geomHash);
          stage = VS_MARCH;
        }
      } else if (best == 3) {
        if (wHit) {  // a collision of the camera ray: bounce() spawns the shadow ray
          ++nrays;
          int ix, iy;
          pixelOf(k, ix, iy);
          const RenderDD rpp = pathParams(k);
          PathRay ray = camera_ray(cam, rpp, ix, iy);
          ray.rng.next();  // as the walk's set-up advanced it (Random2(seed, hash))
          ShadeHit sh;
          sh.hit = true;
          sh.t = tr.t;
          sh.P = tr.org + tr.t * tr.dir;
          sh.N = vec3f(0.f);
          sh.bsdf = phase_bsdf(tr.rgb);
          PathRay sray;
          (void)bounce<kLights>(rpp, ray, sh, sray, 0);  // no fragment at a volume collision
          if (sray.tMax > 0.f) {
            shadowW = sray.mis * sray.tp;  // bounce() of the shadow ray if unoccluded
            tr.org = sray.org;
            tr.dir = fix_dir(sray.dir);
            tr.rayTMax = sray.tMax;
            tr.rng.state = sray.rng.state;
            inShadow = true;
            stage = VS_INIT;
          } else {
            slots[k] = make_float4(0.f, 0.f, 0.f, 1.f);
            stage = VS_GEN;
          }
        }
      } else {
        if (wMiss) {  // an escaped camera ray, or a finished shadow ray
          ++nrays;
          int ix, iy;
          pixelOf(k, ix, iy);
          const RenderDD rpp = pathParams(k);
          vec3f frag;
          float alpha = 1.f;
          if (inShadow) {
            frag = tr.hit ? vec3f(0.f) : shadowW;
          } else {  // bounce() of a primary miss: the background, or the map
            miss_color(rpp, iy, frag, alpha);
            if (kLights && has_map(rpp)) frag = env_radiance(rpp.env, camera_ray(cam, rpp, ix, iy).dir);
          }
          frag = vmin(frag, 10.f * (1 + rpp.accumID));
          slots[k] = make_float4(frag.x > 0.f ? frag.x : 0.f, frag.y > 0.f ? frag.y : 0.f,
                                 frag.z > 0.f ? frag.z : 0.f, alpha);
          stage = VS_GEN;
        }
      }
    }
    // the workgroup's paths are done: each pixel adds its paths in path order
    __syncthreads();
    const int ns = nPaths / PB;
    for (int j = threadIdx.x; j < PB; j += 256) {
      float4 a = (s0 == 0) ? make_float4(0.f, 0.f, 0.f, 0.f) : pixSum;
      for (int si = 0; si < ns; ++si) {
        const float4 r = slots[si * PB + j];
        a.x += r.x;
        a.y += r.y;
        a.z += r.z;
        a.w += r.w;
      }
      pixSum = a;
      const int ix = tx0 + j % TW, iy = ty0 + j / TW;
      if (s0 + kUnit >= spp && ix < rp.width && iy < rp.height) {
        const int pix = ix + iy * rp.width;
        if (rp.accumID == 0) {
          accum[pix] = a;
        } else {
          float4 o = accum[pix];
          o.x += a.x; o.y += a.y; o.z += a.z; o.w += a.w;
          accum[pix] = o;
        }
      }
    }
    __syncthreads();  // slots and the pool counter are reused by the next pool
  }
  wave_add_u64(cnt.samples, counters + 0);
  wave_add_u64(cnt.ddaSteps, counters + 1);
  wave_add_u64((unsigned long long)nrays, counters + 2);
  wave_add_u64(cnt.zeroSteps, counters + 3);
}

template <int PB, class F>
void launch_render_volume(const F &v, const CameraDD &cam, const RenderDD &rp, int spp, float4 *accum,
                          unsigned long long *counters, hipStream_t s) {
  constexpr int TW = (PB >= 1024) ? 32 : (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  const int tiles = div_up(rp.width, TW) * div_up(rp.height, TH);
  // xcd_remap_v hands XCD x the tiles [x*per, (x+1)*per): the grid must
  // hold 8*per workgroups (a multiple of 8) to cover every tile
  const int grid = ((tiles + 7) / 8) * 8;
  if (has_lights(rp))
    k_render_volume<PB, F, true><<<grid, 256, 0, s>>>(v, cam, rp, spp, accum, counters);
  else
    k_render_volume<PB, F, false><<<grid, 256, 0, s>>>(v, cam, rp, spp, accum, counters);
}

// Pixels per workgroup: the largest power of two up to 1024 (a 32x32 tile)
// whose paths fit the pool (all of a pixel's paths in one pool), at least one.
template <class F>
void render_volume_dispatch(const F &v, const CameraDD &cam, const RenderDD &rp, int spp, float4 *accum,
                            unsigned long long *counters, hipStream_t s) {
  int pb = 1024;
  while (pb > 1 && (long long)pb * spp > kPoolPaths) pb >>= 1;
  switch (pb) {
    case 1024: launch_render_volume<1024>(v, cam, rp, spp, accum, counters, s); break;
    case 512: launch_render_volume<512>(v, cam, rp, spp, accum, counters, s); break;
    case 256: launch_render_volume<256>(v, cam, rp, spp, accum, counters, s); break;
    case 128: launch_render_volume<128>(v, cam, rp, spp, accum, counters, s); break;
    case 64: launch_render_volume<64>(v, cam, rp, spp, accum, counters, s); break;
    case 32: launch_render_volume<32>(v, cam, rp, spp, accum, counters, s); break;
    case 16: launch_render_volume<16>(v, cam, rp, spp, accum, counters, s); break;
    case 8: launch_render_volume<8>(v, cam, rp, spp, accum, counters, s); break;
    case 4: launch_render_volume<4>(v, cam, rp, spp, accum, counters, s); break;
    case 2: launch_render_volume<2>(v, cam, rp, spp, accum, counters, s); break;
    default: launch_render_volume<1>(v, cam, rp, spp, accum, counters, s); break;
  }
}

}  // namespace
}  // namespace bnamd
