// Batched multi-view path tracing: a per-pixel megakernel and a wavefront
// renderer over the same per-path computation (camera_ray + traverse +
// bounce, see shade.h), so both produce the same per-pixel values up to the
// order of floating-point accumulation.
//
// Wavefront organisation, per bounce d (all views and paths at once):
//   trace   a resident grid (sized from the occupancy of the traversal
//           kernel) walks the queue in 64-ray slices per wavefront and
//           writes one 16-byte hit record per ray;
//   sort    a stable counting sort of ray indices by BSDF type (miss,
//           Lambert, NVisii, Glass) so each wavefront of the shade kernel
//           runs one material model: per-block key histograms, one scan,
//           then a scatter that ranks keys with wavefront ballots -- no
//           global atomics, so the permutation is deterministic;
//   shade   barney's bounce(): fragments are added to the pixel with float
//           atomics, surviving paths are appended (compacted) to the next
//           queue, shadow rays to the shadow queue, with one counter atomic
//           per workgroup;
//   shadow  persistent occlusion queries; unoccluded contributions added.
// The whole loop, including counter resets (fill kernels, not memsets), is
// captured once into a hipGraph and replayed per batch of paths.
#include "shade.h"
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>

namespace bnamd {

size_t trace_lds_bytes(int depth, int block);
const World &world_of(const bnamd_world *w);

namespace {

constexpr int kMaxViews = 4096;

struct ViewsDD {
  int numViews;
  int pad;                // explicit: the wavefront graph key reads every byte
  const int *pixOffset;   // [numViews + 1]
  const int *tileOffset;  // [numViews + 1] (megakernel tiles)
  const int2 *size;       // [numViews]
  const CameraDD *cams;   // [numViews]
};

__device__ __forceinline__ int find_view(const int *off, int n, int g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <bool TL>
__device__ __forceinline__ ShadeHit shade_hit(const WorldDev &w, vec3f org, vec3f dir, float tmax, const HitRec &h) {
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

__device__ __forceinline__ void add_frag(vec3f &sum, vec3f frag, float clampMax) {
  frag = vmin(frag, clampMax);
  if (frag.x > 0.f) sum.x += frag.x;
  if (frag.y > 0.f) sum.y += frag.y;
  if (frag.z > 0.f) sum.z += frag.z;
}

// Wavefront-aggregated append: one atomic per wavefront on 'ctr', each
// active lane with 'pred' gets base + its rank among those lanes. Must be
// called by all active lanes of the wavefront together.
__device__ __forceinline__ unsigned wave_append(unsigned *ctr, bool pred) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return 0u;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned base = 0u;
  if (lane == leader) base = atomicAdd(ctr, (unsigned)__popcll(mask));
  base = __shfl(base, leader);
  return base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
}

// Workgroup-aggregated append of two predicates to two counters: one atomic
// per counter per 256-thread workgroup. Every thread of the workgroup must
// call it (with pred = false where it has nothing to append). Within the
// workgroup, positions follow thread order.
__device__ __forceinline__ void block_append2(unsigned *ctrA, bool predA, unsigned &posA, unsigned *ctrB,
                                              bool predB, unsigned &posB) {
  __shared__ unsigned cntA[4], cntB[4], baseA, baseB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const unsigned long long mA = __ballot(predA), mB = __ballot(predB);
  if (lane == 0) {
    cntA[wave] = (unsigned)__popcll(mA);
    cntB[wave] = (unsigned)__popcll(mB);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned tA = cntA[0] + cntA[1] + cntA[2] + cntA[3];
    const unsigned tB = cntB[0] + cntB[1] + cntB[2] + cntB[3];
    baseA = tA ? atomicAdd(ctrA, tA) : 0u;
    baseB = tB ? atomicAdd(ctrB, tB) : 0u;
  }
  __syncthreads();
  unsigned oA = baseA, oB = baseB;
  for (int w = 0; w < wave; ++w) {
    oA += cntA[w];
    oB += cntB[w];
  }
  posA = oA + (unsigned)__popcll(mA & lt);
  posB = oB + (unsigned)__popcll(mB & lt);
  __syncthreads();  // counts are rewritten by the next call
}

__device__ __forceinline__ void wave_add(unsigned long long v, unsigned long long *dst) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ===========================================================================
// Megakernel: L lanes per pixel (each looping over its share of the paths),
// 4 wavefronts per workgroup covering a TW x TH pixel tile of one view.
// ===========================================================================
template <int L, bool TL, bool kLights, bool kHit>
__global__ void __launch_bounds__(256) k_batch_mega(WorldDev w, ViewsDD vd, RenderDD rp, int spp,
                                                    float4 *__restrict accum, int4 *__restrict ids,
                                                    unsigned long long *__restrict counters, bool cutouts) {
  extern __shared__ uint2 lds_stack[];
  constexpr int PW = 64 / L, PB = 4 * PW;
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  // XCD-aware: each XCD walks a contiguous range of tiles
  const int numTiles = vd.tileOffset[vd.numViews];
  const int per = (numTiles + 7) >> 3;
  const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if ((blockIdx.x >> 3) >= per || tile >= numTiles) return;
  const int view = find_view(vd.tileOffset, vd.numViews, tile);
  const int t = tile - vd.tileOffset[view];
  const int2 sz = vd.size[view];
  const int tilesX = (sz.x + TW - 1) / TW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane & (L - 1);
  const int q = wave * PW + lane / L;
  const int ix = (t % tilesX) * TW + (q % TW);
  const int iy = (t / tilesX) * TH + (q / TW);
  const bool active = ix < sz.x && iy < sz.y;
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, w.depth + 1};
  unsigned long long nrays = 0, nshadow = 0;
  TraceCount tc;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const CameraDD cam = vd.cams[view];
  for (int r = 0; r * L < spp; ++r) {
    const int p = r * L + sub;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active && p < spp) {
      RenderDD rpp = rp;
      rpp.width = sz.x;
      rpp.height = sz.y;
      rpp.accumID = rp.accumID + p;
      PathRay ray = camera_ray(cam, rpp, ix, iy);
      const float clampMax = 10.f * (1 + rpp.accumID);
      vec3f sum = vec3f(0.f);
      float alpha = 0.f;
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const vec3f dir = fix_dir(ray.dir);
        HitRec h;
        if (kHit)
          world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, h, ctx, &tc,
                                    hit_filter(rpp, ray.org, dir, ray.rng.state));
        else
          world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, h, ctx, &tc);
        ++nrays;
        const ShadeHit sh = shade_hit<TL>(w, ray.org, dir, ray.tMax, h);
        if (gen == 0) {
          alpha = sh.hit ? 1.f : ray.missAlpha;
          if (ids && rpp.accumID == 0) {
            HitRec hid = h;  // closest candidate whatever its opacity
            if (kHit && cutouts)
              world_traverse<TL, false>(w, ray.org, dir, 0.f, ray.tMax, hid, ctx, nullptr,
                                        hit_filter(rpp, ray.org, dir, ray.rng.state, false));
            ids[vd.pixOffset[view] + ix + iy * sz.x] = hit_ids<TL>(w, ray.org, dir, ray.tMax, hid);
          }
        }
        PathRay shadow;
        add_frag(sum, bounce<kLights>(rpp, ray, sh, shadow, gen), clampMax);
        if (shadow.tMax > 0.f) {
          ShadeHit shs;
          HitRec hs;
          const vec3f sdir = fix_dir(shadow.dir);
          if (kHit)
            shs.hit = world_traverse<TL, true>(w, shadow.org, sdir, 0.f, shadow.tMax, hs, ctx, &tc,
                                               hit_filter(rpp, shadow.org, sdir, shadow.rng.state));
          else
            shs.hit = world_traverse<TL, true>(w, shadow.org, sdir, 0.f, shadow.tMax, hs, ctx, &tc);
          ++nshadow;
          PathRay dummy;
          add_frag(sum, bounce<kLights>(rpp, shadow, shs, dummy, gen + 1), clampMax);
        }
      }
      v = make_float4(sum.x, sum.y, sum.z, alpha);
    }
    add_round_in_path_order<L>(acc, v, min(L, spp - r * L));
  }
  if (active && sub == 0) {
    const int pix = vd.pixOffset[view] + ix + iy * sz.x;
    if (rp.accumID == 0) {
      accum[pix] = acc;
    } else {
      float4 o = accum[pix];
      o.x += acc.x; o.y += acc.y; o.z += acc.z; o.w += acc.w;
      accum[pix] = o;
    }
  }
  wave_add(nrays, counters + 0);
  wave_add(nshadow, counters + 1);
  wave_add(tc.nodes, counters + 2);
  wave_add(tc.primBytes, counters + 3);
}

// ===========================================================================
// Wavefront
// ===========================================================================
struct PathRec {
  float4 a;  // org, tMax
  float4 b;  // dir, mis
  float4 c;  // throughput, path id (uint bits)
  uint4 d;   // rng lo, rng hi, flags (numDiffuse | medium << 8), 0
};
struct ShadowRec {
  float4 a;  // org, tMax
  float4 b;  // dir, path id
  float4 c;  // contribution (mis * throughput), 0
  uint4 d;   // rng lo, rng hi (the seed of its anyHit draws), 0, 0
};

// device counters (one slot per bounce depth)
struct WfCounters {
  unsigned int count[32];     // paths in the queue of depth d
  unsigned int shadow[32];    // shadow rays spawned at depth d
  unsigned int work[32];      // persistent trace work counter at depth d
  unsigned int swork[32];     // persistent shadow work counter at depth d
  unsigned int bins[32][4];   // sort histogram
  unsigned int cursor[32][4]; // sort scatter cursors
  unsigned long long rays, shadowRays;
  unsigned long long nodeVisits, primBytes;  // traversal work (TraceCount)
};

// Counting sort geometry: kSortBlocks workgroups, each owning one contiguous
// segment of the queue (a multiple of 256 rays).
constexpr int kSortBlocks = 1024;

struct WfDD {
  WorldDev w;
  ViewsDD vd;
  RenderDD rp;
  int numPixels;  // all views
  int sppBatch;   // paths per pixel in one batch
  int accum0;     // accumID of the first path of the batch
  PathRec *q[2];
  HitRec *hits;
  unsigned int *keys;
  unsigned int *perm;
  unsigned int *blockHist;  // [kSortBlocks][4] key counts, then offsets
  ShadowRec *sq;
  float4 *accum;
  int4 *ids;      // ID buffer or null
  WfCounters *cnt;
  int chunk;      // rays per dynamic work fetch (one atomic per fetch);
                  // 0: kWalkChunk in k_wf_walk, static 64-ray slices in the
                  // two-level passes
  int cutouts;    // the world has cutout triangles (explicit tail: the graph
                  // key reads every byte)
};

__device__ __forceinline__ unsigned sort_segment(unsigned n) {
  const unsigned per = (n + kSortBlocks - 1) / kSortBlocks;
  return (per + 255u) & ~255u;
}

// Static slices: wavefront g of G handles rays [(g + k G) 64, +64).
// Dynamic: wavefronts fetch 'chunk' rays at a time from 'work'.
template <typename F>
__device__ __forceinline__ void for_each_ray(unsigned n, int chunk, unsigned *work, F &&f) {
  const int lane = threadIdx.x & 63;
  if (chunk == 0) {
    const unsigned G = gridDim.x * (blockDim.x >> 6);
    const unsigned g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (unsigned base = g * 64u; base < n; base += G * 64u) {
      const unsigned i = base + lane;
      if (i < n) f(i);
    }
    return;
  }
  while (true) {
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(work, (unsigned)chunk);
    base = __shfl(base, 0);
    if (base >= n) break;
    for (unsigned i = base + lane; i < base + chunk && i < n; i += 64) f(i);
  }
}

// path id = s * numPixels + global pixel
__device__ __forceinline__ void path_pixel(const WfDD &e, unsigned pid, int &s, int &view, int &ix, int &iy,
                                           int &pix) {
  s = (int)(pid / (unsigned)e.numPixels);
  pix = (int)(pid % (unsigned)e.numPixels);
  view = find_view(e.vd.pixOffset, e.vd.numViews, pix);
  const int local = pix - e.vd.pixOffset[view];
  const int w = e.vd.size[view].x;
  ix = local % w;
  iy = local / w;
}

__global__ void k_wf_reset(WfCounters *c) {
  unsigned int *p = (unsigned int *)c;
  const int n = (int)(sizeof(WfCounters) / sizeof(unsigned int));
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void k_wf_generate(WfDD e) {
  const unsigned total = (unsigned)e.numPixels * (unsigned)e.sppBatch;
  const unsigned pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= total) return;
  int s, view, ix, iy, pix;
  path_pixel(e, pid, s, view, ix, iy, pix);
  RenderDD rpp = e.rp;
  rpp.width = e.vd.size[view].x;
  rpp.height = e.vd.size[view].y;
  rpp.accumID = e.accum0 + s;
  const PathRay r = camera_ray(e.vd.cams[view], rpp, ix, iy);
  PathRec rec;
  rec.a = make_float4(r.org.x, r.org.y, r.org.z, r.tMax);
  rec.b = make_float4(r.dir.x, r.dir.y, r.dir.z, r.mis);
  rec.c = make_float4(r.tp.x, r.tp.y, r.tp.z, __uint_as_float(pid));
  rec.d = make_uint4((uint32_t)r.rng.state, (uint32_t)(r.rng.state >> 32), 0u, 0u);
  e.q[0][pid] = rec;
  if (pid == 0) e.cnt->count[0] = total;
}

template <bool TL, bool kHit>
__global__ void __launch_bounds__(256) k_wf_trace(WfDD e, int depth) {
  extern __shared__ uint2 lds_stack[];
  const unsigned n = e.cnt->count[depth];
  const PathRec *__restrict q = e.q[depth & 1];
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, e.w.depth + 1};
  unsigned long long nr = 0;
  TraceCount tc;
  for_each_ray(n, e.chunk, &e.cnt->work[depth], [&](unsigned i) {
    const float4 a = q[i].a, b = q[i].b;
    const vec3f org = {a.x, a.y, a.z};
    const vec3f dir = fix_dir({b.x, b.y, b.z});
    HitRec h;
    if (kHit) {
      const uint4 d = q[i].d;
      world_traverse<TL, false>(e.w, org, dir, 0.f, a.w, h, ctx, &tc,
                                hit_filter(e.rp, org, dir, (uint64_t)d.x | ((uint64_t)d.y << 32)));
    } else {
      world_traverse<TL, false>(e.w, org, dir, 0.f, a.w, h, ctx, &tc);
    }
    e.hits[i] = h;
    ++nr;
  });
  wave_add(nr, &e.cnt->rays);
  wave_add(tc.nodes, &e.cnt->nodeVisits);
  wave_add(tc.primBytes, &e.cnt->primBytes);
}

// The ID channel of a world with cutout triangles: the closest candidate of
// each primary ray whatever its opacity (Triangles.dev.cu:45-60 writes the
// IDs before the opacity test), for the paths with accumID 0.
template <bool TL>
__global__ void __launch_bounds__(256) k_wf_ids(WfDD e) {
  extern __shared__ uint2 lds_stack[];
  const unsigned n = e.cnt->count[0];
  const PathRec *__restrict q = e.q[0];
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, e.w.depth + 1};
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned pid = __float_as_uint(q[i].c.w);
    int sidx, view, ix, iy, pix;
    path_pixel(e, pid, sidx, view, ix, iy, pix);
    if (e.accum0 + sidx != 0) continue;
    const float4 a = q[i].a, b = q[i].b;
    const uint4 d = q[i].d;
    const vec3f org = {a.x, a.y, a.z};
    const vec3f dir = fix_dir({b.x, b.y, b.z});
    HitRec h;
    world_traverse<TL, false>(e.w, org, dir, 0.f, a.w, h, ctx, nullptr,
                              hit_filter(e.rp, org, dir, (uint64_t)d.x | ((uint64_t)d.y << 32), false));
    e.ids[pix] = hit_ids<TL>(e.w, org, dir, a.w, h);
  }
}

template <bool TL>
__device__ __forceinline__ unsigned bsdf_key(const WfDD &e, const HitRec &h) {
  if (h.flat == 0xffffffffu) return 0u;
  const MaterialDev &m = e.w.materials[hit_geom<TL>(e.w, h.flat).materialID];
  if (m.type != BNAMD_MAT_PBR) return 1u;
  return (m.ior != 1.f && m.transmission >= 1e-3f) ? 3u : 2u;
}

// Pass 1: keys and per-block key counts over the block's segment.
template <bool TL>
__global__ void __launch_bounds__(256) k_wf_sort_count(WfDD e, int depth) {
  __shared__ unsigned int hist[4];
  if (threadIdx.x < 4) hist[threadIdx.x] = 0u;
  __syncthreads();
  const unsigned n = e.cnt->count[depth];
  const unsigned seg = sort_segment(n);
  const unsigned lo = blockIdx.x * seg, hi = min(n, lo + seg);
  for (unsigned i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const unsigned k = bsdf_key<TL>(e, e.hits[i]);
    e.keys[i] = k;
    atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 4) e.blockHist[blockIdx.x * 4 + threadIdx.x] = hist[threadIdx.x];
}

// Pass 2 (one workgroup of kSortBlocks threads): blockHist becomes the
// first output position of each (block, key), key-major.
__global__ void __launch_bounds__(kSortBlocks) k_wf_sort_scan(WfDD e) {
  __shared__ unsigned waveTot[kSortBlocks / 64][4];
  __shared__ unsigned keyStart[4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  uint4 c = *(const uint4 *)(e.blockHist + 4 * t);
  unsigned v[4] = {c.x, c.y, c.z, c.w};
  unsigned incl[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    unsigned x = v[k];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    incl[k] = x;
    if (lane == 63) waveTot[wave][k] = x;
  }
  __syncthreads();
  if (t < 4) {
    unsigned acc = 0;
    for (int w = 0; w < kSortBlocks / 64; ++w) {
      const unsigned x = waveTot[w][t];
      waveTot[w][t] = acc;  // exclusive prefix over wavefronts
      acc += x;
    }
    keyStart[t] = acc;  // total of key t
  }
  __syncthreads();
  if (t == 0) {
    unsigned acc = 0;
    for (int k = 0; k < 4; ++k) {
      const unsigned x = keyStart[k];
      keyStart[k] = acc;
      acc += x;
    }
  }
  __syncthreads();
  uint4 o;
  o.x = keyStart[0] + waveTot[wave][0] + incl[0] - v[0];
  o.y = keyStart[1] + waveTot[wave][1] + incl[1] - v[1];
  o.z = keyStart[2] + waveTot[wave][2] + incl[2] - v[2];
  o.w = keyStart[3] + waveTot[wave][3] + incl[3] - v[3];
  *(uint4 *)(e.blockHist + 4 * t) = o;
}

// Pass 3: stable scatter of ray indices; ranks within a 256-ray slice come
// from wavefront ballots, so equal keys keep their queue order.
__global__ void __launch_bounds__(256) k_wf_sort_scatter(WfDD e, int depth) {
  __shared__ unsigned waveCnt[4][4];
  const unsigned n = e.cnt->count[depth];
  const unsigned seg = sort_segment(n);
  const unsigned lo = blockIdx.x * seg, hi = min(n, lo + seg);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const uint4 o = *(const uint4 *)(e.blockHist + 4 * blockIdx.x);
  unsigned run[4] = {o.x, o.y, o.z, o.w};
  for (unsigned base = lo; base < hi; base += 256u) {
    const unsigned i = base + threadIdx.x;
    const unsigned k = i < hi ? e.keys[i] : 4u;
    unsigned long long mine = 0ull;
#pragma unroll
    for (unsigned b = 0; b < 4; ++b) {
      const unsigned long long m = __ballot(k == b);
      if (k == b) mine = m;
      if (lane == 0) waveCnt[wave][b] = (unsigned)__popcll(m);
    }
    __syncthreads();
    if (k < 4u) {
      unsigned pos = run[k] + (unsigned)__popcll(mine & lt);
      for (int w = 0; w < wave; ++w) pos += waveCnt[w][k];
      e.perm[pos] = i;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) run[b] += waveCnt[0][b] + waveCnt[1][b] + waveCnt[2][b] + waveCnt[3][b];
    __syncthreads();
  }
}

__device__ __forceinline__ void accum_add(float4 *accum, int pix, vec3f f, float clampMax) {
  f = vmin(f, clampMax);
  float *p = (float *)(accum + pix);
  if (f.x > 0.f) atomicAdd(p + 0, f.x);
  if (f.y > 0.f) atomicAdd(p + 1, f.y);
  if (f.z > 0.f) atomicAdd(p + 2, f.z);
}

template <bool TL, bool kLights>
__global__ void __launch_bounds__(256) k_wf_shade(WfDD e, int depth) {
  const unsigned n = e.cnt->count[depth];
  const PathRec *__restrict qin = e.q[depth & 1];
  PathRec *__restrict qout = e.q[(depth + 1) & 1];
  // workgroup-uniform trip count: block_append2 needs every thread
  for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const unsigned i = base + threadIdx.x;
    bool wantShadow = false, wantCont = false;
    PathRay ray, shadow;
    unsigned pid = 0;
    if (i < n) {
      const unsigned r = e.perm[i];
      const PathRec rec = qin[r];
      const HitRec h = e.hits[r];
      pid = __float_as_uint(rec.c.w);
      int s, view, ix, iy, pix;
      path_pixel(e, pid, s, view, ix, iy, pix);
      RenderDD rpp = e.rp;
      rpp.width = e.vd.size[view].x;
      rpp.height = e.vd.size[view].y;
      rpp.accumID = e.accum0 + s;
      ray.org = {rec.a.x, rec.a.y, rec.a.z};
      ray.tMax = rec.a.w;
      ray.dir = {rec.b.x, rec.b.y, rec.b.z};
      ray.mis = rec.b.w;
      ray.tp = {rec.c.x, rec.c.y, rec.c.z};
      ray.rng.state = (uint64_t)rec.d.x | ((uint64_t)rec.d.y << 32);
      ray.numDiffuse = (int)(rec.d.z & 0xffu);
      ray.isInMedium = (rec.d.z >> 8) & 1u;
      ray.isShadow = false;
      ray.missColor = vec3f(0.f);
      ray.missAlpha = 0.f;
      if (depth == 0) miss_color(rpp, iy, ray.missColor, ray.missAlpha);  // primary rays' background
      const vec3f dir = fix_dir(ray.dir);
      const ShadeHit sh = shade_hit<TL>(e.w, ray.org, dir, ray.tMax, h);
      const float clampMax = 10.f * (1 + rpp.accumID);
      if (depth == 0) {
        const float alpha = sh.hit ? 1.f : ray.missAlpha;
        if (alpha > 0.f) atomicAdd(((float *)(e.accum + pix)) + 3, alpha);
        // with cutout triangles k_wf_ids wrote the closest candidate instead
        if (e.ids && rpp.accumID == 0 && !e.cutouts) e.ids[pix] = hit_ids<TL>(e.w, ray.org, dir, ray.tMax, h);
      }
      const vec3f frag = bounce<kLights>(rpp, ray, sh, shadow, depth);
      accum_add(e.accum, pix, frag, clampMax);
      wantShadow = shadow.tMax > 0.f;
      wantCont = ray.tMax > 0.f;
    }
    unsigned si, o;
    block_append2(&e.cnt->shadow[depth], wantShadow, si, &e.cnt->count[depth + 1], wantCont, o);
    if (wantShadow) {
      ShadowRec srec;
      const vec3f c = shadow.mis * shadow.tp;  // fragment if unoccluded
      srec.a = make_float4(shadow.org.x, shadow.org.y, shadow.org.z, shadow.tMax);
      srec.b = make_float4(shadow.dir.x, shadow.dir.y, shadow.dir.z, __uint_as_float(pid));
      srec.c = make_float4(c.x, c.y, c.z, 0.f);
      srec.d = make_uint4((uint32_t)shadow.rng.state, (uint32_t)(shadow.rng.state >> 32), 0u, 0u);
      e.sq[si] = srec;
    }
    if (wantCont) {
      PathRec out;
      out.a = make_float4(ray.org.x, ray.org.y, ray.org.z, ray.tMax);
      out.b = make_float4(ray.dir.x, ray.dir.y, ray.dir.z, ray.mis);
      out.c = make_float4(ray.tp.x, ray.tp.y, ray.tp.z, __uint_as_float(pid));
      out.d = make_uint4((uint32_t)ray.rng.state, (uint32_t)(ray.rng.state >> 32),
                         (uint32_t)ray.numDiffuse | ((ray.isInMedium ? 1u : 0u) << 8), 0u);
      qout[o] = out;
    }
  }
}

template <bool TL, bool kHit>
__global__ void __launch_bounds__(256) k_wf_shadow(WfDD e, int depth) {
  extern __shared__ uint2 lds_stack[];
  const unsigned n = e.cnt->shadow[depth];
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, e.w.depth + 1};
  unsigned long long nr = 0;
  TraceCount tc;
  for_each_ray(n, e.chunk, &e.cnt->swork[depth], [&](unsigned i) {
    const ShadowRec s = e.sq[i];
    const vec3f org = {s.a.x, s.a.y, s.a.z};
    const vec3f dir = fix_dir({s.b.x, s.b.y, s.b.z});
    HitRec hs;
    const bool occluded =
        kHit ? world_traverse<TL, true>(e.w, org, dir, 0.f, s.a.w, hs, ctx, &tc,
                                        hit_filter(e.rp, org, dir, (uint64_t)s.d.x | ((uint64_t)s.d.y << 32)))
             : world_traverse<TL, true>(e.w, org, dir, 0.f, s.a.w, hs, ctx, &tc);
    ++nr;
    if (!occluded) {
      const unsigned pid = __float_as_uint(s.b.w);
      const int sidx = (int)(pid / (unsigned)e.numPixels);
      const int pix = (int)(pid % (unsigned)e.numPixels);
      accum_add(e.accum, pix, {s.c.x, s.c.y, s.c.z}, 10.f * (1 + e.accum0 + sidx));
    }
  });
  wave_add(nr, &e.cnt->shadowRays);
  wave_add(tc.nodes, &e.cnt->nodeVisits);
  wave_add(tc.primBytes, &e.cnt->primBytes);
}

// Trace (kAny = false: closest hit into e.hits) or shadow (kAny = true:
// unoccluded contributions into accum) pass over a flattened world with
// per-lane ray replacement: a resident grid of wavefronts walks the queue;
// each wavefront takes e.chunk rays (kWalkChunk when 0) at a time from the
// queue's work counter and hands them to its lanes one by one, giving a lane its next ray
// as soon as its walk is over (once at least kRefillLanes lanes wait, or no
// lane walks), so that a wavefront does not idle behind its longest ray.
// Built for six wavefronts per SIMD (<= 80 VGPRs, no spill; 86 unconstrained
// leave five): the 64-view frame takes 191.5 ms against 204 ms at five and
// 208 ms at eight.
constexpr unsigned kWalkChunk = 128;
constexpr int kRefillLanes = 16;

template <bool kAny, bool kHit>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_wf_walk(WfDD e, int depth) {
  extern __shared__ uint2 lds_stack[];
  const unsigned n = kAny ? e.cnt->shadow[depth] : e.cnt->count[depth];
  unsigned *work = kAny ? &e.cnt->swork[depth] : &e.cnt->work[depth];
  const PathRec *__restrict q = e.q[depth & 1];
  const TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, e.w.depth + 1};
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned chunk = e.chunk > 0 ? (unsigned)e.chunk : kWalkChunk;
  using Filt = typename std::conditional<kHit, HitFilter, NoFilt>::type;
  BvhWalk<kAny, Filt> wk;
  unsigned rid = ~0u;             // this lane's ray (~0u: none)
  bool dry = false;               // the queue is used up (wavefront-uniform)
  unsigned cNext = 0, cEnd = 0;   // rest of the wavefront's chunk
  unsigned long long nr = 0;
  for (;;) {
    const unsigned long long need = __ballot(rid == ~0u);
    const int nNeed = (int)__popcll(need);
    if (!dry && (nNeed >= kRefillLanes || nNeed == 64)) {
      const int rank = (int)__popcll(need & below);
      unsigned mine = ~0u;
      int given = 0;
      while (given < nNeed) {
        if (cNext == cEnd) {
          unsigned base = 0;
          if (lane == 0) base = atomicAdd(work, chunk);
          base = __shfl(base, 0);
          if (base >= n) {
            dry = true;
            break;
          }
          cNext = base;
          cEnd = min(base + chunk, n);
        }
        const int take = min(nNeed - given, (int)(cEnd - cNext));
        if (rid == ~0u && rank >= given && rank < given + take) mine = cNext + (unsigned)(rank - given);
        cNext += (unsigned)take;
        given += take;
      }
      if (mine != ~0u) {
        rid = mine;
        float4 a, b;
        if (kAny) {
          a = e.sq[rid].a;
          b = e.sq[rid].b;
        } else {
          a = q[rid].a;
          b = q[rid].b;
        }
        wk.init({a.x, a.y, a.z}, fix_dir({b.x, b.y, b.z}), 0.f, a.w);
        if constexpr (kHit) {
          const uint4 d = kAny ? e.sq[rid].d : q[rid].d;
          wk.filt = hit_filter(e.rp, wk.org, wk.dir, (uint64_t)d.x | ((uint64_t)d.y << 32));
        }
      }
    }
    if (__ballot(rid != ~0u) == 0ull) {
      if (dry) break;
      continue;
    }
    if (rid != ~0u && wk.step(e.w.nodes, e.w.prims, ctx)) {
      ++nr;
      if (kAny) {
        if (!wk.occluded) {
          const ShadowRec &sr = e.sq[rid];
          const float4 c = sr.c;
          const unsigned pid = __float_as_uint(sr.b.w);
          const int sidx = (int)(pid / (unsigned)e.numPixels);
          const int pix = (int)(pid % (unsigned)e.numPixels);
          accum_add(e.accum, pix, {c.x, c.y, c.z}, 10.f * (1 + e.accum0 + sidx));
        }
      } else {
        e.hits[rid] = wk.best;
      }
      rid = ~0u;
    }
  }
  wave_add(nr, kAny ? &e.cnt->shadowRays : &e.cnt->rays);
  wave_add(wk.count.nodes, &e.cnt->nodeVisits);
  wave_add(wk.count.primBytes, &e.cnt->primBytes);
}

__global__ void k_fill_f4(float4 *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using TuneChoice = BatchTune;

struct BatchState {
  DevBuf<int> pixOffset, tileOffset;
  DevBuf<int2> sizes;
  DevBuf<CameraDD> cams;
  DevBuf<PathRec> q0, q1;
  DevBuf<HitRec> hits;
  DevBuf<unsigned> keys, perm, blockHist;
  DevBuf<ShadowRec> sq;
  DevBuf<WfCounters> cnt;
  DevBuf<unsigned long long> megaCnt;
  // graph cache
  hipGraphExec_t exec = nullptr;
  std::string graphKey;
  ~BatchState() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

BatchState &state() {
  static BatchState s;
  return s;
}

int num_cus() {
  static int n = 0;
  if (!n) {
    int dev = 0;
    BN_HIP(hipGetDevice(&dev));
    BN_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return n;
}

// Trace/shadow grid: blocksPerCU workgroups per CU, capped at what the
// occupancy of the two traversal kernels allows (static slices assume every
// workgroup is resident).
int trace_grid(int blocksPerCU, size_t lds, bool twoLevel) {
  int occT = 0, occS = 0;
  if (twoLevel) {
    BN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occT, k_wf_trace<true, false>, 256, lds));
    BN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occS, k_wf_shadow<true, false>, 256, lds));
  } else {
    BN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occT, k_wf_walk<false, false>, 256, lds));
    BN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occS, k_wf_walk<true, false>, 256, lds));
  }
  int occ = std::max(1, std::min(occT, occS));
  if (blocksPerCU > 0) occ = std::min(occ, blocksPerCU);
  return occ * num_cus();
}

struct HostViews {
  std::vector<int> pixOffset, tileOffset;
  std::vector<int2> sizes;
  std::vector<CameraDD> cams;
  int numPixels = 0;
};

HostViews make_views(const bnamd_views *v, int TW, int TH) {
  BN_REQUIRE(v && v->numViews > 0 && v->numViews <= kMaxViews && v->cameras && v->widths && v->heights,
             "bnamd_render_batch: bad views");
  HostViews h;
  h.pixOffset.push_back(0);
  h.tileOffset.push_back(0);
  long long px = 0, tl = 0;
  for (int i = 0; i < v->numViews; ++i) {
    const int w = v->widths[i], hh = v->heights[i];
    BN_REQUIRE(w > 0 && hh > 0, "bnamd_render_batch: empty view");
    px += (long long)w * hh;
    tl += (long long)div_up(w, TW) * div_up(hh, TH);
    BN_REQUIRE(px < (1ll << 31), "bnamd_render_batch: too many pixels");
    h.pixOffset.push_back((int)px);
    h.tileOffset.push_back((int)tl);
    h.sizes.push_back(make_int2(w, hh));
    h.cams.push_back(make_camera(v->cameras[i]));
  }
  h.numPixels = (int)px;
  return h;
}

ViewsDD upload_views(BatchState &st, const HostViews &h, hipStream_t s) {
  st.pixOffset.upload(h.pixOffset.data(), h.pixOffset.size(), s);
  st.tileOffset.upload(h.tileOffset.data(), h.tileOffset.size(), s);
  st.sizes.upload(h.sizes.data(), h.sizes.size(), s);
  st.cams.upload(h.cams.data(), h.cams.size(), s);
  ViewsDD vd;
  vd.numViews = (int)h.sizes.size();
  vd.pad = 0;
  vd.pixOffset = st.pixOffset.ptr;
  vd.tileOffset = st.tileOffset.ptr;
  vd.size = st.sizes.ptr;
  vd.cams = st.cams.ptr;
  return vd;
}

template <int L>
void launch_mega(const World &w, const bnamd_views *views, const RenderDD &rp, int spp, float4 *accum,
                 int4 *ids, unsigned long long *cnt, hipStream_t s) {
  constexpr int PB = 4 * (64 / L);
  constexpr int TW = (PB >= 256) ? 16 : (PB >= 64 ? 8 : (PB >= 16 ? 4 : (PB >= 4 ? 2 : 1)));
  constexpr int TH = PB / TW;
  BatchState &st = state();
  const HostViews hv = make_views(views, TW, TH);
  const ViewsDD vd = upload_views(st, hv, s);
  const int tiles = hv.tileOffset.back();
  const int grid = ((tiles + 7) / 8) * 8;
  const size_t lds = trace_lds_bytes(w.depth, 256);
  auto go = [&](auto kernel) {
    kernel<<<grid, 256, lds, s>>>(w.dd(), vd, rp, spp, accum, ids, cnt, w.hasCutouts);
  };
  switch ((w.twoLevel ? 4 : 0) | (has_lights(rp) ? 2 : 0) | ((rp.cutOn || w.hasCutouts) ? 1 : 0)) {
    case 0: go(k_batch_mega<L, false, false, false>); break;
    case 1: go(k_batch_mega<L, false, false, true>); break;
    case 2: go(k_batch_mega<L, false, true, false>); break;
    case 3: go(k_batch_mega<L, false, true, true>); break;
    case 4: go(k_batch_mega<L, true, false, false>); break;
    case 5: go(k_batch_mega<L, true, false, true>); break;
    case 6: go(k_batch_mega<L, true, true, false>); break;
    default: go(k_batch_mega<L, true, true, true>); break;
  }
  BN_HIP(hipGetLastError());
}

// lanes per pixel: the tuned value, else the largest power of two <= spp
// (up to 8); a pixel's paths run in rounds of 'lanes'
void render_mega(const World &w, const bnamd_views *views, const RenderDD &rp, int spp, float4 *accum,
                 int4 *ids, unsigned long long *cnt, hipStream_t s, int lanes = 0) {
  if (lanes <= 0) lanes = spp >= 8 ? 8 : spp >= 4 ? 4 : spp >= 2 ? 2 : 1;
  if (lanes >= 8) launch_mega<8>(w, views, rp, spp, accum, ids, cnt, s);
  else if (lanes >= 4) launch_mega<4>(w, views, rp, spp, accum, ids, cnt, s);
  else if (lanes >= 2) launch_mega<2>(w, views, rp, spp, accum, ids, cnt, s);
  else launch_mega<1>(w, views, rp, spp, accum, ids, cnt, s);
}

// Records the kernels of one batch (generate + maxDepth+1 bounces) on s.
// kHit: a cut plane or cutout triangles (the traces filter candidates).
template <bool TL, bool kHit>
void enqueue_wavefront_t(const WfDD &e, int maxDepth, int gridPersist, size_t lds, hipStream_t s) {
  const unsigned total = (unsigned)e.numPixels * (unsigned)e.sppBatch;
  k_wf_reset<<<1, 256, 0, s>>>(e.cnt);
  k_wf_generate<<<div_up(total, 256), 256, 0, s>>>(e);
  const int gridStream = std::min(div_up(total, 256), 8 * num_cus());
  for (int d = 0; d <= maxDepth; ++d) {
    if (TL) k_wf_trace<TL, kHit><<<gridPersist, 256, lds, s>>>(e, d);
    else k_wf_walk<false, kHit><<<gridPersist, 256, lds, s>>>(e, d);
    if (d == 0 && kHit && e.cutouts && e.ids && e.accum0 == 0) k_wf_ids<TL><<<gridStream, 256, lds, s>>>(e);
    k_wf_sort_count<TL><<<kSortBlocks, 256, 0, s>>>(e, d);
    k_wf_sort_scan<<<1, kSortBlocks, 0, s>>>(e);
    k_wf_sort_scatter<<<kSortBlocks, 256, 0, s>>>(e, d);
    if (has_lights(e.rp))
      k_wf_shade<TL, true><<<gridStream, 256, 0, s>>>(e, d);
    else
      k_wf_shade<TL, false><<<gridStream, 256, 0, s>>>(e, d);
    if (TL) k_wf_shadow<TL, kHit><<<gridPersist, 256, lds, s>>>(e, d);
    else k_wf_walk<true, kHit><<<gridPersist, 256, lds, s>>>(e, d);
  }
}

void enqueue_wavefront(const WfDD &e, int maxDepth, int gridPersist, size_t lds, bool twoLevel, hipStream_t s) {
  const bool hit = e.rp.cutOn || e.cutouts;
  if (twoLevel)
    hit ? enqueue_wavefront_t<true, true>(e, maxDepth, gridPersist, lds, s)
        : enqueue_wavefront_t<true, false>(e, maxDepth, gridPersist, lds, s);
  else
    hit ? enqueue_wavefront_t<false, true>(e, maxDepth, gridPersist, lds, s)
        : enqueue_wavefront_t<false, false>(e, maxDepth, gridPersist, lds, s);
}

float render_wavefront(const World &w, const bnamd_views *views, const RenderDD &rp0, int spp, float4 *accum,
                       int4 *ids, hipStream_t s, bnamd_batch_stats *stats, const TuneChoice &tune,
                       bool useGraph) {
  BatchState &st = state();
  const HostViews hv = make_views(views, 16, 16);
  const ViewsDD vd = upload_views(st, hv, s);
  const int maxDepth = std::min(rp0.maxPathDepth, 30);
  int sppBatch = tune.sppBatch > 0 ? std::min(tune.sppBatch, spp) : spp;
  // keep the queues within a memory budget (~24 GB)
  const size_t perPath = 2 * sizeof(PathRec) + sizeof(ShadowRec) + sizeof(HitRec) + 2 * sizeof(unsigned);
  while (sppBatch > 1 && (size_t)hv.numPixels * sppBatch * perPath > (24ull << 30)) --sppBatch;
  const size_t cap = (size_t)hv.numPixels * sppBatch;
  BN_REQUIRE(cap < (1ull << 32), "bnamd_render_batch: too many paths per batch");
  st.q0.ensure(cap);
  st.q1.ensure(cap);
  st.hits.ensure(cap);
  st.keys.ensure(cap);
  st.perm.ensure(cap);
  st.blockHist.ensure(4 * kSortBlocks);
  st.sq.ensure(cap);
  st.cnt.ensure(1);
  const size_t lds = trace_lds_bytes(w.depth, 256);
  BN_REQUIRE(lds <= 64 * 1024, "bnamd_render_batch: BVH too deep for the LDS stack");
  const int gridPersist = trace_grid(tune.traceBlocksPerCU, lds, w.twoLevel);
  if (rp0.accumID == 0)
    k_fill_f4<<<std::min(div_up(hv.numPixels, 256), 4096), 256, 0, s>>>(accum, (size_t)hv.numPixels);
  hipEvent_t e0, e1;
  BN_HIP(hipEventCreate(&e0));
  BN_HIP(hipEventCreate(&e1));
  BN_HIP(hipEventRecord(e0, s));
  unsigned long long rays = 0, shadows = 0, nodeVisits = 0, primBytes = 0;
  int replays = 0;
  for (int p0 = 0; p0 < spp; p0 += sppBatch) {
    WfDD e;
    // zeroed first: padding bytes take part in the graph cache key below
    memset(&e, 0, sizeof e);
    e.w = w.dd();
    e.vd = vd;
    e.rp = rp0;
    e.numPixels = hv.numPixels;
    e.sppBatch = std::min(sppBatch, spp - p0);
    e.accum0 = rp0.accumID + p0;
    e.q[0] = st.q0.ptr;
    e.q[1] = st.q1.ptr;
    e.hits = st.hits.ptr;
    e.keys = st.keys.ptr;
    e.perm = st.perm.ptr;
    e.blockHist = st.blockHist.ptr;
    e.sq = st.sq.ptr;
    e.accum = accum;
    e.ids = ids;
    e.cnt = st.cnt.ptr;
    e.chunk = tune.chunk;
    e.cutouts = w.hasCutouts ? 1 : 0;
    if (useGraph) {
      // The graph bakes every kernel argument (the whole WfDD by value: world
      // pointers and sizes, render parameters such as ambient and background,
      // view buffers, queues) and the launch shapes, so the cache key is all
      // of those bytes: a replay is then exactly the launches enqueue_wavefront
      // would make.
      std::string key((const char *)&e, sizeof e);
      const int shape[4] = {maxDepth, gridPersist, (int)lds, w.twoLevel ? 1 : 0};
      key.append((const char *)shape, sizeof shape);
      if (!st.exec || st.graphKey != key) {
        if (st.exec) {
          (void)hipGraphExecDestroy(st.exec);
          st.exec = nullptr;
        }
        hipStream_t cs;
        BN_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t g;
        BN_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        enqueue_wavefront(e, maxDepth, gridPersist, lds, w.twoLevel, cs);
        BN_HIP(hipStreamEndCapture(cs, &g));
        BN_HIP(hipGraphInstantiate(&st.exec, g, nullptr, nullptr, 0));
        BN_HIP(hipGraphDestroy(g));
        BN_HIP(hipStreamDestroy(cs));
        st.graphKey = key;
      }
      BN_HIP(hipGraphLaunch(st.exec, s));
      ++replays;
    } else {
      enqueue_wavefront(e, maxDepth, gridPersist, lds, w.twoLevel, s);
    }
    WfCounters hc;
    BN_HIP(hipMemcpyAsync(&hc, st.cnt.ptr, sizeof(hc), hipMemcpyDeviceToHost, s));
    BN_HIP(hipStreamSynchronize(s));
    rays += hc.rays;
    shadows += hc.shadowRays;
    nodeVisits += hc.nodeVisits;
    primBytes += hc.primBytes;
  }
  BN_HIP(hipEventRecord(e1, s));
  BN_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  BN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (stats) {
    stats->raysTraced = (long long)(rays + shadows);
    stats->primaryRays = (long long)hv.numPixels * spp;
    stats->shadowRays = (long long)shadows;
    stats->ms = ms;
    stats->graphReplays = replays;
    stats->nodeVisits = (long long)nodeVisits;
    stats->primBytes = (long long)primBytes;
  }
  return ms;
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

extern "C" int bnamd_render_batch(const bnamd_world *world, const bnamd_views *views, const bnamd_render_params *params,
                                  float *accum, int *ids, int mode, void *stream, bnamd_batch_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(world && views && params && accum, "bnamd_render_batch: NULL argument");
    BN_REQUIRE(params->pathsPerPixel >= 1, "bnamd_render_batch: pathsPerPixel must be >= 1");
    BN_REQUIRE(mode == BNAMD_BATCH_MEGAKERNEL || mode == BNAMD_BATCH_WAVEFRONT || mode == BNAMD_BATCH_AUTO,
               "bnamd_render_batch: bad mode");
    const World &w = world_of(world);
    hipStream_t s = (hipStream_t)stream;
    RenderDD rp = make_render_dd(*params);
    if (mode == BNAMD_BATCH_AUTO) mode = w.tune.mode;
    if (mode == BNAMD_BATCH_MEGAKERNEL) {
      BatchState &st = state();
      st.megaCnt.ensure(4);
      BN_HIP(hipMemsetAsync(st.megaCnt.ptr, 0, 4 * sizeof(unsigned long long), s));
      hipEvent_t e0, e1;
      BN_HIP(hipEventCreate(&e0));
      BN_HIP(hipEventCreate(&e1));
      BN_HIP(hipEventRecord(e0, s));
      render_mega(w, views, rp, params->pathsPerPixel, (float4 *)accum, (int4 *)ids, st.megaCnt.ptr, s,
                  w.tune.megaLanes);
      BN_HIP(hipEventRecord(e1, s));
      BN_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BN_HIP(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      unsigned long long h[4];
      BN_HIP(hipMemcpy(h, st.megaCnt.ptr, sizeof(h), hipMemcpyDeviceToHost));
      if (stats) {
        long long px = 0;
        for (int i = 0; i < views->numViews; ++i) px += (long long)views->widths[i] * views->heights[i];
        stats->raysTraced = (long long)(h[0] + h[1]);
        stats->primaryRays = px * params->pathsPerPixel;
        stats->shadowRays = (long long)h[1];
        stats->ms = ms;
        stats->graphReplays = 0;
        stats->nodeVisits = (long long)h[2];
        stats->primBytes = (long long)h[3];
      }
    } else {
      BatchState &st = state();
      const char *ng = getenv("BNAMD_NO_GRAPH");
      render_wavefront(w, views, rp, params->pathsPerPixel, (float4 *)accum, (int4 *)ids, s, stats, w.tune,
                       !(ng && ng[0] == '1'));
    }
  });
}

// Times the batched renderers' launch configurations on this world and views
// and keeps the fastest in the world's BatchTune:
//   1. wavefront trace grid (workgroups per CU) x work-fetch chunk, on one
//      path per pixel;
//   2. wavefront queue size: paths per pixel per batch (sppBatch), on the
//      parameters' paths per pixel;
//   3. megakernel tile: lanes per pixel (1, 2, 4, 8 up to spp), likewise;
//   4. the faster of the two renderers at their best settings.
extern "C" int bnamd_batch_autotune(const bnamd_world *world, const bnamd_views *views,
                                    const bnamd_render_params *params, const char *path) {
  return guarded([&] {
    BN_REQUIRE(world && views && params, "bnamd_batch_autotune: NULL argument");
    BN_REQUIRE(params->pathsPerPixel >= 1, "bnamd_batch_autotune: pathsPerPixel must be >= 1");
    const World &w = world_of(world);
    BatchState &st = state();
    RenderDD rp = make_render_dd(*params);
    rp.accumID = 0;
    const int spp = params->pathsPerPixel;
    long long px = 0;
    for (int i = 0; i < views->numViews; ++i) px += (long long)views->widths[i] * views->heights[i];
    DevBuf<float4> accum((size_t)px);
    // the better of two runs keeps one noisy run from deciding
    auto wave = [&](const TuneChoice &t, int n) {
      float ms = render_wavefront(w, views, rp, n, accum.ptr, nullptr, 0, nullptr, t, false);
      return std::min(ms, render_wavefront(w, views, rp, n, accum.ptr, nullptr, 0, nullptr, t, false));
    };
    st.megaCnt.ensure(4);
    auto mega = [&](int lanes) {
      float best = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        hipEvent_t e0, e1;
        BN_HIP(hipEventCreate(&e0));
        BN_HIP(hipEventCreate(&e1));
        BN_HIP(hipEventRecord(e0, 0));
        render_mega(w, views, rp, spp, accum.ptr, nullptr, st.megaCnt.ptr, 0, lanes);
        BN_HIP(hipEventRecord(e1, 0));
        BN_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        BN_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        best = std::min(best, ms);
      }
      return best;
    };
    const int blocks[] = {2, 3, 4, 0};
    // chunk: rays per work fetch (flattened worlds: the walk passes' fetch,
    // 0 = 128; two-level worlds: 0 = static 64-ray slices)
    const int chunksFlat[] = {0, 64, 256};
    const int chunksTwoLevel[] = {0, 256, 1024};
    const int *chunks = w.twoLevel ? chunksTwoLevel : chunksFlat;
    TuneChoice best;
    float best1 = 1e30f;
    for (int b : blocks)
      for (int ci = 0; ci < 3; ++ci) {
        TuneChoice t;
        t.traceBlocksPerCU = b;
        t.chunk = chunks[ci];
        t.sppBatch = 1;
        const float ms = wave(t, 1);
        if (ms < best1) {
          best1 = ms;
          best = t;
        }
      }
    float bestWave = 1e30f;
    int bestBatch = spp;
    for (int b = spp; b >= 1; b = (b > 1 ? (b + 1) / 2 : 0)) {
      TuneChoice t = best;
      t.sppBatch = b;
      const float ms = wave(t, spp);
      if (ms < bestWave) {
        bestWave = ms;
        bestBatch = b;
      }
      if (b == 1) break;
    }
    best.sppBatch = bestBatch == spp ? 0 : bestBatch;
    float bestMega = 1e30f;
    int bestLanes = 0;
    for (int lanes = 1; lanes <= 8 && lanes <= spp; lanes *= 2) {
      const float ms = mega(lanes);
      if (ms < bestMega) {
        bestMega = ms;
        bestLanes = lanes;
      }
    }
    best.megaLanes = bestLanes;
    best.mode = bestMega < bestWave ? BNAMD_BATCH_MEGAKERNEL : BNAMD_BATCH_WAVEFRONT;
    best.tuned = true;
    w.tune = best;
    if (path) {
      FILE *f = fopen(path, "w");
      BN_REQUIRE(f, std::string("bnamd_batch_autotune: cannot write ") + path);
      fprintf(f,
              "{\"traceBlocksPerCU\": %d, \"chunk\": %d, \"bestMs1spp\": %.4f, \"sppBatch\": %d, "
              "\"wavefrontMs\": %.4f, \"megaLanes\": %d, \"megakernelMs\": %.4f, \"mode\": \"%s\"}\n",
              best.traceBlocksPerCU, best.chunk, best1, bestBatch, bestWave, bestLanes, bestMega,
              best.mode == BNAMD_BATCH_MEGAKERNEL ? "megakernel" : "wavefront");
      fclose(f);
    }
  });
}
