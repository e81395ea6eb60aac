// Emission-absorption volume rendering with a checkpointed adjoint.
//
// Forward: one thread per pixel marches its ray through the density grid at
// a fixed step, accumulating C = sum T_i a_i c_i, and stores the
// transmittance at the start of every K-sample segment (segment-major, so a
// wavefront's stores are coalesced).
// Adjoint: the same thread walks the segments back to front; for each it
// re-marches K samples from the stored checkpoint (keeping T, a, c and the
// density value in registers), then replays them in reverse with the suffix
// radiance S_i = sum_{j>i} T_j a_j c_j:
//   dL/dc_i = g T_i a_i,  dL/ds_i = dt (T_{i+1} <g, c_i> - <g, S_i>).
// Gradients flow to the transfer function (per-block LDS accumulation,
// reduced afterwards) and to the 8 voxels of each trilinear sample (device
// atomics, merged in registers while consecutive samples share a cell).
//
// Sampling follows StructuredDataSampler (barney/volume/StructuredData.h:
// 104-115, embree texel convention rtcore/embree/Texture.cpp:324-416) and the
// transfer-function lookup follows TransferFunction::DD::map
// (barney/volume/TransferFunction.h:67-86).
#include "shade.h"
#include "volume.h"

namespace bnamd {

namespace {

struct EADD {
  VolumeDD vol;     // scalars, dims, origin, spacing, world box
  TFDD tf;
  float step;
  int W, H, numViews;
  int numSeg;       // checkpoints per ray
  long long numRays;
  const CameraDD *cams;
};

constexpr int kTile = 16;  // 16x16 pixels per workgroup

__device__ __forceinline__ bool ea_setup(const EADD &e, long long ray, int ix, int iy, int view, vec3f &org,
                                         vec3f &dir, float &t0, int &n) {
  RenderDD rp;
  rp.width = e.W;
  rp.height = e.H;
  rp.accumID = 0;
  rp.bg[0] = rp.bg[1] = rp.bg[2] = 0.f;
  rp.bg[3] = 1.f;
  const PathRay pr = camera_ray(e.cams[view], rp, ix, iy);
  org = pr.org;
  dir = fix_dir(pr.dir);
  const vec3f a = (e.vol.worldLo - org) * rcp(dir);
  const vec3f b = (e.vol.worldHi - org) * rcp(dir);
  t0 = fmaxf(0.f, reduce_max(vmin(a, b)));
  const float t1 = reduce_min(vmax(a, b));
  n = (t0 <= t1) ? (int)floorf((t1 - t0) / e.step) : 0;
  return n > 0;
}

// Trilinear sample with the corner set and weights needed for the adjoint.
struct Tri {
  float value;
  size_t base;       // index of corner (x0,y0,z0)
  int dx, dy, dz;    // 0 or 1 (clamped edges collapse a corner pair)
  float fx, fy, fz;
  bool valid;
};

// Locates the trilinear cell and weights of P (no loads); valid = inside.
__device__ __forceinline__ Tri tri_locate(const VolumeDD &v, vec3f P) {
#pragma clang fp contract(off)
  Tri r;
  r.valid = false;
  r.value = 0.f;
  const vec3f rel = (P - v.origin) * rcp(v.spacing);
  if (rel.x < 0.f || rel.y < 0.f || rel.z < 0.f) return r;
  if (rel.x >= (float)(v.nx - 1) || rel.y >= (float)(v.ny - 1) || rel.z >= (float)(v.nz - 1)) return r;
  int x0, x1, y0, y1, z0, z1;
  lerp_coords(v.nx, rel.x + .5f, x0, x1, r.fx);
  lerp_coords(v.ny, rel.y + .5f, y0, y1, r.fy);
  lerp_coords(v.nz, rel.z + .5f, z0, z1, r.fz);
  const size_t sy = (size_t)v.nx, sz = (size_t)v.nx * v.ny;
  r.base = (size_t)x0 + y0 * sy + z0 * sz;
  r.dx = x1 - x0;
  r.dy = y1 - y0;
  r.dz = z1 - z0;
  r.valid = true;
  return r;
}

// Trilinear sample (same arithmetic as sample_field) keeping the cell.
__device__ __forceinline__ Tri tri_sample(const VolumeDD &v, vec3f P) {
#pragma clang fp contract(off)
  Tri r = tri_locate(v, P);
  if (!r.valid) return r;
  const size_t sy = (size_t)v.nx, sz = (size_t)v.nx * v.ny;
  const float *s = v.scalars + r.base;
  const size_t ox = r.dx, oy = r.dy * sy, oz = r.dz * sz;
  const float v000 = s[0], v001 = s[ox], v010 = s[oy], v011 = s[ox + oy];
  const float v100 = s[oz], v101 = s[ox + oz], v110 = s[oy + oz], v111 = s[ox + oy + oz];
  const float a = lerp_l(r.fx, v000, v001), b = lerp_l(r.fx, v010, v011);
  const float c = lerp_l(r.fx, v100, v101), d = lerp_l(r.fx, v110, v111);
  r.value = lerp_l(r.fz, lerp_l(r.fy, a, b), lerp_l(r.fy, c, d));
  return r;
}

// TF lookup with index/fraction for the adjoint (TransferFunction.h:67-86).
struct TFLook {
  float r, g, b, s;  // color and extinction (w * baseDensity)
  int idx;
  float f;
  float dfds;        // d(f)/d(scalar): (N-1)/span inside the domain, else 0
};

__device__ __forceinline__ TFLook tf_look(const TFDD &tf, const float4 *__restrict lut, float sv) {
#pragma clang fp contract(off)
  TFLook o;
  const float span = tf.domHi - tf.domLo;
  const float fr = (sv - tf.domLo) / span;
  float f = fminf(fmaxf(fr, 0.f), 1.f);
  f *= (tf.numValues - 1);
  int idx = int(f);
  idx = idx < 0 ? 0 : (idx > tf.numValues - 2 ? tf.numValues - 2 : idx);
  f -= idx;
  const float4 a = lut[idx], b = lut[idx + 1];
  o.r = (1.f - f) * a.x + f * b.x;
  o.g = (1.f - f) * a.y + f * b.y;
  o.b = (1.f - f) * a.z + f * b.z;
  o.s = ((1.f - f) * a.w + f * b.w) * tf.baseDensity;
  o.idx = idx;
  o.f = f;
  o.dfds = (fr > 0.f && fr < 1.f) ? (float)(tf.numValues - 1) / span : 0.f;
  return o;
}

// Tiles over views; each XCD gets a contiguous range of tiles.
__device__ __forceinline__ bool ea_tile(const EADD &e, int tileIdx, int &view, int &ix, int &iy) {
  const int tx = (e.W + kTile - 1) / kTile, ty = (e.H + kTile - 1) / kTile;
  const int per = tx * ty;
  view = tileIdx / per;
  const int t = tileIdx % per;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  ix = (t % tx) * kTile + (wave & 1) * 8 + (lane & 7);
  iy = (t / tx) * kTile + (wave >> 1) * 8 + (lane >> 3);
  return view < e.numViews && ix < e.W && iy < e.H;
}

__device__ __forceinline__ int xcd_tile(int b, int k, int numTiles) {
  // block b runs on XCD b % 8; XCD x owns tiles [x*per, (x+1)*per)
  const int per = (numTiles + 7) >> 3;
  const int nbx = gridDim.x >> 3;  // blocks per XCD (gridDim is a multiple of 8)
  const int local = (b >> 3) + k * nbx;
  if (local >= per) return -1;
  const int t = (b & 7) * per + local;
  return t < numTiles ? t : -1;
}

template <int K>
__global__ void __launch_bounds__(256) k_ea_forward(EADD e, float *__restrict img, float *__restrict ckpt,
                                                    unsigned long long *__restrict counter) {
  __shared__ float4 lut[256];
  for (int i = threadIdx.x; i < e.tf.numValues && i < 256; i += blockDim.x) lut[i] = e.tf.values[i];
  __syncthreads();
  const float4 *tfv = e.tf.numValues <= 256 ? lut : e.tf.values;
  const int numTiles = e.numViews * ((e.W + kTile - 1) / kTile) * ((e.H + kTile - 1) / kTile);
  unsigned long long nsamp = 0;
  for (int k = 0;; ++k) {
    const int tile = xcd_tile(blockIdx.x, k, numTiles);
    if (tile < 0) break;
    int view, ix, iy;
    if (!ea_tile(e, tile, view, ix, iy)) continue;
    const long long ray = ((long long)view * e.H + iy) * e.W + ix;
    vec3f org, dir;
    float t0;
    int n;
    float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f;
    if (ea_setup(e, ray, ix, iy, view, org, dir, t0, n)) {
      for (int i = 0; i < n; ++i) {
        if (ckpt && (i % K) == 0) ckpt[(long long)(i / K) * e.numRays + ray] = T;
        const float t = t0 + ((float)i + 0.5f) * e.step;
        const Tri s = tri_sample(e.vol, org + t * dir);
        if (!s.valid) continue;
        const TFLook m = tf_look(e.tf, tfv, s.value);
        const float alpha = 1.f - expf(-m.s * e.step);
        const float w = T * alpha;
        cr += w * m.r;
        cg += w * m.g;
        cb += w * m.b;
        T *= (1.f - alpha);
      }
      nsamp += (unsigned long long)n;
    }
    img[3 * ray + 0] = cr;
    img[3 * ray + 1] = cg;
    img[3 * ray + 2] = cb;
  }
  for (int off = 32; off > 0; off >>= 1) nsamp += __shfl_xor(nsamp, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(counter, nsamp);
}

// Per-workgroup LDS cache of density gradients: 64-bit entries (value bits
// << 32 | voxel index), updated with compare-and-swap. A colliding voxel
// evicts the resident one to global memory; the cache is drained after each
// tile. Rays of one tile touch overlapping voxels, so most contributions are
// merged in LDS before reaching the device atomics.
constexpr int kCacheSlots = 2048;
constexpr uint32_t kEmpty = 0xffffffffu;

__device__ __forceinline__ void grad_add(unsigned long long *cache, float *__restrict gd, uint32_t v, float g) {
  const uint32_t slot = (v * 2654435761u) >> 21;  // 11-bit hash
  unsigned long long old = cache[slot];
  for (int tries = 0; tries < 8; ++tries) {
    const uint32_t k = (uint32_t)old;
    const float val = __uint_as_float((uint32_t)(old >> 32));
    const float nval = (k == v) ? val + g : g;
    const unsigned long long nw = ((unsigned long long)__float_as_uint(nval) << 32) | v;
    const unsigned long long prev = atomicCAS(&cache[slot], old, nw);
    if (prev == old) {
      if (k != v && k != kEmpty && val != 0.f) atomicAdd(gd + k, val);
      return;
    }
    old = prev;
  }
  atomicAdd(gd + v, g);
}

__device__ __forceinline__ void cache_drain(unsigned long long *cache, float *__restrict gd) {
  for (int i = threadIdx.x; i < kCacheSlots; i += blockDim.x) {
    const unsigned long long e = cache[i];
    const uint32_t k = (uint32_t)e;
    const float val = __uint_as_float((uint32_t)(e >> 32));
    if (k != kEmpty && val != 0.f) atomicAdd(gd + k, val);
    cache[i] = (unsigned long long)kEmpty;
  }
}

// Density-gradient accumulator for the 8 corner voxels of the current
// trilinear cell. When the walk moves to the next cell, corners shared with
// it are carried over (summed into the new cell's slots) and only the others
// are flushed with device atomics, so each voxel is written about once per
// ray instead of once per cell.
struct CellAcc {
  uint32_t v[8];  // absolute voxel index of each corner
  float g[8];
  bool live;
};

__device__ __forceinline__ void cell_corners(const Tri &s, size_t sy, size_t sz, uint32_t *v) {
  const uint32_t b = (uint32_t)s.base, ox = (uint32_t)s.dx, oy = (uint32_t)(s.dy * sy),
                 oz = (uint32_t)(s.dz * sz);
  v[0] = b; v[1] = b + ox; v[2] = b + oy; v[3] = b + ox + oy;
  v[4] = b + oz; v[5] = b + ox + oz; v[6] = b + oy + oz; v[7] = b + ox + oy + oz;
}

__device__ __forceinline__ void cell_flush(CellAcc &c, unsigned long long *cache, float *__restrict gd) {
  if (!c.live) return;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (c.g[k] != 0.f) grad_add(cache, gd, c.v[k], c.g[k]);
  c.live = false;
}

// Switches the accumulator to the cell with corners nv, carrying shared
// corners over and flushing the rest.
__device__ __forceinline__ void cell_move(CellAcc &c, const uint32_t *nv, unsigned long long *cache,
                                          float *__restrict gd) {
  float ng[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) ng[j] = 0.f;
  if (c.live) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      bool kept = false;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool m = !kept && nv[j] == c.v[k];
        ng[j] += m ? c.g[k] : 0.f;
        kept = kept || m;
      }
      if (!kept && c.g[k] != 0.f) grad_add(cache, gd, c.v[k], c.g[k]);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    c.v[j] = nv[j];
    c.g[j] = ng[j];
  }
  c.live = true;
}

// Transfer-function gradient of the current TF bin, merged in registers
// while consecutive samples fall into the same bin.
struct TfAcc {
  int idx;
  float g[8];
};

__device__ __forceinline__ void tf_flush(TfAcc &a, float *gtf) {
  if (a.idx < 0) return;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (a.g[q] != 0.f) atomicAdd(&gtf[4 * a.idx + q], a.g[q]);
  a.idx = -1;
}

// Built for five wavefronts per SIMD (<= 96 VGPRs, ~100-180 bytes of spill
// per lane): the replay waits on field loads and gradient atomics, and the
// 64-view 512x512 adjoint takes 144 ms against 151 ms at the three or four
// the unconstrained allocation (121-128 VGPRs) leaves; at six it is 201 ms.
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_ea_backward(EADD e, const float *__restrict gimg,
                                                     const float *__restrict ckpt, float *__restrict gdens,
                                                     float *__restrict tfPartial,
                                                     unsigned long long *__restrict counter) {
  __shared__ float4 lut[256];
  __shared__ float gtf[256 * 4];
  __shared__ unsigned long long cache[kCacheSlots];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    if (i < e.tf.numValues) lut[i] = e.tf.values[i];
    gtf[4 * i + 0] = gtf[4 * i + 1] = gtf[4 * i + 2] = gtf[4 * i + 3] = 0.f;
  }
  for (int i = threadIdx.x; i < kCacheSlots; i += blockDim.x) cache[i] = (unsigned long long)kEmpty;
  __syncthreads();
  const size_t sy = (size_t)e.vol.nx, sz = (size_t)e.vol.nx * e.vol.ny;
  const int numTiles = e.numViews * ((e.W + kTile - 1) / kTile) * ((e.H + kTile - 1) / kTile);
  unsigned long long nsamp = 0;
  for (int kk = 0;; ++kk) {
    const int tile = xcd_tile(blockIdx.x, kk, numTiles);  // uniform over the workgroup
    if (tile < 0) break;
    do {  // per-ray work; 'break' skips to the tile's drain barrier
    int view, ix, iy;
    if (!ea_tile(e, tile, view, ix, iy)) break;
    const long long ray = ((long long)view * e.H + iy) * e.W + ix;
    vec3f org, dir;
    float t0;
    int n;
    if (!ea_setup(e, ray, ix, iy, view, org, dir, t0, n)) break;
    const float gr = gimg[3 * ray + 0], gg = gimg[3 * ray + 1], gb = gimg[3 * ray + 2];
    if (gr == 0.f && gg == 0.f && gb == 0.f) break;
    nsamp += (unsigned long long)n;
    float Sr = 0.f, Sg = 0.f, Sb = 0.f;  // suffix radiance behind the current sample
    CellAcc cell;
    cell.live = false;
    const int nseg = (n + K - 1) / K;
    TfAcc tfa;
    tfa.idx = -1;
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
      const int i0 = sgi * K;
      // re-march the segment from its checkpoint; keep T and the density
      // of each sample (NaN marks samples outside the grid)
      float Ts[K], Rho[K];
      float T = ckpt[(long long)sgi * e.numRays + ray];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Ts[k] = T;
        Rho[k] = __builtin_nanf("");
        if (i0 + k < n) {
          const float t = t0 + ((float)(i0 + k) + 0.5f) * e.step;
          const Tri s = tri_sample(e.vol, org + t * dir);
          if (s.valid) {
            const TFLook m = tf_look(e.tf, lut, s.value);
            const float alpha = 1.f - expf(-m.s * e.step);
            Rho[k] = s.value;
            T *= (1.f - alpha);
          }
        }
      }
      // replay in reverse
#pragma unroll
      for (int k = K - 1; k >= 0; --k) {
        const float rho = Rho[k];
        if (rho != rho) continue;
        const TFLook m = tf_look(e.tf, lut, rho);
        const float a = 1.f - expf(-m.s * e.step);
        const float Ti = Ts[k];
        const float Tn = Ti * (1.f - a);
        const float w = Ti * a;
        const float dcr = gr * w, dcg = gg * w, dcb = gb * w;
        const float dsig = e.step * (Tn * (gr * m.r + gg * m.g + gb * m.b) - (gr * Sr + gg * Sg + gb * Sb));
        Sr += w * m.r;
        Sg += w * m.g;
        Sb += w * m.b;
        const float f = m.f;
        const float dsw = dsig * e.tf.baseDensity;  // d/d(w entries)
        if (m.idx != tfa.idx) {
          tf_flush(tfa, gtf);
          tfa.idx = m.idx;
#pragma unroll
          for (int q = 0; q < 8; ++q) tfa.g[q] = 0.f;
        }
        tfa.g[0] += dcr * (1.f - f);
        tfa.g[1] += dcg * (1.f - f);
        tfa.g[2] += dcb * (1.f - f);
        tfa.g[3] += dsw * (1.f - f);
        tfa.g[4] += dcr * f;
        tfa.g[5] += dcg * f;
        tfa.g[6] += dcb * f;
        tfa.g[7] += dsw * f;
        if (m.dfds != 0.f) {
          const float4 va = lut[m.idx], vb = lut[m.idx + 1];
          const float drho = m.dfds * (dcr * (vb.x - va.x) + dcg * (vb.y - va.y) + dcb * (vb.z - va.z) +
                                       dsw * (vb.w - va.w));
          if (drho != 0.f) {
            const float t = t0 + ((float)(i0 + k) + 0.5f) * e.step;
            const Tri s = tri_locate(e.vol, org + t * dir);
            uint32_t nv[8];
            cell_corners(s, sy, sz, nv);
            if (!cell.live || nv[0] != cell.v[0] || nv[7] != cell.v[7]) cell_move(cell, nv, cache, gdens);
            const float ax = 1.f - s.fx, ay = 1.f - s.fy, az = 1.f - s.fz;
            cell.g[0] += drho * (ax * ay * az);
            cell.g[1] += drho * (s.fx * ay * az);
            cell.g[2] += drho * (ax * s.fy * az);
            cell.g[3] += drho * (s.fx * s.fy * az);
            cell.g[4] += drho * (ax * ay * s.fz);
            cell.g[5] += drho * (s.fx * ay * s.fz);
            cell.g[6] += drho * (ax * s.fy * s.fz);
            cell.g[7] += drho * (s.fx * s.fy * s.fz);
          }
        }
      }
    }
    tf_flush(tfa, gtf);
    cell_flush(cell, cache, gdens);
    } while (0);
    __syncthreads();
    cache_drain(cache, gdens);
    __syncthreads();
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) tfPartial[(size_t)blockIdx.x * 1024 + i] = gtf[i];
  for (int off = 32; off > 0; off >>= 1) nsamp += __shfl_xor(nsamp, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(counter, nsamp);
}

__global__ void k_reduce_tf(const float *__restrict partial, int numBlocks, int numValues, float *__restrict gtf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 4 * numValues) return;
  float s = 0.f;
  for (int b = 0; b < numBlocks; ++b) s += partial[(size_t)b * 1024 + i];
  gtf[i] += s;
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

namespace {

EADD make_ea(const bnamd_ea_volume *v, const bnamd_ea_views *views) {
  BN_REQUIRE(v && views, "bnamd_ea: NULL argument");
  BN_REQUIRE(v->density && v->tf, "bnamd_ea: density and tf are required");
  BN_REQUIRE(v->dims[0] >= 2 && v->dims[1] >= 2 && v->dims[2] >= 2, "bnamd_ea: dims must be >= 2");
  BN_REQUIRE(v->numTF >= 2 && v->numTF <= 256, "bnamd_ea: numTF must be in [2, 256]");
  BN_REQUIRE((double)v->dims[0] * v->dims[1] * v->dims[2] < 4294967296.0,
             "bnamd_ea: grids are limited to 2^32 voxels (32-bit voxel indices)");
  BN_REQUIRE(v->domainHi > v->domainLo, "bnamd_ea: empty TF domain");
  BN_REQUIRE(v->step > 0.f, "bnamd_ea: step must be positive");
  BN_REQUIRE(views->width > 0 && views->height > 0 && views->numViews > 0 && views->cameras,
             "bnamd_ea: bad views");
  BN_REQUIRE(views->checkpointEvery == 8 || views->checkpointEvery == 16,
             "bnamd_ea: checkpointEvery must be 8 or 16");
  EADD e{};
  VolumeDD &d = e.vol;
  d.scalars = v->density;
  d.nx = v->dims[0]; d.ny = v->dims[1]; d.nz = v->dims[2];
  d.origin = {v->gridOrigin[0], v->gridOrigin[1], v->gridOrigin[2]};
  d.spacing = {v->gridSpacing[0], v->gridSpacing[1], v->gridSpacing[2]};
  d.worldLo = d.origin;
  d.worldHi = d.origin + d.spacing * vec3f{(float)(d.nx - 1), (float)(d.ny - 1), (float)(d.nz - 1)};
  e.tf = TFDD{(const float4 *)v->tf, v->numTF, v->domainLo, v->domainHi, v->baseDensity};
  e.step = v->step;
  e.W = views->width;
  e.H = views->height;
  e.numViews = views->numViews;
  e.numRays = (long long)e.W * e.H * e.numViews;
  // longest possible chord through the box bounds the samples per ray
  const vec3f ext = d.worldHi - d.worldLo;
  const double diag = sqrt((double)ext.x * ext.x + (double)ext.y * ext.y + (double)ext.z * ext.z);
  const long long maxN = (long long)(diag / v->step) + 2;
  e.numSeg = (int)((maxN + views->checkpointEvery - 1) / views->checkpointEvery);
  return e;
}

struct EAScratch {
  DevBuf<CameraDD> cams;
  DevBuf<float> tfPartial;
  DevBuf<unsigned long long> counter;
};
EAScratch &scratch() {
  static EAScratch s;
  return s;
}

int ea_grid(const EADD &e) {
  const long long tiles = (long long)e.numViews * div_up(e.W, kTile) * div_up(e.H, kTile);
  long long g = std::min<long long>(tiles, 8 * 256);
  g = ((g + 7) / 8) * 8;
  return (int)std::max<long long>(g, 8);
}

void upload_cams(EADD &e, const bnamd_ea_views *views, hipStream_t s) {
  std::vector<CameraDD> cams(views->numViews);
  for (int i = 0; i < views->numViews; ++i) cams[i] = make_camera(views->cameras[i]);
  scratch().cams.upload(cams.data(), cams.size(), s);
  e.cams = scratch().cams.ptr;
}

}  // namespace

extern "C" int bnamd_ea_checkpoint_count(const bnamd_ea_volume *vol, const bnamd_ea_views *views, int *count) {
  return guarded([&] {
    BN_REQUIRE(count, "bnamd_ea_checkpoint_count: NULL count");
    *count = make_ea(vol, views).numSeg;
  });
}

extern "C" int bnamd_ea_forward(const bnamd_ea_volume *vol, const bnamd_ea_views *views, float *image,
                                float *checkpoints, void *stream, bnamd_ea_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(image, "bnamd_ea_forward: NULL image");
    EADD e = make_ea(vol, views);
    hipStream_t s = (hipStream_t)stream;
    upload_cams(e, views, s);
    auto &sc = scratch();
    sc.counter.ensure(1);
    BN_HIP(hipMemsetAsync(sc.counter.ptr, 0, 8, s));
    const int grid = ea_grid(e);
    hipEvent_t e0, e1;
    BN_HIP(hipEventCreate(&e0));
    BN_HIP(hipEventCreate(&e1));
    BN_HIP(hipEventRecord(e0, s));
    if (views->checkpointEvery == 8)
      k_ea_forward<8><<<grid, 256, 0, s>>>(e, image, checkpoints, sc.counter.ptr);
    else
      k_ea_forward<16><<<grid, 256, 0, s>>>(e, image, checkpoints, sc.counter.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipEventRecord(e1, s));
    BN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    BN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
      unsigned long long n = 0;
      BN_HIP(hipMemcpy(&n, sc.counter.ptr, 8, hipMemcpyDeviceToHost));
      stats->samples = (long long)n;
      stats->ms = ms;
    }
  });
}

extern "C" int bnamd_ea_backward(const bnamd_ea_volume *vol, const bnamd_ea_views *views, const float *gradImage,
                                 const float *checkpoints, float *gradDensity, float *gradTF, void *stream,
                                 bnamd_ea_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(gradImage && checkpoints && gradDensity && gradTF, "bnamd_ea_backward: NULL argument");
    EADD e = make_ea(vol, views);
    hipStream_t s = (hipStream_t)stream;
    upload_cams(e, views, s);
    auto &sc = scratch();
    const int grid = ea_grid(e);
    sc.tfPartial.ensure((size_t)grid * 1024);
    sc.counter.ensure(1);
    BN_HIP(hipMemsetAsync(sc.counter.ptr, 0, 8, s));
    hipEvent_t e0, e1;
    BN_HIP(hipEventCreate(&e0));
    BN_HIP(hipEventCreate(&e1));
    BN_HIP(hipEventRecord(e0, s));
    if (views->checkpointEvery == 8)
      k_ea_backward<8><<<grid, 256, 0, s>>>(e, gradImage, checkpoints, gradDensity, sc.tfPartial.ptr,
                                            sc.counter.ptr);
    else
      k_ea_backward<16><<<grid, 256, 0, s>>>(e, gradImage, checkpoints, gradDensity, sc.tfPartial.ptr,
                                             sc.counter.ptr);
    BN_HIP(hipGetLastError());
    k_reduce_tf<<<div_up(4 * e.tf.numValues, 256), 256, 0, s>>>(sc.tfPartial.ptr, grid, e.tf.numValues, gradTF);
    BN_HIP(hipGetLastError());
    BN_HIP(hipEventRecord(e1, s));
    BN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    BN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) {
      unsigned long long n = 0;
      BN_HIP(hipMemcpy(&n, sc.counter.ptr, 8, hipMemcpyDeviceToHost));
      stats->samples = (long long)n;
      stats->ms = ms;
    }
  });
}
