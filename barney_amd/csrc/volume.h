// Structured-volume sampling with a macrocell majorant grid and delta
// (Woodcock) tracking, restating:
//   StructuredDataSampler::DD::sample    barney/volume/StructuredData.h:104-115
//   tex3D LINEAR + CLAMP (embree oracle) rtcore/embree/Texture.cpp:324-416
//   TransferFunction::DD::{map,majorant} barney/volume/TransferFunction.h:67-116
//   fastLog / Woodcock::sampleRange      barney/volume/Volume.h:174-223
//   dda::dda3                            barney/volume/DDA.h:65-206
//   MCVolumeAccel::isProg                barney/volume/MCAccelerator.h:569-708
//   boxTest                              barney/render/Ray.h:176-197
#pragma once
#include "common.h"

namespace bnamd {

struct TFDD {
  const float4 *values;
  int numValues;
  float domLo, domHi;
  float baseDensity;
};

// The structured field is sampled from a bricked copy: kBrick^3 voxels per
// brick (2 KB), bricks x-fastest; inside a brick every 128-byte cache line
// holds a 4 x 4 x 2 block of voxels (16 lines: x-halves fastest, then
// y-halves, then z-pairs; x fastest inside a line). A trilinear sample's 2 x
// 2 x 2 corners then lie in about 2.1 lines on average, against 2.6 for
// x-fastest rows of 8 in a brick and 4 in the x-fastest array. Voxel
// (x, y, z) is at X(x) + Y(y) + Z(z), one term per axis.
constexpr int kBrickShift = 3, kBrick = 1 << kBrickShift, kBrickVoxels = kBrick * kBrick * kBrick;

struct VolumeDD {
  const float *scalars;    // the caller's x-fastest array (macrocell build)
  const float *bricks;     // the bricked copy (sampling)
  uint32_t brickStrideY;   // voxels per row of bricks: bricksX * kBrickVoxels
  uint32_t brickStrideZ;   // voxels per slab of bricks: bricksX * bricksY * kBrickVoxels
  int nx, ny, nz;          // numScalars
  vec3f origin, spacing;   // cell grid
  vec3f worldLo, worldHi;  // origin + spacing * numCells
  // macrocell grid
  const float *majorants;
  int mx, my, mz;
  vec3f mcOrigin, mcSpacing;
  TFDD tf;
};

struct VolSample {
  float r, g, b, w;
};

// Volume.h:174-184: 3-term atanh series for log.
BN_HD float fast_log(float f) {
#pragma clang fp contract(off)
  f = (f - 1.f) / (f + 1.f);
  const float f2 = f * f;
  float s = f;
  f *= f2;
  s += (1.f / 3.f) * f;
  f *= f2;
  s += (1.f / 5.f) * f;
  return s + s;
}

// embree Texture.cpp:326-335 (lerpCoords_notNormalized)
BN_HD void lerp_coords(int N, float tc, int &i0, int &i1, float &f) {
#pragma clang fp contract(off)
  if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
  if (tc - .5f >= N - 1) { i0 = i1 = N - 1; f = 0.f; return; }
  f = (tc - .5f) - int(tc - .5f);
  i0 = int(tc - .5f);
  i1 = i0 + 1;
}

BN_HD float lerp_l(float f, float a, float b) {
#pragma clang fp contract(off)
  return (1.f - f) * a + f * b;
}

// Offsets of voxel coordinate x, y, z in the bricked copy (their sum):
// brick, then line within the brick (32 voxels each), then voxel in the line.
BN_HD uint32_t brick_x(int x) {
  const uint32_t u = (uint32_t)x;
  return (u >> 3) * kBrickVoxels + ((u >> 2) & 1u) * 32u + (u & 3u);
}
BN_HD uint32_t brick_y(int y, uint32_t strideY) {
  const uint32_t u = (uint32_t)y;
  return (u >> 3) * strideY + ((u >> 2) & 1u) * 64u + (u & 3u) * 4u;
}
BN_HD uint32_t brick_z(int z, uint32_t strideZ) {
  const uint32_t u = (uint32_t)z;
  return (u >> 3) * strideZ + ((u >> 1) & 3u) * 128u + (u & 1u) * 16u;
}

// Trilinear sample of the scalar field at world point P; NaN outside.
BN_HD float sample_field(const VolumeDD &v, vec3f P) {
#pragma clang fp contract(off)
  const vec3f rel = (P - v.origin) * rcp(v.spacing);
  if (rel.x < 0.f) return __builtin_nanf("");
  if (rel.y < 0.f) return __builtin_nanf("");
  if (rel.z < 0.f) return __builtin_nanf("");
  if (rel.x >= (float)(v.nx - 1)) return __builtin_nanf("");
  if (rel.y >= (float)(v.ny - 1)) return __builtin_nanf("");
  if (rel.z >= (float)(v.nz - 1)) return __builtin_nanf("");
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lerp_coords(v.nx, rel.x + .5f, x0, x1, fx);
  lerp_coords(v.ny, rel.y + .5f, y0, y1, fy);
  lerp_coords(v.nz, rel.z + .5f, z0, z1, fz);
  const uint32_t X0 = brick_x(x0), X1 = brick_x(x1);
  const uint32_t Y0 = brick_y(y0, v.brickStrideY), Y1 = brick_y(y1, v.brickStrideY);
  const uint32_t Z0 = brick_z(z0, v.brickStrideZ), Z1 = brick_z(z1, v.brickStrideZ);
  const float *s = v.bricks;
  const float v000 = s[X0 + Y0 + Z0];
  const float v001 = s[X1 + Y0 + Z0];
  const float v010 = s[X0 + Y1 + Z0];
  const float v011 = s[X1 + Y1 + Z0];
  const float v100 = s[X0 + Y0 + Z1];
  const float v101 = s[X1 + Y0 + Z1];
  const float v110 = s[X0 + Y1 + Z1];
  const float v111 = s[X1 + Y1 + Z1];
  const float v00 = lerp_l(fx, v000, v001);
  const float v01 = lerp_l(fx, v010, v011);
  const float v10 = lerp_l(fx, v100, v101);
  const float v11 = lerp_l(fx, v110, v111);
  const float v0 = lerp_l(fy, v00, v01);
  const float v1 = lerp_l(fy, v10, v11);
  return lerp_l(fz, v0, v1);
}

BN_HD VolSample tf_map(const TFDD &tf, float s) {
#pragma clang fp contract(off)
  float f = (s - tf.domLo) / (tf.domHi - tf.domLo);
  f = fminf(fmaxf(f, 0.f), 1.f);
  f *= (tf.numValues - 1);
  int idx = int(f);
  idx = idx < 0 ? 0 : (idx > tf.numValues - 2 ? tf.numValues - 2 : idx);
  f -= idx;
  const float4 a = tf.values[idx], b = tf.values[idx + 1];
  VolSample r;
  r.r = (1.f - f) * a.x + f * b.x;
  r.g = (1.f - f) * a.y + f * b.y;
  r.b = (1.f - f) * a.z + f * b.z;
  r.w = (1.f - f) * a.w + f * b.w;
  r.w *= tf.baseDensity;
  return r;
}

BN_HD float tf_majorant(const TFDD &tf, float lo, float hi) {
#pragma clang fp contract(off)
  if (lo > hi) return 0.f;
  const float span = tf.domHi - tf.domLo;
  float f_lo = (lo - tf.domLo) / span;
  float f_hi = (hi - tf.domLo) / span;
  f_lo = fminf(fmaxf(f_lo, 0.f), 1.f);
  f_hi = fminf(fmaxf(f_hi, 0.f), 1.f);
  f_lo *= (tf.numValues - 1);
  f_hi *= (tf.numValues - 1);
  auto clampi = [&](int i) { return i < 0 ? 0 : (i > tf.numValues - 2 ? tf.numValues - 2 : i); };
  const int i_lo = clampi(int(f_lo));
  const int i_hi = clampi(int(f_hi));
  f_lo -= i_lo;
  f_hi -= i_hi;
  const float lerp_lo = (1.f - f_lo) * tf.values[i_lo].w + f_lo * tf.values[i_lo + 1].w;
  const float lerp_hi = (1.f - f_hi) * tf.values[i_hi].w + f_hi * tf.values[i_hi + 1].w;
  float m = 0.f;
  m = fmaxf(m, lerp_lo);
  m = fmaxf(m, lerp_hi);
  for (int i = i_lo + 1; i <= i_hi; ++i) m = fmaxf(m, tf.values[i].w);
  return m * tf.baseDensity;
}

BN_HD VolSample sample_and_map(const VolumeDD &v, vec3f P) {
  const float f = sample_field(v, P);
  if (f != f) return VolSample{0.f, 0.f, 0.f, 0.f};
  return tf_map(v.tf, f);
}

struct VolCounters {
  uint32_t samples;
  uint32_t ddaSteps;
  uint32_t zeroSteps;  // macrocells visited whose majorant is zero (no draw)
};

struct VolHit {
  bool hit;
  float t;
  vec3f rgb;
};

BN_HD int vfloor_i(float f) { return (int)floorf(f); }

// MCVolumeAccel::isProg (MCAccelerator.h:569-708) as a resumable walk: the
// box test, dda::dda3 over the macrocells (DDA.h:65-206) and
// Woodcock::sampleRange (Volume.h:188-223) in every cell with a non-zero
// majorant, with the walk's state held here so that one call is one unit
// of work: advance() enters or leaves a macrocell or draws a collision
// distance; sample() evaluates the field at a pending collision. A kernel
// can then keep rays at different stages in one loop and choose, per
// wavefront iteration, which kind of work to run for the lanes that need it,
// instead of nesting a per-ray DDA loop around a per-cell Woodcock loop. The arithmetic, the order of the RNG draws and the
// counters are those of the nested form, operation for operation.
struct DeltaTrack {
  enum : int { DONE = 0, ENTER = 1, WOODCOCK = 2, ADVANCE = 3, SAMPLE = 4 };
  vec3f org, dir;
  float rayTMax;
  Random2 rng;
  vec3f t_next, t_step;
  int cx, cy, cz;
  uint32_t neg;           // bit k: the DDA steps down along axis k; bits 3-5:
                          // the axes the next step crosses; bit 6: the next
                          // cell lies in the grid (both found on entering)
  float ray_t0, tMax;     // DDA origin offset, clipped segment end
  float next_cell_begin;  // where the next cell begins (t_closest on entering)
  float majorant, nextMajorant;
  float t, upper;         // Woodcock position and end of the current cell
  float leaveDepth;       // optical depth beyond which a draw leaves the cell
  float pendingR;         // the draw of a collision distance awaiting sample()
  int mode;
  bool hit;
  vec3f rgb;

  BN_HD bool done() const { return mode == DONE; }

  // Ray set-up up to the first macrocell. Like isProg, the ray's seed is
  // advanced (Random2(seed, geomHash)) only when the box test passes.
  template <class V>
  BN_HD void init(const V &v, vec3f o, vec3f d, float tmin, float tmax, float rtmax, RNGSeed &seed,
                  uint64_t geomHash) {
#pragma clang fp contract(off)
    org = o;
    dir = d;
    rayTMax = rtmax;
    hit = false;
    mode = DONE;
    float t0 = tmin, t1 = tmax;
    {
      const vec3f t_lo = (v.worldLo - org) * rcp(dir);
      const vec3f t_hi = (v.worldHi - org) * rcp(dir);
      const vec3f t_nr = vmin(t_lo, t_hi);
      const vec3f t_fr = vmax(t_lo, t_hi);
      t0 = fmaxf(t0, reduce_max(t_nr));
      t1 = fminf(t1, reduce_min(t_fr));
      if (!(t0 <= t1)) return;
    }
    const vec3f dorg = (org - v.mcOrigin) * rcp(v.mcSpacing);
    const vec3f ddir = dir * rcp(v.mcSpacing);
    rng = Random2(seed, geomHash);
    // dda::dda3(dorg, ddir, t1, mcDims, lambda)
    tMax = t1;
    const vec3f f_size = {(float)v.mx, (float)v.my, (float)v.mz};
    const vec3f rcp_dir = rcp(ddir);
    vec3f t_lo = (vec3f(0.f) - dorg) * rcp(ddir);
    vec3f t_hi = (f_size - dorg) * rcp(ddir);
    vec3f t_nr = vmin(t_lo, t_hi);
    vec3f t_fr = vmax(t_lo, t_hi);
    if (ddir.x == 0.f) {
      if (dorg.x < 0.f || dorg.x > f_size.x) return;
      t_nr.x = -BN_INF; t_fr.x = BN_INF;
    }
    if (ddir.y == 0.f) {
      if (dorg.y < 0.f || dorg.y > f_size.y) return;
      t_nr.y = -BN_INF; t_fr.y = BN_INF;
    }
    if (ddir.z == 0.f) {
      if (dorg.z < 0.f || dorg.z > f_size.z) return;
      t_nr.z = -BN_INF; t_fr.z = BN_INF;
    }
    ray_t0 = fmaxf(0.f, reduce_max(t_nr));
    const float ray_t1 = fminf(tMax, reduce_min(t_fr));
    if (ray_t0 > ray_t1) return;
    const vec3f oiv = dorg + ray_t0 * ddir;
    const vec3f f_cell =
        vmax(vec3f(0.f), vmin(f_size - vec3f(1.f), vec3f(floorf(oiv.x), floorf(oiv.y), floorf(oiv.z))));
    const vec3f f_cell_end = {ddir.x > 0.f ? f_cell.x + 1.f : f_cell.x, ddir.y > 0.f ? f_cell.y + 1.f : f_cell.y,
                              ddir.z > 0.f ? f_cell.z + 1.f : f_cell.z};
    t_step = vabs(rcp_dir);
    t_next = {(ddir.x == 0.f) ? BN_INF : (fabsf(f_cell_end.x - oiv.x) * t_step.x),
              (ddir.y == 0.f) ? BN_INF : (fabsf(f_cell_end.y - oiv.y) * t_step.y),
              (ddir.z == 0.f) ? BN_INF : (fabsf(f_cell_end.z - oiv.z) * t_step.z)};
    neg = (ddir.x > 0.f ? 0u : 1u) | (ddir.y > 0.f ? 0u : 2u) | (ddir.z > 0.f ? 0u : 4u);
    cx = (int)f_cell.x;
    cy = (int)f_cell.y;
    cz = (int)f_cell.z;
    next_cell_begin = 0.f;
    majorant = v.majorants[cx + v.mx * (cy + v.my * cz)];
    mode = ENTER;
  }

  // One cheap unit of work: a collision distance drawn in the current cell,
  // or a macrocell step of the DDA (leaving the cell, then entering the next
  // one, with its majorant fetched one cell ahead).
  //
  // A draw's exact distance (barney's fastLog and correctly rounded
  // divisions) only matters when the collision falls inside the cell. The
  // draw is compared in optical depth: tau = -fastLog(1 - r), evaluated with
  // a hardware reciprocal and within 6e-7 of the exact float value, against
  // leaveDepth = majorant ((upper - t) + 1e-6 (|t| + |upper|)), set when the
  // cell is entered and after each rejected collision. tau (1 - 1e-5) >=
  // leaveDepth implies that the exact t + dt rounds to upper or beyond (the
  // margins exceed the reciprocal, division and addition rounding), so such
  // a draw moves on to the next cell without computing t, which the next
  // cell does not use. Every other draw waits in SAMPLE, where sample()
  // finishes it exactly. The RNG draws keep their order.
  //
  // Entering a cell finds the axes whose crossing ends it (t_next equal to
  // their minimum), the next cell and whether it lies in the grid, for the
  // majorant fetched ahead; leaving the cell then applies exactly that step
  // (the same t_next additions as dda3's loop, which ends when a stepped
  // coordinate leaves the grid).
  template <bool kExactLog, class V>
  BN_HD void advance(const V &v, VolCounters &cnt) {
#pragma clang fp contract(off)
    if (mode == WOODCOCK) {
      const float r = rng();
      bool leaves = false;
      if (!kExactLog) {
        const float f = 1.f - r;
        const float q = (f - 1.f) * __builtin_amdgcn_rcpf(f + 1.f);
        const float q2 = q * q;
        const float tau = -2.f * (q * (1.f + q2 * (1.f / 3.f + q2 * (1.f / 5.f))));
        leaves = tau * (1.f - 1e-5f) >= leaveDepth;
      }
      if (!leaves) {
        pendingR = r;
        mode = SAMPLE;
        return;
      }
      mode = ADVANCE;
    }
    if (mode == ADVANCE) {
      majorant = nextMajorant;
      if (!(neg & 64u)) {
        mode = DONE;
        return;
      }
      if (neg & 8u) {
        t_next.x += t_step.x;
        cx += (neg & 1u) ? -1 : 1;
      }
      if (neg & 16u) {
        t_next.y += t_step.y;
        cy += (neg & 2u) ? -1 : 1;
      }
      if (neg & 32u) {
        t_next.z += t_step.z;
        cz += (neg & 4u) ? -1 : 1;
      }
    }
    // ENTER: the cell's segment, its exit step, and the next cell's majorant
    const float t_closest = reduce_min(t_next);
    const float cell_t0 = ray_t0 + next_cell_begin;
    const float cell_t1 = ray_t0 + fminf(t_closest, tMax);
    next_cell_begin = t_closest;
    cnt.ddaSteps++;
    const bool sx = t_next.x == t_closest, sy = t_next.y == t_closest, sz = t_next.z == t_closest;
    const int ncx = sx ? cx + ((neg & 1u) ? -1 : 1) : cx;
    const int ncy = sy ? cy + ((neg & 2u) ? -1 : 1) : cy;
    const int ncz = sz ? cz + ((neg & 4u) ? -1 : 1) : cz;
    const bool inside = ncx >= 0 && ncx < v.mx && ncy >= 0 && ncy < v.my && ncz >= 0 && ncz < v.mz;
    neg = (neg & 7u) | (sx ? 8u : 0u) | (sy ? 16u : 0u) | (sz ? 32u : 0u) | (inside ? 64u : 0u);
    nextMajorant = inside ? v.majorants[ncx + v.mx * (ncy + v.my * ncz)] : 0.f;
    if (majorant != 0.f) {
      t = cell_t0;
      upper = fminf(cell_t1, rayTMax);
      leaveDepth = majorant * ((upper - t) + 1e-6f * (fabsf(t) + fabsf(upper)));
      mode = WOODCOCK;
    } else {
      cnt.zeroSteps++;
      mode = ADVANCE;
    }
  }

  // The pending collision (mode SAMPLE): the exact distance of the draw
  // advance() left pending (Woodcock::sampleRange, Volume.h:188-223); past
  // the cell's end the walk moves on, else S maps P to the TF-mapped sample,
  // accepted with probability sample.w / majorant.
  template <bool kExactLog, class S>
  BN_HD void sample(const S &sampler, VolCounters &cnt) {
#pragma clang fp contract(off)
    const float dt = kExactLog ? -logf(1.f - pendingR) / majorant : -fast_log(1.f - pendingR) / majorant;
    t += dt;
    if (t >= upper) {
      mode = ADVANCE;
      return;
    }
    const vec3f P = org + t * dir;
    const VolSample s = sampler(P);
    cnt.samples++;
    if (s.w >= rng() * majorant) {
      hit = true;
      rgb = {s.r, s.g, s.b};
      mode = DONE;
    } else {
      leaveDepth = majorant * ((upper - t) + 1e-6f * (fabsf(t) + fabsf(upper)));
      mode = WOODCOCK;
    }
  }

  BN_HD VolHit result() const {
    VolHit r;
    r.hit = hit;
    r.t = t;
    r.rgb = rgb;
    return r;
  }
};

// MCVolumeAccel::isProg for one ray segment [tmin, tmax]: box test, DDA over
// macrocells, Woodcock tracking in each cell with a non-zero majorant. The
// ray's RNG seed is advanced as Random2(seed, hash(inst, geom, 0)) does.
// V supplies the macrocell grid (worldLo/Hi, mcOrigin, mcSpacing,
// majorants, mx/my/mz); sample(P) returns the TF-mapped field sample (the
// structured sampler, or the tet-mesh point location).
template <class V, class S>
BN_HD VolHit volume_intersect_g(const V &v, const S &sample, vec3f org, vec3f dir, float tmin, float tmax,
                                float rayTMax, RNGSeed &seed, bool exactLog, VolCounters &cnt,
                                uint64_t geomHash) {
  DeltaTrack tr;
  tr.init(v, org, dir, tmin, tmax, rayTMax, seed, geomHash);
  for (;;) {
    // cheap steps until the walk needs the field (the lanes of a wavefront
    // then sample together) or ends
    if (exactLog) {
      while (tr.mode != DeltaTrack::DONE && tr.mode != DeltaTrack::SAMPLE) tr.advance<true>(v, cnt);
      if (tr.done()) break;
      tr.sample<true>(sample, cnt);
    } else {
      while (tr.mode != DeltaTrack::DONE && tr.mode != DeltaTrack::SAMPLE) tr.advance<false>(v, cnt);
      if (tr.done()) break;
      tr.sample<false>(sample, cnt);
    }
  }
  return tr.result();
}


BN_HD VolHit volume_intersect(const VolumeDD &v, vec3f org, vec3f dir, float tmin, float tmax, float rayTMax,
                              RNGSeed &seed, bool exactLog, VolCounters &cnt) {
  return volume_intersect_g(
      v, [&](vec3f P) { return sample_and_map(v, P); }, org, dir, tmin, tmax, rayTMax, seed, exactLog, cnt,
      bn_hash3(0u, 0u, 0u));
}

// dda::dda3 (barney/volume/DDA.h:65-206) over an mx x my x mz grid in grid
// space: cb(cx, cy, cz, t0, t1) per visited cell, false stops the walk.
template <class F>
BN_HD void dda3(vec3f dorg, vec3f ddir, float tMax, int mx, int my, int mz, F &&cb) {
#pragma clang fp contract(off)
  const vec3f f_size = {(float)mx, (float)my, (float)mz};
  const vec3f rcp_dir = rcp(ddir);
  vec3f t_lo = (vec3f(0.f) - dorg) * rcp(ddir);
  vec3f t_hi = (f_size - dorg) * rcp(ddir);
  vec3f t_nr = vmin(t_lo, t_hi);
  vec3f t_fr = vmax(t_lo, t_hi);
  if (ddir.x == 0.f) {
    if (dorg.x < 0.f || dorg.x > f_size.x) return;
    t_nr.x = -BN_INF; t_fr.x = BN_INF;
  }
  if (ddir.y == 0.f) {
    if (dorg.y < 0.f || dorg.y > f_size.y) return;
    t_nr.y = -BN_INF; t_fr.y = BN_INF;
  }
  if (ddir.z == 0.f) {
    if (dorg.z < 0.f || dorg.z > f_size.z) return;
    t_nr.z = -BN_INF; t_fr.z = BN_INF;
  }
  const float ray_t0 = fmaxf(0.f, reduce_max(t_nr));
  const float ray_t1 = fminf(tMax, reduce_min(t_fr));
  if (ray_t0 > ray_t1) return;
  const vec3f oiv = dorg + ray_t0 * ddir;
  const vec3f f_cell = vmax(vec3f(0.f), vmin(f_size - vec3f(1.f), vec3f(floorf(oiv.x), floorf(oiv.y), floorf(oiv.z))));
  const vec3f f_cell_end = {ddir.x > 0.f ? f_cell.x + 1.f : f_cell.x, ddir.y > 0.f ? f_cell.y + 1.f : f_cell.y,
                            ddir.z > 0.f ? f_cell.z + 1.f : f_cell.z};
  const vec3f t_step = vabs(rcp_dir);
  vec3f t_next = {(ddir.x == 0.f) ? BN_INF : (fabsf(f_cell_end.x - oiv.x) * t_step.x),
                  (ddir.y == 0.f) ? BN_INF : (fabsf(f_cell_end.y - oiv.y) * t_step.y),
                  (ddir.z == 0.f) ? BN_INF : (fabsf(f_cell_end.z - oiv.z) * t_step.z)};
  const int stopx = ddir.x > 0.f ? mx : -1, stopy = ddir.y > 0.f ? my : -1, stopz = ddir.z > 0.f ? mz : -1;
  const int dx = ddir.x > 0.f ? 1 : -1, dy = ddir.y > 0.f ? 1 : -1, dz = ddir.z > 0.f ? 1 : -1;
  int cx = (int)f_cell.x, cy = (int)f_cell.y, cz = (int)f_cell.z;
  float next_cell_begin = 0.f;
  while (true) {
    const float t_closest = reduce_min(t_next);
    const float cell_t0 = ray_t0 + next_cell_begin;
    const float cell_t1 = ray_t0 + fminf(t_closest, tMax);
    if (!cb(cx, cy, cz, cell_t0, cell_t1)) return;
    next_cell_begin = t_closest;
    if (t_next.x == t_closest) {
      t_next.x += t_step.x;
      cx += dx;
      if (cx == stopx) return;
    }
    if (t_next.y == t_closest) {
      t_next.y += t_step.y;
      cy += dy;
      if (cy == stopy) return;
    }
    if (t_next.z == t_closest) {
      t_next.z += t_step.z;
      cz += dz;
      if (cz == stopz) return;
    }
  }
}

struct IsoHit {
  bool hit;
  float t;
  vec3f P, N;
};

// Whether macrocell idx's value range contains the iso value (the range test
// of MCIsoSurfaceAccel::isProg). Grids with a precomputed per-cell bit for
// the iso value overload it (found by argument-dependent lookup).
template <class V>
BN_HD bool iso_cell_active(const V &v, int idx, float isoValue) {
  const float2 vr = v.ranges[idx];
  return isoValue >= vr.x && isoValue <= vr.y;
}

// MCIsoSurfaceAccel::isProg (barney/volume/MCAccelerator.h:338-567): DDA over
// the macrocells whose value range contains the iso value, 10 sub-steps per
// cell with a linear root, then a central-difference gradient normal.
// V supplies the grid and ranges (float2 lo/hi per macrocell); field(P)
// returns the raw scalar (NaN outside). The ray's RNG seed is advanced as
// the Random construction there does.
template <class V, class F>
BN_HD IsoHit iso_intersect(const V &v, const F &field, float isoValue, vec3f org, vec3f dir, float tmin,
                           float tmax, float rayTMax, RNGSeed &seed, uint64_t geomHash) {
#pragma clang fp contract(off)
  IsoHit res;
  res.hit = false;
  float t0 = tmin, t1 = tmax;
  {
    const vec3f t_lo = (v.worldLo - org) * rcp(dir);
    const vec3f t_hi = (v.worldHi - org) * rcp(dir);
    const vec3f t_nr = vmin(t_lo, t_hi);
    const vec3f t_fr = vmax(t_lo, t_hi);
    t0 = fmaxf(t0, reduce_max(t_nr));
    t1 = fminf(t1, reduce_min(t_fr));
    if (!(t0 <= t1)) return res;
  }
  const vec3f dorg = (org - v.mcOrigin) * rcp(v.mcSpacing);
  const vec3f ddir = dir * rcp(v.mcSpacing);
  Random2 rng(seed, geomHash);
  (void)rng;  // opacity < 1 is not supported: the stream is only advanced
  float tHit = rayTMax;
  dda3(dorg, ddir, t1, v.mx, v.my, v.mz, [&](int cx, int cy, int cz, float ct0, float ct1) -> bool {
    const float _t0 = ct0, _t1 = ct1;
    if (ct0 >= fminf(ct1, rayTMax)) return true;
    if (!iso_cell_active(v, cx + v.mx * (cy + v.my * cz), isoValue)) return true;
    // sub-step i samples tt_i = lerp(i / 10, t0, t1) (tt_0 = t0); one
    // field() call site so the point location is inlined once
    float ff0 = 0.f, ff1 = 0.f, tt0 = 0.f, tt1 = 0.f;
    const int numSteps = 10;
    for (int i = 0; i <= numSteps; ++i) {
      const float tt = (i == 0) ? ct0 : lerp_l(i / float(numSteps), _t0, _t1);
      const float f = field(org + tt * dir);
      tt0 = tt1;
      ff0 = ff1;
      tt1 = tt;
      ff1 = f;
      if (i == 0) continue;
      if (ff0 != ff0 || ff1 != ff1) continue;
      const float vlo = fminf(ff0, ff1), vhi = fmaxf(ff0, ff1);
      if (isoValue >= vlo && isoValue <= vhi) {
        float t = (isoValue - ff0) / (ff1 - ff0);
        t = lerp_l(t, tt0, tt1);
        tHit = fminf(tHit, t);
        if (tHit < rayTMax) return false;
      }
    }
    return true;
  });
  if (tHit >= rayTMax) return res;
  const vec3f osP = org + tHit * dir;
  const vec3f size = v.worldHi - v.worldLo;
  const float delta = length(size) * .1f / float(v.mx + v.my + v.mz);
  // central differences: samples k = 0 (center), then -x, +x, -y, +y, -z, +z
  float fP = 0.f, lo = 0.f, cnt = 0.f;
  float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll 1
  for (int k = 0; k < 7; ++k) {
    const int axis = (k - 1) >> 1;
    const float off = (k & 1) ? -delta : delta;
    const vec3f Q = {osP.x + (axis == 0 ? off : 0.f), osP.y + (axis == 1 ? off : 0.f),
                     osP.z + (axis == 2 ? off : 0.f)};
    float f = field(k == 0 ? osP : Q);
    if (k == 0) {
      fP = f;
      continue;
    }
    const bool miss = f != f;
    if (miss) f = fP;
    if (k & 1) {  // -delta sample of this axis
      lo = f;
      cnt = miss ? 1.f : 2.f;
    } else {      // +delta sample: (f(+) - f(-)) / number of valid samples
      const float dd = cnt - (miss ? 1.f : 0.f);
      const float gv = dd == 0.f ? 0.f : (f - lo) / dd;
      if (axis == 0) gx = gv;
      else if (axis == 1) gy = gv;
      else gz = gv;
    }
  }
  vec3f N = {gx, gy, gz};
  if (N == vec3f(0.f)) N = -normalize(dir);
  res.hit = true;
  res.t = tHit;
  res.P = osP;
  res.N = normalize(N);
  return res;
}

}  // namespace bnamd
