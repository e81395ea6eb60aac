// Compressed 8-wide BVH for gfx950: node layout, primitive records and the
// closest-hit traversal with a per-wavefront LDS stack.
//
// Replaces the acceleration structures built by rtcore/cuda/cudaGroup.cu
// (cuBQL binary BVHs, :144-562) and walked by
// rtcore/cuda/TraceInterface.h:102-261; the closest-hit semantics follow the
// embree backend (rtcore/embree/Device.cpp:107-196, Group.cpp:34-79).
#pragma once
#include "common.h"

namespace bnamd {

enum PrimType : uint32_t {
  PRIM_TRIANGLE = 0, PRIM_SPHERE = 1, PRIM_CYLINDER = 2, PRIM_TET = 3, PRIM_INSTANCE = 4,
  PRIM_USER = 5  // a user geometry's box; its intersect program finds the hit
};
// c.w of a PrimRec: type in bits 0-7, flags above.
// PRIM_SHADOW_TRANSPARENT: a triangle whose material is Glass; barney's
// triangle anyHit ignores it for shadow rays (Glass::getOpacity returns 0
// for shadow rays, Triangles.dev.cu:125-133).
constexpr uint32_t PRIM_SHADOW_TRANSPARENT = 1u << 8;
// PRIM_CUTOUT: a triangle whose material's opacity alpha (b.w of its record)
// is below 1; barney's triangle anyHit ignores a candidate hit when a random
// draw exceeds alpha (Triangles.dev.cu:121-133).
constexpr uint32_t PRIM_CUTOUT = 1u << 9;

// One node = one 128-byte line. Child boxes are quantized to 16 bits per
// plane relative to the node origin with a power-of-two scale per axis.
//   n0: ox oy oz | ex ey ez imask
//   n1: childBase | primBase | meta[0..3] | meta[4..7]
//   n2..n4: slots 0-3, n5..n7: slots 4-7, each as
//           qlo.x[4] qlo.y[4] qlo.z[4] qhi.x[4] qhi.y[4] qhi.z[4] (u16)
// so the traversal tests one half-node (48 bytes) at a time.
// meta[slot]: leaf child -> (count << 5) | offset into the node's prim block;
// internal child -> imask bit set, node = childBase + rank among internal
// slots; empty -> 0 with an inverted (empty) box.
struct alignas(128) BNode8 {
  float ox, oy, oz;
  uint8_t ex, ey, ez, imask;
  uint32_t childBase;
  uint32_t primBase;
  uint8_t meta[8];
  uint16_t q[2][6][4];  // [slot >> 2][plane][slot & 3]
  BN_HD uint16_t &plane(int k, int slot) { return q[slot >> 2][k][slot & 3]; }
  BN_HD uint16_t plane(int k, int slot) const { return q[slot >> 2][k][slot & 3]; }
};
static_assert(sizeof(BNode8) == 128, "BNode8 must be one 128B line");

// Primitive record (48 B, three float4) stored in BVH leaf order.
//   triangle: a = v0 | flat,  b = e1 = v1-v0 | alpha, c = e2 = v2-v0 | type
//   sphere:   a = center | flat, b = (radius,0,0) | -, c = 0 | type
//   cylinder: a = v0 | flat, b = v1 | -, c = (radius,0,0) | type
//   tet:      a = box lo | flat, b = box hi | -, c = 0 | type (point-location
//             BVHs only; the tets themselves are packed separately)
//   instance: a = world box lo | instance index, b = world box hi | -,
//             c = 0 | type (top level of a two-level world)
// 'flat' is the primitive's index in instance-major, geom-major, prim-minor
// order; it doubles as the deterministic tie-break key at equal t.
struct PrimRec {
  float4 a, b, c;
};

struct HitRec {
  float t;
  float u, v;      // triangle barycentrics; sphere: u = hit t in the shifted frame
  uint32_t flat;   // 0xffffffff = no hit
};

// Deterministic acceptance: closest t wins, ties at equal t go to the
// smallest (instance, geom, prim) tuple, i.e. the smallest flat index.
BN_HD bool better(float t, uint32_t flat, const HitRec &best) {
  return t < best.t || (t == best.t && flat < best.flat);
}

// ---------------------------------------------------------------------------
// Candidate filters: what barney's anyHit / intersection programs do with a
// candidate before it is reported. NoFilt accepts every candidate.
// HitFilter applies the renderer's cut plane (OptixGlobals.h:33-42: the
// candidate's world-space point P = o + t d is dropped when
// dot(n, P) + d < 0; Triangles.dev.cu:39-43, Spheres.dev.cu:170-174,
// Cylinders.dev.cu:257-259) and, for PRIM_CUTOUT triangles, stochastic
// opacity (Triangles.dev.cu:121-133: the ray seed advanced by the bits of
// the object-space point, then Random(seed, 290374) > alpha drops it).
// Each candidate's draw depends on the ray's seed at the start of the trace
// and on that candidate only (barney chains the seed through the
// candidates in traversal order), so the closest accepted hit does not
// depend on the order in which the BVH is walked.
// ---------------------------------------------------------------------------
struct NoFilt {
  BN_HD bool any(float) const { return true; }
  BN_HD bool tri(float, float, float, vec3f, vec3f, vec3f, float, uint32_t) const { return true; }
};

struct HitFilter {
  vec3f o, d;       // world-space ray as traced (the cut plane is a world-space test)
  float4 cut;       // (nx, ny, nz, d)
  uint64_t seed;    // the ray's RNG state when the trace starts
  bool cutOn;
  bool stochastic;  // false: opacity ignored (the ID channel's closest candidate)
  BN_HD bool any(float t) const {
#pragma clang fp contract(off)
    if (!cutOn) return true;
    const vec3f P = o + t * d;
    return !(dot(vec3f(cut.x, cut.y, cut.z), P) + cut.w < 0.f);
  }
  BN_HD bool tri(float t, float u, float v, vec3f v0, vec3f e1, vec3f e2, float alpha, uint32_t bits) const {
#pragma clang fp contract(off)
    if (!any(t)) return false;
    if (stochastic && (bits & PRIM_CUTOUT)) {
      const vec3f P = v0 + (u * e1 + v * e2);
      RNGSeed s;
      s.state = seed;
      s.next_hash(float_as_uint(P.x));
      s.next_hash(float_as_uint(P.y));
      s.next_hash(float_as_uint(P.z));
      Random2 rng(s, 290374ull);
      if (rng() > alpha) return false;
    }
    return true;
  }
};

// ---------------------------------------------------------------------------
// Primitive intersectors. Each computes the candidate t exactly as the
// reference program does (same fp32 operation order, no FMA contraction),
// tested against the ray's original [tmin, tmax]; acceptance against the
// current closest hit then uses better().
// ---------------------------------------------------------------------------

// rtcore/cuda/TraceInterface.h:32-98 (barney's own plane + 2D-projection
// triangle test).
template <class Filt = NoFilt>
BN_HD void isect_triangle(vec3f org, vec3f dir, float tmax, vec3f v0, vec3f e1, vec3f e2,
                          uint32_t flat, HitRec &best, const Filt &filt = Filt(), float alpha = 1.f,
                          uint32_t bits = 0u) {
#pragma clang fp contract(off)
  const vec3f N = cross(e1, e2);
  const float dd = dot(dir, N);
  if (fabsf(dd) < 1e-12f) return;
  if (dd == 0.f) return;
  const float t = -dot(org - v0, N) / dd;
  if (t <= 0.f || t >= tmax) return;
  if (t > best.t) return;
  const vec3f P = org - v0 + t * dir;
  float e1u, e2u, Pu, e1v, e2v, Pv;
  if (fabsf(N.x) >= fmaxf(fabsf(N.y), fabsf(N.z))) {
    e1u = e1.y; e2u = e2.y; Pu = P.y;
    e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else if (fabsf(N.y) > fabsf(N.z)) {
    e1u = e1.x; e2u = e2.x; Pu = P.x;
    e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else {
    e1u = e1.x; e2u = e2.x; Pu = P.x;
    e1v = e1.y; e2v = e2.y; Pv = P.y;
  }
  const float den = e1u * e2v - e2u * e1v;
  if (den == 0.f) return;
  const float u = (Pu * e2v - Pv * e2u) / den;
  const float v = (e1u * Pv - e1v * Pu) / den;
  if ((u < 0.f) || (v < 0.f) || ((u + v) >= 1.f)) return;
  if (!better(t, flat, best)) return;
  if (!filt.tri(t, u, v, v0, e1, e2, alpha, bits)) return;
  best.t = t;
  best.u = u;
  best.v = v;
  best.flat = flat;
}

// barney/geometry/Spheres.dev.cu:116-170 ("move the origin" trick).
// Returns the shifted-frame hit in best.u so the closest-hit epilogue can
// rebuild the object-space point bit-identically.
template <class Filt = NoFilt>
BN_HD void isect_sphere(vec3f org, vec3f dir, float tmin, float tmax, vec3f center, float radius,
                        uint32_t flat, HitRec &best, const Filt &filt = Filt()) {
#pragma clang fp contract(off)
  const float t_move = fmaxf(0.f, length(center - org) - 3.f * radius);
  const vec3f o = org + t_move * dir;
  const float t_max = tmax - t_move;
  if (t_max < 0.f) return;
  float hit_t = t_max;
  const float t_min = fmaxf(0.f, tmin - t_move);
  const vec3f oc = o - center;
  const float a = dot(dir, dir);
  const float b = dot(oc, dir);
  const float c = dot(oc, oc) - radius * radius;
  const float disc = b * b - a * c;
  if (disc < 0.f) return;
  const float sq = sqrtf(disc);
  {
    const float tt = (-b - sq) / a;
    if (tt < hit_t && tt > t_min) hit_t = tt;
  }
  {
    const float tt = (-b + sq) / a;
    if (tt < hit_t && tt > t_min) hit_t = tt;
  }
  if (!(hit_t < t_max)) return;
  const float t = hit_t + t_move;
  if (!better(t, flat, best)) return;
  if (!filt.any(t)) return;
  best.t = t;
  best.u = hit_t;
  best.v = t_move;
  best.flat = flat;
}

// barney/geometry/Cylinders.dev.cu:101-216 (capped cylinder). Writes the
// object-space normal of the chosen side when N != nullptr.
BN_HD bool cylinder_hit(vec3f ray_org, vec3f ray_dir, float ray_tmin, float ray_tmax, vec3f v0,
                        vec3f v1, float radius, float &t_out, vec3f *N) {
#pragma clang fp contract(off)
  const vec3f d = ray_dir;
  const vec3f s = v1 - v0;
  const vec3f sxd = cross(s, d);
  const float a = dot(sxd, sxd);
  if (a == 0.f) return false;
  const vec3f f = v0 - ray_org;
  const vec3f sxf = cross(s, f);
  const float ra = 1.f / a;
  const float ts = dot(sxd, sxf) * ra;
  const vec3f fp = f - ts * d;
  const float s2 = dot(s, s);
  const vec3f perp = cross(s, fp);
  const float c = radius * radius * s2 - dot(perp, perp);
  if (c < 0.f) return false;
  float td = sqrtf(c * ra);
  const float tube_t0 = ts - td;
  const float tube_t1 = ts + td;
  const float sf = dot(s, f);
  const float sd = dot(s, d);
  float cap_t0 = -1e20f, cap_t1 = +1e20f;
  float cap_t_v0 = cap_t0, cap_t_v1 = cap_t1;
  if (sd == 0.f) {
    if (dot(ray_org - v0, v1 - v0) < 0.f) return false;
    if (dot(ray_org - v1, v0 - v1) < 0.f) return false;
  } else {
    const float rsd = 1.f / sd;
    cap_t_v0 = sf * rsd;
    cap_t_v1 = cap_t_v0 + s2 * rsd;
    cap_t0 = fminf(cap_t_v0, cap_t_v1);
    cap_t1 = fmaxf(cap_t_v0, cap_t_v1);
  }
  const float t0 = fmaxf(cap_t0, tube_t0);
  const float t1 = fminf(cap_t1, tube_t1);
  if (t0 > t1) return false;
  float thit;
  if (ray_tmin <= t0 && t0 <= ray_tmax) {
    thit = t0;
    if (N) {
      td *= -1.f;
      const float hit_surf_u = (thit * sd - sf) * (1.f / s2);
      if (t0 == cap_t0)
        *N = (cap_t0 == cap_t_v0) ? -s : s;
      else
        *N = td * d - fp - hit_surf_u * s;
    }
  } else if (ray_tmin <= t1 && t1 <= ray_tmax) {
    thit = t1;
    if (N) {
      const float hit_surf_u = (thit * sd - sf) * (1.f / s2);
      // Cylinders.dev.cu:221 compares t0 (not t1) against cap_t1 here;
      // kept as-is so exit-side cap hits get the tube normal formula.
      if (t0 == cap_t1)
        *N = (cap_t0 == cap_t_v0) ? -s : s;
      else
        *N = td * d - fp - hit_surf_u * s;
    }
  } else {
    return false;
  }
  t_out = thit;
  return true;
}

template <class Filt = NoFilt>
BN_HD void isect_cylinder(vec3f org, vec3f dir, float tmin, float tmax, vec3f v0, vec3f v1,
                          float radius, uint32_t flat, HitRec &best, const Filt &filt = Filt()) {
  float t;
  if (!cylinder_hit(org, dir, tmin, tmax, v0, v1, radius, t, nullptr)) return;
  if (!better(t, flat, best)) return;
  if (!filt.any(t)) return;
  best.t = t;
  best.u = 0.f;
  best.v = 0.f;
  best.flat = flat;
}

// barney clamps exactly-zero direction components before tracing
// (barney/kernels/traceRays.dev.cu:42-44).
BN_HD vec3f fix_dir(vec3f d) {
  if (d.x == 0.f) d.x = 1e-6f;
  if (d.y == 0.f) d.y = 1e-6f;
  if (d.z == 0.f) d.z = 1e-6f;
  return d;
}


// keyBase: high bits of the tie-break key (the instance index in a
// two-level world, 0 otherwise), OR-ed onto the record's flat index.
template <class Filt = NoFilt>
BN_D void test_prim(const PrimRec *__restrict prims, uint32_t idx, vec3f org, vec3f dir,
                    float tmin, float tmax, HitRec &best, uint32_t keyBase = 0u, const Filt &filt = Filt()) {
  const float4 a = prims[idx].a;
  const float4 b = prims[idx].b;
  const float4 c = prims[idx].c;
  const uint32_t bits = __float_as_uint(c.w);
  const uint32_t type = bits & 0xffu;
  const uint32_t flat = keyBase | __float_as_uint(a.w);
  const vec3f p0 = {a.x, a.y, a.z};
  if (type == PRIM_TRIANGLE) {
    isect_triangle(org, dir, tmax, p0, {b.x, b.y, b.z}, {c.x, c.y, c.z}, flat, best, filt, b.w, bits);
  } else if (type == PRIM_SPHERE) {
    isect_sphere(org, dir, tmin, tmax, p0, b.x, flat, best, filt);
  } else {
    isect_cylinder(org, dir, tmin, tmax, p0, {b.x, b.y, b.z}, c.x, flat, best, filt);
  }
}

// Occlusion test of one primitive for shadow rays: any accepted hit in the
// ray's range (shadow-transparent primitives skipped).
template <class Filt = NoFilt>
BN_D bool occludes_prim(const PrimRec *__restrict prims, uint32_t idx, vec3f org, vec3f dir, float tmin,
                        float tmax, const Filt &filt = Filt()) {
  const float4 c = prims[idx].c;
  const uint32_t bits = __float_as_uint(c.w);
  if (bits & PRIM_SHADOW_TRANSPARENT) return false;
  HitRec h;
  h.t = tmax;
  h.flat = 0xffffffffu;
  test_prim(prims, idx, org, dir, tmin, tmax, h, 0u, filt);
  return h.flat != 0xffffffffu;
}

// Work counters of a traversal, for the algorithmic-bytes figure of a
// launch: node visits (one 128-byte node each) and primitive tests weighted
// by the bytes a test needs (triangle 36: three vertices; sphere 16: center
// and radius; cylinder 28: two end points and radius).
struct TraceCount {
  uint32_t nodes = 0;
  uint32_t primBytes = 0;
  BN_D void prim(const PrimRec *__restrict prims, uint32_t idx) {
    const uint32_t type = __float_as_uint(prims[idx].c.w) & 0xffu;
    primBytes += type == PRIM_TRIANGLE ? 36u : (type == PRIM_SPHERE ? 16u : 28u);
  }
};

// Stack entry: node group = {childBase, (hitbits | imask << 8)}.
// The per-lane stack lives in LDS as stack[depth][lanes] so consecutive
// lanes hit consecutive banks.
struct TraverseCtx {
  uint2 *stack;    // LDS base for this thread (already offset by lane)
  int stride;      // lanes per stack level
  int maxDepth;
};

// One node visit: slab tests of the 8 quantized child boxes against the
// ray (inv = 1/dir, octant oct), clipped to [tmin, tcap]. Returns the node
// group of hit internal children (bits k = slot ^ oct, imask above) and the
// bit mask of primitives in hit leaf children.
BN_D void bvh8_visit(const BNode8 *__restrict nodes, uint32_t nodeIdx, vec3f org, vec3f inv, uint32_t oct,
                     float tmin, float tcap, uint32_t &grpBase, uint32_t &grpBits, uint32_t &primBase,
                     uint32_t &primMask) {
  const uint4 *np = (const uint4 *)(nodes + nodeIdx);
  const uint4 n0 = np[0];
  const uint4 n1 = np[1];
  const float ox = __uint_as_float(n0.x), oy = __uint_as_float(n0.y), oz = __uint_as_float(n0.z);
  const uint32_t ex = n0.w & 0xffu, ey = (n0.w >> 8) & 0xffu, ez = (n0.w >> 16) & 0xffu;
  const uint32_t nimask = n0.w >> 24;
  const float sx = __uint_as_float(ex << 23) * inv.x;
  const float sy = __uint_as_float(ey << 23) * inv.y;
  const float sz = __uint_as_float(ez << 23) * inv.z;
  // slab offsets from (node origin - ray origin): the rounding error
  // stays relative to the node's extent, which the one-quantum slack of
  // the quantized boxes covers; ox * inv - org * inv would cancel
  // catastrophically far from the world origin
  const float bx = (ox - org.x) * inv.x;
  const float by = (oy - org.y) * inv.y;
  const float bz = (oz - org.z) * inv.z;
  uint32_t hitInner = 0, hitPrims = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint4 A = np[2 + 3 * h], B = np[3 + 3 * h], C = np[4 + 3 * h];
    // u32 word pairs per plane: lo.x A.xy, lo.y A.zw, lo.z B.xy,
    // hi.x B.zw, hi.y C.xy, hi.z C.zw
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = 4 * h + j;
      const uint32_t wlx = (j < 2) ? A.x : A.y, wly = (j < 2) ? A.z : A.w;
      const uint32_t wlz = (j < 2) ? B.x : B.y, whx = (j < 2) ? B.z : B.w;
      const uint32_t why = (j < 2) ? C.x : C.y, whz = (j < 2) ? C.z : C.w;
      const int sh = (j & 1) * 16;
      const float lx = (float)((wlx >> sh) & 0xffffu), ly = (float)((wly >> sh) & 0xffffu),
                  lz = (float)((wlz >> sh) & 0xffffu);
      const float hx = (float)((whx >> sh) & 0xffffu), hy = (float)((why >> sh) & 0xffffu),
                  hz = (float)((whz >> sh) & 0xffffu);
      const float tlx = fmaf(lx, sx, bx), thx = fmaf(hx, sx, bx);
      const float tly = fmaf(ly, sy, by), thy = fmaf(hy, sy, by);
      const float tlz = fmaf(lz, sz, bz), thz = fmaf(hz, sz, bz);
      float tn = fmaxf(fmaxf(fminf(tlx, thx), fminf(tly, thy)), fmaxf(fminf(tlz, thz), tmin));
      float tf = fminf(fminf(fmaxf(tlx, thx), fmaxf(tly, thy)), fmaxf(tlz, thz));
      // conservative widening against rounding in the quantized slab test
      tn = tn - fabsf(tn) * 2e-6f;
      tf = fminf(tf + fabsf(tf) * 2e-6f, tcap);
      const bool hit = (tn <= tf) && (lx <= hx);
      // bookkeeping with selects, not branches: a hit internal child sets
      // its group bit, a hit leaf child its primitives' bits
      const uint32_t meta = (c < 4) ? ((n1.z >> (8 * c)) & 0xffu) : ((n1.w >> (8 * (c - 4))) & 0xffu);
      const bool inner = (nimask >> c) & 1u;
      const uint32_t leafBits = ((1u << (meta >> 5)) - 1u) << (meta & 31u);
      hitInner |= (hit && inner) ? (1u << (c ^ oct)) : 0u;
      hitPrims |= (hit && !inner) ? leafBits : 0u;
    }
  }
  grpBase = n1.x;
  grpBits = hitInner | (nimask << 8);
  primBase = n1.y;
  primMask = hitPrims;
}

BN_D uint32_t ray_octant(vec3f dir) {
  return (dir.x < 0.f ? 1u : 0u) | (dir.y < 0.f ? 2u : 0u) | (dir.z < 0.f ? 4u : 0u);
}

// Traversal of one BVH8 continuing from 'best' (closest hit so far, keys
// with prefix keyBase). kAny: stop at the first occluding primitive (shadow
// rays; returns true when occluded). Otherwise: closest hit into 'best'
// (returns false only when the stack would overflow, which the builder's
// reported depth rules out).
template <bool kAny, class Filt = NoFilt>
BN_D bool traverse_from(const BNode8 *__restrict nodes, const PrimRec *__restrict prims, vec3f org, vec3f dir,
                        float tmin, float tmax, HitRec &best, const TraverseCtx &ctx, uint32_t keyBase,
                        TraceCount *tc = nullptr, const Filt &filt = Filt()) {
  const vec3f inv = rcp(dir);
  const uint32_t oct = ray_octant(dir);
  uint32_t grpBase = 0;
  uint32_t grpBits = (1u << oct) | (1u << 8);  // root = slot 0 (k = 0 ^ oct)
  uint32_t primBase = 0, primMask = 0;
  int sp = 0;
  while (true) {
    if ((grpBits & 0xffu) != 0) {
      const uint32_t k = __builtin_ctz(grpBits);
      grpBits &= ~(1u << k);
      const uint32_t slot = k ^ oct;
      const uint32_t imask = grpBits >> 8;
      const uint32_t nodeIdx = grpBase + __builtin_popcount(imask & ((1u << slot) - 1u));
      if ((grpBits & 0xffu) != 0) {
        if (sp >= ctx.maxDepth) return kAny;
        ctx.stack[sp * ctx.stride] = make_uint2(grpBase, grpBits);
        ++sp;
      }
      bvh8_visit(nodes, nodeIdx, org, inv, oct, tmin, best.t, grpBase, grpBits, primBase, primMask);
      if (tc) ++tc->nodes;
    }
    while (primMask) {
      const uint32_t i = __builtin_ctz(primMask);
      primMask &= primMask - 1u;
      if (tc) tc->prim(prims, primBase + i);
      if (kAny) {
        if (occludes_prim(prims, primBase + i, org, dir, tmin, tmax, filt)) return true;
      } else {
        test_prim(prims, primBase + i, org, dir, tmin, tmax, best, keyBase, filt);
      }
    }
    if ((grpBits & 0xffu) == 0) {
      if (sp == 0) break;
      --sp;
      const uint2 e = ctx.stack[sp * ctx.stride];
      grpBase = e.x;
      grpBits = e.y;
    }
  }
  return !kAny;
}

// The loop of traverse_from as a resumable walk: step() is one iteration
// (a node visit and the primitives of its hit leaves), so that a kernel can
// hand a lane its next ray as soon as the lane's walk is over while the rest
// of the wavefront keeps walking. Closest-hit acceptance (better()) is a
// total order on (t, flat index), so the hit found does not depend on the
// order in which rays or nodes are visited.
template <bool kAny, class Filt = NoFilt>
struct BvhWalk {
  vec3f org, dir, inv;
  uint32_t oct;
  float tmin, tmax;
  HitRec best;
  uint32_t grpBase, grpBits, primBase, primMask;
  int sp;
  bool occluded;  // kAny: an occluder was found (or the stack overflowed)
  TraceCount count;  // over every ray this walk has taken (init keeps it)
  Filt filt;         // the candidate filter of the current ray

  BN_D void init(vec3f o, vec3f d, float t0, float t1) {
    org = o;
    dir = d;
    inv = rcp(d);
    oct = ray_octant(d);
    tmin = t0;
    tmax = t1;
    best.t = t1;
    best.flat = 0xffffffffu;
    best.u = best.v = 0.f;
    grpBase = 0;
    grpBits = (1u << oct) | (1u << 8);  // root = slot 0 (k = 0 ^ oct)
    primBase = 0;
    primMask = 0;
    sp = 0;
    occluded = false;
  }

  // One iteration; true when the walk is over.
  BN_D bool step(const BNode8 *__restrict nodes, const PrimRec *__restrict prims, const TraverseCtx &ctx) {
    if ((grpBits & 0xffu) != 0) {
      const uint32_t k = __builtin_ctz(grpBits);
      grpBits &= ~(1u << k);
      const uint32_t slot = k ^ oct;
      const uint32_t imask = grpBits >> 8;
      const uint32_t nodeIdx = grpBase + __builtin_popcount(imask & ((1u << slot) - 1u));
      if ((grpBits & 0xffu) != 0) {
        if (sp >= ctx.maxDepth) {
          occluded = kAny;
          return true;
        }
        ctx.stack[sp * ctx.stride] = make_uint2(grpBase, grpBits);
        ++sp;
      }
      bvh8_visit(nodes, nodeIdx, org, inv, oct, tmin, best.t, grpBase, grpBits, primBase, primMask);
      ++count.nodes;
    }
    while (primMask) {
      const uint32_t i = __builtin_ctz(primMask);
      primMask &= primMask - 1u;
      count.prim(prims, primBase + i);
      if (kAny) {
        if (occludes_prim(prims, primBase + i, org, dir, tmin, tmax, filt)) {
          occluded = true;
          return true;
        }
      } else {
        test_prim(prims, primBase + i, org, dir, tmin, tmax, best, 0u, filt);
      }
    }
    if ((grpBits & 0xffu) == 0) {
      if (sp == 0) return true;
      --sp;
      const uint2 e = ctx.stack[sp * ctx.stride];
      grpBase = e.x;
      grpBits = e.y;
    }
    return false;
  }
};

// Closest-hit (or occlusion) traversal of one BVH8 from scratch.
template <bool kAny, class Filt = NoFilt>
BN_D bool traverse(const BNode8 *__restrict nodes, const PrimRec *__restrict prims, vec3f org, vec3f dir,
                   float tmin, float tmax, HitRec &best, const TraverseCtx &ctx, TraceCount *tc = nullptr,
                   const Filt &filt = Filt()) {
  best.t = tmax;
  best.flat = 0xffffffffu;
  best.u = best.v = 0.f;
  return traverse_from<kAny>(nodes, prims, org, dir, tmin, tmax, best, ctx, 0u, tc, filt);
}

BN_D bool traverse_closest(const BNode8 *__restrict nodes, const PrimRec *__restrict prims, vec3f org, vec3f dir,
                           float tmin, float tmax, HitRec &best, const TraverseCtx &ctx) {
  return traverse<false>(nodes, prims, org, dir, tmin, tmax, best, ctx);
}

BN_D bool traverse_occluded(const BNode8 *__restrict nodes, const PrimRec *__restrict prims, vec3f org,
                            vec3f dir, float tmin, float tmax, const TraverseCtx &ctx) {
  HitRec h;
  return traverse<true>(nodes, prims, org, dir, tmin, tmax, h, ctx);
}


}  // namespace bnamd
