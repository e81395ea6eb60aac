// Unstructured tetrahedral volume object: BVH8 over the tets (the world
// builder with the tet primitive kind), tets packed in leaf order, the
// macrocell grid of UMeshField::buildInitialMacroCells (UMeshField.cu:173-200)
// rasterized per tet with rasterTet<5> (:39-106, MCGrid.cuh:92-122), majorants
// (mapMCs, MCGrid.cu:94-116), and the volume + iso-surface renderer.
#include "shade.h"
#include "umesh.h"
#include "umesh_render.h"
#include <mutex>
#include <string>
#include <chrono>
#include <cmath>

namespace bnamd {

// The instrumented variants are compiled in umesh_count.hip.
#define BNAMD_UMESH_COUNT_VARIANT(L, M)                                                                        \
  extern template void launch_render_umesh_t<L, M, true>(const UMeshDD &, const UMeshOpts &, const CameraDD &, \
                                                         const RenderDD &, int, float4 *, unsigned long long *, \
                                                         size_t, hipStream_t);
BNAMD_UMESH_COUNT_VARIANT(8, false)
BNAMD_UMESH_COUNT_VARIANT(4, false)
BNAMD_UMESH_COUNT_VARIANT(2, false)
BNAMD_UMESH_COUNT_VARIANT(1, false)
BNAMD_UMESH_COUNT_VARIANT(4, true)
BNAMD_UMESH_COUNT_VARIANT(1, true)
#undef BNAMD_UMESH_COUNT_VARIANT

size_t trace_lds_bytes(int depth, int block);

struct UMesh {
  mutable DevBuf<unsigned long long> counters;  // the render's counters
  World w;                 // point-location BVH (prims freed after packing)
  DevBuf<float4> tets;     // leaf order: 4 per tet, or 8 per general element
  DevBuf<uint8_t> types;   // general elements: ELEM_* per leaf position
  DevBuf<float4> boxes;    // general elements: {lo, cell index bits}, {hi, 0} per leaf position
  DevBuf<int> cellIndices; // general elements: copy of the index stream
  DevBuf<float4> verts;    // general elements: vertices {x, y, z, scalar}
  DevBuf<uint2> rangesOrd; // ordered-uint min/max during rasterization
  DevBuf<float2> ranges;
  DevBuf<float> majorants;
  DevBuf<float4> tf;
  // per-cell bits for isoMaskValue, rebuilt by render calls when the iso
  // value changes (calls on one mesh are not concurrent, as in barney)
  mutable DevBuf<uint32_t> isoMask;
  mutable float isoMaskValue = 0.f;
  mutable bool haveIsoMask = false;
  UMeshDD dd{};
  int numTets = 0;
  bool haveTF = false;
  float buildMs = 0.f, mcMs = 0.f;
};

namespace {

__global__ void k_pack_verts(const float *__restrict v3, const float *__restrict s, int n, float4 *__restrict out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = make_float4(v3[3 * i + 0], v3[3 * i + 1], v3[3 * i + 2], s[i]);
}

// tets in BVH leaf order: position p holds the tet whose flat index the
// builder stored in prims[p]
__global__ void k_pack_tets(const PrimRec *__restrict prims, uint32_t n, const int4 *__restrict tets,
                            const float4 *__restrict v4, float4 *__restrict out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t t = __float_as_uint(prims[p].a.w);
  const int4 ix = tets[t];
  out[4ull * p + 0] = v4[ix.x];
  out[4ull * p + 1] = v4[ix.y];
  out[4ull * p + 2] = v4[ix.z];
  out[4ull * p + 3] = v4[ix.w];
}

// General cells (UMeshField's layout): the element kind, its vertex count
// and its box. Returns false for an unknown type or an index out of range.
__device__ __forceinline__ int elem_kind(int type, int &nv) {
  switch (type) {
    case BNAMD_CELL_TET: case 0: nv = 4; return ELEM_TET;
    case BNAMD_CELL_PYR: case 3: nv = 5; return ELEM_PYR;
    case BNAMD_CELL_PRISM: case 2: nv = 6; return ELEM_PRISM;
    case BNAMD_CELL_HEX: case 1: nv = 8; return ELEM_HEX;
    case BNAMD_CELL_POLYHEDRON: nv = 0; return ELEM_POLY;
  }
  nv = 0;
  return -1;
}

struct CellsIn {
  const int *indices;
  const int *offsets;
  const unsigned char *types;
  int numIndices, numCells, numVertices;
};

// cellBounds (UMeshField.h:269-343): box of the element's vertices; kinds
// into 'kinds'; 'bad' set when a cell cannot be read
__global__ void k_elem_boxes(CellsIn c, const float4 *__restrict v4, float4 *__restrict boxes,
                             uint8_t *__restrict kinds, int *__restrict bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.numCells) return;
  int nv;
  const int k = elem_kind(c.types[i], nv);
  const int off = c.offsets[i];
  bool ok = k >= 0 && off >= 0 && (long long)off + nv <= (long long)c.numIndices;
  vec3f lo = vec3f(BN_INF), hi = vec3f(-BN_INF);
  if (ok && k == ELEM_POLY) {  // FaceStreamReader's checks (UMeshField.h:179-262)
    int pos = off;
    const int nf = pos < c.numIndices ? c.indices[pos++] : 0;
    ok = nf > 0 && nf <= 8192;
    for (int f = 0; ok && f < nf; ++f) {
      const int np = pos < c.numIndices ? c.indices[pos++] : 0;
      ok = np >= 3 && np <= 1024 && (long long)pos + np <= (long long)c.numIndices;
      for (int q = 0; ok && q < np; ++q) {
        const int vi = c.indices[pos + q];
        ok = vi >= 0 && vi < c.numVertices;
        if (ok) {
          const float4 p = v4[vi];
          lo = vmin(lo, vec3f{p.x, p.y, p.z});
          hi = vmax(hi, vec3f{p.x, p.y, p.z});
        }
      }
      pos += np;
    }
    nv = 0;
  }
  for (int q = 0; ok && q < nv; ++q) {
    const int vi = c.indices[off + q];
    if (vi < 0 || vi >= c.numVertices) {
      ok = false;
      break;
    }
    const float4 p = v4[vi];
    lo = vmin(lo, vec3f{p.x, p.y, p.z});
    hi = vmax(hi, vec3f{p.x, p.y, p.z});
  }
  if (!ok) {
    atomicOr(bad, 1);
    lo = hi = vec3f(0.f);
  }
  kinds[i] = (uint8_t)(ok ? k : 0);
  boxes[2 * i] = make_float4(lo.x, lo.y, lo.z, 0.f);
  boxes[2 * i + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
}

// general elements in BVH leaf order: 8 records per element (unused ones 0)
__global__ void k_pack_elems(const PrimRec *__restrict prims, uint32_t n, CellsIn c, const uint8_t *__restrict kinds,
                             const float4 *__restrict v4, float4 *__restrict out, uint8_t *__restrict outKinds,
                             float4 *__restrict outBoxes) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t e = __float_as_uint(prims[p].a.w);
  const int k = kinds[e];
  const int nv = k == ELEM_TET ? 4 : k == ELEM_PYR ? 5 : k == ELEM_PRISM ? 6 : k == ELEM_HEX ? 8 : 0;
  const int off = c.offsets[e];
  for (int q = 0; q < 8; ++q) out[8ull * p + q] = q < nv ? v4[c.indices[off + q]] : make_float4(0.f, 0.f, 0.f, 0.f);
  if (k == ELEM_POLY) out[8ull * p] = make_float4(__int_as_float(off), 0.f, 0.f, 0.f);  // its stream
  outKinds[p] = (uint8_t)k;
  const PrimRec r = prims[p];
  outBoxes[2ull * p] = make_float4(r.a.x, r.a.y, r.a.z, __uint_as_float(e));
  outBoxes[2ull * p + 1] = make_float4(r.b.x, r.b.y, r.b.z, 0.f);
}

__device__ __forceinline__ uint32_t ord_of(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of_ord(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_clear_ranges(uint2 *__restrict r, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    r[i] = make_uint2(ord_of(BN_INF), ord_of(-BN_INF));
}

struct MCRaster {
  uint2 *ranges;
  int mx, my, mz;
  vec3f origin, rcpSpacing;
};

struct V4 {
  float x, y, z, w;
};
__device__ __forceinline__ V4 mid4(const V4 &a, const V4 &b) {
#pragma clang fp contract(off)
  return V4{0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.5f * (a.w + b.w)};
}
__device__ __forceinline__ float len3(const V4 &a, const V4 &b) {
#pragma clang fp contract(off)
  const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// rasterBox(grid, box4f) (MCGrid.cuh:92-122): boxes flat in any dimension
// are skipped; cell range = [min w, max w] via atomic min/max
__device__ void raster_box(const MCRaster &g, const V4 &a, const V4 &b, const V4 &c, const V4 &d) {
#pragma clang fp contract(off)
  const float lx = fminf(fminf(a.x, b.x), fminf(c.x, d.x)), hx = fmaxf(fmaxf(a.x, b.x), fmaxf(c.x, d.x));
  const float ly = fminf(fminf(a.y, b.y), fminf(c.y, d.y)), hy = fmaxf(fmaxf(a.y, b.y), fmaxf(c.y, d.y));
  const float lz = fminf(fminf(a.z, b.z), fminf(c.z, d.z)), hz = fmaxf(fmaxf(a.z, b.z), fmaxf(c.z, d.z));
  const float lw = fminf(fminf(a.w, b.w), fminf(c.w, d.w)), hw = fmaxf(fmaxf(a.w, b.w), fmaxf(c.w, d.w));
  if (lx >= hx || ly >= hy || lz >= hz) return;
  int x0 = (int)((lx - g.origin.x) * g.rcpSpacing.x), x1 = (int)((hx - g.origin.x) * g.rcpSpacing.x);
  int y0 = (int)((ly - g.origin.y) * g.rcpSpacing.y), y1 = (int)((hy - g.origin.y) * g.rcpSpacing.y);
  int z0 = (int)((lz - g.origin.z) * g.rcpSpacing.z), z1 = (int)((hz - g.origin.z) * g.rcpSpacing.z);
  x0 = min(max(x0, 0), g.mx - 1); x1 = min(max(x1, 0), g.mx - 1);
  y0 = min(max(y0, 0), g.my - 1); y1 = min(max(y1, 0), g.my - 1);
  z0 = min(max(z0, 0), g.mz - 1); z1 = min(max(z1, 0), g.mz - 1);
  const uint32_t olo = ord_of(lw), ohi = ord_of(hw);
  for (int iz = z0; iz <= z1; ++iz)
    for (int iy = y0; iy <= y1; ++iy)
      for (int ix = x0; ix <= x1; ++ix) {
        uint2 *cell = g.ranges + ix + (size_t)iy * g.mx + (size_t)iz * g.mx * g.my;
        atomicMin(&cell->x, olo);
        atomicMax(&cell->y, ohi);
      }
}

// rasterTet<D> (UMeshField.cu:39-106): split at the midpoint of the longest
// edge, D levels deep, then raster the 2^D sub-tet boxes
template <int D>
__device__ void raster_tet(const MCRaster &g, const V4 &a, const V4 &b, const V4 &c, const V4 &d) {
  if constexpr (D == 0) {
    raster_box(g, a, b, c, d);
  } else {
    const float lab = len3(a, b), lac = len3(a, c), lad = len3(a, d);
    const float lbc = len3(b, c), lbd = len3(b, d), lcd = len3(c, d);
    const float maxLen = fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(lab, lac), lad), lbc), lbd), lcd);
    V4 oa, ob, oc, od0, od1;
    if (lab >= maxLen) { oa = mid4(a, b); ob = c; oc = d; od0 = a; od1 = b; }
    else if (lac >= maxLen) { oa = mid4(a, c); ob = b; oc = d; od0 = a; od1 = c; }
    else if (lad >= maxLen) { oa = mid4(a, d); ob = b; oc = c; od0 = a; od1 = d; }
    else if (lbc >= maxLen) { oa = mid4(b, c); ob = a; oc = d; od0 = b; od1 = c; }
    else if (lbd >= maxLen) { oa = mid4(b, d); ob = a; oc = c; od0 = b; od1 = d; }
    else { oa = mid4(c, d); ob = a; oc = b; od0 = c; od1 = d; }
    raster_tet<D - 1>(g, oa, ob, oc, od0);
    raster_tet<D - 1>(g, oa, ob, oc, od1);
  }
}

// rasterBox(grid, worldBounds, box4f) (MCGrid.cuh:67-91), the variant
// umeshRasterCells uses for cells other than tets: only an inverted box is
// skipped, so flat boxes are rasterized
__device__ void raster_box_range(const MCRaster &g, float lx, float hx, float ly, float hy, float lz, float hz,
                                 float lw, float hw) {
#pragma clang fp contract(off)
  if (lx > hx || ly > hy || lz > hz) return;
  int x0 = (int)((lx - g.origin.x) * g.rcpSpacing.x), x1 = (int)((hx - g.origin.x) * g.rcpSpacing.x);
  int y0 = (int)((ly - g.origin.y) * g.rcpSpacing.y), y1 = (int)((hy - g.origin.y) * g.rcpSpacing.y);
  int z0 = (int)((lz - g.origin.z) * g.rcpSpacing.z), z1 = (int)((hz - g.origin.z) * g.rcpSpacing.z);
  x0 = min(max(x0, 0), g.mx - 1); x1 = min(max(x1, 0), g.mx - 1);
  y0 = min(max(y0, 0), g.my - 1); y1 = min(max(y1, 0), g.my - 1);
  z0 = min(max(z0, 0), g.mz - 1); z1 = min(max(z1, 0), g.mz - 1);
  const uint32_t olo = ord_of(lw), ohi = ord_of(hw);
  for (int iz = z0; iz <= z1; ++iz)
    for (int iy = y0; iy <= y1; ++iy)
      for (int ix = x0; ix <= x1; ++ix) {
        uint2 *cell = g.ranges + ix + (size_t)iy * g.mx + (size_t)iz * g.mx * g.my;
        atomicMin(&cell->x, olo);
        atomicMax(&cell->y, ohi);
      }
}

// umeshRasterCells (UMeshField.cu:110-160): tets by rasterTet<5>, other
// cells by their box
__global__ void k_raster_cells(MCRaster g, CellsIn c, const uint8_t *__restrict kinds, const float4 *__restrict v4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.numCells) return;
  const int k = kinds[i];
  const int off = c.offsets[i];
  if (k == ELEM_TET) {
    float4 v[4];
    for (int q = 0; q < 4; ++q) v[q] = v4[c.indices[off + q]];
    raster_tet<5>(g, V4{v[0].x, v[0].y, v[0].z, v[0].w}, V4{v[1].x, v[1].y, v[1].z, v[1].w},
                  V4{v[2].x, v[2].y, v[2].z, v[2].w}, V4{v[3].x, v[3].y, v[3].z, v[3].w});
    return;
  }
  // the cell's vertices (a polyhedron: every vertex of its face stream)
  float lx = BN_INF, ly = BN_INF, lz = BN_INF, lw = BN_INF, hx = -BN_INF, hy = -BN_INF, hz = -BN_INF, hw = -BN_INF;
  auto ext = [&](const float4 p) {
    lx = fminf(lx, p.x); hx = fmaxf(hx, p.x);
    ly = fminf(ly, p.y); hy = fmaxf(hy, p.y);
    lz = fminf(lz, p.z); hz = fmaxf(hz, p.z);
    lw = fminf(lw, p.w); hw = fmaxf(hw, p.w);
  };
  if (k == ELEM_POLY) {
    int pos = off;
    const int nf = c.indices[pos++];
    for (int f = 0; f < nf; ++f) {
      const int np = c.indices[pos++];
      for (int q = 0; q < np; ++q) ext(v4[c.indices[pos + q]]);
      pos += np;
    }
  } else {
    const int nv = k == ELEM_PYR ? 5 : k == ELEM_PRISM ? 6 : 8;
    for (int q = 0; q < nv; ++q) ext(v4[c.indices[off + q]]);
  }
  raster_box_range(g, lx, hx, ly, hy, lz, hz, lw, hw);
}

__global__ void k_raster_tets(MCRaster g, const int4 *__restrict tets, const float4 *__restrict v4, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int4 ix = tets[t];
  const float4 p0 = v4[ix.x], p1 = v4[ix.y], p2 = v4[ix.z], p3 = v4[ix.w];
  raster_tet<5>(g, V4{p0.x, p0.y, p0.z, p0.w}, V4{p1.x, p1.y, p1.z, p1.w}, V4{p2.x, p2.y, p2.z, p2.w},
                V4{p3.x, p3.y, p3.z, p3.w});
}

__global__ void k_finish_ranges(const uint2 *__restrict in, float2 *__restrict out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_float2(float_of_ord(in[i].x), float_of_ord(in[i].y));
}

// Bit c of the mask: isoValue lies within macrocell c's value range.
__global__ void k_iso_mask(const float2 *__restrict ranges, size_t n, float isoValue, uint32_t *__restrict mask) {
  const size_t nw = (n + 31) / 32;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < nw; w += (size_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0u;
    for (int b = 0; b < 32; ++b) {
      const size_t c = w * 32 + b;
      if (c < n) {
        const float2 vr = ranges[c];
        if (isoValue >= vr.x && isoValue <= vr.y) bits |= 1u << b;
      }
    }
    mask[w] = bits;
  }
}

__global__ void k_umesh_majorants(TFDD tf, const float2 *__restrict ranges, float *__restrict maj, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    maj[i] = tf_majorant(tf, ranges[i].x, ranges[i].y);
}

template <bool kMixed>
__global__ void __launch_bounds__(256) k_umesh_sample(UMeshDD u, const float *__restrict pts, float *__restrict out,
                                                      int n) {
  extern __shared__ uint2 lds_stack[];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, u.depth + 1};
  uint32_t hint = ~0u;
  out[i] = umesh_locate<kMixed>(u, {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, ctx, hint);
}

template <bool kMixed>
__global__ void __launch_bounds__(256) k_umesh_iso_rays(UMeshDD u, const bnamd_ray *__restrict rays, float iso,
                                                        float4 *__restrict out, int n) {
  extern __shared__ uint2 lds_stack[];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TraverseCtx ctx{lds_stack + threadIdx.x, (int)blockDim.x, u.depth + 1};
  const bnamd_ray r = rays[i];
  RNGSeed seed;
  uint32_t hint = ~0u;
  const IsoHit h = iso_intersect(
      u, [&](vec3f P) { return umesh_locate<kMixed>(u, P, ctx, hint); }, iso, {r.org[0], r.org[1], r.org[2]},
      fix_dir({r.dir[0], r.dir[1], r.dir[2]}), r.tmin, r.tmax, r.tmax, seed, bn_hash3(0u, 1u, 0u));
  out[i] = h.hit ? make_float4(h.t, h.N.x, h.N.y, h.N.z) : make_float4(BN_INF, 0.f, 0.f, 0.f);
}


template <int L, bool kMixed = false>
void launch_render_umesh(const UMeshDD &u, const UMeshOpts &o, const CameraDD &cam, const RenderDD &rp, int spp,
                         float4 *accum, unsigned long long *counters, size_t lds, hipStream_t s, bool count) {
  if (count) launch_render_umesh_t<L, kMixed, true>(u, o, cam, rp, spp, accum, counters, lds, s);
  else launch_render_umesh_t<L, kMixed, false>(u, o, cam, rp, spp, accum, counters, lds, s);
}

void upload_tf(UMesh &m, const float *values, int numValues, float lo, float hi, float base) {
  m.tf.upload((const float4 *)values, numValues, 0);
  m.dd.tf.values = m.tf.ptr;
  m.dd.tf.numValues = numValues;
  m.dd.tf.domLo = lo;
  m.dd.tf.domHi = hi;
  m.dd.tf.baseDensity = base;
  const size_t ncell = (size_t)m.dd.mx * m.dd.my * m.dd.mz;
  k_umesh_majorants<<<(int)std::min<size_t>(div_up(ncell, 256), 8192), 256>>>(m.dd.tf, m.ranges.ptr,
                                                                              m.majorants.ptr, ncell);
  BN_HIP(hipGetLastError());
  BN_HIP(hipDeviceSynchronize());
  m.haveTF = true;
}

// The mesh's device record for an iso-surface walk at isoValue: with the
// macrocell bit mask of that value, rebuilt on s when the value changes.
UMeshDD iso_dd(const UMesh &m, float isoValue, hipStream_t s) {
  const size_t ncell = (size_t)m.dd.mx * m.dd.my * m.dd.mz;
  const bool same = m.haveIsoMask && (m.isoMaskValue == isoValue ||
                                      (m.isoMaskValue != m.isoMaskValue && isoValue != isoValue));
  if (!same) {
    m.isoMask.ensure((ncell + 31) / 32);
    const size_t nw = (ncell + 31) / 32;
    k_iso_mask<<<(int)std::min<size_t>(div_up(nw, 256), 4096), 256, 0, s>>>(m.ranges.ptr, ncell, isoValue,
                                                                            m.isoMask.ptr);
    BN_HIP(hipGetLastError());
    m.isoMaskValue = isoValue;
    m.haveIsoMask = true;
  }
  UMeshDD d = m.dd;
  d.isoMask = m.isoMask.ptr;
  return d;
}

// Lanes per pixel from the paths per pixel (general cells: 4 from 4 spp,
// else 1); the kernel instance for them.
void render_umesh_dispatch(const UMeshDD &dd, const UMeshOpts &o, const CameraDD &cam, const RenderDD &rp, int spp,
                           float4 *accum, unsigned long long *cnt, size_t lds, hipStream_t s, bool count) {
  int lanes = spp >= 8 ? 8 : (spp >= 4 ? 4 : (spp >= 2 ? 2 : 1));
  if (const char *e = getenv("BNAMD_UMESH_LANES")) lanes = atoi(e);
  if (dd.elemTypes) {
    if (lanes >= 4) launch_render_umesh<4, true>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
    else launch_render_umesh<1, true>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
  }
  else if (lanes >= 8) launch_render_umesh<8>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
  else if (lanes >= 4) launch_render_umesh<4>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
  else if (lanes >= 2) launch_render_umesh<2>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
  else launch_render_umesh<1>(dd, o, cam, rp, spp, accum, cnt, lds, s, count);
  BN_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Scratch of bnamd_umesh_render_noalloc, which allocates nothing: one array
// per GPU for the life of the process, shared by every call on that GPU (the
// last call's contents stay; calls that overlap on one GPU race). Words
// [0, 64): the seven 64-bit counters (in 256 bytes); then the iso-surface bit
// mask, one bit per macrocell, at most kNoAllocMaskWords words (32 M
// macrocells, 4 MiB).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocHeaderWords = 64;
constexpr size_t kNoAllocMaskWords = size_t(1) << 20;
static __device__ __align__(256) uint32_t scratch[kNoAllocHeaderWords + kNoAllocMaskWords];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
uint32_t *gScratchPtr[kMaxDevices] = {};

// The device address of this GPU's scratch (the array's name is a host-side
// stand-in and must not be passed to kernels), looked up once per device
// with that device current.
uint32_t *noalloc_scratch() {
  int dev = 0;
  BN_HIP(hipGetDevice(&dev));
  BN_REQUIRE(dev >= 0 && dev < kMaxDevices, "bnamd_umesh_render_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratchPtr[dev]) {
    void *p = nullptr;
    BN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BN_REQUIRE(p, "bnamd_umesh_render_noalloc: no scratch address");
    gScratchPtr[dev] = (uint32_t *)p;
  }
  return gScratchPtr[dev];
}

// Zeroes n 64-bit counters (a kernel, not a memset node, so that graph
// captures of the call replay correctly).
__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

struct bnamd_umesh_s {
  UMesh m;
};

extern "C" int bnamd_umesh_create(const bnamd_umesh_desc *desc, void *stream, bnamd_umesh **out) {
  return guarded([&] {
    BN_REQUIRE(desc && out, "bnamd_umesh_create: NULL argument");
    const bool cells = desc->tets == nullptr;
    BN_REQUIRE(desc->numVertices > 0 && (cells ? desc->numCells > 0 && desc->numIndices > 0 : desc->numTets > 0),
               "bnamd_umesh_create: empty mesh");
    BN_REQUIRE(desc->vertices && desc->scalars && (!cells || (desc->indices && desc->cellOffsets && desc->cellTypes)),
               "bnamd_umesh_create: NULL array");
    auto *um = new bnamd_umesh_s;
    try {
      UMesh &m = um->m;
      hipStream_t s = (hipStream_t)stream;
      const auto t0 = std::chrono::steady_clock::now();
      const int numElems = cells ? desc->numCells : desc->numTets;
      m.numTets = numElems;
      DevBuf<float4> v4(desc->numVertices);
      k_pack_verts<<<div_up(desc->numVertices, 256), 256, 0, s>>>(desc->vertices, desc->scalars, desc->numVertices,
                                                                  v4.ptr);
      BN_HIP(hipGetLastError());
      const CellsIn c{desc->indices, desc->cellOffsets, desc->cellTypes, desc->numIndices, desc->numCells,
                      desc->numVertices};
      DevBuf<float4> boxes;
      DevBuf<uint8_t> kinds;
      GeomDev g{};
      g.primOffset = 0;
      if (cells) {
        boxes.alloc(2ull * numElems);
        kinds.alloc(numElems);
        DevBuf<int> bad(1);
        BN_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
        k_elem_boxes<<<div_up(numElems, 256), 256, 0, s>>>(c, v4.ptr, boxes.ptr, kinds.ptr, bad.ptr);
        BN_HIP(hipGetLastError());
        int hbad = 0;
        BN_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        BN_HIP(hipStreamSynchronize(s));
        BN_REQUIRE(!hbad, "bnamd_umesh_create: a cell has an unknown type, an index out of range or a malformed face stream");
        g.kind = GEOM_ELEM_BOXES;
        g.vertices = (const float *)boxes.ptr;
      } else {
        g.kind = GEOM_TETS;
        g.vertices = (const float *)v4.ptr;
        g.indices = desc->tets;
      }
      g.numPrims = numElems;
      m.w.hgeoms.push_back(g);
      m.w.numPrims = (uint32_t)numElems;
      MaterialDev md{};
      m.w.materials.upload(&md, 1, s);
      build_world_bvh(m.w, s, /*mortonOrder=*/true);
      if (cells) {
        // polyhedra read the index stream and the vertices at sample time
        m.cellIndices.alloc(desc->numIndices);
        BN_HIP(hipMemcpyAsync(m.cellIndices.ptr, desc->indices, sizeof(int) * desc->numIndices,
                              hipMemcpyDeviceToDevice, s));
        m.verts.alloc(desc->numVertices);
        BN_HIP(hipMemcpyAsync(m.verts.ptr, v4.ptr, sizeof(float4) * desc->numVertices, hipMemcpyDeviceToDevice, s));
        m.tets.alloc(8ull * numElems);
        m.types.alloc(numElems);
        m.boxes.alloc(2ull * numElems);
        k_pack_elems<<<div_up(numElems, 256), 256, 0, s>>>(m.w.prims.ptr, (uint32_t)numElems, c, kinds.ptr, v4.ptr,
                                                           m.tets.ptr, m.types.ptr, m.boxes.ptr);
      } else {
        m.tets.alloc(4ull * numElems);
        k_pack_tets<<<div_up(numElems, 256), 256, 0, s>>>(m.w.prims.ptr, (uint32_t)numElems,
                                                          (const int4 *)desc->tets, v4.ptr, m.tets.ptr);
      }
      BN_HIP(hipGetLastError());
      BN_HIP(hipStreamSynchronize(s));
      m.w.prims.release();
      const auto t1 = std::chrono::steady_clock::now();
      m.buildMs = std::chrono::duration<float, std::milli>(t1 - t0).count();
      // macrocell grid (UMeshField.cu:173-200)
      const vec3f lo = {m.w.lo[0], m.w.lo[1], m.w.lo[2]}, hi = {m.w.hi[0], m.w.hi[1], m.w.hi[2]};
      const vec3f size = hi - lo;
      const float maxWidth = reduce_max(size);
      const int MC_GRID_SIZE = 200 + int(sqrtf(numElems / 1000.f));
      const float f = (MC_GRID_SIZE - 1) / maxWidth;
      m.dd.mx = 1 + (int)(size.x * f);
      m.dd.my = 1 + (int)(size.y * f);
      m.dd.mz = 1 + (int)(size.z * f);
      BN_REQUIRE((long long)m.dd.my * m.dd.mz < (1ll << 24), "bnamd_umesh_create: over 2^24 macrocells in y*z");
      m.dd.mcOrigin = lo;
      m.dd.mcSpacing = size * rcp(vec3f((float)m.dd.mx, (float)m.dd.my, (float)m.dd.mz));
      m.dd.worldLo = lo;
      m.dd.worldHi = hi;
      const size_t ncell = (size_t)m.dd.mx * m.dd.my * m.dd.mz;
      m.rangesOrd.alloc(ncell);
      m.ranges.alloc(ncell);
      m.majorants.alloc(ncell);
      const int gs = (int)std::min<size_t>(div_up(ncell, 256), 8192);
      k_clear_ranges<<<gs, 256, 0, s>>>(m.rangesOrd.ptr, ncell);
      MCRaster r;
      r.ranges = m.rangesOrd.ptr;
      r.mx = m.dd.mx; r.my = m.dd.my; r.mz = m.dd.mz;
      r.origin = m.dd.mcOrigin;
      r.rcpSpacing = rcp(m.dd.mcSpacing);
      if (cells)
        k_raster_cells<<<div_up(numElems, 128), 128, 0, s>>>(r, c, kinds.ptr, v4.ptr);
      else
        k_raster_tets<<<div_up(numElems, 128), 128, 0, s>>>(r, (const int4 *)desc->tets, v4.ptr, numElems);
      BN_HIP(hipGetLastError());
      k_finish_ranges<<<gs, 256, 0, s>>>(m.rangesOrd.ptr, m.ranges.ptr, ncell);
      BN_HIP(hipGetLastError());
      BN_HIP(hipStreamSynchronize(s));
      m.rangesOrd.release();
      m.mcMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t1).count();
      m.dd.nodes = m.w.nodes.ptr;
      m.dd.tets = m.tets.ptr;
      m.dd.elemTypes = cells ? m.types.ptr : nullptr;
      m.dd.elemBoxes = cells ? m.boxes.ptr : nullptr;
      m.dd.cellIndices = cells ? m.cellIndices.ptr : nullptr;
      m.dd.verts = cells ? m.verts.ptr : nullptr;
      m.dd.numTets = (uint32_t)numElems;
      m.dd.depth = m.w.depth;
      m.dd.ranges = m.ranges.ptr;
      m.dd.majorants = m.majorants.ptr;
      // default TF: constant white, density 1 (replaced by bnamd_umesh_set_tf)
      const float def[8] = {1, 1, 1, 1, 1, 1, 1, 1};
      upload_tf(m, def, 2, 0.f, 1.f, 1.f);
      m.haveTF = false;
    } catch (...) {
      delete um;
      throw;
    }
    *out = um;
  });
}

extern "C" int bnamd_umesh_destroy(bnamd_umesh *mesh) {
  return guarded([&] { delete mesh; });
}

extern "C" int bnamd_umesh_info_get(const bnamd_umesh *mesh, bnamd_umesh_info *info) {
  return guarded([&] {
    BN_REQUIRE(mesh && info, "bnamd_umesh_info_get: NULL argument");
    const UMesh &m = mesh->m;
    info->numTets = m.numTets;
    info->numNodes = (int)m.w.numNodes;
    info->bvhDepth = m.w.depth;
    info->mcDims[0] = m.dd.mx; info->mcDims[1] = m.dd.my; info->mcDims[2] = m.dd.mz;
    for (int k = 0; k < 3; ++k) {
      info->worldLower[k] = m.w.lo[k];
      info->worldUpper[k] = m.w.hi[k];
    }
    info->buildMs = m.buildMs;
    info->mcMs = m.mcMs;
  });
}

extern "C" int bnamd_umesh_set_tf(bnamd_umesh *mesh, const float *values, int numValues, float domainLo,
                                  float domainHi, float baseDensity) {
  return guarded([&] {
    BN_REQUIRE(mesh && values, "bnamd_umesh_set_tf: NULL argument");
    BN_REQUIRE(numValues >= 2, "bnamd_umesh_set_tf: need at least 2 values");
    BN_REQUIRE(domainHi > domainLo, "bnamd_umesh_set_tf: empty domain");
    upload_tf(mesh->m, values, numValues, domainLo, domainHi, baseDensity);
  });
}

extern "C" int bnamd_umesh_get_mcs(const bnamd_umesh *mesh, float *ranges, float *majorants) {
  return guarded([&] {
    BN_REQUIRE(mesh, "bnamd_umesh_get_mcs: NULL mesh");
    const UMesh &m = mesh->m;
    const size_t n = (size_t)m.dd.mx * m.dd.my * m.dd.mz;
    if (ranges) BN_HIP(hipMemcpy(ranges, m.ranges.ptr, n * sizeof(float2), hipMemcpyDeviceToHost));
    if (majorants) BN_HIP(hipMemcpy(majorants, m.majorants.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int bnamd_umesh_sample_points(const bnamd_umesh *mesh, const float *points, float *out, int numPoints,
                                         void *stream) {
  return guarded([&] {
    BN_REQUIRE(mesh, "bnamd_umesh_sample_points: NULL mesh");
    BN_REQUIRE(numPoints >= 0, "bnamd_umesh_sample_points: negative count");
    if (numPoints == 0) return;
    BN_REQUIRE(points && out, "bnamd_umesh_sample_points: NULL buffer");
    const size_t lds = trace_lds_bytes(mesh->m.dd.depth, 256);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_umesh_sample_points: BVH too deep for the LDS stack");
    hipStream_t s = (hipStream_t)stream;
    if (mesh->m.dd.elemTypes)
      k_umesh_sample<true><<<div_up(numPoints, 256), 256, lds, s>>>(mesh->m.dd, points, out, numPoints);
    else
      k_umesh_sample<false><<<div_up(numPoints, 256), 256, lds, s>>>(mesh->m.dd, points, out, numPoints);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_umesh_iso_rays(const bnamd_umesh *mesh, const bnamd_ray *rays, float isoValue, float *out,
                                    int numRays, void *stream) {
  return guarded([&] {
    BN_REQUIRE(mesh, "bnamd_umesh_iso_rays: NULL mesh");
    BN_REQUIRE(numRays >= 0, "bnamd_umesh_iso_rays: negative count");
    if (numRays == 0) return;
    BN_REQUIRE(rays && out, "bnamd_umesh_iso_rays: NULL buffer");
    const size_t lds = trace_lds_bytes(mesh->m.dd.depth, 256);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_umesh_iso_rays: BVH too deep for the LDS stack");
    hipStream_t s = (hipStream_t)stream;
    const UMeshDD dd = iso_dd(mesh->m, isoValue, s);
    if (dd.elemTypes)
      k_umesh_iso_rays<true><<<div_up(numRays, 256), 256, lds, s>>>(dd, rays, isoValue, (float4 *)out, numRays);
    else
      k_umesh_iso_rays<false><<<div_up(numRays, 256), 256, lds, s>>>(dd, rays, isoValue, (float4 *)out, numRays);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(s));
  });
}

extern "C" int bnamd_umesh_render(const bnamd_umesh *mesh, const bnamd_camera *camera,
                                  const bnamd_render_params *params, const bnamd_umesh_render_opts *opts,
                                  float *accum, void *stream, bnamd_volume_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(mesh && camera && params && opts && accum, "bnamd_umesh_render: NULL argument");
    BN_REQUIRE(params->width > 0 && params->height > 0 && params->pathsPerPixel >= 1,
               "bnamd_umesh_render: bad frame size");
    BN_REQUIRE(!opts->volume || mesh->m.haveTF, "bnamd_umesh_render: volume rendering needs a transfer function");
    const CameraDD cam = make_camera(*camera);
    RenderDD rp = make_render_dd(*params);
    UMeshOpts o;
    o.volume = opts->volume;
    o.iso = opts->iso;
    o.isoValue = opts->isoValue;
    o.isoColor = {opts->isoColor[0], opts->isoColor[1], opts->isoColor[2]};
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = trace_lds_bytes(mesh->m.dd.depth, 256);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_umesh_render: BVH too deep for the LDS stack");
    const bool count = opts->countTraversal != 0;
    unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
    const float ms = render_on_stream(s, mesh->m.counters, 7, stats != nullptr, h, [&] {
      const UMeshDD dd = o.iso ? iso_dd(mesh->m, o.isoValue, s) : mesh->m.dd;
      render_umesh_dispatch(dd, o, cam, rp, params->pathsPerPixel, (float4 *)accum, mesh->m.counters.ptr, lds, s,
                            count);
    });
    if (stats) {
      *stats = bnamd_volume_stats{};
      stats->samples = (long long)h[0];
      stats->ddaSteps = (long long)h[1];
      stats->raysTraced = (long long)h[2];
      stats->ms = ms;
      stats->zeroMajorantSteps = (long long)h[3];
      stats->nodeVisits = (long long)h[4];
      stats->cellTests = (long long)h[5];
      stats->cellBytes = (long long)h[6];
    }
  });
}

extern "C" int bnamd_umesh_render_noalloc(const bnamd_umesh *mesh, const bnamd_camera *camera,
                                          const bnamd_render_params *params, const bnamd_umesh_render_opts *opts,
                                          float *accum, void *stream, bnamd_volume_stats *stats) {
  return guarded([&] {
    BN_REQUIRE(mesh && camera && params && opts && accum, "bnamd_umesh_render_noalloc: NULL argument");
    BN_REQUIRE(params->width > 0 && params->height > 0 && params->pathsPerPixel >= 1,
               "bnamd_umesh_render_noalloc: bad frame size");
    BN_REQUIRE(!opts->volume || mesh->m.haveTF,
               "bnamd_umesh_render_noalloc: volume rendering needs a transfer function");
    const UMesh &m = mesh->m;
    const size_t ncell = (size_t)m.dd.mx * m.dd.my * m.dd.mz;
    const size_t maskWords = (ncell + 31) / 32;
    BN_REQUIRE(!opts->iso || maskWords <= kNoAllocMaskWords,
               "bnamd_umesh_render_noalloc: the iso-surface mask of this mesh's " + std::to_string(ncell) +
                   " macrocells exceeds the scratch maximum of " + std::to_string(kNoAllocMaskWords * 32) +
                   " macrocells; use bnamd_umesh_render");
    const size_t lds = trace_lds_bytes(m.dd.depth, 256);
    BN_REQUIRE(lds <= 64 * 1024, "bnamd_umesh_render_noalloc: BVH too deep for the LDS stack");
    const CameraDD cam = make_camera(*camera);
    const RenderDD rp = make_render_dd(*params);
    UMeshOpts o;
    o.volume = opts->volume;
    o.iso = opts->iso;
    o.isoValue = opts->isoValue;
    o.isoColor = {opts->isoColor[0], opts->isoColor[1], opts->isoColor[2]};
    hipStream_t s = (hipStream_t)stream;
    uint32_t *sp = noalloc_scratch();
    unsigned long long *cnt = (unsigned long long *)sp;
    const bool count = opts->countTraversal != 0;
    k_zero_u64<<<1, 64, 0, s>>>(cnt, 7);
    UMeshDD dd = m.dd;
    if (o.iso) {  // this call's mask, whatever an earlier call left there
      uint32_t *mask = sp + kNoAllocHeaderWords;
      k_iso_mask<<<(int)std::min<size_t>(div_up(maskWords, 256), 4096), 256, 0, s>>>(m.ranges.ptr, ncell,
                                                                                    o.isoValue, mask);
      dd.isoMask = mask;
    }
    render_umesh_dispatch(dd, o, cam, rp, params->pathsPerPixel, (float4 *)accum, cnt, lds, s, count);
    if (stats) {  // read back on the stream (not for graph capture: pass NULL there)
      unsigned long long h[7];
      BN_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
      BN_HIP(hipStreamSynchronize(s));
      *stats = bnamd_volume_stats{};
      stats->samples = (long long)h[0];
      stats->ddaSteps = (long long)h[1];
      stats->raysTraced = (long long)h[2];
      stats->ms = -1.f;
      stats->zeroMajorantSteps = (long long)h[3];
      stats->nodeVisits = (long long)h[4];
      stats->cellTests = (long long)h[5];
      stats->cellBytes = (long long)h[6];
    }
  });
}
