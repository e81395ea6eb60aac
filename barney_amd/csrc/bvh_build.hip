// GPU BVH8 builder: primitive gather -> 63-bit Morton codes -> radix sort ->
// binary tree by PLOC clustering of the sorted primitives (or the Karras
// radix tree with bottom-up boxes) -> top-down collapse into 128-byte 8-wide
// nodes with 16-bit quantized child boxes.
//
// Replaces the cuBQL builds of rtcore/cuda/cudaGroup.cu:261-413 (triangles),
// :418-562 (user geoms) and :144-258 (instances); the primitive bounds follow
// the geometry programs' bounds() (Spheres.dev.cu:29-40,
// Cylinders.dev.cu:71-87) and the triangle group's box of three vertices.
#include "world.h"
#include <hipcub/hipcub.hpp>
#include <chrono>
#include <cfloat>

namespace bnamd {

namespace {

constexpr int kLeafMax = 3;  // max primitives per leaf child

struct BinNode {
  float lo[3];
  int left;   // >= 0 internal index, < 0: ~leafIndex
  float hi[3];
  int right;
};

__global__ void k_gather(const GeomDev *__restrict geoms, int numGeoms, const MaterialDev *__restrict mats,
                         uint32_t numPrims, PrimRec *__restrict recs, float4 *__restrict boxes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numPrims) return;
  const int g = find_geom(geoms, numGeoms, i);
  const GeomDev &gd = geoms[g];
  const int primID = (int)(i - gd.primOffset);
  PrimRec r;
  vec3f lo, hi;
  if (gd.kind == BNAMD_GEOM_TRIANGLES) {
    const int *ix = gd.indices + 3 * primID;
    const vec3f v0 = load3(gd.vertices + 3 * ix[0]);
    const vec3f v1 = load3(gd.vertices + 3 * ix[1]);
    const vec3f v2 = load3(gd.vertices + 3 * ix[2]);
    const vec3f e1 = v1 - v0, e2 = v2 - v0;
    const MaterialDev &m = mats[gd.materialID];
    const bool glass = m.type == BNAMD_MAT_PBR && m.ior != 1.f && m.transmission >= 1e-3f;
    const bool cutout = !glass && m.alpha < 1.f;
    r.a = make_float4(v0.x, v0.y, v0.z, __uint_as_float(i));
    r.b = make_float4(e1.x, e1.y, e1.z, cutout ? m.alpha : 1.f);
    r.c = make_float4(e2.x, e2.y, e2.z, __uint_as_float(PRIM_TRIANGLE | (glass ? PRIM_SHADOW_TRANSPARENT : 0u) |
                                                       (cutout ? PRIM_CUTOUT : 0u)));
    lo = vmin(v0, vmin(v1, v2));
    hi = vmax(v0, vmax(v1, v2));
  } else if (gd.kind == BNAMD_GEOM_SPHERES) {
    const vec3f c = load3(gd.vertices + 3 * primID);
    const float rad = gd.radii ? gd.radii[primID] : gd.defaultRadius;
    r.a = make_float4(c.x, c.y, c.z, __uint_as_float(i));
    r.b = make_float4(rad, 0.f, 0.f, 0.f);
    r.c = make_float4(0.f, 0.f, 0.f, __uint_as_float(PRIM_SPHERE));
    lo = c - vec3f(rad);
    hi = c + vec3f(rad);
  } else if (gd.kind == GEOM_INSTANCES || gd.kind == GEOM_ELEM_BOXES || gd.kind == BNAMD_GEOM_USER) {
    const float4 *bx = (const float4 *)gd.vertices;
    const float4 l = bx[2 * primID], h = bx[2 * primID + 1];
    lo = {l.x, l.y, l.z};
    hi = {h.x, h.y, h.z};
    r.a = make_float4(lo.x, lo.y, lo.z, __uint_as_float(i));
    r.b = make_float4(hi.x, hi.y, hi.z, 0.f);
    const uint32_t type = gd.kind == GEOM_INSTANCES ? PRIM_INSTANCE : (gd.kind == BNAMD_GEOM_USER ? PRIM_USER : PRIM_TET);
    r.c = make_float4(0.f, 0.f, 0.f, __uint_as_float(type));
  } else if (gd.kind == GEOM_TETS) {
    const int4 ix = ((const int4 *)gd.indices)[primID];
    const float4 *v4 = (const float4 *)gd.vertices;
    const float4 p0 = v4[ix.x], p1 = v4[ix.y], p2 = v4[ix.z], p3 = v4[ix.w];
    lo = {fminf(fminf(p0.x, p1.x), fminf(p2.x, p3.x)), fminf(fminf(p0.y, p1.y), fminf(p2.y, p3.y)),
          fminf(fminf(p0.z, p1.z), fminf(p2.z, p3.z))};
    hi = {fmaxf(fmaxf(p0.x, p1.x), fmaxf(p2.x, p3.x)), fmaxf(fmaxf(p0.y, p1.y), fmaxf(p2.y, p3.y)),
          fmaxf(fmaxf(p0.z, p1.z), fmaxf(p2.z, p3.z))};
    r.a = make_float4(lo.x, lo.y, lo.z, __uint_as_float(i));
    r.b = make_float4(hi.x, hi.y, hi.z, 0.f);
    r.c = make_float4(0.f, 0.f, 0.f, __uint_as_float(PRIM_TET));
  } else {
    const int *ix = gd.indices + 2 * primID;
    const vec3f a = load3(gd.vertices + 3 * ix[0]);
    const vec3f b = load3(gd.vertices + 3 * ix[1]);
    const float rad = gd.radii ? gd.radii[primID] : gd.defaultRadius;
    r.a = make_float4(a.x, a.y, a.z, __uint_as_float(i));
    r.b = make_float4(b.x, b.y, b.z, 0.f);
    r.c = make_float4(rad, 0.f, 0.f, __uint_as_float(PRIM_CYLINDER));
    lo = vmin(a - vec3f(rad), b - vec3f(rad));
    hi = vmax(a + vec3f(rad), b + vec3f(rad));
  }
  recs[i] = r;
  boxes[2 * i + 0] = make_float4(lo.x, lo.y, lo.z, 0.f);
  boxes[2 * i + 1] = make_float4(hi.x, hi.y, hi.z, 0.f);
}

// Order-preserving float <-> uint mapping for atomic min/max.
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t u) {
  const uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  memcpy(&f, &v, 4);
  return f;
}

// bounds[0..2] = min centroid, [3..5] = max centroid, [6..8] = min box lo,
// [9..11] = max box hi (ordered-uint encoded).
__global__ void k_bounds(const float4 *__restrict boxes, uint32_t n, uint32_t *__restrict bounds) {
  float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float bmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; ++k) {
      cmin[k] = fminf(cmin[k], c[k]);
      cmax[k] = fmaxf(cmax[k], c[k]);
      bmin[k] = fminf(bmin[k], l[k]);
      bmax[k] = fmaxf(bmax[k], h[k]);
    }
  }
  // wave reduction, then one atomic per wave
  for (int k = 0; k < 3; ++k) {
    for (int off = 32; off > 0; off >>= 1) {
      cmin[k] = fminf(cmin[k], __shfl_xor(cmin[k], off));
      cmax[k] = fmaxf(cmax[k], __shfl_xor(cmax[k], off));
      bmin[k] = fminf(bmin[k], __shfl_xor(bmin[k], off));
      bmax[k] = fmaxf(bmax[k], __shfl_xor(bmax[k], off));
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; ++k) {
      atomicMin(&bounds[k], f2ord(cmin[k]));
      atomicMax(&bounds[3 + k], f2ord(cmax[k]));
      atomicMin(&bounds[6 + k], f2ord(bmin[k]));
      atomicMax(&bounds[9 + k], f2ord(bmax[k]));
    }
  }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_morton(const float4 *__restrict boxes, uint32_t n, const uint32_t *__restrict bounds,
                         uint64_t *__restrict keys, uint32_t *__restrict vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 lo = boxes[2 * i], hi = boxes[2 * i + 1];
  const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float mn = ord2f(bounds[k]), mx = ord2f(bounds[3 + k]);
    const float ext = mx - mn;
    float f = ext > 0.f ? (c[k] - mn) / ext : 0.f;
    f = fminf(fmaxf(f, 0.f), 1.f);
    q[k] = min((uint32_t)(f * 2097152.f), 2097151u);
  }
  keys[i] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
  vals[i] = i;
}

__device__ __forceinline__ int kdelta(const uint64_t *__restrict keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll((long long)(a ^ b));
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs,
// octrees, and k-d trees": one thread per internal node.
__global__ void k_karras(const uint64_t *__restrict keys, int n, BinNode *__restrict nodes,
                         int *__restrict parent, int2 *__restrict range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(keys, n, i, i - d);
  int lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + min(d, 0);
  const int first = min(i, j), last = max(i, j);
  const int left = (first == gamma) ? ~gamma : gamma;
  const int right = (last == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  nodes[i].left = left;
  nodes[i].right = right;
  range[i] = make_int2(first, last - first + 1);
  // parent links: internal nodes at [0, n-1), leaves at [n-1, 2n-1)
  parent[left >= 0 ? left : (n - 1 + ~left)] = i;
  parent[right >= 0 ? right : (n - 1 + ~right)] = i;
}

__global__ void k_refit(const float4 *__restrict boxes, const uint32_t *__restrict sortedIdx, int n,
                        BinNode *__restrict nodes, const int *__restrict parent,
                        uint32_t *__restrict flags) {
  const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
  if (leaf >= n) return;
  int p = parent[n - 1 + leaf];
  while (p >= 0) {
    // the first child to arrive stops; the second goes on with both boxes.
    // acquire-release at device scope: it publishes the box this thread
    // wrote one level down and makes the sibling's box visible here
    if (__hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    const BinNode nd = nodes[p];
    float l[3], h[3];
    for (int side = 0; side < 2; ++side) {
      const int c = side ? nd.right : nd.left;
      float cl[3], ch[3];
      if (c < 0) {
        const uint32_t pi = sortedIdx[~c];
        const float4 lo = boxes[2 * pi], hi = boxes[2 * pi + 1];
        cl[0] = lo.x; cl[1] = lo.y; cl[2] = lo.z;
        ch[0] = hi.x; ch[1] = hi.y; ch[2] = hi.z;
      } else {
        for (int k = 0; k < 3; ++k) { cl[k] = nodes[c].lo[k]; ch[k] = nodes[c].hi[k]; }
      }
      for (int k = 0; k < 3; ++k) {
        l[k] = side ? fminf(l[k], cl[k]) : cl[k];
        h[k] = side ? fmaxf(h[k], ch[k]) : ch[k];
      }
    }
    for (int k = 0; k < 3; ++k) { nodes[p].lo[k] = l[k]; nodes[p].hi[k] = h[k]; }
    p = parent[p];
  }
}

// ---------------------------------------------------------------------------
// PLOC: parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018). The clusters start as the Morton-sorted primitives; each
// round every cluster finds its nearest neighbour among the kPlocRadius
// clusters on either side (the smallest surface area of the union, ties to
// the smaller index, a strict order on pairs, so the closest pair in any
// window is mutual and every round merges), mutual nearest neighbours merge
// into a new binary node, and the survivors are compacted in order. The
// binary tree it makes has a far lower SAH cost than the radix tree of the
// Morton codes alone. A node's leaves are then renumbered in depth-first
// order so that every subtree's primitives are contiguous (range).
constexpr int kPlocRadius = 16;

__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}

__global__ void k_ploc_init(const float4 *__restrict boxes, const uint32_t *__restrict sortedIdx, int n,
                            int *__restrict cid, float4 *__restrict cbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t pi = sortedIdx[i];
  cid[i] = ~i;
  cbox[2 * i] = boxes[2 * pi];
  cbox[2 * i + 1] = boxes[2 * pi + 1];
}

// nearest neighbour of each cluster within the window; the window's boxes
// are staged in LDS (256 clusters and the radius on either side)
__global__ void __launch_bounds__(256) k_ploc_nn(const float4 *__restrict cbox, int nc, int *__restrict nn) {
  __shared__ float4 lo[256 + 2 * kPlocRadius], hi[256 + 2 * kPlocRadius];
  const int base = blockIdx.x * 256 - kPlocRadius;
  for (int t = threadIdx.x; t < 256 + 2 * kPlocRadius; t += 256) {
    const int c = base + t;
    if (c >= 0 && c < nc) {
      lo[t] = cbox[2 * c];
      hi[t] = cbox[2 * c + 1];
    }
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  const int li = threadIdx.x + kPlocRadius;
  const float4 a = lo[li], b = hi[li];
  float best = FLT_MAX;
  int bj = -1;
  const int j0 = max(0, i - kPlocRadius), j1 = min(nc - 1, i + kPlocRadius);
  for (int j = j0; j <= j1; ++j) {
    if (j == i) continue;
    const int lj = j - base;
    const float4 ul = make_float4(fminf(a.x, lo[lj].x), fminf(a.y, lo[lj].y), fminf(a.z, lo[lj].z), 0.f);
    const float4 uh = make_float4(fmaxf(b.x, hi[lj].x), fmaxf(b.y, hi[lj].y), fmaxf(b.z, hi[lj].z), 0.f);
    const float d = half_area(ul, uh);
    if (d < best) {  // j ascends, so a tie keeps the smaller index
      best = d;
      bj = j;
    }
  }
  nn[i] = bj;
}

// Mutual nearest neighbours i < j merge into a new node held by cluster i;
// cluster j is dropped (keep[j] = 0). count: leaves under each new node.
__global__ void k_ploc_merge(int *__restrict cid, float4 *__restrict cbox, const int *__restrict nn, int nc,
                             BinNode *__restrict nodes, int *__restrict count, uint32_t *__restrict nodeCounter,
                             uint32_t *__restrict keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const int j = nn[i];
  if (j < 0 || nn[j] != i) {
    keep[i] = 1u;
    return;
  }
  if (j < i) {
    keep[i] = 0u;
    return;
  }
  const uint32_t node = atomicAdd(nodeCounter, 1u);
  const int a = cid[i], b = cid[j];
  const float4 al = cbox[2 * i], ah = cbox[2 * i + 1], bl = cbox[2 * j], bh = cbox[2 * j + 1];
  BinNode nd;
  nd.lo[0] = fminf(al.x, bl.x); nd.lo[1] = fminf(al.y, bl.y); nd.lo[2] = fminf(al.z, bl.z);
  nd.hi[0] = fmaxf(ah.x, bh.x); nd.hi[1] = fmaxf(ah.y, bh.y); nd.hi[2] = fmaxf(ah.z, bh.z);
  nd.left = a;
  nd.right = b;
  nodes[node] = nd;
  count[node] = (a < 0 ? 1 : count[a]) + (b < 0 ? 1 : count[b]);
  cid[i] = (int)node;
  cbox[2 * i] = make_float4(nd.lo[0], nd.lo[1], nd.lo[2], 0.f);
  cbox[2 * i + 1] = make_float4(nd.hi[0], nd.hi[1], nd.hi[2], 0.f);
  keep[i] = 1u;
}

__global__ void k_ploc_compact(const int *__restrict cid, const float4 *__restrict cbox,
                               const uint32_t *__restrict keep, const uint32_t *__restrict pos, int nc,
                               int *__restrict cid2, float4 *__restrict cbox2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc || !keep[i]) return;
  const uint32_t o = pos[i];
  cid2[o] = cid[i];
  cbox2[2 * o] = cbox[2 * i];
  cbox2[2 * o + 1] = cbox[2 * i + 1];
}

// Depth-first leaf numbering, one round of nodes at a time from the root's
// round down (a node's parent is made in a later round): a node's leaves
// start at first[node], its left subtree's first; leaves get their new
// sorted position (newPos), nodes their range {first, count}.
__global__ void k_ploc_first(const BinNode *__restrict nodes, const int *__restrict count, uint32_t begin,
                             uint32_t end, int *__restrict first, int *__restrict newPos, int2 *__restrict range) {
  const uint32_t p = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= end) return;
  const int f = first[p];
  const BinNode nd = nodes[p];
  const int cl = nd.left < 0 ? 1 : count[nd.left];
  if (nd.left < 0) newPos[~nd.left] = f; else first[nd.left] = f;
  if (nd.right < 0) newPos[~nd.right] = f + cl; else first[nd.right] = f + cl;
  range[p] = make_int2(f, count[p]);
}

// leaves renumbered: node child references and the sorted primitive order
__global__ void k_ploc_renumber(BinNode *__restrict nodes, int numNodes, const int *__restrict newPos,
                                const uint32_t *__restrict sortedIdx, int n, uint32_t *__restrict sorted2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sorted2[newPos[i]] = sortedIdx[i];
  if (i < numNodes) {
    if (nodes[i].left < 0) nodes[i].left = ~newPos[~nodes[i].left];
    if (nodes[i].right < 0) nodes[i].right = ~newPos[~nodes[i].right];
  }
}

struct CollapseTask {
  int bin;      // binary node (>= 0 internal, < 0 ~leaf)
  uint32_t out; // BVH8 node index
};

__device__ void child_box(int c, const BinNode *__restrict nodes, const float4 *__restrict boxes,
                          const uint32_t *__restrict sortedIdx, float *lo, float *hi) {
  if (c < 0) {
    const uint32_t pi = sortedIdx[~c];
    const float4 l = boxes[2 * pi], h = boxes[2 * pi + 1];
    lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
    hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
  } else {
    for (int k = 0; k < 3; ++k) { lo[k] = nodes[c].lo[k]; hi[k] = nodes[c].hi[k]; }
  }
}

__device__ __forceinline__ int child_count(int c, const int2 *__restrict range) {
  return c < 0 ? 1 : range[c].y;
}
__device__ __forceinline__ int child_first(int c, const int2 *__restrict range) {
  return c < 0 ? ~c : range[c].x;
}

__global__ void k_collapse(const CollapseTask *__restrict tasks, uint32_t numTasks,
                           const BinNode *__restrict bnodes, const int2 *__restrict range,
                           const float4 *__restrict boxes, const uint32_t *__restrict sortedIdx,
                           const PrimRec *__restrict srcRecs, BNode8 *__restrict out,
                           PrimRec *__restrict outRecs, uint32_t *__restrict counters,
                           CollapseTask *__restrict nextTasks) {
  const uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= numTasks) return;
  const CollapseTask task = tasks[ti];
  int ch[8];
  float clo[8][3], chi[8][3];
  int nc = 0;
  if (task.bin < 0) {
    ch[nc++] = task.bin;
  } else {
    ch[nc++] = bnodes[task.bin].left;
    ch[nc++] = bnodes[task.bin].right;
  }
  for (int i = 0; i < nc; ++i) child_box(ch[i], bnodes, boxes, sortedIdx, clo[i], chi[i]);
  // Greedy expansion: open the child with the largest surface area until
  // all eight slots are used.
  while (nc < 8) {
    int best = -1;
    float bestArea = -1.f;
    for (int i = 0; i < nc; ++i) {
      if (ch[i] < 0) continue;
      const float dx = chi[i][0] - clo[i][0], dy = chi[i][1] - clo[i][1], dz = chi[i][2] - clo[i][2];
      const float area = dx * dy + dy * dz + dz * dx;
      if (area > bestArea) { bestArea = area; best = i; }
    }
    if (best < 0) break;
    const int b = ch[best];
    ch[best] = bnodes[b].left;
    ch[nc] = bnodes[b].right;
    child_box(ch[best], bnodes, boxes, sortedIdx, clo[best], chi[best]);
    child_box(ch[nc], bnodes, boxes, sortedIdx, clo[nc], chi[nc]);
    ++nc;
  }
  // node box
  float nlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, nhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < 3; ++k) {
      nlo[k] = fminf(nlo[k], clo[i][k]);
      nhi[k] = fmaxf(nhi[k], chi[i][k]);
    }
  // Slot assignment: slot s (bit k set = +axis k) gets the child whose
  // centroid lies furthest toward that corner, so that visiting slots in
  // order k ^ octant approximates front-to-back order.
  int slotOf[8];
  for (int i = 0; i < 8; ++i) slotOf[i] = -1;
  {
    const float cx = 0.5f * (nlo[0] + nhi[0]), cy = 0.5f * (nlo[1] + nhi[1]), cz = 0.5f * (nlo[2] + nhi[2]);
    float off[8][3];
    for (int i = 0; i < nc; ++i) {
      off[i][0] = 0.5f * (clo[i][0] + chi[i][0]) - cx;
      off[i][1] = 0.5f * (clo[i][1] + chi[i][1]) - cy;
      off[i][2] = 0.5f * (clo[i][2] + chi[i][2]) - cz;
    }
    uint32_t usedSlots = 0;
    for (int it = 0; it < nc; ++it) {
      float bestScore = -FLT_MAX;
      int bi = -1, bs = -1;
      for (int i = 0; i < nc; ++i) {
        if (slotOf[i] >= 0) continue;
        for (int s = 0; s < 8; ++s) {
          if (usedSlots & (1u << s)) continue;
          const float score = ((s & 1) ? off[i][0] : -off[i][0]) + ((s & 2) ? off[i][1] : -off[i][1]) +
                              ((s & 4) ? off[i][2] : -off[i][2]);
          if (score > bestScore) { bestScore = score; bi = i; bs = s; }
        }
      }
      slotOf[bi] = bs;
      usedSlots |= 1u << bs;
    }
  }
  int childOfSlot[8];
  for (int s = 0; s < 8; ++s) childOfSlot[s] = -1;
  for (int i = 0; i < nc; ++i) childOfSlot[slotOf[i]] = i;

  int numInner = 0, numLeafPrims = 0;
  for (int i = 0; i < nc; ++i) {
    const int cnt = child_count(ch[i], range);
    if (cnt > kLeafMax) ++numInner; else numLeafPrims += cnt;
  }
  const uint32_t childBase = numInner ? atomicAdd(&counters[0], (uint32_t)numInner) : 0u;
  const uint32_t primBase = numLeafPrims ? atomicAdd(&counters[1], (uint32_t)numLeafPrims) : 0u;
  const uint32_t taskBase = numInner ? atomicAdd(&counters[2], (uint32_t)numInner) : 0u;

  BNode8 node;
  node.ox = nlo[0]; node.oy = nlo[1]; node.oz = nlo[2];
  uint8_t ebits[3];
  double scale[3];
  for (int k = 0; k < 3; ++k) {
    const double ext = (double)nhi[k] - (double)nlo[k];
    int e;
    (void)frexp(ext / 65000.0, &e);  // ext/65000 < 2^e
    e = max(e, -100);
    scale[k] = ldexp(1.0, e);
    ebits[k] = (uint8_t)(e + 127);
  }
  node.ex = ebits[0]; node.ey = ebits[1]; node.ez = ebits[2];
  node.childBase = childBase;
  node.primBase = primBase;
  uint8_t imask = 0;
  int rank = 0, poff = 0;
  for (int s = 0; s < 8; ++s) {
    const int i = childOfSlot[s];
    node.meta[s] = 0;
    if (i < 0) {
      for (int k = 0; k < 3; ++k) { node.plane(k, s) = 0xffff; node.plane(3 + k, s) = 0; }
      continue;
    }
    const double o[3] = {nlo[0], nlo[1], nlo[2]};
    for (int k = 0; k < 3; ++k) {
      double ql = floor(((double)clo[i][k] - o[k]) / scale[k]);
      double qh = ceil(((double)chi[i][k] - o[k]) / scale[k]);
      ql = fmax(0.0, ql - 1.0);
      qh = fmin(65535.0, qh + 1.0);
      node.plane(k, s) = (uint16_t)ql;
      node.plane(3 + k, s) = (uint16_t)qh;
    }
    const int cnt = child_count(ch[i], range);
    if (cnt > kLeafMax) {
      imask |= (uint8_t)(1u << s);
      CollapseTask nt;
      nt.bin = ch[i];
      nt.out = childBase + rank;
      nextTasks[taskBase + rank] = nt;
      ++rank;
    } else {
      node.meta[s] = (uint8_t)((cnt << 5) | poff);
      const int first = child_first(ch[i], range);
      for (int j = 0; j < cnt; ++j) outRecs[primBase + poff + j] = srcRecs[sortedIdx[first + j]];
      poff += cnt;
    }
  }
  node.imask = imask;
  out[task.out] = node;
}

__global__ void k_empty_root(BNode8 *out) {
  BNode8 node;
  node.ox = node.oy = node.oz = 0.f;
  node.ex = node.ey = node.ez = 127;
  node.imask = 0;
  node.childBase = 0;
  node.primBase = 0;
  for (int s = 0; s < 8; ++s) {
    node.meta[s] = 0;
    for (int k = 0; k < 3; ++k) { node.plane(k, s) = 0xffff; node.plane(3 + k, s) = 0; }
  }
  out[0] = node;
}

// Decoded (float) box of one slot of a node.
void slot_box(const BNode8 &nd, int s, float lo[3], float hi[3]) {
  const float o[3] = {nd.ox, nd.oy, nd.oz};
  const int e[3] = {nd.ex, nd.ey, nd.ez};
  for (int k = 0; k < 3; ++k) {
    const float sc = ldexpf(1.f, e[k] - 127);
    lo[k] = o[k] + (float)nd.plane(k, s) * sc;
    hi[k] = o[k] + (float)nd.plane(3 + k, s) * sc;
  }
}

void prim_box(const PrimRec &r, uint32_t type, float lo[3], float hi[3]) {
  const float a[3] = {r.a.x, r.a.y, r.a.z}, b[3] = {r.b.x, r.b.y, r.b.z}, c[3] = {r.c.x, r.c.y, r.c.z};
  for (int k = 0; k < 3; ++k) {
    if (type == PRIM_TRIANGLE) {
      lo[k] = std::min(a[k], std::min(a[k] + b[k], a[k] + c[k]));
      hi[k] = std::max(a[k], std::max(a[k] + b[k], a[k] + c[k]));
    } else if (type == PRIM_SPHERE) {
      lo[k] = a[k] - r.b.x;
      hi[k] = a[k] + r.b.x;
    } else if (type == PRIM_TET || type == PRIM_INSTANCE) {
      lo[k] = a[k];
      hi[k] = b[k];
    } else {
      lo[k] = std::min(a[k], b[k]) - r.c.x;
      hi[k] = std::max(a[k], b[k]) + r.c.x;
    }
  }
}

// inner box within outer box up to float rounding of the decode
bool box_inside(const float ilo[3], const float ihi[3], const float olo[3], const float ohi[3]) {
  for (int k = 0; k < 3; ++k) {
    const float tol = 1e-6f * (1.f + std::max(fabsf(olo[k]), fabsf(ohi[k])));
    if (ilo[k] < olo[k] - tol || ihi[k] > ohi[k] + tol) return false;
  }
  return true;
}

// Host-side check of a finished BVH8: every child / primitive reference in
// range, every primitive referenced exactly once, and every decoded child
// box containing what it bounds (the primitives of a leaf slot, all slots of
// an inner child node). Guards the traversal kernels against out-of-bounds
// fetches and the quantization against dropping geometry. It copies the
// whole tree to the host, so builds run it only with BNAMD_VALIDATE_BVH=1;
// bnamd_world_validate runs it on demand (the tests do, on every kind of
// world they build).
void validate_bvh(const World &w) {
  std::vector<BNode8> nodes(w.numNodes);
  BN_HIP(hipMemcpy(nodes.data(), w.nodes.ptr, nodes.size() * sizeof(BNode8), hipMemcpyDeviceToHost));
  std::vector<PrimRec> prims(w.numPrims);
  BN_HIP(hipMemcpy(prims.data(), w.prims.ptr, prims.size() * sizeof(PrimRec), hipMemcpyDeviceToHost));
  std::vector<uint8_t> seen(w.numPrims, 0);
  for (uint32_t i = 0; i < w.numNodes; ++i) {
    const BNode8 &nd = nodes[i];
    int rank = 0;
    for (int s = 0; s < 8; ++s) {
      if (nd.imask & (1u << s)) {
        const uint32_t ci = nd.childBase + rank;
        BN_REQUIRE(ci < w.numNodes && ci > i, "BVH8 check: bad child index");
        ++rank;
      } else {
        const int cnt = nd.meta[s] >> 5, off = nd.meta[s] & 31;
        for (int j = 0; j < cnt; ++j) {
          const uint64_t p = (uint64_t)nd.primBase + off + j;
          BN_REQUIRE(p < w.numPrims, "BVH8 check: primitive index out of range");
          BN_REQUIRE(seen[p] == 0, "BVH8 check: primitive referenced twice");
          seen[p] = 1;
        }
      }
    }
    BN_REQUIRE(nd.ex >= 1 && nd.ey >= 1 && nd.ez >= 1 && nd.ex <= 254 && nd.ey <= 254 && nd.ez <= 254,
               "BVH8 check: bad quantization exponent");
  }
  // Bottom-up (children have larger indices than their parent): the union
  // of the primitive boxes below each node must lie inside the slot box
  // that leads to it.
  std::vector<float> content(6 * (size_t)w.numNodes);
  for (int64_t i = (int64_t)w.numNodes - 1; i >= 0; --i) {
    const BNode8 &nd = nodes[i];
    float *u = &content[6 * i];
    for (int k = 0; k < 3; ++k) {
      u[k] = INFINITY;
      u[3 + k] = -INFINITY;
    }
    int rank = 0;
    for (int s = 0; s < 8; ++s) {
      float blo[3], bhi[3], lo[3], hi[3];
      slot_box(nd, s, blo, bhi);
      bool any = false;
      if (nd.imask & (1u << s)) {
        const float *c = &content[6 * (size_t)(nd.childBase + rank)];
        ++rank;
        for (int k = 0; k < 3; ++k) {
          lo[k] = c[k];
          hi[k] = c[3 + k];
        }
        any = true;
      } else {
        const int cnt = nd.meta[s] >> 5, off = nd.meta[s] & 31;
        for (int k = 0; k < 3; ++k) {
          lo[k] = INFINITY;
          hi[k] = -INFINITY;
        }
        for (int j = 0; j < cnt; ++j) {
          const PrimRec &r = prims[nd.primBase + off + j];
          uint32_t type;
          memcpy(&type, &r.c.w, 4);
          float plo[3], phi[3];
          prim_box(r, type & 0xffu, plo, phi);
          for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], plo[k]);
            hi[k] = std::max(hi[k], phi[k]);
          }
          any = true;
        }
      }
      if (!any) continue;
      BN_REQUIRE(box_inside(lo, hi, blo, bhi), "BVH8 check: geometry outside the slot box that leads to it");
      for (int k = 0; k < 3; ++k) {
        u[k] = std::min(u[k], lo[k]);
        u[3 + k] = std::max(u[3 + k], hi[k]);
      }
    }
  }
  for (uint32_t p = 0; p < w.numPrims; ++p) {
    BN_REQUIRE(seen[p], "BVH8 check: unreferenced primitive");
    uint32_t flat, type;
    memcpy(&flat, &prims[p].a.w, 4);
    memcpy(&type, &prims[p].c.w, 4);
    BN_REQUIRE(flat < w.numPrims && (type & 0xffu) <= 4u, "BVH8 check: corrupt primitive record");
  }
}

}  // namespace

void check_world_bvh(const World &w) { validate_bvh(w); }

// Binary tree builder under the 8-wide collapse (see world.h).
enum { kBuildPloc = 0, kBuildRadix = 1 };
int build_mode(bool mortonOrder) {
  const char *m = getenv("BNAMD_BVH_BUILD");
  if (m && std::string(m) == "radix") return kBuildRadix;
  if (m && std::string(m) == "ploc") return kBuildPloc;
  return mortonOrder ? kBuildRadix : kBuildPloc;
}

void build_world_bvh(World &w, hipStream_t stream, bool mortonOrder) {
  const auto t0 = std::chrono::steady_clock::now();
  const uint32_t n = w.numPrims;
  const int numGeoms = (int)w.hgeoms.size();
  w.geoms.upload(w.hgeoms.data(), w.hgeoms.size(), stream);
  if (n == 0) {
    w.nodes.alloc(1);
    w.prims.alloc(1);
    k_empty_root<<<1, 1, 0, stream>>>(w.nodes.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipStreamSynchronize(stream));
    w.numNodes = 1;
    w.depth = 1;
    return;
  }
  DevBuf<PrimRec> recs(n);
  DevBuf<float4> boxes(2ull * n);
  k_gather<<<div_up(n, 256), 256, 0, stream>>>(w.geoms.ptr, numGeoms, w.materials.ptr, n, recs.ptr, boxes.ptr);
  BN_HIP(hipGetLastError());

  DevBuf<uint32_t> bounds(12);
  {
    uint32_t init[12];
    for (int k = 0; k < 3; ++k) {
      init[k] = 0xffffffffu; init[3 + k] = 0u; init[6 + k] = 0xffffffffu; init[9 + k] = 0u;
    }
    BN_HIP(hipMemcpyAsync(bounds.ptr, init, sizeof(init), hipMemcpyHostToDevice, stream));
  }
  k_bounds<<<std::min(div_up(n, 256), 2048), 256, 0, stream>>>(boxes.ptr, n, bounds.ptr);
  BN_HIP(hipGetLastError());

  DevBuf<uint64_t> keys(n), keysSorted(n);
  DevBuf<uint32_t> vals(n), valsSorted(n);
  k_morton<<<div_up(n, 256), 256, 0, stream>>>(boxes.ptr, n, bounds.ptr, keys.ptr, vals.ptr);
  BN_HIP(hipGetLastError());
  size_t tempBytes = 0;
  BN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tempBytes, keys.ptr, keysSorted.ptr, vals.ptr,
                                            valsSorted.ptr, (int)n, 0, 63, stream));
  DevBuf<uint8_t> temp(tempBytes);
  BN_HIP(hipcub::DeviceRadixSort::SortPairs(temp.ptr, tempBytes, keys.ptr, keysSorted.ptr, vals.ptr,
                                            valsSorted.ptr, (int)n, 0, 63, stream));

  const int nInner = (int)n - 1;
  DevBuf<BinNode> bnodes(std::max(nInner, 1));
  DevBuf<int2> range(std::max(nInner, 1));
  int rootBin = ~0;
  uint32_t *sortedPrims = valsSorted.ptr;
  DevBuf<uint32_t> sorted2;
  if (nInner > 0 && build_mode(mortonOrder) == kBuildRadix) {
    // the radix tree of the Morton codes with bottom-up boxes
    DevBuf<int> parent(2ull * n);
    DevBuf<uint32_t> flags(nInner);
    BN_HIP(hipMemsetAsync(parent.ptr, 0xff, parent.bytes(), stream));
    BN_HIP(hipMemsetAsync(flags.ptr, 0, flags.bytes(), stream));
    k_karras<<<div_up(nInner, 256), 256, 0, stream>>>(keysSorted.ptr, (int)n, bnodes.ptr, parent.ptr,
                                                     range.ptr);
    BN_HIP(hipGetLastError());
    k_refit<<<div_up(n, 256), 256, 0, stream>>>(boxes.ptr, valsSorted.ptr, (int)n, bnodes.ptr, parent.ptr,
                                               flags.ptr);
    BN_HIP(hipGetLastError());
    rootBin = 0;
  } else if (nInner > 0) {
    // PLOC over the Morton-sorted primitives
    DevBuf<int> cidA(n), cidB(n), nn(n), count(nInner), first(nInner), newPos(n);
    DevBuf<float4> boxA(2ull * n), boxB(2ull * n);
    DevBuf<uint32_t> keep(n), pos(n), nodeCounter(1);
    BN_HIP(hipMemsetAsync(nodeCounter.ptr, 0, 4, stream));
    k_ploc_init<<<div_up(n, 256), 256, 0, stream>>>(boxes.ptr, valsSorted.ptr, (int)n, cidA.ptr, boxA.ptr);
    BN_HIP(hipGetLastError());
    size_t scanBytes = 0;
    BN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, keep.ptr, pos.ptr, (int)n, stream));
    DevBuf<uint8_t> scanTemp(scanBytes);
    std::vector<uint32_t> roundEnd;  // node counter after each round
    int nc = (int)n;
    int *cid = cidA.ptr, *cid2 = cidB.ptr;
    float4 *cb = boxA.ptr, *cb2 = boxB.ptr;
    while (nc > 1) {
      BN_REQUIRE(roundEnd.size() < 4096, "PLOC made no progress");
      k_ploc_nn<<<div_up(nc, 256), 256, 0, stream>>>(cb, nc, nn.ptr);
      k_ploc_merge<<<div_up(nc, 256), 256, 0, stream>>>(cid, cb, nn.ptr, nc, bnodes.ptr, count.ptr,
                                                        nodeCounter.ptr, keep.ptr);
      BN_HIP(hipGetLastError());
      BN_HIP(hipcub::DeviceScan::ExclusiveSum(scanTemp.ptr, scanBytes, keep.ptr, pos.ptr, nc, stream));
      k_ploc_compact<<<div_up(nc, 256), 256, 0, stream>>>(cid, cb, keep.ptr, pos.ptr, nc, cid2, cb2);
      BN_HIP(hipGetLastError());
      uint32_t tail[3];
      BN_HIP(hipMemcpyAsync(&tail[0], pos.ptr + nc - 1, 4, hipMemcpyDeviceToHost, stream));
      BN_HIP(hipMemcpyAsync(&tail[1], keep.ptr + nc - 1, 4, hipMemcpyDeviceToHost, stream));
      BN_HIP(hipMemcpyAsync(&tail[2], nodeCounter.ptr, 4, hipMemcpyDeviceToHost, stream));
      BN_HIP(hipStreamSynchronize(stream));
      nc = (int)(tail[0] + tail[1]);
      roundEnd.push_back(tail[2]);
      std::swap(cid, cid2);
      std::swap(cb, cb2);
    }
    BN_REQUIRE(roundEnd.back() == (uint32_t)nInner, "PLOC node count");
    // the root is the last node made; depth-first leaf numbering from it
    rootBin = nInner - 1;
    const int zero = 0;
    BN_HIP(hipMemcpyAsync(first.ptr + rootBin, &zero, 4, hipMemcpyHostToDevice, stream));
    for (size_t r = roundEnd.size(); r-- > 0;) {
      const uint32_t begin = r ? roundEnd[r - 1] : 0u, end = roundEnd[r];
      if (end > begin)
        k_ploc_first<<<div_up(end - begin, 256), 256, 0, stream>>>(bnodes.ptr, count.ptr, begin, end, first.ptr,
                                                                  newPos.ptr, range.ptr);
    }
    sorted2.alloc(n);
    k_ploc_renumber<<<div_up(n, 256), 256, 0, stream>>>(bnodes.ptr, nInner, newPos.ptr, valsSorted.ptr, (int)n,
                                                        sorted2.ptr);
    BN_HIP(hipGetLastError());
    sortedPrims = sorted2.ptr;
  }

  // top-down collapse, one launch per BVH8 level
  w.nodes.alloc(std::max<uint32_t>(n, 1));
  w.prims.alloc(n);
  DevBuf<CollapseTask> taskA(std::max<uint32_t>(n, 1)), taskB(std::max<uint32_t>(n, 1));
  DevBuf<uint32_t> counters(3);
  uint32_t hc[3] = {1u, 0u, 0u};
  BN_HIP(hipMemcpyAsync(counters.ptr, hc, sizeof(hc), hipMemcpyHostToDevice, stream));
  CollapseTask root;
  root.bin = rootBin;
  root.out = 0;
  BN_HIP(hipMemcpyAsync(taskA.ptr, &root, sizeof(root), hipMemcpyHostToDevice, stream));
  uint32_t numTasks = 1;
  int depth = 0;
  CollapseTask *cur = taskA.ptr, *nxt = taskB.ptr;
  while (numTasks > 0) {
    ++depth;
    BN_REQUIRE(depth <= 64, "BVH8 collapse exceeded 64 levels");
    const uint32_t zero = 0;
    BN_HIP(hipMemcpyAsync(counters.ptr + 2, &zero, 4, hipMemcpyHostToDevice, stream));
    k_collapse<<<div_up(numTasks, 64), 64, 0, stream>>>(cur, numTasks, bnodes.ptr, range.ptr, boxes.ptr,
                                                       sortedPrims, recs.ptr, w.nodes.ptr, w.prims.ptr,
                                                       counters.ptr, nxt);
    BN_HIP(hipGetLastError());
    BN_HIP(hipMemcpyAsync(hc, counters.ptr, sizeof(hc), hipMemcpyDeviceToHost, stream));
    BN_HIP(hipStreamSynchronize(stream));
    numTasks = hc[2];
    std::swap(cur, nxt);
  }
  BN_REQUIRE(hc[1] == n, "BVH8 collapse lost primitives");
  w.numNodes = hc[0];
  w.depth = depth;
  const char *check = getenv("BNAMD_VALIDATE_BVH");
  if (check && check[0] == '1') validate_bvh(w);
  uint32_t hb[12];
  BN_HIP(hipMemcpy(hb, bounds.ptr, sizeof(hb), hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; ++k) {
    w.lo[k] = ord2f(hb[6 + k]);
    w.hi[k] = ord2f(hb[9 + k]);
  }
  BN_HIP(hipStreamSynchronize(stream));
  w.buildMs = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace bnamd
