// NanoVDB scalar fields on the GPU: barney's NanoVDBData and
// NanoVDBDataSampler (barney/volume/NanoVDB.h:37-127, NanoVDB.cu:28-102,
// 204-235) over a NanoVDB 32.x grid buffer kept in HBM as given. The buffer's
// layout (3rdParty/nanovdb/include/nanovdb/NanoVDB.h):
//
//   GridData 672 B (map at 296, world box at 560, grid type at 636)
//   TreeData 64 B (node offsets; the root follows at +64)
//   RootData 64 B (index box, tile count at 24, background at 28), then
//            32-byte tiles {key u64, child i64 (from the root), state, value}
//   upper internal node 32^3 (270400 B): child mask at 4128, 8-byte table
//            at 8256 (child offset from the node, or a tile value)
//   lower internal node 16^3 (33856 B): child mask at 544, table at 1088
//   leaf 8^3: flags at 15; float values at 96, or {min at 80, quantum at 84,
//            codes at 96} for Fp4 / Fp8 / Fp16 and FpN (1 << (flags >> 5)
//            bits per code); value = code * quantum + min
//
// A voxel lookup is ReadAccessor::getValue's walk (root tile by key, then
// the tables by the coordinate's bits); a sample is SampleFromVoxels<1>'s
// trilinear stencil at the index position barney computes. The device
// record also carries the macrocell grid and transfer function under the
// names DeltaTrack reads, so the volume path tracer renders it.
#pragma once
#include "volume.h"

namespace bnamd {

enum : int { kNvFloat = 1, kNvFp4 = 13, kNvFp8 = 14, kNvFp16 = 15, kNvFpN = 16 };

struct NvdbDD {
  // macrocell grid + transfer function (as in VolumeDD)
  vec3f worldLo, worldHi;
  const float *majorants;
  int mx, my, mz;
  vec3f mcOrigin, mcSpacing;
  TFDD tf;
  // the grid
  const uint8_t *root;  // RootData (the buffer is 32-byte aligned throughout)
  uint32_t numTiles;
  float background;
  int type;
  int ilo[3], ihi[3];         // index box (the root's)
  float invMatF[9], vecF[3];  // Map::mInvMatF, mVecF
  // leaf lookup: per 8^3 block over the index box (grown by one voxel),
  // the block's leaf offset from the grid start in 32-byte units, 0 where
  // no leaf holds it (tiles, background: the tree walk answers); NULL when
  // the grid has no table
  const uint8_t *base;
  const uint32_t *leafTable;
  int tlo[3], tdim[3];
};

template <class T>
BN_HD T nv_load(const uint8_t *p) {
  return *reinterpret_cast<const T *>(p);
}

BN_HD bool nv_bit(const uint8_t *mask, uint32_t n) {
  return (nv_load<uint64_t>(mask + 8 * (n >> 6)) >> (n & 63)) & 1ull;
}

// LeafData<T>::getValue at leaf-local offset n
BN_HD float nv_leaf_value(int type, const uint8_t *leaf, uint32_t n) {
#pragma clang fp contract(off)
  if (type == kNvFloat) return nv_load<float>(leaf + 96 + 4 * n);
  const float mn = nv_load<float>(leaf + 80), q = nv_load<float>(leaf + 84);
  const uint8_t *code = leaf + 96;
  uint32_t c;
  if (type == kNvFp8) {
    c = code[n];
  } else if (type == kNvFp16) {
    c = nv_load<uint16_t>(code + 2 * n);
  } else {
    // Fp4 is FpN at 4 bits: 2^b bits per code, 8 >> b codes per byte below 16
    const int b = (type == kNvFp4) ? 2 : (leaf[15] >> 5);
    if (b >= 4) {
      c = nv_load<uint16_t>(code + 2 * n);
    } else {
      const uint32_t perByteLog = 3 - b;  // log2 of codes per byte
      const uint32_t byte = code[n >> perByteLog];
      c = (byte >> ((n & ((1u << perByteLog) - 1)) << b)) & ((1u << (1 << b)) - 1);
    }
  }
  return (float)c * q + mn;
}

// Where voxel (x, y, z) lives: a leaf (its address and the voxel's offset in
// it), or a constant (a tile value or the background).
struct NvSpot {
  const uint8_t *leaf;
  uint32_t n;
  float value;
};

BN_HD NvSpot nv_locate(const NvdbDD &g, int x, int y, int z) {
  const uint64_t key = ((uint64_t)((uint32_t)z >> 12)) | ((uint64_t)((uint32_t)y >> 12) << 21) |
                       ((uint64_t)((uint32_t)x >> 12) << 42);
  const uint8_t *tile = nullptr;
  for (uint32_t i = 0; i < g.numTiles; ++i) {
    const uint8_t *t = g.root + 64 + 32 * (size_t)i;
    if (nv_load<uint64_t>(t) == key) {
      tile = t;
      break;
    }
  }
  if (!tile) return NvSpot{nullptr, 0u, g.background};
  const int64_t child = nv_load<int64_t>(tile + 8);
  if (child == 0) return NvSpot{nullptr, 0u, nv_load<float>(tile + 20)};
  const uint8_t *up = g.root + child;
  uint32_t n = (((uint32_t)x & 4095u) >> 7 << 10) | (((uint32_t)y & 4095u) >> 7 << 5) | (((uint32_t)z & 4095u) >> 7);
  if (!nv_bit(up + 4128, n)) return NvSpot{nullptr, 0u, nv_load<float>(up + 8256 + 8 * (size_t)n)};
  const uint8_t *lo = up + nv_load<int64_t>(up + 8256 + 8 * (size_t)n);
  n = (((uint32_t)x & 127u) >> 3 << 8) | (((uint32_t)y & 127u) >> 3 << 4) | (((uint32_t)z & 127u) >> 3);
  if (!nv_bit(lo + 544, n)) return NvSpot{nullptr, 0u, nv_load<float>(lo + 1088 + 8 * (size_t)n)};
  const uint8_t *leaf = lo + nv_load<int64_t>(lo + 1088 + 8 * (size_t)n);
  return NvSpot{leaf, ((uint32_t)x & 7u) << 6 | ((uint32_t)y & 7u) << 3 | ((uint32_t)z & 7u), 0.f};
}

// nv_locate through the leaf table where it covers the block: one L2-held
// load instead of the root / upper / lower walk (the same leaf either way)
BN_HD NvSpot nv_find(const NvdbDD &g, int x, int y, int z) {
  if (g.leafTable) {
    const uint32_t bx = (uint32_t)((x >> 3) - g.tlo[0]), by = (uint32_t)((y >> 3) - g.tlo[1]),
                   bz = (uint32_t)((z >> 3) - g.tlo[2]);
    if (bx < (uint32_t)g.tdim[0] && by < (uint32_t)g.tdim[1] && bz < (uint32_t)g.tdim[2]) {
      const uint32_t e = g.leafTable[bx + (size_t)g.tdim[0] * (by + (size_t)g.tdim[1] * bz)];
      if (e) return NvSpot{g.base + 32 * (size_t)e, ((uint32_t)x & 7u) << 6 | ((uint32_t)y & 7u) << 3 | ((uint32_t)z & 7u),
                           0.f};
    }
  }
  return nv_locate(g, x, y, z);
}

BN_HD float nv_value(const NvdbDD &g, int x, int y, int z) {
  const NvSpot s = nv_find(g, x, y, z);
  return s.leaf ? nv_leaf_value(g.type, s.leaf, s.n) : s.value;
}

// NanoVDBDataSampler::DD::sample (NanoVDB.h:98-127): Map::applyInverseMapF
// of the position in double (the difference in double, the rows as fp32
// fmaf), minus half a voxel, clamped to the index box in double, then
// SampleFromVoxels<1>: floor, the 8-voxel stencil, and nested lerps
// a + float(w) * (b - a) over z, then y, then x. When the stencil lies in
// one 8^3 block (7 of 8 positions per axis) the block is located once.
BN_HD float nvdb_sample(const NvdbDD &g, vec3f P) {
#pragma clang fp contract(off)
  const float f0 = (float)((double)P.x - (double)g.vecF[0]);
  const float f1 = (float)((double)P.y - (double)g.vecF[1]);
  const float f2 = (float)((double)P.z - (double)g.vecF[2]);
  const float *m = g.invMatF;
  double ip[3] = {(double)fmaf(f0, m[0], fmaf(f1, m[1], f2 * m[2])),
                  (double)fmaf(f0, m[3], fmaf(f1, m[4], f2 * m[5])),
                  (double)fmaf(f0, m[6], fmaf(f1, m[7], f2 * m[8]))};
  int c[3];
  float w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double q = ip[k] - 0.5;
    q = q < (double)g.ilo[k] ? (double)g.ilo[k] : (q > (double)g.ihi[k] ? (double)g.ihi[k] : q);
    const double fl = floor(q);
    c[k] = (int)fl;
    w[k] = (float)(q - fl);
  }
  float v[8];  // v[dx * 4 + dy * 2 + dz]
  const NvSpot s = nv_find(g, c[0], c[1], c[2]);
  if ((c[0] & 7) != 7 && (c[1] & 7) != 7 && (c[2] & 7) != 7) {
    if (s.leaf) {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        v[i] = nv_leaf_value(g.type, s.leaf, s.n + ((i >> 2) << 6) + (((i >> 1) & 1) << 3) + (i & 1));
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = s.value;
    }
  } else {
    v[0] = s.leaf ? nv_leaf_value(g.type, s.leaf, s.n) : s.value;
#pragma unroll
    for (int i = 1; i < 8; ++i) v[i] = nv_value(g, c[0] + (i >> 2), c[1] + ((i >> 1) & 1), c[2] + (i & 1));
  }
  auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
  return lerp(lerp(lerp(v[0], v[1], w[2]), lerp(v[2], v[3], w[2]), w[1]),
              lerp(lerp(v[4], v[5], w[2]), lerp(v[6], v[7], w[2]), w[1]), w[0]);
}

BN_HD VolSample sample_and_map(const NvdbDD &g, vec3f P) {
  const float f = nvdb_sample(g, P);
  if (f != f) return VolSample{0.f, 0.f, 0.f, 0.f};
  return tf_map(g.tf, f);
}

}  // namespace bnamd
