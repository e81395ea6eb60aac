// Block-structured AMR scalar fields (barney/amr): BlockStructuredField
// (BlockStructuredField.h, .cu) sampled as BlockStructuredCuBQLSampler does
// (BlockStructuredCuBQLSampler.h:66-80): every block whose filter domain
// (its cells grown by half a cell) contains P adds its tent-basis
// contributions (addBasisFunctions, BlockStructuredField.h:100-200); the
// value is the weighted mean, NaN where no block contributes. The device
// record also carries the macrocell grid and transfer function, under the
// names DeltaTrack reads, so the structured-volume path tracer renders it.
#pragma once
#include "volume.h"

namespace bnamd {

struct AmrDD {
  // macrocell grid + transfer function (as in VolumeDD)
  vec3f worldLo, worldHi;  // union of the block domains
  const float *majorants;
  int mx, my, mz;
  vec3f mcOrigin, mcSpacing;
  TFDD tf;
  // blocks: domain {lower.xyz, cellSize}, {upper.xyz, 0}; cell counts; the
  // first scalar of each block
  const float4 *dom;
  const int4 *dims;
  const unsigned long long *offsets;
  const float *scalars;
  // block lookup grid over the world box: the blocks whose domain overlaps
  // cell c are cellBlocks[cellStart[c] .. cellStart[c + 1]), ascending
  const uint32_t *cellStart;
  const uint32_t *cellBlocks;
  int gx, gy, gz;
  vec3f gOrigin, gRcp;
};

// addBasisFunctions of block b at P (BlockStructuredField.h:100-200): the
// eight cell centers around P, each with its trilinear tent weight, cells
// outside the block skipped.
BN_HD void amr_add_block(const AmrDD &a, uint32_t b, vec3f P, float &sumW, float &sumWV) {
#pragma clang fp contract(off)
  const float4 lo = a.dom[2 * b], hi = a.dom[2 * b + 1];
  if (!(lo.x <= P.x && lo.y <= P.y && lo.z <= P.z && P.x <= hi.x && P.y <= hi.y && P.z <= hi.z)) return;
  const float cs = lo.w;
  const int4 d = a.dims[b];
  const float *s = a.scalars + a.offsets[b];
  const vec3f c000 = {lo.x + cs, lo.y + cs, lo.z + cs};
  const vec3f lp = {(P.x - c000.x) / cs, (P.y - c000.y) / cs, (P.z - c000.z) / cs};
  const vec3f fl = {floorf(lp.x), floorf(lp.y), floorf(lp.z)};
  const int lx = max(-1, (int)fl.x), ly = max(-1, (int)fl.y), lz = max(-1, (int)fl.z);
  const int hx = lx + 1, hy = ly + 1, hz = lz + 1;
  const vec3f f = lp - fl;
  const vec3f nf = vec3f(1.f) - f;
  auto add = [&](int x, int y, int z, float w) {
    sumW += w;
    sumWV += w * s[x + y * d.x + z * d.x * d.y];
  };
  const bool inLx = lx >= 0 && lx < d.x, inHx = hx >= 0 && hx < d.x;
  const bool inLy = ly >= 0 && ly < d.y, inHy = hy >= 0 && hy < d.y;
  if (lz >= 0 && lz < d.z) {
    if (inLy) {
      if (inLx) add(lx, ly, lz, nf.z * nf.y * nf.x);
      if (inHx) add(hx, ly, lz, nf.z * nf.y * f.x);
    }
    if (inHy) {
      if (inLx) add(lx, hy, lz, nf.z * f.y * nf.x);
      if (inHx) add(hx, hy, lz, nf.z * f.y * f.x);
    }
  }
  if (hz >= 0 && hz < d.z) {
    if (inLy) {
      if (inLx) add(lx, ly, hz, f.z * nf.y * nf.x);
      if (inHx) add(hx, ly, hz, f.z * nf.y * f.x);
    }
    if (inHy) {
      if (inLx) add(lx, hy, hz, f.z * f.y * nf.x);
      if (inHx) add(hx, hy, hz, f.z * f.y * f.x);
    }
  }
}

// BlockStructuredCuBQLSampler::DD::sample: the blocks are visited in
// ascending order (the reference sums in its BVH's traversal order, so
// values where domains overlap may differ from it in the last bits).
BN_HD float amr_sample(const AmrDD &a, vec3f P) {
#pragma clang fp contract(off)
  if (!(a.worldLo.x <= P.x && a.worldLo.y <= P.y && a.worldLo.z <= P.z && P.x <= a.worldHi.x &&
        P.y <= a.worldHi.y && P.z <= a.worldHi.z))
    return __builtin_nanf("");
  const int cx = min(max((int)((P.x - a.gOrigin.x) * a.gRcp.x), 0), a.gx - 1);
  const int cy = min(max((int)((P.y - a.gOrigin.y) * a.gRcp.y), 0), a.gy - 1);
  const int cz = min(max((int)((P.z - a.gOrigin.z) * a.gRcp.z), 0), a.gz - 1);
  const uint32_t c = (uint32_t)cx + (uint32_t)a.gx * ((uint32_t)cy + (uint32_t)a.gy * (uint32_t)cz);
  float sumW = 0.f, sumWV = 0.f;
  for (uint32_t i = a.cellStart[c], e = a.cellStart[c + 1]; i < e; ++i) amr_add_block(a, a.cellBlocks[i], P, sumW, sumWV);
  return sumW == 0.f ? __builtin_nanf("") : sumWV / sumW;
}

BN_HD VolSample sample_and_map(const AmrDD &a, vec3f P) {
  const float f = amr_sample(a, P);
  if (f != f) return VolSample{0.f, 0.f, 0.f, 0.f};
  return tf_map(a.tf, f);
}

}  // namespace bnamd
