// Golden-vector generator for NanoVDB fields (test infrastructure only).
//
// Built from this file and the NanoVDB 32.7.0 headers that the reference
// vendors (3rdParty/nanovdb/include), by oracle/Makefile into oracle/_ref/.
// It uses the NanoVDB library itself, not a restatement:
//   * grids are made by nanovdb::tools::createNanoGrid (float, Fp4, Fp8,
//     Fp16, FpN) from a seeded sparse field built with tools::build::Grid,
//     with pruned leaf tiles, an upper-node tile and a non-unit map;
//   * samples follow barney's NanoVDBDataSampler::DD::sample
//     (barney/volume/NanoVDB.h:98-127): worldToIndexF, minus half a voxel,
//     clamp to the index box, nanovdb::math::createSampler<1>;
//   * macrocell ranges follow NanoVDBData_computeMCs and buildMCs
//     (barney/volume/NanoVDB.cu:28-102) through the library's ReadAccessor;
//   * the field's boxes follow NanoVDBData::commit (NanoVDB.cu:204-235).
//
// usage: nvdb_ref <outdir> <float|fp4|fp8|fp16|fpn> <seed> <numPoints> [tiles]
//          writes <outdir>/{grid.nvdb, points.f32, samples.f32, ranges.f32, meta.txt}
//        nvdb_ref --sample <grid.nvdb> <points.f32> <out.f32>
//          barney's samples of any grid buffer (float32 xyz points)
#include <nanovdb/NanoVDB.h>
#include <nanovdb/tools/CreateNanoGrid.h>
#include <nanovdb/math/SampleFromVoxels.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

uint64_t g_state;
uint32_t rnd_u32() {  // PCG32 (XSH-RR)
  const uint64_t old = g_state;
  g_state = old * 6364136223846793005ull + 1442695040888963407ull;
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31));
}
float rnd01() { return (float)(rnd_u32() >> 8) * (1.f / 16777216.f); }

bool write_file(const std::string &path, const void *p, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(p, 1, n, f) == n;
  std::fclose(f);
  return ok;
}

// barney's sampler (NanoVDB.h:98-127) through the library
template <class BuildT>
float barney_sample(const nanovdb::NanoGrid<BuildT> *grid, const nanovdb::CoordBBox &ib, const float P[3]) {
  const auto loc = nanovdb::Vec3d(P[0], P[1], P[2]);
  auto idx = grid->worldToIndexF(loc);
  idx[0] -= 0.5;
  idx[1] -= 0.5;
  idx[2] -= 0.5;
  idx[0] = nanovdb::math::Clamp(idx[0], (double)ib.min()[0], (double)ib.max()[0]);
  idx[1] = nanovdb::math::Clamp(idx[1], (double)ib.min()[1], (double)ib.max()[1]);
  idx[2] = nanovdb::math::Clamp(idx[2], (double)ib.min()[2], (double)ib.max()[2]);
  auto acc = grid->getAccessor();
  auto smp = nanovdb::math::createSampler<1>(acc);
  return smp(nanovdb::math::Vec3d(idx));
}

template <class BuildT>
int emit(const std::string &dir, nanovdb::GridHandle<nanovdb::HostBuffer> &h, int numPoints) {
  const auto *grid = h.template grid<BuildT>();
  if (!grid) return 2;
  const nanovdb::GridMetaData *md = h.gridMetaData();
  if (!write_file(dir + "/grid.nvdb", h.data(), h.size())) return 3;
  // NanoVDBData::commit
  const auto wmin = md->worldBBox().min(), wmax = md->worldBBox().max();
  const float wlo[3] = {(float)wmin[0], (float)wmin[1], (float)wmin[2]};
  const float whi[3] = {(float)wmax[0], (float)wmax[1], (float)wmax[2]};
  const nanovdb::CoordBBox ib = md->indexBBox();
  const auto dim = ib.dim();
  const int gs[3] = {dim[0], dim[1], dim[2]};
  // buildMCs: divRoundUp(gridSize, 16) cells over the world box
  const int mc[3] = {(gs[0] + 15) / 16, (gs[1] + 15) / 16, (gs[2] + 15) / 16};
  float spacing[3];
  for (int k = 0; k < 3; ++k) spacing[k] = (whi[k] - wlo[k]) * (1.f / (float)mc[k]);
  // NanoVDBData_computeMCs
  const size_t nmc = (size_t)mc[0] * mc[1] * mc[2];
  std::vector<float> ranges(2 * nmc);
  auto acc = grid->getAccessor();
  for (size_t tid = 0; tid < nmc; ++tid) {
    const int id[3] = {(int)(tid % mc[0]), (int)((tid / mc[0]) % mc[1]), (int)(tid / ((size_t)mc[0] * mc[1]))};
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = ib.min()[k] + (id[k] * gs[k]) / mc[k];
      hi[k] = ib.min()[k] + ((id[k] + 1) * gs[k]) / mc[k];
    }
    const nanovdb::CoordBBox box({lo[0], lo[1], lo[2]}, {hi[0] + 1, hi[1] + 1, hi[2] + 1});
    float rlo = INFINITY, rhi = -INFINITY;
    for (auto it = box.begin(); it; ++it) {
      const float v = acc.getValue(*it);
      rlo = std::fmin(rlo, v);
      rhi = std::fmax(rhi, v);
    }
    ranges[2 * tid] = rlo;
    ranges[2 * tid + 1] = rhi;
  }
  if (!write_file(dir + "/ranges.f32", ranges.data(), ranges.size() * 4)) return 3;
  // points: uniform over the world box grown by 10 % on each side, then
  // voxel centers, voxel corners and the box's own corners
  std::vector<float> pts;
  for (int i = 0; i < numPoints; ++i)
    for (int k = 0; k < 3; ++k) {
      const float ext = whi[k] - wlo[k];
      pts.push_back(wlo[k] - 0.1f * ext + 1.2f * ext * rnd01());
    }
  const double vs = md->voxelSize()[0];
  for (int i = 0; i < numPoints / 4; ++i) {
    const int c[3] = {ib.min()[0] + (int)(rnd_u32() % (uint32_t)gs[0]), ib.min()[1] + (int)(rnd_u32() % (uint32_t)gs[1]),
                      ib.min()[2] + (int)(rnd_u32() % (uint32_t)gs[2])};
    const double off = (i & 1) ? 0.5 : 0.0;
    const auto w = grid->indexToWorld(nanovdb::Vec3d(c[0] + off, c[1] + off, c[2] + off));
    for (int k = 0; k < 3; ++k) pts.push_back((float)w[k]);
  }
  for (int c = 0; c < 8; ++c)
    for (int k = 0; k < 3; ++k) pts.push_back((c >> k) & 1 ? whi[k] : wlo[k]);
  const size_t np = pts.size() / 3;
  std::vector<float> out(np);
  for (size_t i = 0; i < np; ++i) out[i] = barney_sample(grid, ib, &pts[3 * i]);
  if (!write_file(dir + "/points.f32", pts.data(), pts.size() * 4)) return 3;
  if (!write_file(dir + "/samples.f32", out.data(), out.size() * 4)) return 3;
  FILE *f = std::fopen((dir + "/meta.txt").c_str(), "w");
  if (!f) return 3;
  std::fprintf(f, "gridType %d\nvoxelSize %.17g\n", (int)md->gridType(), vs);
  std::fprintf(f, "indexLo %d %d %d\nindexHi %d %d %d\n", ib.min()[0], ib.min()[1], ib.min()[2], ib.max()[0],
               ib.max()[1], ib.max()[2]);
  std::fprintf(f, "worldLo %a %a %a\nworldHi %a %a %a\n", wlo[0], wlo[1], wlo[2], whi[0], whi[1], whi[2]);
  std::fprintf(f, "mcDims %d %d %d\nmcSpacing %a %a %a\n", mc[0], mc[1], mc[2], spacing[0], spacing[1], spacing[2]);
  std::fprintf(f, "nodes %u %u %u\nrootTiles %u\n", md->nodeCount(0), md->nodeCount(1), md->nodeCount(2),
               md->rootTableSize());
  std::fclose(f);
  return 0;
}

// samples of a grid buffer read from a file (any of the five value types)
int sample_file(const char *gridPath, const char *pointsPath, const char *outPath) {
  auto read = [](const char *path, std::vector<char> &v) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    v.resize((size_t)std::ftell(f));
    std::fseek(f, 0, SEEK_SET);
    const bool ok = std::fread(v.data(), 1, v.size(), f) == v.size();
    std::fclose(f);
    return ok;
  };
  std::vector<char> g, p;
  if (!read(gridPath, g) || !read(pointsPath, p)) return 3;
  auto buffer = nanovdb::HostBuffer::create(g.size());
  std::memcpy(buffer.data(), g.data(), g.size());
  nanovdb::GridHandle<nanovdb::HostBuffer> h(std::move(buffer));
  const nanovdb::GridMetaData *md = h.gridMetaData();
  if (!md) return 2;
  const nanovdb::CoordBBox ib = md->indexBBox();
  const size_t n = p.size() / 12;
  std::vector<float> out(n);
  const float *pts = reinterpret_cast<const float *>(p.data());
  for (size_t i = 0; i < n; ++i) {
    switch (md->gridType()) {
      case nanovdb::GridType::Float: out[i] = barney_sample(h.grid<float>(), ib, pts + 3 * i); break;
      case nanovdb::GridType::Fp4: out[i] = barney_sample(h.grid<nanovdb::Fp4>(), ib, pts + 3 * i); break;
      case nanovdb::GridType::Fp8: out[i] = barney_sample(h.grid<nanovdb::Fp8>(), ib, pts + 3 * i); break;
      case nanovdb::GridType::Fp16: out[i] = barney_sample(h.grid<nanovdb::Fp16>(), ib, pts + 3 * i); break;
      case nanovdb::GridType::FpN: out[i] = barney_sample(h.grid<nanovdb::FpN>(), ib, pts + 3 * i); break;
      default: return 2;
    }
  }
  return write_file(outPath, out.data(), out.size() * 4) ? 0 : 3;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 5 && std::string(argv[1]) == "--sample") return sample_file(argv[2], argv[3], argv[4]);
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s <outdir> <float|fp4|fp8|fp16|fpn> <seed> <numPoints> [tiles]\n", argv[0]);
    return 1;
  }
  const std::string dir = argv[1], type = argv[2];
  g_state = std::strtoull(argv[3], nullptr, 0) * 2654435761ull + 0x9E3779B97F4A7C15ull;
  const int numPoints = std::atoi(argv[4]);
  const bool tiles = argc > 5 && std::string(argv[5]) == "tiles";

  // a sparse field: three seeded blobs over a box that straddles x = 0 (two
  // root tiles, several lower nodes), background 0 where it is small
  nanovdb::tools::build::Grid<float> src(0.f, "density", nanovdb::GridClass::FogVolume);
  src.setTransform(0.75, nanovdb::Vec3d(1.5, -2.25, 0.5));
  float bc[3][3], br[3];
  for (int b = 0; b < 3; ++b) {
    for (int k = 0; k < 3; ++k) bc[b][k] = (k ? 6.f : -8.f) + 20.f * rnd01();
    br[b] = 4.f + 6.f * rnd01();
  }
  auto field = [&](const nanovdb::Coord &c) -> float {
    // an 8^3-aligned cube of one value: createNanoGrid keeps it as a tile
    if (c[0] >= 8 && c[0] < 16 && c[1] >= 8 && c[1] < 16 && c[2] >= 8 && c[2] < 16) return 0.625f;
    float v = 0.f;
    for (int b = 0; b < 3; ++b) {
      const float dx = c[0] - bc[b][0], dy = c[1] - bc[b][1], dz = c[2] - bc[b][2];
      v += std::exp(-(dx * dx + dy * dy + dz * dz) / (2.f * br[b] * br[b]));
    }
    v += 0.05f * std::sin(0.7f * c[0] + 0.3f * c[1]) * std::cos(0.5f * c[2]);
    return v < 0.08f ? 0.f : v;
  };
  src(field, nanovdb::CoordBBox(nanovdb::Coord(-13, 3, 2), nanovdb::Coord(18, 34, 25)));
  if (tiles) {
    // an active upper-node tile (one 128^3 block of 0.25) away from the voxels
    src.tree().root().addTile<2>(nanovdb::Coord(-256, 0, 0), 0.25f, true);
  }
  using SrcT = nanovdb::tools::build::Grid<float>;
  if (type == "float") {
    auto h = nanovdb::tools::createNanoGrid<SrcT, float>(src);
    return emit<float>(dir, h, numPoints);
  } else if (type == "fp4") {
    auto h = nanovdb::tools::createNanoGrid<SrcT, nanovdb::Fp4>(src);
    return emit<nanovdb::Fp4>(dir, h, numPoints);
  } else if (type == "fp8") {
    auto h = nanovdb::tools::createNanoGrid<SrcT, nanovdb::Fp8>(src);
    return emit<nanovdb::Fp8>(dir, h, numPoints);
  } else if (type == "fp16") {
    auto h = nanovdb::tools::createNanoGrid<SrcT, nanovdb::Fp16>(src);
    return emit<nanovdb::Fp16>(dir, h, numPoints);
  } else if (type == "fpn") {
    // AbsDiff tolerance 0.002: leaves get different bit widths
    auto h = nanovdb::tools::createNanoGrid<SrcT, nanovdb::FpN>(src, nanovdb::tools::StatsMode::Default,
                                                                 nanovdb::CheckMode::Default, false, 0,
                                                                 nanovdb::tools::AbsDiff(0.002f));
    return emit<nanovdb::FpN>(dir, h, numPoints);
  }
  std::fprintf(stderr, "unknown type %s\n", type.c_str());
  return 1;
}
