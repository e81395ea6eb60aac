// Frame-buffer tiles and readback for MI355X (gfx950).
//
// Row-major tile ownership and the tiled -> linear readback of barney's
// TiledFB (barney/fb/TiledFB.cu:13-130, 226-264), the resolve of the linear
// accumulation buffer that this library's fused renderers write, and the
// clamp-and-quantize pass of barney's denoised read path
// (barney/fb/FrameBuffer.cu:105-128).
//
// All of them are streaming kernels bound by HBM bandwidth: per pixel one
// float4 read (plus 12 B of normal when asked for) and one float4 or uint32
// write. A tile is one 256-thread workgroup that walks its 1024 pixels in
// four passes, so a 64-wide wave reads 1 KiB of accumulation contiguously and
// writes two 32-pixel row segments.
#include "../../include/bnamd.h"
#include "common.h"
#include "runtime.h"

namespace bnamd {
namespace {

constexpr int kTile = BNAMD_TILE_SIZE;
constexpr int kTilePixels = BNAMD_PIXELS_PER_TILE;
constexpr int kThreads = 256;
static_assert(kTilePixels % kThreads == 0, "tile passes");
static_assert(sizeof(bnamd_accum_tile) == kTilePixels * 28, "AccumTile layout: float4 accum + float3 normal");

// make_8bit (barney/common/math.h:94-97): min(255, max(0, int(f * 256))).
// The GPU's float -> int conversion saturates and maps NaN to 0; it is
// spelled out here so no out-of-range conversion is left to the compiler.
BN_D uint32_t to_8bit(float f) {
  const float s = f * 256.f;
  if (s >= 256.f) return 255u;
  if (s > 0.f) return (uint32_t)(int)s;
  return 0u;  // negative, zero or NaN
}

// linear_to_srgb (barney/common/math.h:52-58). The final multiply-add is
// fused, as nvcc's default contraction (--fmad=true) compiles it.
BN_D float linear_to_srgb(float x) {
  if (x <= 0.0031308f) return 12.92f * x;
  return fmaf(1.055f, powf(x, 1.f / 2.4f), -0.055f);
}

// saturate (barney/common/math.h:39-46) with fminf/fmaxf semantics: NaN -> 1.
BN_D float saturate(float f) { return fmaxf(0.f, fminf(f, 1.f)); }

// make_rgba(vec4f) (barney/common/math.h:109-116), after the sRGB transfer
// on R, G, B for BNAMD_FB_UFIXED8_RGBA_SRGB (math.h:63-66: alpha unchanged).
template <int F>
BN_D void store_color(void *__restrict out, size_t idx, float4 c) {
  if constexpr (F == BNAMD_FB_FLOAT4) {
    ((float4 *)out)[idx] = c;
  } else {
    if constexpr (F == BNAMD_FB_UFIXED8_RGBA_SRGB) {
      c.x = linear_to_srgb(c.x);
      c.y = linear_to_srgb(c.y);
      c.z = linear_to_srgb(c.z);
    }
    ((uint32_t *)out)[idx] = to_8bit(c.x) | (to_8bit(c.y) << 8) | (to_8bit(c.z) << 16) | (to_8bit(c.w) << 24);
  }
}

// setTileCoordsKernel (barney/fb/TiledFB.cu:226-238): the tid-th tile this
// device owns is global tile tid * globalSize + globalRank.
__global__ void k_tile_descs(bnamd_tile_desc *__restrict descs, int numActive, int tilesX, int rank, int size) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= numActive) return;
  const long long id = (long long)tid * size + rank;
  descs[tid].lower[0] = (int)(id % tilesX) * kTile;
  descs[tid].lower[1] = (int)(id / tilesX) * kTile;
}

// linearizeColorAndNormalKernel (barney/fb/TiledFB.cu:13-48).
template <int F>
__global__ void __launch_bounds__(kThreads)
    k_linearize(const bnamd_accum_tile *__restrict tiles, const bnamd_tile_desc *__restrict descs, int W, int H,
                float scale, void *__restrict outColor, float *__restrict outNormal) {
  const bnamd_accum_tile *t = tiles + blockIdx.x;
  const int lx = descs[blockIdx.x].lower[0], ly = descs[blockIdx.x].lower[1];
  const float4 *acc = (const float4 *)t->accum;
#pragma unroll
  for (int k = 0; k < kTilePixels / kThreads; ++k) {
    const int sub = threadIdx.x + k * kThreads;
    const int ix = lx + sub % kTile, iy = ly + sub / kTile;
    if (ix >= W || iy >= H) continue;
    const size_t idx = (size_t)ix + (size_t)W * iy;
    float4 c = acc[sub];
    c.x *= scale;
    c.y *= scale;
    c.z *= scale;
    c.w *= scale;
    store_color<F>(outColor, idx, c);
    if (outNormal) {
      outNormal[3 * idx + 0] = t->normal[sub][0];
      outNormal[3 * idx + 1] = t->normal[sub][1];
      outNormal[3 * idx + 2] = t->normal[sub][2];
    }
  }
}

// linearizeAuxTilesKernel (barney/fb/TiledFB.cu:88-105).
__global__ void __launch_bounds__(kThreads)
    k_linearize_aux(const uint32_t *__restrict aux, const bnamd_tile_desc *__restrict descs, int W, int H,
                    uint32_t *__restrict out) {
  const uint32_t *t = aux + (size_t)blockIdx.x * kTilePixels;
  const int lx = descs[blockIdx.x].lower[0], ly = descs[blockIdx.x].lower[1];
#pragma unroll
  for (int k = 0; k < kTilePixels / kThreads; ++k) {
    const int sub = threadIdx.x + k * kThreads;
    const int ix = lx + sub % kTile, iy = ly + sub / kTile;
    if (ix >= W || iy >= H) continue;
    out[(size_t)ix + (size_t)W * iy] = t[sub];
  }
}

// Linear float4 frame -> output format: scaled by `scale` (resolve) or
// clamped to [0,1] (LinearToFixed8, barney/fb/FrameBuffer.cu:116-128).
template <int F, bool kSaturate>
__global__ void __launch_bounds__(kThreads)
    k_resolve(const float4 *__restrict in, size_t n, float scale, void *__restrict out) {
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
    float4 c = in[i];
    if constexpr (kSaturate) {
      c = make_float4(saturate(c.x), saturate(c.y), saturate(c.z), saturate(c.w));
    } else {
      c.x *= scale;
      c.y *= scale;
      c.z *= scale;
      c.w *= scale;
    }
    store_color<F>(out, i, c);
  }
}

void check_format(int format, const char *who) {
  BN_REQUIRE(format == BNAMD_FB_FLOAT4 || format == BNAMD_FB_UFIXED8_RGBA || format == BNAMD_FB_UFIXED8_RGBA_SRGB,
             std::string(who) + ": unsupported color format " + std::to_string(format));
}

void check_frame(int W, int H, const char *who) {
  BN_REQUIRE(W >= 0 && H >= 0, std::string(who) + ": negative frame size");
}

int resolve_blocks(size_t n) {
  // a few waves per SIMD on 256 CUs; the loop strides over the rest
  const size_t want = (n + kThreads - 1) / kThreads;
  return (int)(want < 8192 ? want : 8192);
}

void finish(void *stream) {
  BN_HIP(hipGetLastError());
  if (!stream) BN_HIP(hipStreamSynchronize(nullptr));
}

}  // namespace
}  // namespace bnamd

using namespace bnamd;

extern "C" int bnamd_fb_tile_descs(int width, int height, int globalRank, int globalSize, bnamd_tile_desc *descs,
                                   int *numActiveTiles, void *stream) {
  return guarded([&] {
    check_frame(width, height, "bnamd_fb_tile_descs");
    BN_REQUIRE(numActiveTiles, "bnamd_fb_tile_descs: NULL numActiveTiles");
    BN_REQUIRE(globalSize >= 1 && globalRank >= 0 && globalRank < globalSize,
               "bnamd_fb_tile_descs: need 0 <= globalRank < globalSize");
    // TiledFB::resize (barney/fb/TiledFB.cu:243-252)
    const long long tilesX = (width + kTile - 1) / kTile, tilesY = (height + kTile - 1) / kTile;
    const long long total = tilesX * tilesY;
    BN_REQUIRE(total <= (1ll << 30), "bnamd_fb_tile_descs: frame too large");
    const long long left = total - globalRank > 0 ? total - globalRank : 0;
    const int numActive = (int)((left + globalSize - 1) / globalSize);
    *numActiveTiles = numActive;
    if (!descs || numActive == 0) return;
    k_tile_descs<<<(numActive + 127) / 128, 128, 0, (hipStream_t)stream>>>(descs, numActive, (int)tilesX,
                                                                          globalRank, globalSize);
    finish(stream);
  });
}

extern "C" int bnamd_fb_linearize(const bnamd_accum_tile *tiles, const bnamd_tile_desc *descs, int numTiles, int width,
                                  int height, float accumScale, int format, void *outColor, float *outNormal,
                                  void *stream) {
  return guarded([&] {
    check_frame(width, height, "bnamd_fb_linearize");
    check_format(format, "bnamd_fb_linearize");
    BN_REQUIRE(numTiles >= 0, "bnamd_fb_linearize: negative tile count");
    if (numTiles == 0 || width == 0 || height == 0) return;
    BN_REQUIRE(tiles && descs && outColor, "bnamd_fb_linearize: NULL tiles, descs or output");
    hipStream_t s = (hipStream_t)stream;
    switch (format) {
      case BNAMD_FB_FLOAT4:
        k_linearize<BNAMD_FB_FLOAT4><<<numTiles, kThreads, 0, s>>>(tiles, descs, width, height, accumScale, outColor,
                                                                  outNormal);
        break;
      case BNAMD_FB_UFIXED8_RGBA:
        k_linearize<BNAMD_FB_UFIXED8_RGBA><<<numTiles, kThreads, 0, s>>>(tiles, descs, width, height, accumScale,
                                                                        outColor, outNormal);
        break;
      default:
        k_linearize<BNAMD_FB_UFIXED8_RGBA_SRGB><<<numTiles, kThreads, 0, s>>>(tiles, descs, width, height,
                                                                             accumScale, outColor, outNormal);
    }
    finish(stream);
  });
}

extern "C" int bnamd_fb_linearize_aux(const uint32_t *auxTiles, const bnamd_tile_desc *descs, int numTiles, int width,
                                      int height, uint32_t *out, void *stream) {
  return guarded([&] {
    check_frame(width, height, "bnamd_fb_linearize_aux");
    BN_REQUIRE(numTiles >= 0, "bnamd_fb_linearize_aux: negative tile count");
    if (numTiles == 0 || width == 0 || height == 0) return;
    BN_REQUIRE(auxTiles && descs && out, "bnamd_fb_linearize_aux: NULL tiles, descs or output");
    k_linearize_aux<<<numTiles, kThreads, 0, (hipStream_t)stream>>>(auxTiles, descs, width, height, out);
    finish(stream);
  });
}

extern "C" int bnamd_fb_resolve(const float *accum, int width, int height, float accumScale, int format, void *out,
                                void *stream) {
  return guarded([&] {
    check_frame(width, height, "bnamd_fb_resolve");
    check_format(format, "bnamd_fb_resolve");
    const size_t n = (size_t)width * height;
    if (n == 0) return;
    BN_REQUIRE(accum && out, "bnamd_fb_resolve: NULL accum or output");
    BN_REQUIRE(((uintptr_t)accum & 15) == 0, "bnamd_fb_resolve: accum must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    const float4 *in = (const float4 *)accum;
    const int blocks = resolve_blocks(n);
    switch (format) {
      case BNAMD_FB_FLOAT4:
        k_resolve<BNAMD_FB_FLOAT4, false><<<blocks, kThreads, 0, s>>>(in, n, accumScale, out);
        break;
      case BNAMD_FB_UFIXED8_RGBA:
        k_resolve<BNAMD_FB_UFIXED8_RGBA, false><<<blocks, kThreads, 0, s>>>(in, n, accumScale, out);
        break;
      default:
        k_resolve<BNAMD_FB_UFIXED8_RGBA_SRGB, false><<<blocks, kThreads, 0, s>>>(in, n, accumScale, out);
    }
    finish(stream);
  });
}

extern "C" int bnamd_fb_to_fixed8(const float *in, int width, int height, int srgb, uint32_t *out, void *stream) {
  return guarded([&] {
    check_frame(width, height, "bnamd_fb_to_fixed8");
    const size_t n = (size_t)width * height;
    if (n == 0) return;
    BN_REQUIRE(in && out, "bnamd_fb_to_fixed8: NULL input or output");
    BN_REQUIRE(((uintptr_t)in & 15) == 0, "bnamd_fb_to_fixed8: input must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    const int blocks = resolve_blocks(n);
    if (srgb)
      k_resolve<BNAMD_FB_UFIXED8_RGBA_SRGB, true><<<blocks, kThreads, 0, s>>>((const float4 *)in, n, 1.f, out);
    else
      k_resolve<BNAMD_FB_UFIXED8_RGBA, true><<<blocks, kThreads, 0, s>>>((const float4 *)in, n, 1.f, out);
    finish(stream);
  });
}
