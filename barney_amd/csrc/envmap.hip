// 2D textures and the HDRI environment light: creation, importance CDFs and
// probes (sampling math in envmap.h).
//
//   bnamd_texture2d_*   rtc::TextureData / rtc::Texture for 2D images
//                       (rtcore/cudaCommon/TextureData.h:15-31,
//                       Texture.h:12-24); texels widened to float4 as the
//                       embree backend's getTexel reads them
//                       (rtcore/embree/Texture.cpp:166-214)
//   bnamd_envmap_*      EnvMapLight::commit + computeCDFs
//                       (barney/light/EnvMap.cu:59-150, 178-230)
//
// The CDF sums run in the order EnvMap.cu adds them (one serial running sum
// per row, then one over the rows), so the tables equal the oracle's bit
// for bit; they are built once per map.
#include "envmap.h"
#include "runtime.h"
#include "../../include/bnamd.h"
#include <memory>

namespace bnamd {
namespace {

constexpr int kThreads = 256;

// widening of the texel formats (getTexel): float -> (v,v,v,v),
// uchar -> v / 255
__global__ void k_widen(const void *__restrict in, int format, size_t n, float4 *__restrict out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 t;
  if (format == 33) {
    t = ((const float4 *)in)[i];
  } else if (format == 30) {
    const float v = ((const float *)in)[i];
    t = make_float4(v, v, v, v);
  } else if (format == 1) {
    const uchar4 c = ((const uchar4 *)in)[i];
    t = make_float4((float)c.x * 1.f / 255.f, (float)c.y * 1.f / 255.f, (float)c.z * 1.f / 255.f,
                    (float)c.w * 1.f / 255.f);
  } else {
    const float v = (float)((const uint8_t *)in)[i] * 1.f / 255.f;
    t = make_float4(v, v, v, v);
  }
  out[i] = t;
}

__global__ void k_tex_sample(Tex2DDD t, const float2 *__restrict uv, float4 *__restrict out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 c = uv[i];
  out[i] = tex2d(t, c.x, c.y);
}

// ComputeWeights_xy (EnvMap.cu:88-116): the largest channel over a 3x3
// stencil of half-texel offsets
__global__ void k_env_weights(Tex2DDD t, float *__restrict w) {
#pragma clang fp contract(off)
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y;
  if (ix >= t.w) return;
  float weight = 0.f;
  for (int jy = 0; jy <= 2; ++jy)
    for (int jx = 0; jx <= 2; ++jx) {
      const float4 v = tex2d(t, (ix + jx * .5f) / t.w, (iy + jy * .5f) / t.h);
      const float m0 = v.x < v.y ? v.y : v.x;
      const float m = m0 < v.z ? v.z : m0;
      weight = weight < m ? m : weight;
    }
  w[ix + (size_t)t.w * iy] = weight;
}

// ComputeCDFs_doLine (EnvMap.cu:118-146): one lane per row, serial sums
__global__ void k_env_rows(float *__restrict cdfX, float *__restrict cdfY, int W, int H) {
#pragma clang fp contract(off)
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= H) return;
  float *line = cdfX + (size_t)y * W;
  float sum = 0.f;
  for (int x = 0; x < W; ++x) sum += line[x];
  const float rs = 1.f / sum;
  sum = 0.f;
  for (int x = 0; x < W; ++x) {
    sum += line[x];
    line[x] = sum * rs;
  }
  line[W - 1] = 1.f;
  cdfY[y] = sum * bnf::sin(BN_PI * ((y + .5f) / H));
}

// Normalize_cdf_y (EnvMap.cu:64-80): one lane
__global__ void k_env_cdfy(float *__restrict cdfY, int H) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float sum = 0.f;
  for (int i = 0; i < H; ++i) sum += cdfY[i];
  const float rs = 1.f / sum;
  sum = 0.f;
  for (int i = 0; i < H; ++i) {
    sum += cdfY[i];
    cdfY[i] = sum * rs;
  }
  cdfY[H - 1] = 1.f;
}

__global__ void k_env_eval(EnvDD e, const float *__restrict dirs, float *__restrict pdf, float *__restrict rad, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const vec3f d = load3(dirs + 3 * (size_t)i);
  pdf[i] = env_pdf(e, d);
  const vec3f r = env_radiance(e, d);
  rad[3 * (size_t)i] = r.x;
  rad[3 * (size_t)i + 1] = r.y;
  rad[3 * (size_t)i + 2] = r.z;
}

__global__ void k_env_sample(EnvDD e, uint32_t seed, float *__restrict out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Random2 r((uint32_t)i * 0x9E3779B9u, seed);
  vec3f d, rad;
  float pdf;
  env_sample(e, r, d, rad, pdf);
  float *o = out + 7 * (size_t)i;
  o[0] = d.x; o[1] = d.y; o[2] = d.z;
  o[3] = rad.x; o[4] = rad.y; o[5] = rad.z;
  o[6] = pdf;
}

size_t texel_bytes(int format) {
  switch (format) {
    case 33: return 16;
    case 30: return 4;
    case 1: return 4;
    case 0: return 1;
  }
  throw Error("bnamd_texture2d_create: texel format must be rtc::FLOAT4 (33), FLOAT (30), UCHAR4 (1) or UCHAR (0)");
}

}  // namespace
}  // namespace bnamd

struct bnamd_texture2d_s {
  bnamd::Tex2DDD dd;
  bnamd::DevBuf<float4> texels;
};

namespace {
struct EnvMapImpl : bnamd_envmap_s {
  bnamd::DevBuf<float> cdfY, cdfX;
};

void check_tex(const bnamd_texture2d_s *t, const char *who) {
  BN_REQUIRE(t && t->dd.texels, std::string(who) + ": NULL texture");
}
}  // namespace

using namespace bnamd;

extern "C" int bnamd_texture2d_create(const bnamd_texture2d_desc *desc, void *stream, bnamd_texture2d **out) {
  return guarded([&] {
    BN_REQUIRE(desc && out, "bnamd_texture2d_create: NULL argument");
    *out = nullptr;
    BN_REQUIRE(desc->width > 0 && desc->height > 0, "bnamd_texture2d_create: size must be positive");
    BN_REQUIRE((long long)desc->width * desc->height <= (1ll << 30), "bnamd_texture2d_create: over 2^30 texels");
    BN_REQUIRE(desc->texels, "bnamd_texture2d_create: NULL texels");
    BN_REQUIRE(desc->filterMode == TEX_POINT || desc->filterMode == TEX_LINEAR,
               "bnamd_texture2d_create: filterMode must be 0 (point) or 1 (linear)");
    for (int k = 0; k < 2; ++k)
      BN_REQUIRE(desc->addressMode[k] >= TEX_WRAP && desc->addressMode[k] <= TEX_MIRROR,
                 "bnamd_texture2d_create: addressMode must be 0..3 (wrap, clamp, border, mirror)");
    const size_t n = (size_t)desc->width * desc->height;
    const size_t bytes = texel_bytes(desc->format) * n;
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<bnamd_texture2d_s> t(new bnamd_texture2d_s);
    t->texels.alloc(n);
    DevBuf<uint8_t> raw(bytes);
    BN_HIP(hipMemcpyAsync(raw.ptr, desc->texels, bytes, hipMemcpyHostToDevice, s));
    k_widen<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(raw.ptr, desc->format, n, t->texels.ptr);
    BN_HIP(hipGetLastError());
    BN_HIP(hipStreamSynchronize(s));  // the staging buffer is freed on return
    t->dd.texels = t->texels.ptr;
    t->dd.w = desc->width;
    t->dd.h = desc->height;
    t->dd.filter = desc->filterMode;
    t->dd.addr[0] = desc->addressMode[0];
    t->dd.addr[1] = desc->addressMode[1];
    t->dd.pad = 0;
    *out = t.release();
  });
}

extern "C" int bnamd_texture2d_destroy(bnamd_texture2d *tex) {
  return guarded([&] { delete tex; });
}

extern "C" int bnamd_texture2d_device_view(const bnamd_texture2d *tex, const void **texels4, int *width,
                                           int *height) {
  return guarded([&] {
    BN_REQUIRE(tex && texels4 && width && height, "bnamd_texture2d_device_view: NULL argument");
    *texels4 = tex->dd.texels;
    *width = tex->dd.w;
    *height = tex->dd.h;
  });
}

extern "C" int bnamd_texture2d_sample(const bnamd_texture2d *tex, const float *uv, float *out, int n, void *stream) {
  return guarded([&] {
    check_tex(tex, "bnamd_texture2d_sample");
    BN_REQUIRE(n >= 0, "bnamd_texture2d_sample: negative count");
    if (n == 0) return;
    BN_REQUIRE(uv && out, "bnamd_texture2d_sample: NULL buffers");
    hipStream_t s = (hipStream_t)stream;
    k_tex_sample<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(tex->dd, (const float2 *)uv, (float4 *)out, n);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(nullptr));
  });
}

extern "C" int bnamd_envmap_create(const bnamd_texture2d *tex, const bnamd_envmap_desc *desc, void *stream,
                                   bnamd_envmap **out) {
  return guarded([&] {
#pragma clang fp contract(off)
    BN_REQUIRE(desc && out, "bnamd_envmap_create: NULL argument");
    *out = nullptr;
    check_tex(tex, "bnamd_envmap_create");
    const vec3f dir = load3(desc->direction), up = load3(desc->up);
    BN_REQUIRE(dot(up, up) > 0.f && dot(cross(up, dir), cross(up, dir)) > 0.f,
               "bnamd_envmap_create: direction and up must be non-zero and not parallel");
    std::unique_ptr<EnvMapImpl> e(new EnvMapImpl);
    EnvDD &d = e->dd;
    d.tex = tex->dd;
    d.scale = desc->scale;
    // EnvMapLight::commit (EnvMap.cu:157-171): vz = -up, vy = -(vz x dir),
    // vx = vy x vz, normalized; toLocal = its inverse as the adjugate
    // (cross products of the columns, transposed) times 1 / determinant
    const vec3f vz = -normalize(up);
    const vec3f vy = -normalize(cross(vz, dir));
    const vec3f vx = normalize(cross(vy, vz));
    d.toWorld[0] = vx;
    d.toWorld[1] = vy;
    d.toWorld[2] = vz;
    const vec3f a = cross(vy, vz), b = cross(vz, vx), c = cross(vx, vy);
    const float r = 1.f / dot(vx, a);
    d.toLocal[0] = r * vec3f(a.x, b.x, c.x);
    d.toLocal[1] = r * vec3f(a.y, b.y, c.y);
    d.toLocal[2] = r * vec3f(a.z, b.z, c.z);
    const int W = d.tex.w, H = d.tex.h;
    e->cdfX.alloc((size_t)W * H);
    e->cdfY.alloc(H);
    hipStream_t s = (hipStream_t)stream;
    k_env_weights<<<dim3((W + kThreads - 1) / kThreads, H), kThreads, 0, s>>>(d.tex, e->cdfX.ptr);
    k_env_rows<<<(H + 63) / 64, 64, 0, s>>>(e->cdfX.ptr, e->cdfY.ptr, W, H);
    k_env_cdfy<<<1, 64, 0, s>>>(e->cdfY.ptr, H);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(nullptr));
    d.cdfX = e->cdfX.ptr;
    d.cdfY = e->cdfY.ptr;
    *out = e.release();
  });
}

extern "C" int bnamd_envmap_destroy(bnamd_envmap *env) {
  return guarded([&] { delete static_cast<EnvMapImpl *>(env); });
}

extern "C" int bnamd_envmap_get_cdfs(const bnamd_envmap *env, float *cdfY, float *cdfX) {
  return guarded([&] {
    BN_REQUIRE(env && cdfY && cdfX, "bnamd_envmap_get_cdfs: NULL argument");
    const EnvDD &d = env->dd;
    BN_HIP(hipDeviceSynchronize());
    BN_HIP(hipMemcpy(cdfY, d.cdfY, sizeof(float) * d.tex.h, hipMemcpyDeviceToHost));
    BN_HIP(hipMemcpy(cdfX, d.cdfX, sizeof(float) * d.tex.h * (size_t)d.tex.w, hipMemcpyDeviceToHost));
  });
}

extern "C" int bnamd_envmap_get_frame(const bnamd_envmap *env, float *toWorld, float *toLocal) {
  return guarded([&] {
    BN_REQUIRE(env && toWorld && toLocal, "bnamd_envmap_get_frame: NULL argument");
    for (int k = 0; k < 3; ++k) {
      const vec3f w = env->dd.toWorld[k], l = env->dd.toLocal[k];
      toWorld[3 * k] = w.x; toWorld[3 * k + 1] = w.y; toWorld[3 * k + 2] = w.z;
      toLocal[3 * k] = l.x; toLocal[3 * k + 1] = l.y; toLocal[3 * k + 2] = l.z;
    }
  });
}

extern "C" int bnamd_envmap_eval(const bnamd_envmap *env, const float *dirs, float *pdf, float *radiance, int n,
                                 void *stream) {
  return guarded([&] {
    BN_REQUIRE(env, "bnamd_envmap_eval: NULL map");
    BN_REQUIRE(n >= 0, "bnamd_envmap_eval: negative count");
    if (n == 0) return;
    BN_REQUIRE(dirs && pdf && radiance, "bnamd_envmap_eval: NULL buffers");
    hipStream_t s = (hipStream_t)stream;
    k_env_eval<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(env->dd, dirs, pdf, radiance, n);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(nullptr));
  });
}

extern "C" int bnamd_envmap_sample(const bnamd_envmap *env, uint32_t seed, float *out, int n, void *stream) {
  return guarded([&] {
    BN_REQUIRE(env, "bnamd_envmap_sample: NULL map");
    BN_REQUIRE(n >= 0, "bnamd_envmap_sample: negative count");
    if (n == 0) return;
    BN_REQUIRE(out, "bnamd_envmap_sample: NULL output");
    hipStream_t s = (hipStream_t)stream;
    k_env_sample<<<(n + kThreads - 1) / kThreads, kThreads, 0, s>>>(env->dd, seed, out, n);
    BN_HIP(hipGetLastError());
    if (!stream) BN_HIP(hipStreamSynchronize(nullptr));
  });
}
