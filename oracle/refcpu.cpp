// refcpu.cpp -- CPU oracle (test infrastructure only; see refcpu.h).
//
// A multithreaded C++ restatement of barney's embree-backend path: camera
// rays, closest-hit tracing with barney's geometry programs, the env-light
// shade/bounce step, and structured-volume delta tracking. Compiled with
// -ffp-contract=off so every fp32 expression rounds exactly as written.
#include "refcpu.h"

// the sine/cosine/x^5 sequences the kernels use (plain C++, no HIP)
#include "../barney_amd/csrc/fmath.h"

#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

const float kInf = std::numeric_limits<float>::infinity();
const float kPi = (float)M_PI;

struct V3 {
  float x, y, z;
};
inline V3 mk(float a, float b, float c) { return V3{a, b, c}; }
inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 scl(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline V3 nrm(V3 a) { return scl(1.f / std::sqrt(dot(a, a)), a); }
inline float len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 vmin(V3 a, V3 b) { return mk(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return mk(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
inline V3 vmins(V3 a, float b) { return mk(std::fmin(a.x, b), std::fmin(a.y, b), std::fmin(a.z, b)); }
inline float maxc(V3 a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }
inline float minc(V3 a) { return std::fmin(a.x, std::fmin(a.y, a.z)); }
inline V3 inv(V3 a) { return mk(1.f / a.x, 1.f / a.y, 1.f / a.z); }
inline V3 vabs(V3 a) { return mk(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
inline bool eq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline V3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
inline uint32_t fbits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// ---- RNG: barney-common.h:65-111 -------------------------------------------
struct Rng {
  uint64_t s = 0xcafef00dd15ea5e5ull;
  uint32_t next() {
    uint64_t x = s;
    unsigned c = (unsigned)(x >> 61);
    s = x * 6364136223846793005ull;
    x ^= x >> 22;
    return (uint32_t)(x >> (22 + c));
  }
  uint32_t mix(uint64_t h) {
    s ^= (h << 1);
    return next();
  }
  void seed(uint32_t a, uint32_t b) {
    s ^= 2ull * a + 1ull;
    next();
    s ^= 2ull * b;
    next();
  }
  float uni() { return (float)(next() & 0xFFFFFFu) * (1.f / 16777216.f); }
};
// Random2(RNGSeed &parent, b)
inline Rng child(Rng &parent, uint64_t b) {
  Rng r;
  r.s = parent.s;
  parent.next();
  r.mix(b ? b : 290374ull);
  return r;
}
inline uint64_t fnv3(uint32_t a, uint32_t b, uint32_t c) {
  uint64_t h = 0xcbf29ce484222325ull ^ (0x100000001b3ull * (uint64_t)a);
  h = h * 0x100000001b3ull ^ b;
  h = h * 0x100000001b3ull ^ c;
  return h;
}

template <typename F>
void parallel_for(int n, int threads, F f) {
  if (threads <= 1 || n < 64) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      while (true) {
        const int b = next.fetch_add(64);
        if (b >= n) break;
        const int e = std::min(n, b + 64);
        for (int i = b; i < e; ++i) f(i);
      }
    });
  for (auto &th : pool) th.join();
}
int default_threads(int t) { return t > 0 ? t : (int)std::max(1u, std::thread::hardware_concurrency()); }

// ---- primitives --------------------------------------------------------------
struct Prim {
  int geom, prim;
  uint32_t key;  // tie-break: position in (geom, prim) order
  float lo[3], hi[3];
};

struct Best {
  float t;
  float u, v;       // tri barycentrics; sphere: shifted-frame t, t_move
  uint32_t key;
  int geom, prim;
};
inline bool better(float t, uint32_t key, const Best &b) { return t < b.t || (t == b.t && key < b.key); }

}  // namespace

// ---- 2D textures and the HDRI environment light ----------------------------
// A 2D texture as the embree backend samples it (rtcore/embree/Texture.cpp):
// texels widened to float4 at creation (getTexel, :166-214: float -> (v,v,v,v),
// uchar -> v / 255), nearest (:232-240) or bilinear (:266-289) filtering with
// the per-axis addressing of computeAddress (:26-84).
struct ref_envmap_s {
  std::vector<float> tex;  // float4 texels, row-major
  int w = 0, h = 0, filter = 1, addr[2] = {1, 1};
  V3 toWorld[3], toLocal[3];  // columns
  float scale = 1.f;
  std::vector<float> cdfY, cdfX;
};

namespace {

struct V4 {
  float x, y, z, w;
};

struct LerpAddr {
  int i0, i1;
  float f;
};

// rtcore/embree/Texture.cpp:26-84 (modes: 0 wrap, 1 clamp, 2 border, 3 mirror)
LerpAddr tex_address(int mode, float tc, int N) {
  LerpAddr o{0, 0, 0.f};
  if (mode == 0) {
    tc = tc * (float)N - .5f;
    tc = std::fmod(tc, (float)N);
    if (tc < 0.f) tc += (float)N;
    const float fc = std::floor(tc);
    const int ic = (int)fc;
    o.f = tc - ic;
    // tc + N can round up to N itself; the texel index then wraps to 0
    // (the embree sampler would read one texel past the row)
    o.i0 = ic >= N ? ic - N : ic;
    o.i1 = (int)std::fmod(tc + 1.f, (float)N);
  } else if (mode == 3) {
    tc = tc * N + .5f;
    const float fc = std::floor(tc);
    const int ic = (int)fc;
    o.f = tc - ic;
    int i0 = ic - 1, i1 = ic;
    if (i0 < 0) i0 = -i0;
    if (i1 < 0) i1 = -i1;
    i0 %= 2 * N;
    i1 %= 2 * N;
    if (i0 >= N) i0 = 2 * N - 1 - i0;
    if (i1 >= N) i1 = 2 * N - 1 - i1;
    o.i0 = i0;
    o.i1 = i1;
  } else if (mode == 1) {
    tc = tc * N;
    if (tc - .5f <= 0.f) {
      o.i0 = o.i1 = 0;
      o.f = 0.f;
    } else if (tc - .5f >= N - 1) {
      o.i0 = o.i1 = N - 1;
      o.f = 0.f;
    } else {
      o.f = (tc - .5f) - (int)(tc - .5f);
      o.i0 = (int)(tc - .5f);
      o.i1 = o.i0 + 1;
    }
  }
  // border: the embree sampler leaves the record zeroed (texel 0, weight 0)
  return o;
}

V4 texel(const ref_envmap_s *e, int i) {
  const float *t = e->tex.data() + 4 * (size_t)i;
  return V4{t[0], t[1], t[2], t[3]};
}

V4 lerp4(float f, V4 a, V4 b) {
  return V4{(1.f - f) * a.x + f * b.x, (1.f - f) * a.y + f * b.y, (1.f - f) * a.z + f * b.z,
            (1.f - f) * a.w + f * b.w};
}

V4 tex2d(const ref_envmap_s *e, float u, float v) {
  if (e->filter == 0) {  // Texture.cpp:232-240
    const int ix = (int)((uint32_t)(std::fabs(u) * e->w) % (uint32_t)e->w);
    const int iy = (int)((uint32_t)(std::fabs(v) * e->h) % (uint32_t)e->h);
    return texel(e, ix + iy * e->w);
  }
  const LerpAddr lx = tex_address(e->addr[0], u, e->w), ly = tex_address(e->addr[1], v, e->h);
  const V4 v00 = texel(e, lx.i0 + ly.i0 * e->w), v01 = texel(e, lx.i1 + ly.i0 * e->w);
  const V4 v10 = texel(e, lx.i0 + ly.i1 * e->w), v11 = texel(e, lx.i1 + ly.i1 * e->w);
  return lerp4(ly.f, lerp4(lx.f, v00, v01), lerp4(lx.f, v10, v11));
}

// linear3f * vector: columns (common/math, xfmVector)
V3 xfm_lin(const V3 *m, V3 v) { return add(add(scl(v.x, m[0]), scl(v.y, m[1])), scl(v.z, m[2])); }

const float kTwoPi = 2.f * kPi;
const float kOneOverPi = 1.f / kPi;
const float kOneOverTwoPi = 1.f / kTwoPi;

// render/DG.h:97-106 with the shared acos / atan2
float sph_theta(V3 v) { return bnf::acos(std::fmax(std::fmin(v.z, 1.f), -1.f)); }
float sph_phi(V3 v) {
  const float p = bnf::atan2(v.y, v.x);
  return (p < 0.f) ? (p + float(2.f * M_PI)) : p;
}
float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

// EnvMap.h:97-133
float cdf_pdf(int pos, const float *cdf, int N) { return N * (cdf[pos] - (pos ? cdf[pos - 1] : 0.f)); }
int sample_cdf(const float *cdf, int N, float v, float &pdf) {
  int begin = 0, end = N;
  while (end - begin > 1) {
    const int mid = (begin + end) / 2;
    if (v < cdf[mid - 1])
      end = mid;
    else
      begin = mid;
  }
  pdf = cdf_pdf(begin, cdf, N);
  return begin;
}

// EnvMapLight::DD::worldToPixel (EnvMap.h:161-177)
void env_pixel(const ref_envmap_s *e, V3 dir, int &ix, int &iy) {
  const V3 l = xfm_lin(e->toLocal, dir);
  const float u = sph_phi(l) * kOneOverTwoPi, v = sph_theta(l) * kOneOverPi;
  ix = (int)clampf(u * (float)e->w, 0.f, e->w - 1.f);
  iy = (int)clampf(v * (float)e->h, 0.f, e->h - 1.f);
}

// EnvMapLight::DD::pdf (EnvMap.h:135-159)
float env_pdf(const ref_envmap_s *e, V3 dir) {
  int ix, iy;
  env_pixel(e, dir, ix, iy);
  const float py = cdf_pdf(iy, e->cdfY.data(), e->h);
  const float px = cdf_pdf(ix, e->cdfX.data() + (size_t)iy * e->w, e->w);
  const float theta = kPi * ((iy + .5f) / e->h);
  return px * py * 1.f / (kTwoPi * kPi * bnf::sin(theta));
}

// EnvMapLight::DD::uvToWorld (EnvMap.h:195-206)
V3 env_uv_dir(const ref_envmap_s *e, float fx, float fy) {
  const float phi = kTwoPi * fx, theta = kPi * fy;
  float sp, cp, st, ct;
  bnf::sincos(phi, sp, cp);
  bnf::sincos(theta, st, ct);
  return xfm_lin(e->toWorld, mk(cp * st, sp * st, ct));
}

// EnvMapLight::DD::sample (EnvMap.h:208-246): direction, radiance, pdf
void env_sample(const ref_envmap_s *e, Rng &r, V3 &dir, V3 &rad, float &pdf) {
  const float ry = r.uni();
  const float rx = r.uni();
  float py, px;
  const int iy = sample_cdf(e->cdfY.data(), e->h, ry, py);
  const int ix = sample_cdf(e->cdfX.data() + (size_t)e->w * iy, e->w, rx, px);
  const float sx = (ix + r.uni()) / e->w;
  const float sy = (iy + r.uni()) / e->h;
  const V4 c = tex2d(e, sx, sy);
  rad = mk(e->scale * c.x, e->scale * c.y, e->scale * c.z);
  dir = env_uv_dir(e, sx, sy);
  pdf = px * py * 1.f / (kTwoPi * kPi * bnf::sin(kPi * sy));
}

// radianceFromEnv with a texture (shadeRays.cu:437-456)
V3 env_radiance(const ref_envmap_s *e, V3 rayDir) {
  const V3 d = xfm_lin(e->toLocal, nrm(rayDir));
  const float theta = sph_theta(d), phi = sph_phi(d);
  const V4 c = tex2d(e, phi * kOneOverTwoPi, theta * kOneOverPi);
  return mk(e->scale * c.x, e->scale * c.y, e->scale * c.z);
}

// EnvMapLight::commit + computeCDFs (EnvMap.cu:59-150, 178-230): the frame
// from direction and up, its inverse (adjugate times the reciprocal
// determinant), per-texel importance (max channel over a 3x3 half-texel
// stencil), per-row CDFs and the row CDF weighted by sin(theta).
void env_build(ref_envmap_s *e, const float *direction, const float *up) {
  const V3 vz = neg(nrm(ld3(up)));
  const V3 vy = neg(nrm(cross(vz, ld3(direction))));
  const V3 vx = nrm(cross(vy, vz));
  e->toWorld[0] = vx;
  e->toWorld[1] = vy;
  e->toWorld[2] = vz;
  const V3 a = cross(vy, vz), b = cross(vz, vx), c = cross(vx, vy);
  const float r = 1.f / dot(vx, a);
  e->toLocal[0] = scl(r, mk(a.x, b.x, c.x));
  e->toLocal[1] = scl(r, mk(a.y, b.y, c.y));
  e->toLocal[2] = scl(r, mk(a.z, b.z, c.z));
  const int W = e->w, H = e->h;
  e->cdfX.assign((size_t)W * H, 0.f);
  e->cdfY.assign(H, 0.f);
  for (int iy = 0; iy < H; ++iy)
    for (int ix = 0; ix < W; ++ix) {
      float wgt = 0.f;
      for (int jy = 0; jy <= 2; ++jy)
        for (int jx = 0; jx <= 2; ++jx) {
          const V4 t = tex2d(e, (ix + jx * .5f) / W, (iy + jy * .5f) / H);
          wgt = std::max(wgt, std::max(std::max(t.x, t.y), t.z));
        }
      e->cdfX[ix + (size_t)W * iy] = wgt;
    }
  for (int y = 0; y < H; ++y) {
    float *line = e->cdfX.data() + (size_t)y * W;
    float sum = 0.f;
    for (int x = 0; x < W; ++x) sum += line[x];
    const float rs = 1.f / sum;
    sum = 0.f;
    for (int x = 0; x < W; ++x) {
      sum += line[x];
      line[x] = sum * rs;
    }
    line[W - 1] = 1.f;
    e->cdfY[y] = sum * bnf::sin(kPi * ((y + .5f) / H));
  }
  float sum = 0.f;
  for (int i = 0; i < H; ++i) sum += e->cdfY[i];
  const float rs = 1.f / sum;
  sum = 0.f;
  for (int i = 0; i < H; ++i) {
    sum += e->cdfY[i];
    e->cdfY[i] = sum * rs;
  }
  e->cdfY[H - 1] = 1.f;
}

}  // namespace

// ---- quad, directional and point lights ------------------------------------
// Each light's DD as its getDD leaves it (light/QuadLight.cu:9-18,
// DirLight.cu:10-20, PointLight.cu:9-17); the quad's normal, which getDD
// leaves unset, is normalize(edge0 x edge1).
struct RefQuad {
  V3 corner, edge0, edge1, emission, normal;
  float area;
};
struct RefDir {
  V3 direction, color;
  float radiance;
};
struct RefPoint {
  V3 position, color;
  float power, intensity;
};
struct ref_lights_s {
  std::vector<RefQuad> quad;
  std::vector<RefDir> dir;
  std::vector<RefPoint> point;
};

namespace {

// Light::Sample (light/Light.h); pdf 0 = nothing sampled
struct LSample {
  V3 dir = mk(0, 0, 0), rad = mk(0, 0, 0);
  float dist = 0.f, pdf = 0.f;
};

// sampleAreaLights (kernels/shadeRays.cu:46-139)
bool sample_area(const ref_lights_s *L, V3 P, V3 N, Rng &rnd, LSample &ls) {
  const int n = L ? (int)L->quad.size() : 0;
  if (n == 0) return false;
  const int R = 8;
  int id[R];
  float u[R], v[R], w[R];
  float sum = 0.f;
  for (int i = 0; i < R; ++i) {
    id[i] = std::min((int)(rnd.uni() * (float)n), n - 1);
    w[i] = 0.f;
    const RefQuad &q = L->quad[id[i]];
    u[i] = rnd.uni();
    v[i] = rnd.uni();
    const V3 LP = add(add(q.corner, scl(u[i], q.edge0)), scl(v[i], q.edge1));
    V3 d = sub(LP, P);
    const float dist = len(d);
    if (dist < 1e-3f) continue;
    d = scl(1.f / dist, d);
    float weight = eq(N, mk(0, 0, 0)) ? 1.f : dot(d, N);
    if (weight <= 1e-3f) continue;
    weight *= -dot(d, q.normal);
    if (weight <= 1e-3f) continue;
    weight *= (1.f / (dist * dist)) * q.area * maxc(q.emission);
    sum += weight;
    w[i] = weight;
  }
  if (sum == 0.f) return false;
  float r = rnd.uni() * sum;
  int i = 0;
  while (i < R && r >= w[i]) {
    r -= w[i];
    ++i;
  }
  if (i == R) return false;
  const RefQuad &q = L->quad[id[i]];
  const V3 LD = sub(add(add(q.corner, scl(u[i], q.edge0)), scl(v[i], q.edge1)), P);
  ls.dir = nrm(LD);
  ls.dist = len(LD);
  ls.rad = scl(q.area * -dot(q.normal, ls.dir) / (ls.dist * ls.dist), q.emission);
  ls.pdf = w[i] / sum * ((float)R / (float)n);
  return true;
}

// sampleDirLights (shadeRays.cu:141-193)
bool sample_dir(const ref_lights_s *L, V3 N, Rng &rnd, LSample &ls) {
  const int n = L ? (int)L->dir.size() : 0;
  if (n == 0) return false;
  const int R = 2;
  int id[R];
  float w[R];
  float sum = 0.f;
  for (int i = 0; i < R; ++i) {
    id[i] = std::min((int)(rnd.uni() * (float)n), n - 1);
    w[i] = 0.f;
    const RefDir &d = L->dir[id[i]];
    const V3 lrad = scl(d.radiance, d.color);
    float weight = eq(N, mk(0, 0, 0)) ? 1.f : dot(neg(d.direction), N);
    if (weight <= 1e-3f) continue;
    weight *= maxc(lrad);
    if (weight <= 1e-3f) continue;
    w[i] = weight;
    sum += weight;
  }
  if (sum == 0.f) return false;
  float r = rnd.uni() * sum;
  int i = 0;
  while (i < R && r >= w[i]) {
    r -= w[i];
    ++i;
  }
  if (i == R) return false;
  const RefDir &d = L->dir[id[i]];
  ls.dir = neg(d.direction);
  ls.dist = kInf;
  const float q = w[i] / sum * ((float)R / (float)n);
  const V3 c = scl(1.f, scl(d.radiance, d.color));
  ls.rad = mk(c.x / q, c.y / q, c.z / q);
  ls.pdf = kInf;
  return w[i] != 0.f;
}

// PointLight::DD::radianceTowards (PointLight.h:14-20)
V3 point_radiance(const RefPoint &p) {
  return std::isnan(p.intensity) ? scl(1.f / (4.f * kPi), scl(p.power, p.color)) : scl(p.intensity, p.color);
}

// samplePointLights (shadeRays.cu:195-265): falloff c0 + c1 d + c2 d^2 with
// (0, 0, 1)
bool sample_point(const ref_lights_s *L, V3 P, V3 N, Rng &rnd, LSample &ls) {
  const int n = L ? (int)L->point.size() : 0;
  if (n == 0) return false;
  const int R = 2;
  const float c0 = 0.f, c1 = 0.f, c2 = 1.f;
  int id[R];
  float w[R];
  float sum = 0.f;
  for (int i = 0; i < R; ++i) {
    id[i] = std::min((int)(rnd.uni() * (float)n), n - 1);
    w[i] = 0.f;
    const RefPoint &p = L->point[id[i]];
    V3 d = sub(p.position, P);
    const float dist = len(d);
    const float falloff = c0 + c1 * dist + c2 * dist * dist;
    d = nrm(d);
    const V3 lrad = scl(1.f / falloff, point_radiance(p));
    float weight = eq(N, mk(0, 0, 0)) ? 1.f : dot(d, N);
    if (weight <= 1e-3f) continue;
    weight *= maxc(lrad);
    if (weight <= 1e-3f) continue;
    w[i] = weight;
    sum += weight;
  }
  if (sum == 0.f) return false;
  float r = rnd.uni() * sum;
  int i = 0;
  while (i < R && r >= w[i]) {
    r -= w[i];
    ++i;
  }
  if (i == R) return false;
  const RefPoint &p = L->point[id[i]];
  V3 d = sub(p.position, P);
  const float dist = len(d);
  ls.dir = nrm(d);
  const float falloff = c0 + c1 * dist + c2 * dist * dist;
  ls.dist = dist * .9999f;
  const float q = w[i] / sum * ((float)R / (float)n);
  const V3 c = scl(1.f, scl(1.f / falloff, point_radiance(p)));
  ls.rad = mk(c.x / q, c.y / q, c.z / q);
  ls.pdf = kInf;
  return w[i] != 0.f;
}

}  // namespace

struct ref_scene_s {
  std::vector<ref_geom> geoms;
  // triangles whose material is Glass: Glass::getOpacity returns 0 for
  // shadow rays and the triangle anyHit ignores them (Triangles.dev.cu:119-133)
  std::vector<uint8_t> shadowTransparent;
  // per geom: the BSDF alpha its triangle anyHit tests (1 = opaque)
  std::vector<float> alpha;
  bool cutouts = false;
  std::vector<Prim> prims;  // reordered into BVH order
  struct Node {
    float lo[3], hi[3];
    int left;  // internal: left child index (right = left+1); leaf: -1
    int first, count;
  };
  std::vector<Node> nodes;
};

namespace {

// What barney's anyHit / intersection programs do with a candidate before
// reporting it: the renderer's cut plane (OptixGlobals.h:33-42, the world
// point o + t d on the invisible side is dropped) and, for triangles whose
// material alpha is below 1, stochastic opacity (Triangles.dev.cu:121-133:
// the seed advanced by the bits of the object-space point, then
// Random(seed, 290374) > alpha drops the candidate). Each candidate's draw
// uses the ray's seed at the start of the trace (barney chains the seed in
// traversal order), the point v0 + (u e1 + v e2) with e1 = v1 - v0,
// e2 = v2 - v0 as the kernels' primitive records hold them.
struct Filt {
  bool cutOn;
  float cut[4];
  V3 o, d;
  uint64_t seed;
  bool stochastic;
  bool any(float t) const {
    if (!cutOn) return true;
    const V3 P = add(o, scl(t, d));
    return !(dot(mk(cut[0], cut[1], cut[2]), P) + cut[3] < 0.f);
  }
  bool tri(float t, float u, float v, V3 v0, V3 e1, V3 e2, float alpha) const {
    if (!any(t)) return false;
    if (stochastic && alpha < 1.f) {
      const V3 P = add(v0, add(scl(u, e1), scl(v, e2)));
      Rng s;
      s.s = seed;
      s.mix(fbits(P.x));
      s.mix(fbits(P.y));
      s.mix(fbits(P.z));
      Rng r = child(s, 290374ull);
      if (r.uni() > alpha) return false;
    }
    return true;
  }
};

bool cut_enabled(const float *c) {
  return c[3] > -1e28f && !(c[0] == 0.f && c[1] == 0.f && c[2] == 0.f && c[3] == 0.f);
}

Filt make_filt(const ref_render_params *p, V3 o, V3 d, uint64_t seed, bool stochastic = true) {
  Filt f;
  f.cutOn = cut_enabled(p->cutPlane);
  for (int k = 0; k < 4; ++k) f.cut[k] = p->cutPlane[k];
  f.o = o;
  f.d = d;
  f.seed = seed;
  f.stochastic = stochastic;
  return f;
}

// rtcore/cuda/TraceInterface.h:32-98 (barney's plane + 2D-projection test)
void hit_triangle(V3 org, V3 dir, float tmax, V3 v0, V3 v1, V3 v2, uint32_t key, int g, int p, Best &b,
                  const Filt *f = nullptr, float alpha = 1.f) {
  const V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  const V3 N = cross(e1, e2);
  const float dd = dot(dir, N);
  if (std::fabs(dd) < 1e-12f) return;
  if (dd == 0.f) return;
  const float t = -dot(sub(org, v0), N) / dd;
  if (t <= 0.f || t >= tmax) return;
  const V3 P = add(sub(org, v0), scl(t, dir));
  float e1u, e2u, Pu, e1v, e2v, Pv;
  const float ax = std::fabs(N.x), ay = std::fabs(N.y), az = std::fabs(N.z);
  if (ax >= std::fmax(ay, az)) {
    e1u = e1.y; e2u = e2.y; Pu = P.y; e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else if (ay > az) {
    e1u = e1.x; e2u = e2.x; Pu = P.x; e1v = e1.z; e2v = e2.z; Pv = P.z;
  } else {
    e1u = e1.x; e2u = e2.x; Pu = P.x; e1v = e1.y; e2v = e2.y; Pv = P.y;
  }
  // det(a,b,c,d) = a*d - c*b
  if (e1u * e2v - e2u * e1v == 0.f) return;
  const float den = e1u * e2v - e1v * e2u;
  const float u = (Pu * e2v - Pv * e2u) / den;
  const float v = (e1u * Pv - e1v * Pu) / den;
  if (u < 0.f || v < 0.f || (u + v) >= 1.f) return;
  if (!better(t, key, b)) return;
  if (f && !f->tri(t, u, v, v0, e1, e2, alpha)) return;
  b.t = t; b.u = u; b.v = v; b.key = key; b.geom = g; b.prim = p;
}

// barney/geometry/Spheres.dev.cu:116-170
void hit_sphere(V3 org, V3 dir, float tmin, float tmax, V3 c, float r, uint32_t key, int g, int p, Best &b,
                const Filt *f = nullptr) {
  const float t_move = std::fmax(0.f, len(sub(c, org)) - 3.f * r);
  const V3 o = add(org, scl(t_move, dir));
  const float t_max = tmax - t_move;
  if (t_max < 0.f) return;
  float hit_t = t_max;
  const float t_min = std::fmax(0.f, tmin - t_move);
  const V3 oc = sub(o, c);
  const float A = dot(dir, dir), B = dot(oc, dir), C = dot(oc, oc) - r * r;
  const float disc = B * B - A * C;
  if (disc < 0.f) return;
  float tt = (-B - std::sqrt(disc)) / A;
  if (tt < hit_t && tt > t_min) hit_t = tt;
  tt = (-B + std::sqrt(disc)) / A;
  if (tt < hit_t && tt > t_min) hit_t = tt;
  if (!(hit_t < t_max)) return;
  const float t = hit_t + t_move;
  if (!better(t, key, b)) return;
  if (f && !f->any(t)) return;
  b.t = t; b.u = hit_t; b.v = t_move; b.key = key; b.geom = g; b.prim = p;
}

// barney/geometry/Cylinders.dev.cu:101-243
bool cyl(V3 ro, V3 rd, float tmin, float tmax, V3 v0, V3 v1, float radius, float &tout, V3 *Nout) {
  const V3 d = rd, s = sub(v1, v0);
  const V3 sxd = cross(s, d);
  const float a = dot(sxd, sxd);
  if (a == 0.f) return false;
  const V3 f = sub(v0, ro);
  const V3 sxf = cross(s, f);
  const float ra = 1.f / a;
  const float ts = dot(sxd, sxf) * ra;
  const V3 fp = sub(f, scl(ts, d));
  const float s2 = dot(s, s);
  const V3 perp = cross(s, fp);
  const float c = radius * radius * s2 - dot(perp, perp);
  if (c < 0.f) return false;
  float td = std::sqrt(c * ra);
  const float tube0 = ts - td, tube1 = ts + td;
  const float sf = dot(s, f), sd = dot(s, d);
  float cap0 = -1e20f, cap1 = 1e20f, capv0 = cap0, capv1 = cap1;
  if (sd == 0.f) {
    if (dot(sub(ro, v0), sub(v1, v0)) < 0.f) return false;
    if (dot(sub(ro, v1), sub(v0, v1)) < 0.f) return false;
  } else {
    const float rsd = 1.f / sd;
    capv0 = sf * rsd;
    capv1 = capv0 + s2 * rsd;
    cap0 = std::fmin(capv0, capv1);
    cap1 = std::fmax(capv0, capv1);
  }
  const float t0 = std::fmax(cap0, tube0), t1 = std::fmin(cap1, tube1);
  if (t0 > t1) return false;
  V3 N = mk(0, 0, 0);
  float th;
  if (tmin <= t0 && t0 <= tmax) {
    th = t0;
    td *= -1.f;
    const float u = (th * sd - sf) * (1.f / s2);
    if (t0 == cap0) N = (cap0 == capv0) ? neg(s) : s;
    else N = sub(sub(scl(td, d), fp), scl(u, s));
  } else if (tmin <= t1 && t1 <= tmax) {
    th = t1;
    const float u = (th * sd - sf) * (1.f / s2);
    if (t0 == cap1) N = (cap0 == capv0) ? neg(s) : s;  // as in Cylinders.dev.cu:221
    else N = sub(sub(scl(td, d), fp), scl(u, s));
  } else {
    return false;
  }
  tout = th;
  if (Nout) *Nout = N;
  return true;
}

void test_prim(const ref_scene *sc, const Prim &pr, V3 org, V3 dir, float tmin, float tmax, Best &b,
               const Filt *f = nullptr) {
  const ref_geom &g = sc->geoms[pr.geom];
  const int p = pr.prim;
  if (g.kind == 0) {
    const int *ix = g.indices + 3 * p;
    hit_triangle(org, dir, tmax, ld3(g.vertices + 3 * ix[0]), ld3(g.vertices + 3 * ix[1]),
                 ld3(g.vertices + 3 * ix[2]), pr.key, pr.geom, p, b, f, sc->alpha[pr.geom]);
  } else if (g.kind == 1) {
    const float r = g.radii ? g.radii[p] : g.defaultRadius;
    hit_sphere(org, dir, tmin, tmax, ld3(g.vertices + 3 * p), r, pr.key, pr.geom, p, b, f);
  } else {
    const int *ix = g.indices + 2 * p;
    float t;
    if (cyl(org, dir, tmin, tmax, ld3(g.vertices + 3 * ix[0]), ld3(g.vertices + 3 * ix[1]), g.radii[p], t, nullptr) &&
        better(t, pr.key, b) && (!f || f->any(t))) {
      b.t = t; b.u = 0.f; b.v = 0.f; b.key = pr.key; b.geom = pr.geom; b.prim = p;
    }
  }
}

// Conservative slab test (double precision, widened) so that no primitive
// whose box the ray touches is culled.
inline bool box_hit(const float *lo, const float *hi, const double *o, const double *id, double tmin, double tmax) {
  double t0 = tmin, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double a = ((double)lo[k] - o[k]) * id[k], c = ((double)hi[k] - o[k]) * id[k];
    if (a > c) std::swap(a, c);
    t0 = std::max(t0, a);
    t1 = std::min(t1, c);
  }
  return t0 <= t1 * (1.0 + 1e-9) + 1e-9;
}

Best closest(const ref_scene *sc, V3 org, V3 dir, float tmin, float tmax, bool brute, long long *tests,
             bool shadowRay = false, const Filt *f = nullptr) {
  Best b;
  b.t = tmax; b.u = b.v = 0.f; b.key = 0xffffffffu; b.geom = -1; b.prim = -1;
  auto test = [&](const Prim &p) {
    if (shadowRay && sc->shadowTransparent[p.geom]) return;
    test_prim(sc, p, org, dir, tmin, tmax, b, f);
  };
  if (brute) {
    for (const Prim &p : sc->prims) test(p);
    if (tests) *tests += (long long)sc->prims.size();
    return b;
  }
  if (sc->nodes.empty()) return b;
  const double o[3] = {org.x, org.y, org.z};
  const double id[3] = {1.0 / dir.x, 1.0 / dir.y, 1.0 / dir.z};
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const auto &n = sc->nodes[stack[--sp]];
    if (!box_hit(n.lo, n.hi, o, id, tmin, b.t)) continue;
    if (n.left < 0) {
      for (int i = 0; i < n.count; ++i) test(sc->prims[n.first + i]);
      if (tests) *tests += n.count;
    } else {
      stack[sp++] = n.left + 1;
      stack[sp++] = n.left;
    }
  }
  return b;
}

// Binned SAH build (16 bins), leaves of <= 4 prims.
void build_bvh(ref_scene *s) {
  auto &P = s->prims;
  auto &N = s->nodes;
  N.clear();
  if (P.empty()) return;
  N.reserve(2 * P.size());
  N.push_back({});
  struct Task { int node, first, count; };
  std::vector<Task> st{{0, 0, (int)P.size()}};
  while (!st.empty()) {
    Task t = st.back();
    st.pop_back();
    auto &nd = N[t.node];
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    float clo[3] = {kInf, kInf, kInf}, chi[3] = {-kInf, -kInf, -kInf};
    for (int i = t.first; i < t.first + t.count; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], P[i].lo[k]);
        hi[k] = std::max(hi[k], P[i].hi[k]);
        const float c = 0.5f * (P[i].lo[k] + P[i].hi[k]);
        clo[k] = std::min(clo[k], c);
        chi[k] = std::max(chi[k], c);
      }
    for (int k = 0; k < 3; ++k) { N[t.node].lo[k] = lo[k]; N[t.node].hi[k] = hi[k]; }
    (void)nd;
    if (t.count <= 4) {
      N[t.node].left = -1; N[t.node].first = t.first; N[t.node].count = t.count;
      continue;
    }
    const int B = 16;
    double bestCost = 1e300;
    int bestAxis = -1, bestBin = -1;
    for (int k = 0; k < 3; ++k) {
      const float ext = chi[k] - clo[k];
      if (!(ext > 0.f)) continue;
      int cnt[B] = {};
      float blo[B][3], bhi[B][3];
      for (int b = 0; b < B; ++b)
        for (int j = 0; j < 3; ++j) { blo[b][j] = kInf; bhi[b][j] = -kInf; }
      for (int i = t.first; i < t.first + t.count; ++i) {
        const float c = 0.5f * (P[i].lo[k] + P[i].hi[k]);
        int b = std::min(B - 1, (int)((c - clo[k]) / ext * B));
        ++cnt[b];
        for (int j = 0; j < 3; ++j) {
          blo[b][j] = std::min(blo[b][j], P[i].lo[j]);
          bhi[b][j] = std::max(bhi[b][j], P[i].hi[j]);
        }
      }
      auto area = [](const float *l, const float *h) {
        const double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
        return (dx < 0) ? 0.0 : dx * dy + dy * dz + dz * dx;
      };
      double rA[B];
      int rC[B];
      float rl[3] = {kInf, kInf, kInf}, rh[3] = {-kInf, -kInf, -kInf};
      int c = 0;
      for (int b = B - 1; b > 0; --b) {
        c += cnt[b];
        for (int j = 0; j < 3; ++j) { rl[j] = std::min(rl[j], blo[b][j]); rh[j] = std::max(rh[j], bhi[b][j]); }
        rA[b] = area(rl, rh);
        rC[b] = c;
      }
      float ll[3] = {kInf, kInf, kInf}, lh[3] = {-kInf, -kInf, -kInf};
      int lc = 0;
      for (int b = 0; b < B - 1; ++b) {
        lc += cnt[b];
        for (int j = 0; j < 3; ++j) { ll[j] = std::min(ll[j], blo[b][j]); lh[j] = std::max(lh[j], bhi[b][j]); }
        if (lc == 0 || rC[b + 1] == 0) continue;
        const double cost = area(ll, lh) * lc + rA[b + 1] * rC[b + 1];
        if (cost < bestCost) { bestCost = cost; bestAxis = k; bestBin = b; }
      }
    }
    int mid;
    if (bestAxis < 0) {
      mid = t.first + t.count / 2;
    } else {
      const int k = bestAxis;
      const float ext = chi[k] - clo[k];
      auto it = std::partition(P.begin() + t.first, P.begin() + t.first + t.count, [&](const Prim &p) {
        const float c = 0.5f * (p.lo[k] + p.hi[k]);
        return std::min(B - 1, (int)((c - clo[k]) / ext * B)) <= bestBin;
      });
      mid = (int)(it - P.begin());
      if (mid == t.first || mid == t.first + t.count) mid = t.first + t.count / 2;
    }
    const int l = (int)N.size();
    N.push_back({});
    N.push_back({});
    N[t.node].left = l;
    N[t.node].first = t.first;
    N[t.node].count = t.count;
    st.push_back({l, t.first, mid - t.first});
    st.push_back({l + 1, mid, t.first + t.count - mid});
  }
}

V3 fix_dir(V3 d) {
  if (d.x == 0.f) d.x = 1e-6f;
  if (d.y == 0.f) d.y = 1e-6f;
  if (d.z == 0.f) d.z = 1e-6f;
  return d;
}

// Closest-hit epilogue: Triangles.dev.cu:57-74, Spheres.dev.cu:57-90,
// Cylinders.dev.cu:208-216.
void hit_point_normal(const ref_scene *sc, const Best &b, V3 org, V3 dir, float tmin, float tmax, V3 &P, V3 &N) {
  const ref_geom &g = sc->geoms[b.geom];
  const int p = b.prim;
  if (g.kind == 0) {
    const int *ix = g.indices + 3 * p;
    const V3 v0 = ld3(g.vertices + 3 * ix[0]), v1 = ld3(g.vertices + 3 * ix[1]), v2 = ld3(g.vertices + 3 * ix[2]);
    // Triangles.dev.cu:65-75: n = normalize(cross) or the normalized
    // interpolated vertex normal; the object normal is normalize(n)
    V3 n = nrm(cross(sub(v1, v0), sub(v2, v0)));
    if (g.normals)
      n = nrm(add(add(scl(1.f - b.u - b.v, ld3(g.normals + 3 * ix[0])), scl(b.u, ld3(g.normals + 3 * ix[1]))),
                  scl(b.v, ld3(g.normals + 3 * ix[2]))));
    N = nrm(n);
    P = add(add(scl(1.f - b.u - b.v, v0), scl(b.u, v1)), scl(b.v, v2));
  } else if (g.kind == 1) {
    const V3 c = ld3(g.vertices + 3 * p);
    const V3 o = add(org, scl(b.v, dir));
    P = add(o, scl(b.u, dir));
    N = nrm(eq(P, c) ? mk(1, 0, 0) : sub(P, c));
  } else {
    const int *ix = g.indices + 2 * p;
    float t;
    V3 n = mk(0, 0, 0);
    cyl(org, dir, tmin, tmax, ld3(g.vertices + 3 * ix[0]), ld3(g.vertices + 3 * ix[1]), g.radii[p], t, &n);
    P = add(add(org, scl(t, dir)), scl(1e-4f, nrm(n)));
    N = nrm(n);
  }
}

// ---- camera (Camera.cu:81-100, generateRays.cu:73-129) ----------------------
struct Cam {
  V3 org, d00, du, dv;
};
Cam make_cam(const ref_camera *c) {
  Cam r;
  r.org = ld3(c->position);
  V3 d = nrm(nrm(ld3(c->direction)));
  r.du = nrm(cross(d, ld3(c->up)));
  r.dv = nrm(cross(r.du, d));
  r.d00 = scl((float)(1.f / (2.0f * std::tan((0.5f * c->fovy) * (float)M_PI / 180.0f))), d);
  return r;
}

struct PRay {
  V3 org, dir;
  float tMax;
  Rng rng;
  V3 tp;
  float mis;
  int nDiff;
  bool shadow;
  bool inMedium;
  V3 miss;
  float missA;
};

PRay cam_ray(const Cam &c, const ref_render_params *p, int ix, int iy, int accumID) {
  PRay r;
  Rng rnd;
  rnd.seed((uint32_t)(ix + p->width * accumID), (uint32_t)(iy + p->height * accumID));
  r.rng = Rng();
  r.rng.seed((uint32_t)(ix + accumID * p->width), (uint32_t)iy);
  const float pu = accumID == 0 ? .5f : rnd.uni();
  const float pv = accumID == 0 ? .5f : rnd.uni();
  const float iu = (ix + pu) / float(p->width), iv = (iy + pv) / float(p->height);
  const float aspect = p->width / float(p->height);
  const V3 d = add(add(c.d00, scl(1.f * aspect * (iu - .5f), c.du)), scl(1.f * (iv - .5f), c.dv));
  r.org = c.org;
  r.dir = nrm(d);
  r.tMax = 1e30f;
  if (cut_enabled(p->cutPlane)) {  // generateRays.cu:171-193
    const V3 N = ld3(p->cutPlane);
    const float distOrg = dot(N, r.org) + p->cutPlane[3];
    const float denom = dot(N, r.dir);
    if (distOrg >= 0.f) {
      if (denom < 0.f) r.tMax = std::fmin(r.tMax, -distOrg / denom);
    } else if (denom > 0.f) {
      r.org = add(r.org, scl(-distOrg / denom, r.dir));
    } else {
      r.tMax = 0.f;
    }
  }
  r.tp = mk(1, 1, 1);
  r.mis = 0.f;
  r.nDiff = 0;
  r.shadow = false;
  r.inMedium = false;
  const float t = (iy + .5f) / float(p->height);
  if (p->bgColor[3] >= 0.f) {
    r.miss = ld3(p->bgColor);
    r.missA = p->bgColor[3];
  } else {
    r.miss = add(scl(1.0f - t, mk(0.9f, 0.9f, 0.9f)), scl(t, mk(0.15f, 0.25f, .8f)));
    r.missA = (1.0f - t) * 1.f + t * 1.f;
  }
  return r;
}

// ---- materials --------------------------------------------------------------
V3 frame_xfm(V3 N, V3 l) {  // published embree/ospray frame(N)
  const V3 dx0 = mk(0.f, N.z, -N.y), dx1 = mk(-N.z, 0.f, N.x);
  const V3 dx = nrm(dot(dx0, dx0) > dot(dx1, dx1) ? dx0 : dx1);
  const V3 dy = nrm(cross(N, dx));
  return add(add(scl(l.x, dx), scl(l.y, dy)), scl(l.z, N));
}

// float -> binary16 -> float with round-to-nearest-even: barney keeps the
// NVisii parameters as halfs (NVisii.h:1062-1083).
float half_round(float x) {
  if (!std::isfinite(x) || x == 0.f) return x;
  const float ax = std::fabs(x);
  if (ax >= 65520.f) return std::copysign(kInf, x);
  int e;
  std::frexp(ax, &e);                      // ax in [2^(e-1), 2^e)
  const float q = std::ldexp(1.f, std::max(e - 11, -24));  // 11 significant bits; subnormal step 2^-24
  return std::copysign(std::nearbyint(ax / q) * q, x);
}

enum { M_LAMBERT = 1, M_PHASE = 2, M_NVISII = 3, M_GLASS = 4 };

// what Ray::setHit leaves for bounce(): point, geometric normal, BSDF
struct HitInfo {
  bool hit;
  V3 P, N;
  float t;
  int type;
  V3 albedo;       // Lambert/phase albedo, NVisii base color
  float metallic, roughness, ior;
};

// DeviceMaterial::createBSDF for one geometry's material at one primitive:
// AnariMatte.h:51-63 (albedo .85 color / pi), AnariPBR.h:68-94 (Glass or NVisii).
void material_hit(const ref_geom &g, V3 col, HitInfo &h) {
  if (g.matType == 1) {
    if (g.ior != 1.f && g.transmission >= 1e-3f) {
      h.type = M_GLASS;
      h.albedo = mk(1, 1, 1);
      h.ior = g.ior;
      h.metallic = h.roughness = 0.f;
      return;
    }
    h.type = M_NVISII;
    h.albedo = mk(half_round(col.x), half_round(col.y), half_round(col.z));
    h.metallic = half_round(g.metallic);
    h.roughness = half_round(std::fmin(std::fmax(g.roughness, .001f), 1.f - .1f));
    h.ior = half_round(g.ior);
    return;
  }
  h.type = M_LAMBERT;
  h.albedo = mk(.85f * col.x * (1.f / kPi), .85f * col.y * (1.f / kPi), .85f * col.z * (1.f / kPi));
  h.metallic = h.roughness = h.ior = 0.f;
}

// ---- NVisii (barney/packedBSDF/NVisii.h), all lobes as written there ------
namespace nv {
const float kMinAlpha = .002f, kSmallEps = 2e-10f;
const float k1Pi = (float)M_1_PI;
struct Mat {
  V3 base, subsurf;
  float metallic, specular, roughness, specTint, aniso, sheen, sheenTint, clearcoat, clearcoatGloss, ior,
      specTrans, transRough, flatness;
};
// NVisii::setDefaults (NVisii.h:1040-1060) stored as halfs, then the
// AnariPBR overrides (base color, metallic, roughness, ior)
Mat unpack(const HitInfo &h) {
  Mat m;
  m.base = h.albedo;
  const float ss = half_round(.8f);
  m.subsurf = mk(ss, ss, ss);
  m.metallic = h.metallic;
  m.specular = half_round(.5f);
  m.roughness = h.roughness;
  m.specTint = 0.f;
  m.aniso = 0.f;
  m.sheen = 0.f;
  m.sheenTint = half_round(.5f);
  m.clearcoat = 0.f;
  m.clearcoatGloss = half_round(1.f - .03f * .03f);
  m.ior = h.ior;
  m.specTrans = 0.f;
  m.transRough = half_round(.04f);
  m.flatness = 0.f;
  return m;
}
float saturate(float f) { return std::fmax(0.f, std::fmin(f, 1.f)); }
float lerp_r(float a, float b, float t) { return (1.f - t) * a + t * b; }
V3 lerp_r(V3 a, V3 b, float t) { return add(scl(1.f - t, a), scl(t, b)); }
float schlick(float c) { return bnf::pow5(saturate(1.f - c)); }
bool same_hemi(V3 wo, V3 wi, V3 n) { return dot(wo, n) * dot(wi, n) > 0.f; }
float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }  // DG.h:109
V3 reflect(V3 v, V3 n) { return sub(v, scl(2.f * dot(v, n), n)); }                    // DG.h:115
// DG.h:121-128; '(eta*d + sqrtf(k)*n)' adds a scalar to a vector there,
// i.e. vec3f(eta*d) + sqrtf(k)*n, which is kept
V3 refract(V3 v, V3 n, float eta) {
  const float d = dot(n, v);
  const float k = 1.f - eta * eta * (1.f - d * d);
  if (!(k >= 0.f)) return mk(0, 0, 0);
  const float ed = eta * d;
  return sub(scl(eta, v), add(mk(ed, ed, ed), scl(std::sqrt(k), n)));
}
float gtr_1(float c, float a) {  // :120-128
  if (a >= 1.f) return k1Pi;
  const float a2 = a * a;
  return k1Pi * (a2 - 1.f) / (std::log(a2) * (1.f + (a2 - 1.f) * c * c));
}
float gtr_2(float c, float a) {  // :132-141 (no a^2 numerator)
  const float a2 = a * a;
  const float den1 = 1.f + (a2 - 1.f) * c * c;
  const float den2 = std::fmax(den1, kSmallEps);
  return k1Pi * 1.f / den2;
}
float smith(float n_dot_o, float a_g) {  // :150-154
  const float a = a_g * a_g, b = n_dot_o * n_dot_o;
  return 1.f / (n_dot_o + std::sqrt(a + b - a * b));
}
float lambertian_pdf(V3 wi, V3 n) {  // :204-210
  const float d = dot(wi, n);
  return d > 0.f ? d * k1Pi : 0.f;
}
float gtr_1_pdf(V3 wo, V3 wi, V3 wh, V3 n, float a) {  // :213-220
  if (!same_hemi(wo, wi, n)) return 0.f;
  const float c = dot(n, wh);
  return gtr_1(c, a) * c / (4.f * dot(wo, wh));
}
float gtr_2_pdf(V3 wo, V3 wi, V3 wh, V3 n, float a) {  // :223-230
  if (!same_hemi(wo, wi, n)) return 0.f;
  const float c = std::fabs(dot(n, wh));
  return gtr_2(c, a) * c / (4.f * std::fabs(dot(wo, wh)));
}
// cos_sample_hemisphere (:73-91): concentric disk
V3 cos_hemi(float ux, float uy) {
  const float sx = 2.f * ux - 1.f, sy = 2.f * uy - 1.f;
  float radius = 0.f, theta = 0.f;
  if (!(sx == 0.f && sy == 0.f)) {
    if (std::fabs(sx) > std::fabs(sy)) {
      radius = sx;
      theta = (float)M_PI / 4.f * (sy / sx);
    } else {
      radius = sy;
      theta = (float)M_PI / 2.f - (float)M_PI / 4.f * (sx / sy);
    }
  }
  const float dx = radius * bnf::cos(theta), dy = radius * bnf::sin(theta);
  return mk(dx, dy, std::sqrt(std::fmax(0.f, 1.f - dx * dx - dy * dy)));
}
V3 on_frame(V3 h, V3 vx, V3 vy, V3 n) { return add(add(scl(h.x, vx), scl(h.y, vy)), scl(h.z, n)); }
// sample_gtr_2_h (:186-194); sin_theta_h = 1 - cos^2 as written
V3 sample_gtr_2_h(V3 n, V3 vx, V3 vy, float a, float sx, float sy) {
  const float phi = 2.f * (float)M_PI * sx;
  const float c2 = (1.f - sy) / (1.f + (a * a - 1.f) * sy);
  const float ch = std::sqrt(c2), sh = 1.f - c2;
  const V3 hd = nrm(mk(sh * bnf::cos(phi), sh * bnf::sin(phi), ch));
  return on_frame(hd, vx, vy, n);
}
// disney_brdf (:655-700) with the isotropic microfacet lobe (:480-503)
V3 brdf(const Mat &m, V3 n, V3 wo, V3 wi, V3 wh) {
  const float n_dot_o = std::fabs(dot(wo, n)), n_dot_i = std::fabs(dot(wi, n));
  const float i_dot_h = dot(wi, wh);
  // clear coat (:605-613)
  const float cca = lerp_r(0.1f, kMinAlpha, m.clearcoatGloss);
  const float coat = m.clearcoat * gtr_1(std::fabs(dot(n, wh)), cca) * lerp_r(0.04f, 1.f, schlick(dot(wi, n))) *
                     (smith(std::fabs(dot(n, wi)), 0.25f) * smith(std::fabs(dot(n, wo)), 0.25f));
  // sheen (:617-624)
  const float lum = luminance(m.base);
  const V3 tint = lum > 0.f ? scl(1.f / lum, m.base) : mk(1, 1, 1);
  const V3 sheen = scl(schlick(dot(wi, n)) * m.sheen, lerp_r(mk(1, 1, 1), tint, m.sheenTint));
  // diffuse (:338-352, no 1/pi)
  const float fd90 = 0.5f + 1.0f * m.roughness * i_dot_h * i_dot_h;
  const float diff = lerp_r(1.f, fd90, schlick(n_dot_i)) * lerp_r(1.f, fd90, schlick(n_dot_o));
  // subsurface (:355-367)
  const float FL = schlick(n_dot_i), FV = schlick(n_dot_o);
  const float Fss90 = i_dot_h * i_dot_h * m.roughness;
  const float Fss = lerp_r(1.0f, Fss90, FL) * lerp_r(1.0f, Fss90, FV);
  const float ssv = k1Pi * (1.25f * (Fss * (1.f / (n_dot_i + n_dot_o) - .5f) + .5f));
  // microfacet (:480-503)
  const V3 spec = lerp_r(scl(m.specular * 0.08f, lerp_r(mk(1, 1, 1), tint, m.specTint)), m.base, m.metallic);
  const float alpha = std::fmax(kMinAlpha, m.roughness * m.roughness);
  const float d = gtr_2(std::fabs(dot(n, wh)), alpha);
  const float fw = schlick(std::fabs(dot(wi, wh))) * lerp_r(.5f, 1.f, std::fmax(m.metallic, alpha));
  const V3 f = lerp_r(spec, mk(1, 1, 1), fw);
  const float g = smith(std::fabs(dot(n, wi)), alpha) * smith(std::fabs(dot(n, wo)), alpha);
  const V3 gloss = scl(g, scl(d, f));
  const V3 flat = lerp_r(scl(diff, m.base), scl(ssv, m.subsurf), m.flatness);
  V3 r = scl((1.f - m.specTrans), scl((1.f - m.metallic), flat));
  r = add(r, sheen);
  r = add(r, mk(coat, coat, coat));
  return add(r, gloss);
}
// disney_pdf (:717-810), IMPORTANCE_SAMPLE_BRDF weights, unnormalized
float pdf(const Mat &m, V3 n, V3 wo, V3 wi, V3 wh) {
  const float dw = (1.f - m.metallic) * (1.f - m.specTrans);
  const float gw = 1.f + m.metallic + m.sheen + m.specular + m.roughness;
  const float cw = m.clearcoat, tw = 0.f;
  if (dw + gw + cw + tw == 0.f) return 0.f;
  const float alpha = std::fmax(kMinAlpha, m.roughness * m.roughness);
  const float cca = lerp_r(0.1f, kMinAlpha, m.clearcoatGloss);
  const float diffuse = lambertian_pdf(wi, n);
  const float cc = gtr_1_pdf(wo, wi, wh, n, cca);
  const float mf = gtr_2_pdf(wo, wi, wh, n, alpha);
  return dw * diffuse + gw * mf + cw * cc + tw * 0.f;
}
}  // namespace nv

// ---- Glass: RobustDielectric::sample (barney/packedBSDF/Glass.h:85-190) -----
V3 glass_sample(float eta, V3 wo, V3 Ns, V3 Ng, Rng &rnd, bool &transmit) {
  const V3 incoming = neg(wo);
  const float geps = 1.e-6f;
  const bool canReflect = (eta < 1.f - geps || eta > 1.f + geps);
  if (dot(incoming, Ns) > 0.f) Ns = neg(Ns);
  if (dot(incoming, Ng) > 0.f) Ng = neg(Ng);
  // reflectionDirection (:72-80)
  const V3 par = scl(dot(incoming, Ns), Ns);
  const V3 dirR = sub(sub(incoming, par), par);
  // refractionDirection (:20-68)
  V3 dirT;
  float fres;
  bool mustReflect = false;
  {
    const float cos1 = dot(incoming, Ns);
    float c1a = cos1;
    if (c1a < 0.0) c1a = -c1a;
    const bool headOn = (1.f - 1e-6f < c1a && c1a < 1.f - 1e-6f);
    if (headOn) {
      const float q = (eta - 1.f) / (eta + 1.f);
      fres = q * q;
      dirT = incoming;
    } else {
      const V3 dpar = scl(cos1, Ns);
      const V3 dperp = sub(incoming, dpar);
      const V3 b = nrm(dperp);
      const float sin1 = dot(dperp, b);
      const float sin2 = sin1 / eta;
      if (sin2 > 1.f) {
        dirT = mk(0, 0, 0);
        fres = 1.f;
        mustReflect = true;
      } else {
        const float cos2 = std::sqrt(std::fmin(std::fmax(1.f - sin2 * sin2, 0.f), 1.f));
        const V3 N2 = nrm(dpar);
        dirT = add(scl(cos2, N2), scl(sin2, b));
        const float r1 = (eta * c1a - cos2) / (eta * c1a + cos2);
        const float r2 = (c1a - eta * cos2) / (c1a + eta * cos2);
        fres = 0.5f * (r1 * r1 + r2 * r2);
      }
    }
    fres = std::fmin(std::fmax(fres, 0.f), 1.f);
  }
  V3 wi;
  if ((rnd.uni() < fres && canReflect) || mustReflect) {
    wi = dirR;
    transmit = false;
    const float gc = dot(wi, Ng);
    if (gc >= 0.f) {
      if (gc < geps) wi = nrm(add(wi, scl(geps, Ng)));
    } else {
      wi = sub(wi, scl(gc, scl(2.f, Ng)));
    }
  } else {
    wi = dirT;
    transmit = true;
    const float gc = dot(wi, Ng);
    if (gc <= 0.f) {
      if (gc > -geps) wi = nrm(sub(wi, scl(geps, Ng)));
    } else {
      transmit = false;
    }
  }
  return wi;
}

enum { S_DIFFUSE = 0, S_GLOSSY = 1, S_SPECULAR = 2, S_VOLUME = 3 };

// BSDF::eval / pdf / scatter dispatch (PackedBSDF.h), dg.Ns = dg.Ng = Ng
float bsdf_eval(const HitInfo &h, V3 Ng, V3 wo, V3 wi, V3 &val) {
  if (h.type == M_LAMBERT) {  // Lambertian.h:55-64
    const float c = std::fmax(dot(wi, Ng), 0.f);
    val = mk(h.albedo.x * (1.f / kPi) * c, h.albedo.y * (1.f / kPi) * c, h.albedo.z * (1.f / kPi) * c);
    return c * (1.f / kPi);
  }
  if (h.type == M_NVISII) {  // NVisii.h:1167-1195
    const nv::Mat m = nv::unpack(h);
    const V3 wh = nrm(add(wi, wo));
    val = nv::brdf(m, Ng, wo, wi, wh);
    return nv::pdf(m, Ng, wo, wi, wh);
  }
  if (h.type == M_GLASS) {
    val = mk(0, 0, 0);
    return 0.f;
  }
  const float d = 1.f / (4.f * kPi);  // Phase.h:120-141
  val = scl(d, h.albedo);
  return d;
}

float bsdf_pdf(const HitInfo &h, V3 Ng, V3 wo, V3 wi) {
  if (h.type == M_LAMBERT) {
    V3 N = Ng;
    if (dot(wi, N) < 0.f) N = neg(N);
    return std::fmax(dot(wi, N), 0.f) * (1.f / kPi);
  }
  if (h.type == M_NVISII) {
    const nv::Mat m = nv::unpack(h);
    return nv::pdf(m, Ng, wo, wi, nrm(add(wi, wo)));
  }
  if (h.type == M_GLASS) return 0.f;
  return 1.f / (4.f * kPi);
}

struct Scat {
  V3 fr, dir;
  float pdf;
  int type;
  float offDir;
  bool medium;
};

Scat bsdf_scatter(const HitInfo &h, V3 Ng, V3 wo, bool inMedium, Rng &rnd) {
  Scat s;
  s.offDir = 1.f;
  s.medium = false;
  if (h.type == M_LAMBERT) {  // Lambertian.h:66-87, DG.h:160-166
    V3 N = Ng;
    if (dot(wo, N) < 0.f) N = neg(N);
    const float sx = rnd.uni(), sy = rnd.uni();
    const float phi = (2.f * kPi) * sx;
    const float ct = std::sqrt(sy), st = std::sqrt(1.0f - sy);
    const V3 l = mk(bnf::cos(phi) * st, bnf::sin(phi) * st, ct);
    s.dir = frame_xfm(N, l);
    s.pdf = l.z * (1.f / kPi);
    s.fr = h.albedo;
    s.type = S_DIFFUSE;
  } else if (h.type == M_NVISII) {  // NVisii.h:1092-1163 + sample_disney_brdf :820-975
    const nv::Mat m = nv::unpack(h);
    const V3 n = Ng;
    const V3 vy = nrm(cross(Ng, wo)), vx = nrm(cross(Ng, vy));
    float dw = (1.f - m.metallic) * (1.f - m.specTrans);
    float gw = 1.f + m.metallic + m.sheen + m.specular + m.roughness;
    float cw = m.clearcoat, tw = 0.f;
    const float sum = dw + gw + cw + tw;
    if (sum == 0.f) {
      s.pdf = 0.f;
      s.fr = s.dir = mk(0, 0, 0);
      s.type = S_GLOSSY;
      return s;
    }
    const float sc = 1.f / sum;
    dw *= sc; gw *= sc; cw *= sc; tw *= sc;
    const float tr = rnd.uni();
    int lobe;
    if (tr < dw) lobe = 0;
    else if ((tr - dw) < gw) lobe = 1;
    else if ((tr - dw - gw) < cw) lobe = 2;
    else lobe = 3;
    const float ux = rnd.uni(), uy = rnd.uni();
    V3 wi;
    s.type = lobe == 0 ? S_DIFFUSE : S_GLOSSY;
    if (lobe == 0) {
      wi = nv::on_frame(nrm(nv::cos_hemi(ux, uy)), vx, vy, n);
    } else if (lobe == 1) {
      const float alpha = std::fmax(nv::kMinAlpha, m.roughness * m.roughness);
      const V3 wh = nv::sample_gtr_2_h(n, vx, vy, alpha, ux, uy);
      wi = nv::reflect(neg(wo), wh);
      if (!nv::same_hemi(wo, wi, n)) {
        s.pdf = 0.f;
        s.fr = s.dir = mk(0, 0, 0);
        return s;
      }
    } else if (lobe == 2) {
      // clear coat lobe: weight 0 here (clearcoat = 0), never selected
      s.pdf = 0.f;
      s.fr = s.dir = mk(0, 0, 0);
      return s;
    } else {
      // transmission lobe: reachable only when the normalized weights
      // round to a sum below the drawn number
      const float alpha = std::fmax(nv::kMinAlpha, m.transRough * m.transRough);
      const V3 wh = nv::sample_gtr_2_h(n, vx, vy, alpha, ux, uy);
      const bool entering = dot(wo, wh) > 0.f;
      const float eta_i = entering ? 1.f : m.ior, eta_o = entering ? m.ior : 1.f;
      const V3 hn = entering ? wh : neg(wh);
      wi = nv::refract(neg(wo), hn, eta_i / eta_o);
      if (wi.x == 0.f && wi.y == 0.f && wi.z == 0.f) {
        wi = nv::reflect(neg(wo), hn);
        s.pdf = 1.f;
        s.fr = mk(1, 1, 1);
        s.dir = nrm(wi);
        return s;
      }
    }
    const V3 wh = nrm(add(wi, wo));
    s.pdf = nv::pdf(m, n, wo, wi, wh);
    s.fr = nv::brdf(m, n, wo, wi, wh);
    s.dir = nrm(wi);
  } else if (h.type == M_GLASS) {  // Glass.h:219-241
    const float eta = !inMedium ? 1.f / h.ior : h.ior;
    bool tr = false;
    s.dir = glass_sample(1.f / eta, wo, Ng, Ng, rnd, tr);
    s.fr = mk(1, 1, 1);
    s.pdf = kInf;
    s.offDir = tr ? -1.f : 1.f;
    s.medium = tr;
    s.type = S_SPECULAR;
  } else {  // Phase.h:143-172
    const float r1 = rnd.uni(), r2 = rnd.uni();
    const float x = bnf::cos(2.f * kPi * r1) * std::sqrt(r2 * (1.f - r2));
    const float y = bnf::sin(2.f * kPi * r1) * std::sqrt(r2 * (1.f - r2));
    const float z = (1.f - 2.f * r2);
    s.pdf = 1.f / (4.f * kPi);
    s.fr = scl(s.pdf, h.albedo);
    s.dir = mk(x, y, z);
    s.type = S_VOLUME;
  }
  return s;
}

// shadeRays.cu:476-878: the environment light (an HDRI map, EnvMapLight, or
// the uniform ambientRadiance) and the quad, directional and point lights
V3 bounce(const ref_render_params *p, PRay &ray, const HitInfo &h, PRay &sh, int depth, int maxDepth, int maxDiff) {
  sh.tMax = -1.f;
  V3 frag = mk(0, 0, 0);
  const V3 tp0 = ray.tp;
  const float amb = p->ambientRadiance;
  const ref_envmap_s *env = p->envMap;
  if (ray.shadow) {
    if (!h.hit) frag = scl(ray.mis, ray.tp);
    ray.tMax = -1.f;
    return frag;
  }
  V3 Ng = h.hit ? h.N : mk(0, 0, 0);
  const bool vol = eq(Ng, mk(0, 0, 0));
  if (!vol) Ng = nrm(Ng);
  const V3 Ngff = dot(ray.dir, Ng) < 0.f ? Ng : neg(Ng);
  if (!h.hit) {
    // primaryRayMissColor / radianceFromEnv (shadeRays.cu:437-470, 556-575)
    if (depth == 0)
      frag = env ? env_radiance(env, ray.dir) : ray.miss;
    else
      frag = scl(ray.mis, mul(ray.tp, env ? env_radiance(env, ray.dir) : mk(amb, amb, amb)));
    ray.tMax = -1.f;
    return frag;
  }
  Rng rnd = child(ray.rng, (uint64_t)fbits(h.t));
  const V3 P = h.P;
  const V3 wo = neg(nrm(ray.dir));
  const float eps = std::fmax(1e-4f, 1e-6f * maxc(vabs(P)));
  const V3 off = vol ? wo : Ngff;
  // sampleEnvLight (shadeRays.cu:267-293): the map's importance sample, or
  // randomDirection (DG.h:184-194) into the hemisphere of the normal
  V3 ld, lrad;
  float lpdf;
  if (env) {
    env_sample(env, rnd, ld, lrad, lpdf);
  } else {
    while (true) {
      V3 v;
      v.x = 1.f - 2.f * rnd.uni();
      v.y = 1.f - 2.f * rnd.uni();
      v.z = 1.f - 2.f * rnd.uni();
      if (dot(v, v) <= 1.f) { ld = nrm(v); break; }
    }
    if (!eq(Ngff, mk(0, 0, 0))) {
      lrad = mk(2.f * kPi * amb, 2.f * kPi * amb, 2.f * kPi * amb);
      if (dot(ld, Ngff) < 0.f) ld = neg(ld);
      lpdf = 1.f / (2.f * kPi);
    } else {
      const float r4 = 4.f * kPi * amb;
      lrad = mk(r4, r4, r4);
      lpdf = 1.f / (4.f * kPi);
    }
  }
  // sampleLights (shadeRays.cu:295-404): the environment's sample (above),
  // then the area, directional and point lights' samples; one is picked by
  // its share of the summed weights, its radiance divided by that share.
  // Picking a light whose sampler returned false gives Light::Sample's
  // default pdf 0, i.e. no light.
  bool needMIS = false, have = false;
  float ldist = kInf;
  {
    const float elsW0 = maxc(lrad);
    LSample als, dls, pls;
    const float alsW0 = sample_area(p->lights, P, Ngff, rnd, als) ? maxc(als.rad) / als.pdf : 0.f;
    const float dlsW0 = sample_dir(p->lights, Ngff, rnd, dls) ? maxc(dls.rad) : 0.f;
    const float plsW0 = sample_point(p->lights, P, Ngff, rnd, pls) ? maxc(pls.rad) : 0.f;
    const float sum = alsW0 + dlsW0 + elsW0 + plsW0;
    if (sum != 0.f) {
      const float alsW = alsW0 * (1.f / sum), elsW = elsW0 * (1.f / sum);
      const float dlsW = dlsW0 * (1.f / sum), plsW = plsW0 * (1.f / sum);
      const float r = rnd.uni();
      LSample ls;
      if (r <= alsW) {
        ls = als;
        ls.rad = scl(1.f / alsW, ls.rad);
      } else if (r <= alsW + elsW) {
        ls.dir = ld;
        ls.rad = scl(1.f / elsW, lrad);
        ls.dist = kInf;
        ls.pdf = lpdf;
        needMIS = true;
      } else if (r <= alsW + elsW + plsW) {
        ls = pls;
        ls.rad = scl(1.f / plsW, ls.rad);
      } else {
        ls = dls;
        ls.rad = scl(1.f / dlsW, ls.rad);
      }
      have = !(std::isnan(ls.pdf) || ls.pdf <= 0.f);
      ld = ls.dir;
      lrad = ls.rad;
      lpdf = ls.pdf;
      ldist = ls.dist;
    }
  }
  if (have) {
    V3 fv;
    const float fp = bsdf_eval(h, Ng, wo, ld, fv);
    if (!(fp > 0.f) || maxc(fv) < 1e-4f) {
      sh.tMax = -1.f;
    } else {
      fv = vmins(fv, 13.f);
      const float s1 = std::isinf(lpdf) ? 1.f : 1.f / lpdf;
      const float s2 = vol ? 1.f : ((1.f / kPi) * std::fabs(dot(Ng, ld)));
      const V3 tsr = scl(s2, mul(mul(scl(s1, tp0), fv), lrad));
      sh.shadow = true;
      sh.dir = ld;
      sh.org = add(P, scl(eps, off));
      sh.tMax = ldist * (1.f - 2.f * eps);
      sh.tp = tsr;
      sh.rng = ray.rng;
      ray.rng.mix((uint64_t)fbits(h.t));
      sh.mis = 1.f;
      sh.nDiff = ray.nDiff;
      sh.inMedium = ray.inMedium;
      if (needMIS) {
        const float pl = env ? env_pdf(env, ld) : 1.f / (4.f * kPi);
        const float pb = bsdf_pdf(h, Ng, wo, ld);
        sh.mis = pl / (pl + pb + 1e-10f);
        if (sh.mis < 1e-5f) sh.tMax = -1.f;
      }
    }
  }
  if (depth >= maxDepth) {
    ray.tMax = -1.f;
    return frag;
  }
  ray.tMax = kInf;
  const Scat sr = bsdf_scatter(h, Ng, wo, ray.inMedium, rnd);
  if (!(sr.pdf > 0.f) || sr.pdf <= 1e-6f) return frag;
  if (sr.type == S_VOLUME) {  // SCI_VIS_MODE
    ray.tMax = -1.f;
    return frag;
  }
  if (sr.type == S_DIFFUSE) {
    if (ray.nDiff >= maxDiff) {
      ray.tMax = -1.f;
      return frag;
    }
    ray.nDiff += 1;
  }
  ray.org = add(P, scl(sr.offDir * eps, off));
  ray.dir = nrm(sr.dir);
  const float cosf_ = vol ? 1.f : std::fabs(dot(Ng, nrm(ray.dir)));
  const float den = std::isinf(sr.pdf) ? 1.f : (sr.pdf + 1e-10f);
  V3 sf = mk(sr.fr.x * cosf_ / den, sr.fr.y * cosf_ / den, sr.fr.z * cosf_ / den);
  sf = vmins(sf, 2.5f);
  ray.tp = mul(ray.tp, sf);
  if (sr.medium) ray.inMedium = !ray.inMedium;
  if (needMIS && !std::isinf(sr.pdf)) {
    const float pb = bsdf_pdf(h, Ng, wo, ray.dir), pl = env ? env_pdf(env, ray.dir) : 1.f / (4.f * kPi);
    ray.mis = pb / (pb + pl);
  } else {
    ray.mis = 1.f;
  }
  if (p->russianRoulette > 0 && depth + 1 >= p->russianRoulette) {
    const float q = std::fmin(maxc(ray.tp), 0.95f);
    if (rnd.uni() >= q) {
      ray.tMax = -1.f;
      return frag;
    }
    ray.tp = scl(1.f / q, ray.tp);
  }
  return frag;
}

void add_frag(V3 &sum, V3 f, float cmax) {
  f = vmins(f, cmax);
  if (f.x > 0.f) sum.x += f.x;
  if (f.y > 0.f) sum.y += f.y;
  if (f.z > 0.f) sum.z += f.z;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

uint64_t ref_rng_seed(uint64_t state, uint32_t a, uint32_t b) {
  Rng r;
  r.s = state;
  r.seed(a, b);
  return r.s;
}
void ref_rng_sequence(uint64_t *state, uint32_t *out, int n) {
  Rng r;
  r.s = *state;
  for (int i = 0; i < n; ++i) out[i] = r.next();
  *state = r.s;
}
void ref_random2_floats(uint32_t a, uint32_t b, float *out, int n) {
  Rng r;
  r.seed(a, b);
  for (int i = 0; i < n; ++i) out[i] = r.uni();
}
uint64_t ref_hash3(uint32_t a, uint32_t b, uint32_t c) { return fnv3(a, b, c); }
// ---- environment map (EnvMap.cu, EnvMap.h; textures: embree Texture.cpp) ----
ref_envmap *ref_envmap_create(const void *texels, int format, int w, int h, int filter, const int *addr,
                              const float *direction, const float *up, float scale) {
  if (w <= 0 || h <= 0 || !texels) return nullptr;
  auto *e = new ref_envmap_s;
  e->w = w;
  e->h = h;
  e->filter = filter;
  e->addr[0] = addr[0];
  e->addr[1] = addr[1];
  e->scale = scale;
  const size_t n = (size_t)w * h;
  e->tex.resize(4 * n);
  for (size_t i = 0; i < n; ++i) {
    float *t = e->tex.data() + 4 * i;
    if (format == 33) {  // rtc::FLOAT4
      for (int k = 0; k < 4; ++k) t[k] = ((const float *)texels)[4 * i + k];
    } else if (format == 30) {  // rtc::FLOAT
      t[0] = t[1] = t[2] = t[3] = ((const float *)texels)[i];
    } else if (format == 1) {  // rtc::UCHAR4
      for (int k = 0; k < 4; ++k) t[k] = (float)((const uint8_t *)texels)[4 * i + k] * 1.f / 255.f;
    } else if (format == 0) {  // rtc::UCHAR
      t[0] = t[1] = t[2] = t[3] = (float)((const uint8_t *)texels)[i] * 1.f / 255.f;
    } else {
      delete e;
      return nullptr;
    }
  }
  env_build(e, direction, up);
  return e;
}
void ref_envmap_destroy(ref_envmap *e) { delete e; }

ref_lights *ref_lights_create(const float *quads, int nq, const float *dirs, int nd, const float *points, int np) {
  auto *L = new ref_lights_s;
  for (int i = 0; i < nq; ++i) {
    const float *q = quads + 12 * i;
    RefQuad d;
    d.corner = ld3(q);
    d.edge0 = ld3(q + 3);
    d.edge1 = ld3(q + 6);
    d.emission = ld3(q + 9);
    const V3 c = cross(d.edge0, d.edge1);
    d.area = len(c);  // QuadLight::getDD: length(cross(edge0, edge1))
    d.normal = nrm(c);
    L->quad.push_back(d);
  }
  for (int i = 0; i < nd; ++i) {
    const float *q = dirs + 7 * i;
    RefDir d;
    d.direction = nrm(ld3(q));  // DirLight::getDD normalizes
    d.color = ld3(q + 3);
    d.radiance = q[6];
    L->dir.push_back(d);
  }
  for (int i = 0; i < np; ++i) {
    const float *q = points + 8 * i;
    RefPoint d;
    d.position = ld3(q);
    d.color = ld3(q + 3);
    d.power = q[6];
    d.intensity = q[7];
    L->point.push_back(d);
  }
  return L;
}

void ref_lights_destroy(ref_lights *l) { delete l; }
void ref_envmap_get_cdfs(const ref_envmap *e, float *cdfY, float *cdfX) {
  std::copy(e->cdfY.begin(), e->cdfY.end(), cdfY);
  std::copy(e->cdfX.begin(), e->cdfX.end(), cdfX);
}
void ref_envmap_frame(const ref_envmap *e, float *toWorld, float *toLocal) {
  for (int k = 0; k < 3; ++k) {
    toWorld[3 * k] = e->toWorld[k].x; toWorld[3 * k + 1] = e->toWorld[k].y; toWorld[3 * k + 2] = e->toWorld[k].z;
    toLocal[3 * k] = e->toLocal[k].x; toLocal[3 * k + 1] = e->toLocal[k].y; toLocal[3 * k + 2] = e->toLocal[k].z;
  }
}
void ref_envmap_tex2d(const ref_envmap *e, const float *uv, float *out, int n) {
  for (int i = 0; i < n; ++i) {
    const V4 v = tex2d(e, uv[2 * i], uv[2 * i + 1]);
    out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
  }
}
void ref_envmap_eval(const ref_envmap *e, const float *dirs, float *pdf, float *rad, int n) {
  for (int i = 0; i < n; ++i) {
    const V3 d = ld3(dirs + 3 * i);
    pdf[i] = env_pdf(e, d);
    const V3 r = env_radiance(e, d);
    rad[3 * i] = r.x; rad[3 * i + 1] = r.y; rad[3 * i + 2] = r.z;
  }
}
void ref_envmap_sample(const ref_envmap *e, uint32_t seed, float *out, int n) {
  for (int i = 0; i < n; ++i) {
    Rng r;
    r.seed((uint32_t)i * 0x9E3779B9u, seed);
    V3 d, rad;
    float pdf;
    env_sample(e, r, d, rad, pdf);
    float *o = out + 7 * (size_t)i;
    o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = rad.x; o[4] = rad.y; o[5] = rad.z; o[6] = pdf;
  }
}


float ref_half_round(float x) { return half_round(x); }

// BSDF probes for the oracle's own tests: the material of 'g' at 'color',
// dg.Ng = dg.Ns = Ng. Sample i draws from Rng::seed(i, seed).
void ref_bsdf_sample(const ref_geom *g, const float *color, const float *Ng, const float *wo, int inMedium,
                     uint32_t seed, int n, float *out) {
  HitInfo h;
  h.hit = true;
  material_hit(*g, ld3(color), h);
  for (int i = 0; i < n; ++i) {
    Rng r;
    r.seed((uint32_t)i, seed);
    const Scat sc = bsdf_scatter(h, ld3(Ng), ld3(wo), inMedium != 0, r);
    float *o = out + 10 * i;
    o[0] = sc.dir.x; o[1] = sc.dir.y; o[2] = sc.dir.z;
    o[3] = sc.fr.x; o[4] = sc.fr.y; o[5] = sc.fr.z;
    o[6] = sc.pdf;
    o[7] = (float)sc.type;
    o[8] = sc.offDir;
    o[9] = sc.medium ? 1.f : 0.f;
  }
}

// out: eval value[3], eval pdf, BSDF::pdf
void ref_bsdf_eval(const ref_geom *g, const float *color, const float *Ng, const float *wo, const float *wi,
                   float *out) {
  HitInfo h;
  h.hit = true;
  material_hit(*g, ld3(color), h);
  V3 v;
  out[3] = bsdf_eval(h, ld3(Ng), ld3(wo), ld3(wi), v);
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
  out[4] = bsdf_pdf(h, ld3(Ng), ld3(wo), ld3(wi));
}

// the shared sine / cosine / x^5 (fmath.h), for their own pins
void ref_fmath_eval(const float *x, float *sinOut, float *cosOut, float *pow5Out, int n) {
  for (int i = 0; i < n; ++i) {
    bnf::sincos(x[i], sinOut[i], cosOut[i]);
    pow5Out[i] = bnf::pow5(x[i]);
  }
}

// the shared arc cosine and two-argument arctangent (fmath.h)
void ref_fmath_eval_inv(const float *y, const float *x, float *atan2Out, float *acosOut, int n) {
  for (int i = 0; i < n; ++i) {
    atan2Out[i] = bnf::atan2(y[i], x[i]);
    acosOut[i] = bnf::acos(x[i]);
  }
}

// Volume.h:174-184
float ref_fast_log(float f) {
  f = (f - 1.f) / (f + 1.f);
  const float f2 = f * f;
  float s = f;
  f *= f2;
  s += (1.f / 3.f) * f;
  f *= f2;
  s += (1.f / 5.f) * f;
  return s + s;
}

ref_scene *ref_scene_create(const ref_geom *geoms, int numGeoms) {
  auto *s = new ref_scene_s;
  s->geoms.assign(geoms, geoms + numGeoms);
  for (int g = 0; g < numGeoms; ++g) {
    const ref_geom &gd = geoms[g];
    const bool glass = gd.kind == 0 && gd.matType == 1 && gd.ior != 1.f && gd.transmission >= 1e-3f;
    s->shadowTransparent.push_back(glass);
    // Lambertian alpha = baseColor.w * opacity (AnariMatte.h:63); NVisii
    // alpha = (1 - transmission) * baseColor.w * opacity as a half
    // (AnariPBR.h:88-91, NVisii.h:1080); glass: the flag above
    float a = 1.f;
    if (gd.kind == 0 && !glass)
      a = gd.matType == 1 ? half_round(((1.f - gd.transmission) * 1.f) * gd.opacity) : 1.f * gd.opacity;
    s->alpha.push_back(a);
    s->cutouts = s->cutouts || a < 1.f;
  }
  uint32_t key = 0;
  for (int g = 0; g < numGeoms; ++g) {
    const ref_geom &gd = geoms[g];
    for (int p = 0; p < gd.numPrims; ++p, ++key) {
      Prim pr;
      pr.geom = g;
      pr.prim = p;
      pr.key = key;
      V3 lo, hi;
      if (gd.kind == 0) {
        const int *ix = gd.indices + 3 * p;
        const V3 a = ld3(gd.vertices + 3 * ix[0]), b = ld3(gd.vertices + 3 * ix[1]), c = ld3(gd.vertices + 3 * ix[2]);
        lo = vmin(a, vmin(b, c));
        hi = vmax(a, vmax(b, c));
      } else if (gd.kind == 1) {
        const V3 c = ld3(gd.vertices + 3 * p);
        const float r = gd.radii ? gd.radii[p] : gd.defaultRadius;
        lo = sub(c, mk(r, r, r));
        hi = add(c, mk(r, r, r));
      } else {
        const int *ix = gd.indices + 2 * p;
        const V3 a = ld3(gd.vertices + 3 * ix[0]), b = ld3(gd.vertices + 3 * ix[1]);
        const float r = gd.radii[p];
        lo = vmin(sub(a, mk(r, r, r)), sub(b, mk(r, r, r)));
        hi = vmax(add(a, mk(r, r, r)), add(b, mk(r, r, r)));
      }
      pr.lo[0] = lo.x; pr.lo[1] = lo.y; pr.lo[2] = lo.z;
      pr.hi[0] = hi.x; pr.hi[1] = hi.y; pr.hi[2] = hi.z;
      s->prims.push_back(pr);
    }
  }
  build_bvh(s);
  return s;
}

void ref_scene_destroy(ref_scene *s) { delete s; }

// Closest hit (shadow = 0) or occlusion (shadow = 1: glass skipped, hit =
// an accepted occluder) with the renderer's candidate filter: p's cut plane
// and stochastic opacity drawn from seeds[i] (stochastic = 0: opacity
// ignored, as for the ID channel).
void ref_trace_filtered(const ref_scene *s, const ref_render_params *p, const float *rays, const uint64_t *seeds,
                        ref_hit *hits, int n, int brute, int shadow, int stochastic, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * i;
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    const Filt f = make_filt(p, org, dir, seeds[i], stochastic != 0);
    const Best b = closest(s, org, dir, r[3], r[7], brute != 0, nullptr, shadow != 0, &f);
    ref_hit &h = hits[i];
    h.t = b.t; h.u = b.u; h.v = b.v;
    if (b.geom < 0) {
      h.primID = h.geomID = h.instID = h.instIndex = -1;
      h.N[0] = h.N[1] = h.N[2] = 0.f;
    } else {
      V3 P, N;
      hit_point_normal(s, b, org, dir, r[3], r[7], P, N);
      h.primID = b.prim; h.geomID = b.geom; h.instID = 0; h.instIndex = 0;
      h.N[0] = N.x; h.N[1] = N.y; h.N[2] = N.z;
    }
  });
}

void ref_trace(const ref_scene *s, const float *rays, ref_hit *hits, int n, int brute, int threads, long long *primTests) {
  std::atomic<long long> tests(0);
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * i;
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    long long tc = 0;
    const Best b = closest(s, org, dir, r[3], r[7], brute != 0, &tc);
    tests += tc;
    ref_hit &h = hits[i];
    h.t = b.t; h.u = b.u; h.v = b.v;
    if (b.geom < 0) {
      h.primID = h.geomID = h.instID = h.instIndex = -1;
      h.N[0] = h.N[1] = h.N[2] = 0.f;
    } else {
      V3 P, N;
      hit_point_normal(s, b, org, dir, r[3], r[7], P, N);
      h.primID = b.prim; h.geomID = b.geom; h.instID = 0; h.instIndex = 0;
      h.N[0] = N.x; h.N[1] = N.y; h.N[2] = N.z;
    }
  });
  if (primTests) *primTests = tests.load();
}

double ref_triangle_t_f64(const double *o, const double *d, const double *a, const double *b, const double *c) {
  // Moller-Trumbore in double precision
  double e1[3], e2[3], p[3], q[3], s[3];
  for (int k = 0; k < 3; ++k) { e1[k] = b[k] - a[k]; e2[k] = c[k] - a[k]; s[k] = o[k] - a[k]; }
  p[0] = d[1] * e2[2] - d[2] * e2[1]; p[1] = d[2] * e2[0] - d[0] * e2[2]; p[2] = d[0] * e2[1] - d[1] * e2[0];
  const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (std::fabs(det) < 1e-300) return -1.0;
  const double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) / det;
  if (u < 0.0 || u > 1.0) return -1.0;
  q[0] = s[1] * e1[2] - s[2] * e1[1]; q[1] = s[2] * e1[0] - s[0] * e1[2]; q[2] = s[0] * e1[1] - s[1] * e1[0];
  const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
  if (v < 0.0 || u + v > 1.0) return -1.0;
  return (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
}

void ref_camera_ray(const ref_camera *cam, const ref_render_params *p, int ix, int iy, float *org, float *dir,
                    uint64_t *seed) {
  const Cam c = make_cam(cam);
  const PRay r = cam_ray(c, p, ix, iy, p->accumID);
  org[0] = r.org.x; org[1] = r.org.y; org[2] = r.org.z;
  dir[0] = r.dir.x; dir[1] = r.dir.y; dir[2] = r.dir.z;
  *seed = r.rng.s;
}

}  // extern "C"

struct ref_volume_s;

namespace {

// the structured volume's MCVolumeAccel::isProg (defined with the volumes)
bool vol_isect_fwd(const ref_volume_s *v, V3 org, V3 dir, float tmin, float tmax, float rayTMax, Rng &seed,
                   long long &samples, float &t, V3 &rgb);

// The surface path tracer; with a volume (a model holding geometry and a
// structured volume), each ray's closest surface hit is found first and the
// volume is then delta-tracked up to it (MCVolumeAccel::isProg with the
// surface hit as the ray's tMax): a collision before the surface is a volume
// hit. Shadow rays not blocked by a surface are tracked through the volume.
long long render_surface_g(const ref_scene *s, const ref_volume_s *vol, const ref_camera *cam,
                           const ref_render_params *p, float *accum, int *ids, int y0, int y1, int threads,
                           long long *samples) {
  const Cam c = make_cam(cam);
  std::atomic<long long> nsamples(0);
  const int W = p->width;
  const int rows = y1 - y0;
  std::atomic<long long> nrays(0);
  const int maxDepth = p->maxPathDepth > 0 ? p->maxPathDepth : 10;
  const int maxDiff = p->maxDiffuseBounces >= 0 ? p->maxDiffuseBounces : 1;
  parallel_for(rows * W, default_threads(threads), [&](int li) {
    const int ix = li % W, iy = y0 + li / W;
    const int pix = ix + iy * W;
    float acc[4] = {0, 0, 0, 0};
    long long nr = 0, vcnt = 0;
    for (int pp = 0; pp < p->pathsPerPixel; ++pp) {
      const int a = p->accumID + pp;
      PRay ray = cam_ray(c, p, ix, iy, a);
      const float cmax = 10.f * (1 + a);
      V3 sum = mk(0, 0, 0);
      float alpha = 0.f;
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const V3 d = fix_dir(ray.dir);
        const Filt f = make_filt(p, ray.org, d, ray.rng.s);
        const Best b0 = closest(s, ray.org, d, 0.f, ray.tMax, false, nullptr, false, &f);
        const Best &b = b0;
        ++nr;
        HitInfo hi;
        hi.hit = b.geom >= 0;
        if (hi.hit) {
          hit_point_normal(s, b, ray.org, d, 0.f, ray.tMax, hi.P, hi.N);
          hi.t = b.t;
          const ref_geom &g = s->geoms[b.geom];
          const V3 col = g.colors ? ld3(g.colors + 3 * b.prim) : ld3(g.color);
          material_hit(g, col, hi);
        }
        if (vol) {
          const float tS = hi.hit ? hi.t : ray.tMax;
          float vt;
          V3 rgb;
          if (vol_isect_fwd(vol, ray.org, d, 0.f, tS, tS, ray.rng, vcnt, vt, rgb)) {
            hi.hit = true;
            hi.P = add(ray.org, scl(vt, d));
            hi.N = mk(0, 0, 0);
            hi.t = vt;
            hi.type = 2;
            hi.albedo = scl(.7f, rgb);  // Phase(color), Phase.h:137-141
          }
        }
        if (gen == 0) {
          alpha = hi.hit ? 1.f : ray.missA;
          if (ids && pp == 0 && p->accumID == 0) {
            int *rec = ids + 4 * pix;
            // the closest candidate whatever its opacity (Triangles.dev.cu:
            // 45-60 writes the IDs before the opacity test)
            Best b = b0;
            if (s->cutouts) {
              const Filt fid = make_filt(p, ray.org, d, ray.rng.s, false);
              b = closest(s, ray.org, d, 0.f, ray.tMax, false, nullptr, false, &fid);
            }
            if (b.geom >= 0) {
              const ref_geom &g = s->geoms[b.geom];
              rec[0] = b.prim;
              rec[1] = g.kind == 1 ? 13 : 0;
              rec[2] = g.userID;
              std::memcpy(&rec[3], &b.t, 4);
            } else {
              rec[0] = rec[1] = rec[2] = -1;
              const float inf = kInf;
              std::memcpy(&rec[3], &inf, 4);
            }
          }
        }
        PRay sh;
        add_frag(sum, bounce(p, ray, hi, sh, gen, maxDepth, maxDiff), cmax);
        if (sh.tMax > 0.f) {
          const V3 sd = fix_dir(sh.dir);
          const Filt fs = make_filt(p, sh.org, sd, sh.rng.s);
          const Best bs = closest(s, sh.org, sd, 0.f, sh.tMax, false, nullptr, true, &fs);
          ++nr;
          HitInfo hs;
          hs.hit = bs.geom >= 0;
          if (vol && !hs.hit) {
            float vt;
            V3 rgb;
            hs.hit = vol_isect_fwd(vol, sh.org, sd, 0.f, sh.tMax, sh.tMax, sh.rng, vcnt, vt, rgb);
          }
          PRay dummy;
          add_frag(sum, bounce(p, sh, hs, dummy, gen + 1, maxDepth, maxDiff), cmax);
        }
      }
      acc[0] += sum.x; acc[1] += sum.y; acc[2] += sum.z; acc[3] += alpha;
    }
    float *o = accum + 4 * pix;
    if (p->accumID == 0) {
      o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3];
    } else {
      o[0] += acc[0]; o[1] += acc[1]; o[2] += acc[2]; o[3] += acc[3];
    }
    nrays += nr;
    nsamples += vcnt;
  });
  if (samples) *samples = nsamples.load();
  return nrays.load();
}

}  // namespace

extern "C" {

long long ref_render_surface(const ref_scene *s, const ref_camera *cam, const ref_render_params *p, float *accum,
                             int *ids, int y0, int y1, int threads) {
  return render_surface_g(s, nullptr, cam, p, accum, ids, y0, y1, threads, nullptr);
}

long long ref_render_mixed(const ref_scene *s, const ref_volume *v, const ref_camera *cam,
                           const ref_render_params *p, float *accum, int *ids, int y0, int y1, int threads,
                           long long *samples) {
  return render_surface_g(s, v, cam, p, accum, ids, y0, y1, threads, samples);
}

}  // extern "C"

// ============================================================================
// Structured volumes (barney/volume)
// ============================================================================
struct ref_volume_s {
  std::vector<float> scalars;
  int n[3];
  V3 origin, spacing, wlo, whi;
  int cpm;
  int m[3];
  V3 mcOrigin, mcSpacing;
  std::vector<float> rlo, rhi, maj;
  std::vector<float> tf;  // float4
  int numTF = 0;
  float dlo = 0.f, dhi = 1.f, base = 1.f;
};

namespace {

// min / max into a cell shared between threads (compare-and-swap on the
// float's bits); fmin / fmax are order-independent, so the grid does not
// depend on the thread schedule.
inline void atomic_fmin(float *p, float v) {
  uint32_t *u = reinterpret_cast<uint32_t *>(p);
  uint32_t old = __atomic_load_n(u, __ATOMIC_RELAXED);
  for (;;) {
    float of;
    std::memcpy(&of, &old, 4);
    const float nv = std::fmin(of, v);
    if (nv == of) return;
    uint32_t nu;
    std::memcpy(&nu, &nv, 4);
    if (__atomic_compare_exchange_n(u, &old, nu, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
  }
}
inline void atomic_fmax(float *p, float v) {
  uint32_t *u = reinterpret_cast<uint32_t *>(p);
  uint32_t old = __atomic_load_n(u, __ATOMIC_RELAXED);
  for (;;) {
    float of;
    std::memcpy(&of, &old, 4);
    const float nv = std::fmax(of, v);
    if (nv == of) return;
    uint32_t nu;
    std::memcpy(&nu, &nv, 4);
    if (__atomic_compare_exchange_n(u, &old, nu, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
  }
}


// embree Texture.cpp:326-335
void lcoords(int N, float tc, int &i0, int &i1, float &f) {
  if (tc - .5f <= 0.f) { i0 = i1 = 0; f = 0.f; return; }
  if (tc - .5f >= N - 1) { i0 = i1 = N - 1; f = 0.f; return; }
  f = (tc - .5f) - int(tc - .5f);
  i0 = int(tc - .5f);
  i1 = i0 + 1;
}
inline float lerpl(float f, float a, float b) { return (1.f - f) * a + f * b; }

// StructuredData.h:104-115 + embree Texture.cpp:338-409
float vsample(const ref_volume_s *v, V3 P) {
  const V3 rel = mul(sub(P, v->origin), inv(v->spacing));
  if (rel.x < 0.f || rel.y < 0.f || rel.z < 0.f) return NAN;
  if (rel.x >= v->n[0] - 1 || rel.y >= v->n[1] - 1 || rel.z >= v->n[2] - 1) return NAN;
  int x0, x1, y0, y1, z0, z1;
  float fx, fy, fz;
  lcoords(v->n[0], rel.x + .5f, x0, x1, fx);
  lcoords(v->n[1], rel.y + .5f, y0, y1, fy);
  lcoords(v->n[2], rel.z + .5f, z0, z1, fz);
  auto at = [&](int x, int y, int z) { return v->scalars[(size_t)x + (size_t)v->n[0] * ((size_t)y + (size_t)v->n[1] * z)]; };
  const float a = lerpl(fx, at(x0, y0, z0), at(x1, y0, z0));
  const float b = lerpl(fx, at(x0, y1, z0), at(x1, y1, z0));
  const float c = lerpl(fx, at(x0, y0, z1), at(x1, y0, z1));
  const float d = lerpl(fx, at(x0, y1, z1), at(x1, y1, z1));
  return lerpl(fz, lerpl(fy, a, b), lerpl(fy, c, d));
}

// TransferFunction.h:67-86
void tfmap(const ref_volume_s *v, float s, float *o) {
  float f = (s - v->dlo) / (v->dhi - v->dlo);
  f = std::min(std::max(f, 0.f), 1.f);
  f *= (v->numTF - 1);
  const int i = std::min(std::max(int(f), 0), v->numTF - 2);
  f -= i;
  for (int k = 0; k < 4; ++k) o[k] = (1.f - f) * v->tf[4 * i + k] + f * v->tf[4 * (i + 1) + k];
  o[3] *= v->base;
}

// TransferFunction.h:92-116
float tfmaj(const ref_volume_s *v, float lo, float hi) {
  if (lo > hi) return 0.f;
  const float span = v->dhi - v->dlo;
  float a = std::min(std::max((lo - v->dlo) / span, 0.f), 1.f) * (v->numTF - 1);
  float b = std::min(std::max((hi - v->dlo) / span, 0.f), 1.f) * (v->numTF - 1);
  const int ia = std::min(std::max(int(a), 0), v->numTF - 2);
  const int ib = std::min(std::max(int(b), 0), v->numTF - 2);
  a -= ia;
  b -= ib;
  const float la = (1.f - a) * v->tf[4 * ia + 3] + a * v->tf[4 * (ia + 1) + 3];
  const float lb = (1.f - b) * v->tf[4 * ib + 3] + b * v->tf[4 * (ib + 1) + 3];
  float m = std::fmax(std::fmax(0.f, la), lb);
  for (int i = ia + 1; i <= ib; ++i) m = std::fmax(m, v->tf[4 * i + 3]);
  return m * v->base;
}

// DDA.h:65-206 (callback returns false to stop)
template <typename F>
void dda3(V3 org, V3 dir, float tMax, const int *g, F &&cb) {
  const V3 fs = mk((float)g[0], (float)g[1], (float)g[2]);
  const V3 rd = inv(dir);
  V3 tlo = mul(sub(mk(0, 0, 0), org), inv(dir)), thi = mul(sub(fs, org), inv(dir));
  V3 tnr = vmin(tlo, thi), tfr = vmax(tlo, thi);
  const float *o = &org.x, *d = &dir.x, *fsz = &fs.x;
  float *pn = &tnr.x, *pf = &tfr.x;
  for (int k = 0; k < 3; ++k)
    if (d[k] == 0.f) {
      if (o[k] < 0.f || o[k] > fsz[k]) return;
      pn[k] = -kInf;
      pf[k] = kInf;
    }
  const float t0 = std::fmax(0.f, maxc(tnr));
  const float t1 = std::fmin(tMax, minc(tfr));
  if (t0 > t1) return;
  const V3 oiv = add(org, scl(t0, dir));
  V3 fc = vmax(mk(0, 0, 0), vmin(sub(fs, mk(1, 1, 1)), mk(std::floor(oiv.x), std::floor(oiv.y), std::floor(oiv.z))));
  const float fce[3] = {d[0] > 0.f ? fc.x + 1.f : fc.x, d[1] > 0.f ? fc.y + 1.f : fc.y, d[2] > 0.f ? fc.z + 1.f : fc.z};
  const V3 ts = vabs(rd);
  const float *pts = &ts.x, *po = &oiv.x;
  float tn[3];
  int stop[3], delta[3], cell[3] = {(int)fc.x, (int)fc.y, (int)fc.z};
  for (int k = 0; k < 3; ++k) {
    tn[k] = d[k] == 0.f ? kInf : std::fabs(fce[k] - po[k]) * pts[k];
    stop[k] = d[k] > 0.f ? g[k] : -1;
    delta[k] = d[k] > 0.f ? 1 : -1;
  }
  float ncb = 0.f;
  while (true) {
    const float tc = std::fmin(tn[0], std::fmin(tn[1], tn[2]));
    const float c0 = t0 + ncb;
    const float c1 = t0 + std::fmin(tc, tMax);
    if (!cb(cell, c0, c1)) return;
    ncb = tc;
    for (int k = 0; k < 3; ++k)
      if (tn[k] == tc) {
        tn[k] += pts[k];
        cell[k] += delta[k];
        if (cell[k] == stop[k]) return;
      }
  }
}

struct VHit {
  bool hit;
  float t;
  float rgb[3];
};

// MCAccelerator.h:569-708 + Volume.h:188-223. 'v' supplies the macrocell
// grid and transfer function, field(P) the raw scalar (NaN outside).
template <class Field>
VHit vol_isect_g(const ref_volume_s *v, const Field &field, V3 org, V3 dir, float tmin, float tmax, float rayTMax,
                 Rng &seed, bool exactLog, long long &samples, uint64_t geomHash) {
  VHit res;
  res.hit = false;
  float t0 = tmin, t1 = tmax;
  {
    const V3 a = mul(sub(v->wlo, org), inv(dir)), b = mul(sub(v->whi, org), inv(dir));
    t0 = std::fmax(t0, maxc(vmin(a, b)));
    t1 = std::fmin(t1, minc(vmax(a, b)));
    if (!(t0 <= t1)) return res;
  }
  const V3 dorg = mul(sub(org, v->mcOrigin), inv(v->mcSpacing));
  const V3 ddir = mul(dir, inv(v->mcSpacing));
  Rng rng = child(seed, geomHash);
  dda3(dorg, ddir, t1, v->m, [&](const int *c, float c0, float c1) -> bool {
    const float maj = v->maj[(size_t)c[0] + (size_t)v->m[0] * ((size_t)c[1] + (size_t)v->m[1] * c[2])];
    if (maj == 0.f) return true;
    const float upper = std::fmin(c1, rayTMax);
    float t = c0;
    while (true) {
      const float r = rng.uni();
      const float dt = exactLog ? -std::log(1.f - r) / maj : -ref_fast_log(1.f - r) / maj;
      t += dt;
      if (t >= upper) return true;
      const V3 P = add(org, scl(t, dir));
      const float f = field(P);
      float s[4] = {0, 0, 0, 0};
      if (!std::isnan(f)) tfmap(v, f, s);
      ++samples;
      if (s[3] >= rng.uni() * maj) {
        res.hit = true;
        res.t = t;
        res.rgb[0] = s[0]; res.rgb[1] = s[1]; res.rgb[2] = s[2];
        return false;
      }
    }
  });
  return res;
}

VHit vol_isect(const ref_volume_s *v, V3 org, V3 dir, float tmin, float tmax, float rayTMax, Rng &seed, bool exactLog,
               long long &samples) {
  return vol_isect_g(v, [&](V3 P) { return vsample(v, P); }, org, dir, tmin, tmax, rayTMax, seed, exactLog, samples,
                     fnv3(0, 0, 0));
}

bool vol_isect_fwd(const ref_volume_s *v, V3 org, V3 dir, float tmin, float tmax, float rayTMax, Rng &seed,
                   long long &samples, float &t, V3 &rgb) {
  const VHit h = vol_isect(v, org, dir, tmin, tmax, rayTMax, seed, false, samples);
  t = h.t;
  rgb = mk(h.rgb[0], h.rgb[1], h.rgb[2]);
  return h.hit;
}

}  // namespace

extern "C" {

ref_volume *ref_volume_create(const float *scalars, const int *dims, const float *origin, const float *spacing,
                              int cpm) {
  auto *v = new ref_volume_s;
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  v->scalars.assign(scalars, scalars + n);
  for (int k = 0; k < 3; ++k) v->n[k] = dims[k];
  v->origin = ld3(origin);
  v->spacing = ld3(spacing);
  v->wlo = v->origin;
  v->whi = add(v->origin, mul(v->spacing, mk((float)(dims[0] - 1), (float)(dims[1] - 1), (float)(dims[2] - 1))));
  v->cpm = cpm;
  for (int k = 0; k < 3; ++k) v->m[k] = (dims[k] - 1 + cpm - 1) / cpm;
  v->mcOrigin = v->wlo;
  v->mcSpacing = mul(mk((float)cpm, (float)cpm, (float)cpm), v->spacing);
  const size_t nm = (size_t)v->m[0] * v->m[1] * v->m[2];
  v->rlo.assign(nm, kInf);
  v->rhi.assign(nm, -kInf);
  v->maj.assign(nm, 0.f);
  // StructuredData.cu:30-61: min/max over the (cpm+1)^3 voxels of each MC
  for (int mz = 0; mz < v->m[2]; ++mz)
    for (int my = 0; my < v->m[1]; ++my)
      for (int mx = 0; mx < v->m[0]; ++mx) {
        float lo = kInf, hi = -kInf;
        for (int iz = 0; iz <= cpm; ++iz)
          for (int iy = 0; iy <= cpm; ++iy)
            for (int ix = 0; ix <= cpm; ++ix) {
              const int x = mx * cpm + ix, y = my * cpm + iy, z = mz * cpm + iz;
              if (x >= dims[0] || y >= dims[1] || z >= dims[2]) continue;
              const float f = scalars[(size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * z)];
              lo = std::min(lo, f);
              hi = std::max(hi, f);
            }
        const size_t i = (size_t)mx + (size_t)v->m[0] * ((size_t)my + (size_t)v->m[1] * mz);
        v->rlo[i] = lo;
        v->rhi[i] = hi;
      }
  return v;
}
void ref_volume_destroy(ref_volume *v) { delete v; }
void ref_volume_set_tf(ref_volume *v, const float *values, int numValues, float lo, float hi, float base) {
  v->tf.assign(values, values + 4 * numValues);
  v->numTF = numValues;
  v->dlo = lo;
  v->dhi = hi;
  v->base = base;
  for (size_t i = 0; i < v->maj.size(); ++i) v->maj[i] = tfmaj(v, v->rlo[i], v->rhi[i]);
}
void ref_volume_mc_dims(const ref_volume *v, int *dims) {
  for (int k = 0; k < 3; ++k) dims[k] = v->m[k];
}
void ref_volume_get_mcs(const ref_volume *v, float *ranges, float *maj) {
  for (size_t i = 0; i < v->maj.size(); ++i) {
    if (ranges) { ranges[2 * i] = v->rlo[i]; ranges[2 * i + 1] = v->rhi[i]; }
    if (maj) maj[i] = v->maj[i];
  }
}
float ref_volume_sample(const ref_volume *v, const float *P) { return vsample(v, ld3(P)); }
void ref_tf_map(const ref_volume *v, float s, float *o) { tfmap(v, s, o); }
float ref_tf_majorant(const ref_volume *v, float lo, float hi) { return tfmaj(v, lo, hi); }

void ref_volume_sample_rays(const ref_volume *v, const float *rays, const uint64_t *seeds, float *out, int n,
                            int exactLog, int threads, long long *samples) {
  std::atomic<long long> ns(0);
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * i;
    Rng seed;
    seed.s = seeds[i];
    long long cnt = 0;
    const VHit h = vol_isect(v, ld3(r), fix_dir(ld3(r + 4)), r[3], r[7], r[7], seed, exactLog != 0, cnt);
    ns += cnt;
    float *o = out + 4 * i;
    if (h.hit) { o[0] = h.t; o[1] = h.rgb[0]; o[2] = h.rgb[1]; o[3] = h.rgb[2]; }
    else { o[0] = kInf; o[1] = o[2] = o[3] = 0.f; }
  });
  if (samples) *samples = ns.load();
}

int ref_dda_cells(const float *org, const float *dir, float tMax, const int *g, int *cells, float *ts, int maxCells) {
  int n = 0;
  dda3(ld3(org), ld3(dir), tMax, g, [&](const int *c, float c0, float c1) -> bool {
    if (n < maxCells) {
      cells[3 * n] = c[0]; cells[3 * n + 1] = c[1]; cells[3 * n + 2] = c[2];
      ts[2 * n] = c0; ts[2 * n + 1] = c1;
    }
    ++n;
    return n < 100000;
  });
  return n;
}

}  // extern "C"

namespace {

// The delta-tracking render of a field over v's macrocell grid and transfer
// function (generate / trace / shade with the SCI_VIS_MODE bounce).
template <class Field>
long long render_volume_g(const ref_volume_s *v, const Field &field, const ref_camera *cam,
                          const ref_render_params *p, float *accum, int y0, int y1, int threads, long long *samples) {
  auto vol_isect = [&](const ref_volume_s *vv, V3 org, V3 dir, float tmin, float tmax, float rayTMax, Rng &seed,
                       bool exactLog, long long &cnt) {
    return vol_isect_g(vv, field, org, dir, tmin, tmax, rayTMax, seed, exactLog, cnt, fnv3(0, 0, 0));
  };
  const Cam c = make_cam(cam);
  const int W = p->width;
  std::atomic<long long> nrays(0), ns(0);
  parallel_for((y1 - y0) * W, default_threads(threads), [&](int li) {
    const int ix = li % W, iy = y0 + li / W, pix = ix + iy * W;
    float acc[4] = {0, 0, 0, 0};
    long long nr = 0, cnt = 0;
    for (int pp = 0; pp < p->pathsPerPixel; ++pp) {
      const int a = p->accumID + pp;
      PRay ray = cam_ray(c, p, ix, iy, a);
      const float cmax = 10.f * (1 + a);
      V3 sum = mk(0, 0, 0);
      float alpha = 0.f;
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const V3 d = fix_dir(ray.dir);
        const VHit vh = vol_isect(v, ray.org, d, 0.f, ray.tMax, ray.tMax, ray.rng, false, cnt);
        ++nr;
        HitInfo hi;
        hi.hit = vh.hit;
        if (vh.hit) {
          hi.P = add(ray.org, scl(vh.t, d));
          hi.N = mk(0, 0, 0);
          hi.t = vh.t;
          hi.type = 2;
          hi.albedo = scl(.7f, mk(vh.rgb[0], vh.rgb[1], vh.rgb[2]));
        }
        if (gen == 0) alpha = vh.hit ? 1.f : ray.missA;
        PRay sh;
        add_frag(sum, bounce(p, ray, hi, sh, gen, 10, 1), cmax);
        if (sh.tMax > 0.f) {
          const VHit sv = vol_isect(v, sh.org, fix_dir(sh.dir), 0.f, sh.tMax, sh.tMax, sh.rng, false, cnt);
          ++nr;
          HitInfo hs;
          hs.hit = sv.hit;
          PRay dummy;
          add_frag(sum, bounce(p, sh, hs, dummy, gen + 1, 10, 1), cmax);
        }
      }
      acc[0] += sum.x; acc[1] += sum.y; acc[2] += sum.z; acc[3] += alpha;
    }
    float *o = accum + 4 * pix;
    if (p->accumID == 0) { o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3]; }
    else { o[0] += acc[0]; o[1] += acc[1]; o[2] += acc[2]; o[3] += acc[3]; }
    nrays += nr;
    ns += cnt;
  });
  if (samples) *samples = ns.load();
  return nrays.load();
}

}  // namespace

extern "C" {

long long ref_render_volume(const ref_volume *v, const ref_camera *cam, const ref_render_params *p, float *accum, int y0,
                            int y1, int threads, long long *samples) {
  return render_volume_g(v, [&](V3 P) { return vsample(v, P); }, cam, p, accum, y0, y1, threads, samples);
}

}  // extern "C"

// ============================================================================
// Block-structured AMR (barney/amr): BlockStructuredField.h:100-248,
// BlockStructuredField.cu:33-105, BlockStructuredCuBQLSampler.h:66-80
// ============================================================================
struct ref_amr_s {
  ref_volume_s grid;  // macrocell grid, world box and transfer function
  std::vector<float> dom;  // per block: lower.xyz, cellSize, upper.xyz, 0
  std::vector<int> dims, origins;
  std::vector<unsigned long long> offsets;
  std::vector<float> scalars;
  int g[3];
  V3 gOrigin, gRcp;
  std::vector<uint32_t> start, list;  // ascending block ids per lookup cell
};

namespace {

// addBasisFunctions (BlockStructuredField.h:100-200)
void amr_add_block(const ref_amr_s *a, int b, V3 P, float &sumW, float &sumWV) {
  const float *lo = &a->dom[8 * (size_t)b], *hi = lo + 4;
  if (!(lo[0] <= P.x && lo[1] <= P.y && lo[2] <= P.z && P.x <= hi[0] && P.y <= hi[1] && P.z <= hi[2])) return;
  const float cs = lo[3];
  const int *d = &a->dims[3 * (size_t)b];
  const float *s = &a->scalars[a->offsets[b]];
  const float c0[3] = {lo[0] + cs, lo[1] + cs, lo[2] + cs};
  const float lp[3] = {(P.x - c0[0]) / cs, (P.y - c0[1]) / cs, (P.z - c0[2]) / cs};
  float fl[3], f[3], nf[3];
  int l[3], h[3];
  for (int k = 0; k < 3; ++k) {
    fl[k] = std::floor(lp[k]);
    l[k] = std::max(-1, (int)fl[k]);
    h[k] = l[k] + 1;
    f[k] = lp[k] - fl[k];
    nf[k] = 1.f - f[k];
  }
  auto in = [&](int i, int k) { return i >= 0 && i < d[k]; };
  auto add = [&](int x, int y, int z, float w) {
    sumW += w;
    sumWV += w * s[x + y * d[0] + z * d[0] * d[1]];
  };
  for (int zi = 0; zi < 2; ++zi) {
    const int z = zi ? h[2] : l[2];
    if (!in(z, 2)) continue;
    const float wz = zi ? f[2] : nf[2];
    for (int yi = 0; yi < 2; ++yi) {
      const int y = yi ? h[1] : l[1];
      if (!in(y, 1)) continue;
      const float wy = yi ? f[1] : nf[1];
      if (in(l[0], 0)) add(l[0], y, z, wz * wy * nf[0]);
      if (in(h[0], 0)) add(h[0], y, z, wz * wy * f[0]);
    }
  }
}

// BlockStructuredCuBQLSampler::DD::sample with the blocks in ascending order
float amr_sample(const ref_amr_s *a, V3 P) {
  const ref_volume_s &g = a->grid;
  if (!(g.wlo.x <= P.x && g.wlo.y <= P.y && g.wlo.z <= P.z && P.x <= g.whi.x && P.y <= g.whi.y && P.z <= g.whi.z))
    return NAN;
  const int cx = std::min(std::max((int)((P.x - a->gOrigin.x) * a->gRcp.x), 0), a->g[0] - 1);
  const int cy = std::min(std::max((int)((P.y - a->gOrigin.y) * a->gRcp.y), 0), a->g[1] - 1);
  const int cz = std::min(std::max((int)((P.z - a->gOrigin.z) * a->gRcp.z), 0), a->g[2] - 1);
  const size_t c = (size_t)cx + (size_t)a->g[0] * ((size_t)cy + (size_t)a->g[1] * cz);
  float sumW = 0.f, sumWV = 0.f;
  for (uint32_t i = a->start[c]; i < a->start[c + 1]; ++i) amr_add_block(a, (int)a->list[i], P, sumW, sumWV);
  return sumW == 0.f ? NAN : sumWV / sumW;
}

// the same without the lookup grid: every block, ascending
float amr_sample_brute(const ref_amr_s *a, V3 P) {
  float sumW = 0.f, sumWV = 0.f;
  for (size_t b = 0; b < a->offsets.size(); ++b) amr_add_block(a, (int)b, P, sumW, sumWV);
  return sumW == 0.f ? NAN : sumWV / sumW;
}

}  // namespace

extern "C" {

ref_amr *ref_amr_create(const int *origins, const int *dims, const int *levels, const unsigned long long *offsets,
                        int numBlocks, const int *refinements, int numLevels, const float *scalars,
                        unsigned long long numScalars) {
  auto *a = new ref_amr_s;
  a->scalars.assign(scalars, scalars + numScalars);
  a->offsets.assign(offsets, offsets + numBlocks);
  a->dims.assign(dims, dims + 3 * (size_t)numBlocks);
  a->origins.assign(origins, origins + 3 * (size_t)numBlocks);
  a->dom.resize(8 * (size_t)numBlocks);
  V3 wlo = mk(kInf, kInf, kInf), whi = mk(-kInf, -kInf, -kInf);
  for (int b = 0; b < numBlocks; ++b) {
    const int L = levels[b];
    const int *o = origins + 3 * (size_t)b, *d = dims + 3 * (size_t)b;
    if (L < 0 || L >= numLevels || d[0] <= 0 || d[1] <= 0 || d[2] <= 0 ||
        offsets[b] + (unsigned long long)d[0] * d[1] * d[2] > numScalars) {
      delete a;
      return nullptr;
    }
    // Block::getFrom / getDomain (BlockStructuredField.h:227-248)
    const float cs = (float)(int)(std::pow((float)refinements[L], (float)L));
    float *q = &a->dom[8 * (size_t)b];
    for (int k = 0; k < 3; ++k) {
      q[k] = ((float)o[k] - .5f) * cs;
      q[4 + k] = ((float)(o[k] + d[k]) + .5f) * cs;
    }
    q[3] = cs;
    q[7] = 0.f;
    wlo = vmin(wlo, mk(q[0], q[1], q[2]));
    whi = vmax(whi, mk(q[4], q[5], q[6]));
  }
  ref_volume_s &g = a->grid;
  g.wlo = wlo;
  g.whi = whi;
  const V3 size = sub(whi, wlo);
  const float maxWidth = maxc(size);
  {  // buildMCs (BlockStructuredField.cu:82-105)
    const float f = (256 - 1) / maxWidth;
    g.m[0] = 1 + (int)(size.x * f);
    g.m[1] = 1 + (int)(size.y * f);
    g.m[2] = 1 + (int)(size.z * f);
    g.mcOrigin = wlo;
    g.mcSpacing = mul(size, inv(mk((float)g.m[0], (float)g.m[1], (float)g.m[2])));
  }
  {  // block lookup grid
    const int G = std::min(256, std::max(16, 4 * (int)std::ceil(std::cbrt((double)numBlocks))));
    const float f = (G - 1) / maxWidth;
    a->g[0] = 1 + (int)(size.x * f);
    a->g[1] = 1 + (int)(size.y * f);
    a->g[2] = 1 + (int)(size.z * f);
    a->gOrigin = wlo;
    a->gRcp = inv(mul(size, inv(mk((float)a->g[0], (float)a->g[1], (float)a->g[2]))));
    const size_t ng = (size_t)a->g[0] * a->g[1] * a->g[2];
    a->start.assign(ng + 1, 0u);
    const float go[3] = {wlo.x, wlo.y, wlo.z}, gr[3] = {a->gRcp.x, a->gRcp.y, a->gRcp.z};
    std::vector<int> r(6 * (size_t)numBlocks);
    for (int b = 0; b < numBlocks; ++b)
      for (int k = 0; k < 3; ++k) {
        r[6 * (size_t)b + k] = std::min(std::max((int)((a->dom[8 * (size_t)b + k] - go[k]) * gr[k]), 0), a->g[k] - 1);
        r[6 * (size_t)b + 3 + k] =
            std::min(std::max((int)((a->dom[8 * (size_t)b + 4 + k] - go[k]) * gr[k]), 0), a->g[k] - 1);
      }
    auto each = [&](int b, auto &&fn) {
      const int *q = &r[6 * (size_t)b];
      for (int z = q[2]; z <= q[5]; ++z)
        for (int y = q[1]; y <= q[4]; ++y)
          for (int x = q[0]; x <= q[3]; ++x) fn((size_t)x + (size_t)a->g[0] * ((size_t)y + (size_t)a->g[1] * z));
    };
    for (int b = 0; b < numBlocks; ++b) each(b, [&](size_t c) { ++a->start[c + 1]; });
    for (size_t c = 0; c < ng; ++c) a->start[c + 1] += a->start[c];
    a->list.resize(a->start[ng]);
    std::vector<uint32_t> fill(a->start.begin(), a->start.end() - 1);
    for (int b = 0; b < numBlocks; ++b) each(b, [&](size_t c) { a->list[fill[c]++] = (uint32_t)b; });
  }
  // BSField_rasterGrids (BlockStructuredField.cu:33-57) through rasterBox
  const size_t nc = (size_t)g.m[0] * g.m[1] * g.m[2];
  g.rlo.assign(nc, kInf);
  g.rhi.assign(nc, -kInf);
  g.maj.assign(nc, 0.f);
  const V3 rs = inv(g.mcSpacing);
  const float mo[3] = {g.mcOrigin.x, g.mcOrigin.y, g.mcOrigin.z}, mr[3] = {rs.x, rs.y, rs.z};
  parallel_for(numBlocks, default_threads(0), [&](int b) {
    const float cs = a->dom[8 * (size_t)b + 3];
    const int *o = &a->origins[3 * (size_t)b], *d = &a->dims[3 * (size_t)b];
    const float *s = &a->scalars[a->offsets[b]];
    for (int z = 0; z < d[2]; ++z)
      for (int y = 0; y < d[1]; ++y)
        for (int x = 0; x < d[0]; ++x) {
          const int c[3] = {o[0] + x, o[1] + y, o[2] + z};
          int lo[3], hi[3];
          for (int k = 0; k < 3; ++k) {
            const float l = ((float)c[k] - .5f) * cs, h = l + 2.f * cs;
            lo[k] = std::min(std::max((int)((l - mo[k]) * mr[k]), 0), g.m[k] - 1);
            hi[k] = std::min(std::max((int)((h - mo[k]) * mr[k]), 0), g.m[k] - 1);
          }
          const float v = s[x + y * d[0] + z * d[0] * d[1]];
          for (int iz = lo[2]; iz <= hi[2]; ++iz)
            for (int iy = lo[1]; iy <= hi[1]; ++iy)
              for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                const size_t id = (size_t)ix + (size_t)g.m[0] * ((size_t)iy + (size_t)g.m[1] * iz);
                atomic_fmin(&g.rlo[id], v);
                atomic_fmax(&g.rhi[id], v);
              }
        }
  });
  const float one[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  g.tf.assign(one, one + 8);
  g.numTF = 2;
  return a;
}

void ref_amr_destroy(ref_amr *a) { delete a; }

void ref_amr_set_tf(ref_amr *a, const float *values, int numValues, float lo, float hi, float base) {
  ref_volume_s &g = a->grid;
  g.tf.assign(values, values + 4 * (size_t)numValues);
  g.numTF = numValues;
  g.dlo = lo;
  g.dhi = hi;
  g.base = base;
  for (size_t i = 0; i < g.maj.size(); ++i) g.maj[i] = tfmaj(&g, g.rlo[i], g.rhi[i]);
}

void ref_amr_mc_dims(const ref_amr *a, int *dims) {
  for (int k = 0; k < 3; ++k) dims[k] = a->grid.m[k];
}

void ref_amr_get_mcs(const ref_amr *a, float *ranges, float *maj) {
  const ref_volume_s &g = a->grid;
  for (size_t i = 0; i < g.maj.size(); ++i) {
    if (ranges) { ranges[2 * i] = g.rlo[i]; ranges[2 * i + 1] = g.rhi[i]; }
    if (maj) maj[i] = g.maj[i];
  }
}

void ref_amr_sample(const ref_amr *a, const float *points, float *out, int n, int brute, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const V3 P = ld3(points + 3 * (size_t)i);
    out[i] = brute ? amr_sample_brute(a, P) : amr_sample(a, P);
  });
}

long long ref_amr_render(const ref_amr *a, const ref_camera *cam, const ref_render_params *p, float *accum, int y0,
                         int y1, int threads, long long *samples) {
  return render_volume_g(&a->grid, [&](V3 P) { return amr_sample(a, P); }, cam, p, accum, y0, y1, threads, samples);
}

}  // extern "C"

// ============================================================================
// NanoVDB fields (barney/volume/NanoVDB.h:98-127, NanoVDB.cu:28-102,
// 204-235) over the grid buffer of NanoVDB 32.x (the layout of
// 3rdParty/nanovdb/include/nanovdb/NanoVDB.h: GridData 672 B, TreeData 64 B,
// RootData 64 B + 32-byte tiles, upper 32^3 and lower 16^3 internal nodes,
// 8^3 leaves of float or of Fp4/Fp8/Fp16/FpN codes)
// ============================================================================
struct ref_nvdb_s {
  ref_volume_s grid;           // macrocell grid, world box and transfer function
  std::vector<uint64_t> words; // the grid buffer, 8-byte aligned
  const uint8_t *buf = nullptr;
  size_t bytes = 0;
  int type = 0;
  size_t root = 0;             // byte offset of RootData
  int ilo[3], ihi[3];          // index bounding box (the root's)
  float invMatF[9], vecF[3];   // Map::mInvMatF, mVecF
};

namespace {

const size_t kNvUpperSize = 270400, kNvLowerSize = 33856;

template <class T>
inline T nv_ld(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

// leaf byte size by grid type (FpN: from the leaf's flags)
inline size_t nv_leaf_size(int type, const uint8_t *leaf) {
  switch (type) {
    case 1: return 2144;   // float
    case 13: return 352;   // Fp4
    case 14: return 608;   // Fp8
    case 15: return 1120;  // Fp16
    default: return 96 + ((size_t)1 << (leaf[15] >> 5)) * 64;  // FpN
  }
}

// LeafData<T>::getValue for the leaf-local offset n
inline float nv_leaf_value(int type, const uint8_t *leaf, uint32_t n) {
  if (type == 1) return nv_ld<float>(leaf + 96 + 4 * n);
  const float mn = nv_ld<float>(leaf + 80), q = nv_ld<float>(leaf + 84);
  const uint8_t *code = leaf + 96;
  uint32_t c;
  switch (type) {
    case 13: c = (code[n >> 1] >> ((n & 1) << 2)) & 15u; break;
    case 14: c = code[n]; break;
    case 15: c = nv_ld<uint16_t>(code + 2 * n); break;
    default: {
      const int b = leaf[15] >> 5;  // bit width 1 << b
      switch (b) {
        case 0: c = (code[n >> 3] >> (n & 7)) & 1u; break;
        case 1: c = (code[n >> 2] >> ((n & 3) << 1)) & 3u; break;
        case 2: c = (code[n >> 1] >> ((n & 1) << 2)) & 15u; break;
        case 3: c = code[n]; break;
        default: c = nv_ld<uint16_t>(code + 2 * n); break;
      }
    }
  }
  return (float)c * q + mn;
}

// ReadAccessor::getValue: root tile by key, then upper, lower, leaf
float nv_value(const ref_nvdb_s *g, int x, int y, int z) {
  const uint8_t *root = g->buf + g->root;
  const uint64_t key = ((uint64_t)((uint32_t)z >> 12)) | ((uint64_t)((uint32_t)y >> 12) << 21) |
                       ((uint64_t)((uint32_t)x >> 12) << 42);
  const uint32_t nt = nv_ld<uint32_t>(root + 24);
  const uint8_t *tile = nullptr;
  for (uint32_t i = 0; i < nt; ++i)
    if (nv_ld<uint64_t>(root + 64 + 32 * (size_t)i) == key) {
      tile = root + 64 + 32 * (size_t)i;
      break;
    }
  if (!tile) return nv_ld<float>(root + 28);  // background
  const int64_t child = nv_ld<int64_t>(tile + 8);
  if (child == 0) return nv_ld<float>(tile + 20);
  const uint8_t *up = root + child;
  uint32_t n = (((uint32_t)x & 4095u) >> 7 << 10) | (((uint32_t)y & 4095u) >> 7 << 5) | (((uint32_t)z & 4095u) >> 7);
  if (!((nv_ld<uint64_t>(up + 4128 + 8 * (n >> 6)) >> (n & 63)) & 1u)) return nv_ld<float>(up + 8256 + 8 * (size_t)n);
  const uint8_t *lo = up + nv_ld<int64_t>(up + 8256 + 8 * (size_t)n);
  n = (((uint32_t)x & 127u) >> 3 << 8) | (((uint32_t)y & 127u) >> 3 << 4) | (((uint32_t)z & 127u) >> 3);
  if (!((nv_ld<uint64_t>(lo + 544 + 8 * (n >> 6)) >> (n & 63)) & 1u)) return nv_ld<float>(lo + 1088 + 8 * (size_t)n);
  const uint8_t *leaf = lo + nv_ld<int64_t>(lo + 1088 + 8 * (size_t)n);
  return nv_leaf_value(g->type, leaf, ((uint32_t)x & 7u) << 6 | ((uint32_t)y & 7u) << 3 | ((uint32_t)z & 7u));
}

// NanoVDBDataSampler::DD::sample (NanoVDB.h:98-127): Map::applyInverseMapF
// on the double position (double difference, then fp32 fmaf rows), minus
// half a voxel, clamped to the index box, then SampleFromVoxels<1>'s floor,
// stencil and nested lerps a + float(w) * (b - a) (z innermost, then y, x).
float nv_sample(const ref_nvdb_s *g, V3 P) {
  const double d[3] = {(double)P.x - (double)g->vecF[0], (double)P.y - (double)g->vecF[1],
                       (double)P.z - (double)g->vecF[2]};
  const float f[3] = {(float)d[0], (float)d[1], (float)d[2]};
  const float *m = g->invMatF;
  double ip[3] = {(double)std::fma(f[0], m[0], std::fma(f[1], m[1], f[2] * m[2])),
                  (double)std::fma(f[0], m[3], std::fma(f[1], m[4], f[2] * m[5])),
                  (double)std::fma(f[0], m[6], std::fma(f[1], m[7], f[2] * m[8]))};
  int c[3];
  double w[3];
  for (int k = 0; k < 3; ++k) {
    ip[k] -= 0.5;
    ip[k] = std::min(std::max(ip[k], (double)g->ilo[k]), (double)g->ihi[k]);
    const double fl = std::floor(ip[k]);
    c[k] = (int32_t)fl;
    w[k] = ip[k] - fl;
  }
  auto lerp = [](float a, float b, double t) { return a + (float)t * (b - a); };
  const float v000 = nv_value(g, c[0], c[1], c[2]), v001 = nv_value(g, c[0], c[1], c[2] + 1);
  const float v010 = nv_value(g, c[0], c[1] + 1, c[2]), v011 = nv_value(g, c[0], c[1] + 1, c[2] + 1);
  const float v100 = nv_value(g, c[0] + 1, c[1], c[2]), v101 = nv_value(g, c[0] + 1, c[1], c[2] + 1);
  const float v110 = nv_value(g, c[0] + 1, c[1] + 1, c[2]), v111 = nv_value(g, c[0] + 1, c[1] + 1, c[2] + 1);
  return lerp(lerp(lerp(v000, v001, w[2]), lerp(v010, v011, w[2]), w[1]),
              lerp(lerp(v100, v101, w[2]), lerp(v110, v111, w[2]), w[1]), w[0]);
}

// every node reached from the root lies inside the buffer
bool nv_check(const ref_nvdb_s *g) {
  const uint8_t *b = g->buf;
  auto in = [&](const uint8_t *p, size_t n) { return p >= b && (size_t)(p - b) + n <= g->bytes; };
  const uint8_t *root = b + g->root;
  const uint32_t nt = nv_ld<uint32_t>(root + 24);
  if (!in(root, 64 + 32 * (size_t)nt)) return false;
  for (uint32_t i = 0; i < nt; ++i) {
    const int64_t ch = nv_ld<int64_t>(root + 64 + 32 * (size_t)i + 8);
    if (!ch) continue;
    const uint8_t *up = root + ch;
    if (!in(up, kNvUpperSize)) return false;
    for (uint32_t n = 0; n < 32768; ++n) {
      if (!((nv_ld<uint64_t>(up + 4128 + 8 * (n >> 6)) >> (n & 63)) & 1u)) continue;
      const uint8_t *lo = up + nv_ld<int64_t>(up + 8256 + 8 * (size_t)n);
      if (!in(lo, kNvLowerSize)) return false;
      for (uint32_t m = 0; m < 4096; ++m) {
        if (!((nv_ld<uint64_t>(lo + 544 + 8 * (m >> 6)) >> (m & 63)) & 1u)) continue;
        const uint8_t *leaf = lo + nv_ld<int64_t>(lo + 1088 + 8 * (size_t)m);
        if (!in(leaf, 96) || !in(leaf, nv_leaf_size(g->type, leaf))) return false;
        if (g->type == 16 && (leaf[15] >> 5) > 4) return false;
      }
    }
  }
  return true;
}

}  // namespace

extern "C" {

ref_nvdb *ref_nvdb_create(const void *data, unsigned long long bytes) {
  if (!data || bytes < 672 + 64 + 64) return nullptr;
  auto *g = new ref_nvdb_s;
  g->words.resize((bytes + 7) / 8);
  std::memcpy(g->words.data(), data, bytes);
  g->buf = reinterpret_cast<const uint8_t *>(g->words.data());
  g->bytes = bytes;
  const uint8_t *b = g->buf;
  const uint64_t magic = nv_ld<uint64_t>(b);
  const uint32_t version = nv_ld<uint32_t>(b + 16);
  g->type = (int)nv_ld<uint32_t>(b + 636);
  const int64_t rootOff = nv_ld<int64_t>(b + 672 + 24);
  const bool typeOk = g->type == 1 || (g->type >= 13 && g->type <= 16);
  if ((magic & 0x00FFFFFFFFFFFFFFull) != 0x004244566f6e614eull || (version >> 21) != 32u || !typeOk ||
      rootOff != 64 || nv_ld<uint64_t>(b + 32) > bytes) {
    delete g;
    return nullptr;
  }
  g->root = 672 + (size_t)rootOff;
  if (!nv_check(g)) {
    delete g;
    return nullptr;
  }
  for (int k = 0; k < 3; ++k) {
    g->ilo[k] = nv_ld<int32_t>(b + g->root + 4 * k);
    g->ihi[k] = nv_ld<int32_t>(b + g->root + 12 + 4 * k);
  }
  std::memcpy(g->invMatF, b + 296 + 36, 36);
  std::memcpy(g->vecF, b + 296 + 72, 12);
  // NanoVDBData::commit: the world box as floats, the grid size
  ref_volume_s &v = g->grid;
  const double *wb = reinterpret_cast<const double *>(b + 560);
  v.wlo = mk((float)wb[0], (float)wb[1], (float)wb[2]);
  v.whi = mk((float)wb[3], (float)wb[4], (float)wb[5]);
  const int gs[3] = {g->ihi[0] - g->ilo[0] + 1, g->ihi[1] - g->ilo[1] + 1, g->ihi[2] - g->ilo[2] + 1};
  // buildMCs (NanoVDB.cu:82-102): divRoundUp(gridSize, 16) cells over the
  // world box, spacing = size * rcp(dims)
  for (int k = 0; k < 3; ++k) v.m[k] = (gs[k] + 15) / 16;
  v.mcOrigin = v.wlo;
  v.mcSpacing = mul(sub(v.whi, v.wlo), inv(mk((float)v.m[0], (float)v.m[1], (float)v.m[2])));
  // NanoVDBData_computeMCs (NanoVDB.cu:28-62): the range of the voxels
  // lower + id * gridSize / dims .. lower + (id + 1) * gridSize / dims + 1
  const size_t nc = (size_t)v.m[0] * v.m[1] * v.m[2];
  v.rlo.assign(nc, kInf);
  v.rhi.assign(nc, -kInf);
  v.maj.assign(nc, 0.f);
  parallel_for((int)nc, default_threads(0), [&](int tid) {
    const int id[3] = {tid % v.m[0], (tid / v.m[0]) % v.m[1], tid / (v.m[0] * v.m[1])};
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = g->ilo[k] + (id[k] * gs[k]) / v.m[k];
      hi[k] = g->ilo[k] + ((id[k] + 1) * gs[k]) / v.m[k] + 1;
    }
    float rl = kInf, rh = -kInf;
    for (int x = lo[0]; x <= hi[0]; ++x)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int z = lo[2]; z <= hi[2]; ++z) {
          const float s = nv_value(g, x, y, z);
          rl = std::fmin(rl, s);
          rh = std::fmax(rh, s);
        }
    v.rlo[tid] = rl;
    v.rhi[tid] = rh;
  });
  const float one[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  v.tf.assign(one, one + 8);
  v.numTF = 2;
  return g;
}

void ref_nvdb_destroy(ref_nvdb *g) { delete g; }

void ref_nvdb_set_tf(ref_nvdb *g, const float *values, int numValues, float lo, float hi, float base) {
  ref_volume_s &v = g->grid;
  v.tf.assign(values, values + 4 * (size_t)numValues);
  v.numTF = numValues;
  v.dlo = lo;
  v.dhi = hi;
  v.base = base;
  for (size_t i = 0; i < v.maj.size(); ++i) v.maj[i] = tfmaj(&v, v.rlo[i], v.rhi[i]);
}

void ref_nvdb_mc_dims(const ref_nvdb *g, int *dims) {
  for (int k = 0; k < 3; ++k) dims[k] = g->grid.m[k];
}

void ref_nvdb_get_mcs(const ref_nvdb *g, float *ranges, float *maj) {
  const ref_volume_s &v = g->grid;
  for (size_t i = 0; i < v.maj.size(); ++i) {
    if (ranges) { ranges[2 * i] = v.rlo[i]; ranges[2 * i + 1] = v.rhi[i]; }
    if (maj) maj[i] = v.maj[i];
  }
}

void ref_nvdb_boxes(const ref_nvdb *g, float *worldLoHi, int *indexLoHi, float *mcOriginSpacing) {
  const ref_volume_s &v = g->grid;
  const float w[6] = {v.wlo.x, v.wlo.y, v.wlo.z, v.whi.x, v.whi.y, v.whi.z};
  std::memcpy(worldLoHi, w, sizeof(w));
  for (int k = 0; k < 3; ++k) {
    indexLoHi[k] = g->ilo[k];
    indexLoHi[3 + k] = g->ihi[k];
  }
  const float m[6] = {v.mcOrigin.x, v.mcOrigin.y, v.mcOrigin.z, v.mcSpacing.x, v.mcSpacing.y, v.mcSpacing.z};
  std::memcpy(mcOriginSpacing, m, sizeof(m));
}

void ref_nvdb_sample(const ref_nvdb *g, const float *points, float *out, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) { out[i] = nv_sample(g, ld3(points + 3 * (size_t)i)); });
}

void ref_nvdb_values(const ref_nvdb *g, const int *coords, float *out, int n) {
  for (int i = 0; i < n; ++i) out[i] = nv_value(g, coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
}

long long ref_nvdb_render(const ref_nvdb *g, const ref_camera *cam, const ref_render_params *p, float *accum, int y0,
                          int y1, int threads, long long *samples) {
  return render_volume_g(&g->grid, [&](V3 P) { return nv_sample(g, P); }, cam, p, accum, y0, y1, threads, samples);
}

}  // extern "C"

// ============================================================================
// Unstructured tetrahedral volumes (barney/umesh)
// ============================================================================
struct ref_umesh_s {
  std::vector<float> v4;  // x, y, z, scalar per vertex
  std::vector<int> tets;  // 4 vertex indices per tet
  // general cells (UMeshField's layout) when 'cells' is set
  bool cells = false;
  std::vector<int> indices, offsets;
  std::vector<int> kinds;  // 0 tet, 1 pyramid, 2 prism, 3 hex
  ref_scene_s bvh;        // binned-SAH BVH over the element boxes (prim = element)
  ref_volume_s grid;      // macrocell grid + transfer function
};

namespace {

// UMeshField.h:346-352 evalToImplicitPlane + :390-434 tetScalar
float implicit_plane(V3 P, V3 a, V3 b, V3 c) { return dot(sub(P, a), cross(sub(b, a), sub(c, a))); }

bool tet_value_ix(const ref_umesh_s *u, const int *ix, V3 P, float &val) {
  const float *a = &u->v4[4 * (size_t)ix[0]], *b = &u->v4[4 * (size_t)ix[1]];
  const float *c = &u->v4[4 * (size_t)ix[2]], *d = &u->v4[4 * (size_t)ix[3]];
  const V3 v0 = ld3(a), v1 = ld3(b), v2 = ld3(c), v3 = ld3(d);
  const float t3 = implicit_plane(P, v0, v1, v2);
  if (t3 < 0.f) return false;
  const float t2 = implicit_plane(P, v0, v3, v1);
  if (t2 < 0.f) return false;
  const float t1 = implicit_plane(P, v0, v2, v3);
  if (t1 < 0.f) return false;
  const float t0 = implicit_plane(P, v1, v3, v2);
  if (t0 < 0.f) return false;
  const float scale = 1.f / (t0 + t1 + t2 + t3);
  val = scale * (t0 * a[3] + t1 * b[3] + t2 * c[3] + t3 * d[3]);
  return true;
}

bool tet_value(const ref_umesh_s *u, int t, V3 P, float &val) { return tet_value_ix(u, &u->tets[4 * (size_t)t], P, val); }

// ElementIntersection.h (adapted there from OpenVKL): intersectPyrEXT,
// intersectPrismEXT, intersectHexEXT. Newton iteration for the parametric
// coordinates from (.5, .5, .5): at most 10 steps; converged when all three
// updates are below 1e-4; diverged when a coordinate exceeds 1e6 in
// magnitude; |det| below 1e-6 (hex: 1e-10 x the diagonal of the vertices'
// box) rejects; inside within 1e-6 of [0, 1] (prism: also r + s); the value
// uses the shape functions evaluated at the start of the last step.
void pyr_shape(const float *pc, float *sf, float *d) {
  const float rm = 1.f - pc[0], sm = 1.f - pc[1], tm = 1.f - pc[2];
  sf[0] = rm * sm * tm;
  sf[1] = pc[0] * sm * tm;
  sf[2] = pc[0] * pc[1] * tm;
  sf[3] = rm * pc[1] * tm;
  sf[4] = pc[2];
  d[0] = -(pc[1] - 1.f) * (pc[2] - 1.f);
  d[1] = (pc[1] - 1.f) * (pc[2] - 1.f);
  d[2] = pc[1] - pc[1] * pc[2];
  d[3] = pc[1] * (pc[2] - 1.f);
  d[4] = 0.f;
  d[5] = -(pc[0] - 1.f) * (pc[2] - 1.f);
  d[6] = pc[0] * (pc[2] - 1.f);
  d[7] = pc[0] - pc[0] * pc[2];
  d[8] = (pc[0] - 1.f) * (pc[2] - 1.f);
  d[9] = 0.f;
  d[10] = -(pc[0] - 1.f) * (pc[1] - 1.f);
  d[11] = pc[0] * (pc[1] - 1.f);
  d[12] = -pc[0] * pc[1];
  d[13] = (pc[0] - 1.f) * pc[1];
  d[14] = 1.f;
}

void prism_shape(const float *pc, float *sf, float *d) {
  sf[0] = (1.f - pc[0] - pc[1]) * (1.f - pc[2]);
  sf[1] = pc[0] * (1.f - pc[2]);
  sf[2] = pc[1] * (1.f - pc[2]);
  sf[3] = (1.f - pc[0] - pc[1]) * pc[2];
  sf[4] = pc[0] * pc[2];
  sf[5] = pc[1] * pc[2];
  d[0] = -1.f + pc[2];
  d[1] = 1.f - pc[2];
  d[2] = 0.f;
  d[3] = -pc[2];
  d[4] = pc[2];
  d[5] = 0.f;
  d[6] = -1.f + pc[2];
  d[7] = 0.f;
  d[8] = 1.f - pc[2];
  d[9] = -pc[2];
  d[10] = 0.f;
  d[11] = pc[2];
  d[12] = -1.f + pc[0] + pc[1];
  d[13] = -pc[0];
  d[14] = -pc[1];
  d[15] = 1.f - pc[0] - pc[1];
  d[16] = pc[0];
  d[17] = pc[1];
}

void hex_shape(const float *pc, float *sf, float *d) {
  const float rm = 1.f - pc[0], sm = 1.f - pc[1], tm = 1.f - pc[2];
  sf[0] = rm * sm * tm;
  sf[1] = pc[0] * sm * tm;
  sf[2] = pc[0] * pc[1] * tm;
  sf[3] = rm * pc[1] * tm;
  sf[4] = rm * sm * pc[2];
  sf[5] = pc[0] * sm * pc[2];
  sf[6] = pc[0] * pc[1] * pc[2];
  sf[7] = rm * pc[1] * pc[2];
  d[0] = -sm * tm;
  d[1] = sm * tm;
  d[2] = pc[1] * tm;
  d[3] = -pc[1] * tm;
  d[4] = -sm * pc[2];
  d[5] = sm * pc[2];
  d[6] = pc[1] * pc[2];
  d[7] = -pc[1] * pc[2];
  d[8] = -rm * tm;
  d[9] = -pc[0] * tm;
  d[10] = pc[0] * tm;
  d[11] = rm * tm;
  d[12] = -rm * pc[2];
  d[13] = -pc[0] * pc[2];
  d[14] = pc[0] * pc[2];
  d[15] = rm * pc[2];
  d[16] = -rm * sm;
  d[17] = -pc[0] * sm;
  d[18] = -pc[0] * pc[1];
  d[19] = -rm * pc[1];
  d[20] = rm * sm;
  d[21] = pc[0] * sm;
  d[22] = pc[0] * pc[1];
  d[23] = rm * pc[1];
}

inline float det3(V3 x, V3 y, V3 z) { return dot(x, cross(y, z)); }

// kind: 1 pyramid (5 vertices), 2 prism (6), 3 hex (8); V[i] -> x, y, z, s
bool newton_value(int kind, const float *const *V, V3 P, float &value) {
  const int N = kind == 1 ? 5 : kind == 2 ? 6 : 8;
  float tol = 1e-6f;
  if (kind == 3) {
    V3 lo = ld3(V[0]), hi = lo;
    for (int i = 1; i < N; ++i) {
      lo = vmin(lo, ld3(V[i]));
      hi = vmax(hi, ld3(V[i]));
    }
    const V3 e = sub(hi, lo);
    tol = 1e-10f * std::sqrt(dot(e, e));
  }
  float pc[3] = {.5f, .5f, .5f};
  float w[8], d[24];
  bool converged = false;
  for (int it = 0; !converged && it < 10; ++it) {
    if (kind == 1) pyr_shape(pc, w, d);
    else if (kind == 2) prism_shape(pc, w, d);
    else hex_shape(pc, w, d);
    V3 fcol = mk(0.f, 0.f, 0.f), rcol = fcol, scol = fcol, tcol = fcol;
    for (int i = 0; i < N; ++i) {
      const V3 pt = ld3(V[i]);
      fcol = add(fcol, scl(w[i], pt));
      rcol = add(rcol, scl(d[i], pt));
      scol = add(scol, scl(d[i + N], pt));
      tcol = add(tcol, scl(d[i + 2 * N], pt));
    }
    fcol = sub(fcol, P);
    const float D = det3(rcol, scol, tcol);
    if (std::fabs(D) < tol) return false;
    const float d0 = det3(fcol, scol, tcol) / D;
    const float d1 = det3(rcol, fcol, tcol) / D;
    const float d2 = det3(rcol, scol, fcol) / D;
    pc[0] = pc[0] - d0;
    pc[1] = pc[1] - d1;
    pc[2] = pc[2] - d2;
    if (std::fabs(d0) < 1e-4f && std::fabs(d1) < 1e-4f && std::fabs(d2) < 1e-4f)
      converged = true;
    else if (std::fabs(pc[0]) > 1e6f || std::fabs(pc[1]) > 1e6f || std::fabs(pc[2]) > 1e6f)
      return false;
  }
  if (!converged) return false;
  const float lo = 0.f - 1e-6f, hi = 1.f + 1e-6f;
  if (!(pc[0] >= lo && pc[0] <= hi && pc[1] >= lo && pc[1] <= hi && pc[2] >= lo && pc[2] <= hi)) return false;
  if (kind == 2 && !(pc[0] + pc[1] <= hi)) return false;
  float val = 0.f;
  for (int i = 0; i < N; ++i) val += w[i] * V[i][3];
  value = val;
  return true;
}

// polyScalar (UMeshField.h:503-620): the face stream at ix (numFaces, then
// per face numPts and its vertex indices). Inside when a ray from P along +x
// crosses the fan-triangulated faces an odd number of times (Moller-Trumbore,
// |det| < 1e-12 skipped; the ray direction's zero components written out);
// the value is the inverse-squared-distance weighted mean over the stream's
// vertices.
bool poly_value(const ref_umesh_s *u, const int *ix, V3 P, float &val) {
  auto vert = [&](int i) { return ld3(&u->v4[4 * (size_t)i]); };
  int pos = 0;
  const int nf = ix[pos++];
  int crossings = 0;
  for (int f = 0; f < nf; ++f) {
    const int np = ix[pos++];
    const V3 a = vert(ix[pos]);
    for (int t = 0; t < np - 2; ++t) {
      const V3 e1 = sub(vert(ix[pos + 1 + t]), a), e2 = sub(vert(ix[pos + 2 + t]), a);
      const V3 h = mk(0.f, -e2.z, e2.y);  // cross((1, 0, 0), e2)
      const float det = dot(e1, h);
      if (std::fabs(det) < 1e-12f) continue;
      const float inv = 1.f / det;
      const V3 s = sub(P, a);
      const float uu = inv * dot(s, h);
      if (uu < 0.f || uu > 1.f) continue;
      const V3 q = cross(s, e1);
      const float vv = inv * q.x;  // dot((1, 0, 0), q)
      if (vv < 0.f || uu + vv > 1.f) continue;
      if (inv * dot(e2, q) > 0.f) ++crossings;
    }
    pos += np;
  }
  if ((crossings & 1) == 0) return false;
  float ws = 0.f, vs = 0.f;
  pos = 1;
  for (int f = 0; f < nf; ++f) {
    const int np = ix[pos++];
    for (int k = 0; k < np; ++k) {
      const float *v = &u->v4[4 * (size_t)ix[pos + k]];
      const V3 d = sub(ld3(v), P);
      const float w = 1.f / std::fmax(dot(d, d), 1e-20f);
      ws += w;
      vs += w * v[3];
    }
    pos += np;
  }
  if (ws <= 0.f) return false;
  val = vs / ws;
  return true;
}

// eltScalar (UMeshField.h:362-388)
bool elem_value(const ref_umesh_s *u, int e, V3 P, float &val) {
  if (!u->cells) return tet_value(u, e, P, val);
  const int *ix = &u->indices[(size_t)u->offsets[e]];
  const int k = u->kinds[e];
  if (k == 0) return tet_value_ix(u, ix, P, val);
  if (k == 4) return poly_value(u, ix, P, val);
  const float *V[8];
  const int N = k == 1 ? 5 : k == 2 ? 6 : 8;
  for (int i = 0; i < N; ++i) V[i] = &u->v4[4 * (size_t)ix[i]];
  return newton_value(k, V, P, val);
}

// UMeshCuBQLSampler.h:51-67: box query with the degenerate box {P, P}; the
// first element whose inside test passes gives the value, NaN if none. For
// mixed meshes, whose inside tests overlap near shared faces, the element
// taken is the one with the smallest cell index whose own box holds P and
// whose inside test passes (a rule independent of the BVH; the reference
// takes the first its cuBQL query meets).
float um_sample(const ref_umesh_s *u, V3 P) {
  const auto &nodes = u->bvh.nodes;
  if (nodes.empty()) return NAN;
  int stack[128];
  int sp = 0;
  stack[sp++] = 0;
  const float p[3] = {P.x, P.y, P.z};
  int bestE = -1;
  float best = NAN;
  while (sp) {
    const auto &n = nodes[stack[--sp]];
    bool in = true;
    for (int k = 0; k < 3; ++k) in = in && p[k] >= n.lo[k] && p[k] <= n.hi[k];
    if (!in) continue;
    if (n.left < 0) {
      for (int i = 0; i < n.count; ++i) {
        const Prim &pr = u->bvh.prims[n.first + i];
        const int e = pr.prim;
        float v;
        if (!u->cells) {
          if (elem_value(u, e, P, v)) return v;
        } else if ((bestE < 0 || e < bestE) && p[0] >= pr.lo[0] && p[1] >= pr.lo[1] && p[2] >= pr.lo[2] &&
                   p[0] <= pr.hi[0] && p[1] <= pr.hi[1] && p[2] <= pr.hi[2] && elem_value(u, e, P, v)) {
          bestE = e;
          best = v;
        }
      }
    } else {
      stack[sp++] = n.left + 1;
      stack[sp++] = n.left;
    }
  }
  return best;
}

// rasterBox(grid, box4f) (MCGrid.cuh:92-122) and rasterTet<D>
// (UMeshField.cu:39-106)
struct P4 {
  float x, y, z, w;
};
P4 mid(const P4 &a, const P4 &b) { return P4{0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), 0.5f * (a.w + b.w)}; }
float edge_len(const P4 &a, const P4 &b) {
  const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}


void raster_box4(ref_volume_s *g, const P4 *q) {
  float lo[4], hi[4];
  for (int k = 0; k < 4; ++k) {
    lo[k] = kInf;
    hi[k] = -kInf;
  }
  for (int i = 0; i < 4; ++i) {
    const float c[4] = {q[i].x, q[i].y, q[i].z, q[i].w};
    for (int k = 0; k < 4; ++k) {
      lo[k] = std::fmin(lo[k], c[k]);
      hi[k] = std::fmax(hi[k], c[k]);
    }
  }
  if (lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]) return;
  const float org[3] = {g->mcOrigin.x, g->mcOrigin.y, g->mcOrigin.z};
  const float sp[3] = {g->mcSpacing.x, g->mcSpacing.y, g->mcSpacing.z};
  int a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = std::min(std::max((int)((lo[k] - org[k]) * (1.f / sp[k])), 0), g->m[k] - 1);
    b[k] = std::min(std::max((int)((hi[k] - org[k]) * (1.f / sp[k])), 0), g->m[k] - 1);
  }
  for (int z = a[2]; z <= b[2]; ++z)
    for (int y = a[1]; y <= b[1]; ++y)
      for (int x = a[0]; x <= b[0]; ++x) {
        const size_t id = (size_t)x + (size_t)g->m[0] * ((size_t)y + (size_t)g->m[1] * z);
        atomic_fmin(&g->rlo[id], lo[3]);
        atomic_fmax(&g->rhi[id], hi[3]);
      }
}

void raster_tet(ref_volume_s *g, int depth, P4 a, P4 b, P4 c, P4 d) {
  if (depth == 0) {
    const P4 q[4] = {a, b, c, d};
    raster_box4(g, q);
    return;
  }
  const float lab = edge_len(a, b), lac = edge_len(a, c), lad = edge_len(a, d);
  const float lbc = edge_len(b, c), lbd = edge_len(b, d), lcd = edge_len(c, d);
  const float mx = std::fmax(std::fmax(std::fmax(std::fmax(std::fmax(lab, lac), lad), lbc), lbd), lcd);
  P4 oa, ob, oc, od0, od1;
  if (lab >= mx) { oa = mid(a, b); ob = c; oc = d; od0 = a; od1 = b; }
  else if (lac >= mx) { oa = mid(a, c); ob = b; oc = d; od0 = a; od1 = c; }
  else if (lad >= mx) { oa = mid(a, d); ob = b; oc = c; od0 = a; od1 = d; }
  else if (lbc >= mx) { oa = mid(b, c); ob = a; oc = d; od0 = b; od1 = c; }
  else if (lbd >= mx) { oa = mid(b, d); ob = a; oc = c; od0 = b; od1 = d; }
  else { oa = mid(c, d); ob = a; oc = b; od0 = c; od1 = d; }
  raster_tet(g, depth - 1, oa, ob, oc, od0);
  raster_tet(g, depth - 1, oa, ob, oc, od1);
}

struct IHit {
  bool hit;
  float t;
  V3 P, N;
};

// MCIsoSurfaceAccel::isProg (MCAccelerator.h:338-567)
template <class Field>
IHit iso_isect(const ref_volume_s *g, const Field &field, float iso, V3 org, V3 dir, float tmin, float tmax,
               float rayTMax, Rng &seed, uint64_t geomHash) {
  IHit res;
  res.hit = false;
  float t0 = tmin, t1 = tmax;
  {
    const V3 a = mul(sub(g->wlo, org), inv(dir)), b = mul(sub(g->whi, org), inv(dir));
    t0 = std::fmax(t0, maxc(vmin(a, b)));
    t1 = std::fmin(t1, minc(vmax(a, b)));
    if (!(t0 <= t1)) return res;
  }
  const V3 dorg = mul(sub(org, g->mcOrigin), inv(g->mcSpacing));
  const V3 ddir = mul(dir, inv(g->mcSpacing));
  Rng rng = child(seed, geomHash);
  (void)rng;
  float tHit = rayTMax;
  dda3(dorg, ddir, t1, g->m, [&](const int *c, float c0, float c1) -> bool {
    float rlo = c0, rhi = std::fmin(c1, rayTMax);
    if (rlo >= rhi) return true;
    const size_t id = (size_t)c[0] + (size_t)g->m[0] * ((size_t)c[1] + (size_t)g->m[1] * c[2]);
    float vlo = g->rlo[id], vhi = g->rhi[id];
    if (!(iso >= vlo && iso <= vhi)) return true;
    float ff0 = 0.f, ff1 = field(add(org, scl(c0, dir)));
    float tt1 = c0;
    for (int i = 1; i <= 10; ++i) {
      const float tt0 = tt1;
      ff0 = ff1;
      tt1 = lerpl(i / 10.f, c0, c1);
      ff1 = field(add(org, scl(tt1, dir)));
      if (std::isnan(ff0) || std::isnan(ff1)) continue;
      vlo = std::fmin(ff0, ff1);
      vhi = std::fmax(ff0, ff1);
      rlo = tt0;
      rhi = tt1;
      if (iso >= vlo && iso <= vhi) {
        float t = (iso - ff0) / (ff1 - ff0);
        t = lerpl(t, rlo, rhi);
        tHit = std::fmin(tHit, t);
        if (tHit < rayTMax) return false;
      }
    }
    return true;
  });
  if (tHit >= rayTMax) return res;
  const V3 P = add(org, scl(tHit, dir));
  const V3 size = sub(g->whi, g->wlo);
  const float delta = len(size) * .1f / float(g->m[0] + g->m[1] + g->m[2]);
  const float fP = field(P);
  float f[6] = {field(add(P, mk(-delta, 0.f, 0.f))), field(add(P, mk(+delta, 0.f, 0.f))),
                field(add(P, mk(0.f, -delta, 0.f))), field(add(P, mk(0.f, +delta, 0.f))),
                field(add(P, mk(0.f, 0.f, -delta))), field(add(P, mk(0.f, 0.f, +delta)))};
  float dd[3] = {2.f, 2.f, 2.f};
  for (int k = 0; k < 6; ++k)
    if (std::isnan(f[k])) {
      dd[k / 2] -= 1.f;
      f[k] = fP;
    }
  V3 N = mk(dd[0] == 0.f ? 0.f : (f[1] - f[0]) / dd[0], dd[1] == 0.f ? 0.f : (f[3] - f[2]) / dd[1],
            dd[2] == 0.f ? 0.f : (f[5] - f[4]) / dd[2]);
  if (eq(N, mk(0, 0, 0))) N = neg(nrm(dir));
  res.hit = true;
  res.t = tHit;
  res.P = P;
  res.N = nrm(N);
  return res;
}

// volume (geometry 0) first, then the iso-surface (geometry 1); each
// intersection advances the ray seed once
HitInfo um_trace(const ref_umesh_s *u, bool vol, bool iso, float isoValue, V3 isoColor, V3 org, V3 dir, float tMax,
                 Rng &seed, long long &samples) {
  HitInfo h;
  h.hit = false;
  float tc = tMax;
  auto field = [&](V3 P) { return um_sample(u, P); };
  if (vol) {
    const VHit vh = vol_isect_g(&u->grid, field, org, dir, 0.f, tMax, tMax, seed, false, samples, fnv3(0, 0, 0));
    if (vh.hit) {
      h.hit = true;
      h.P = add(org, scl(vh.t, dir));
      h.N = mk(0, 0, 0);
      h.t = vh.t;
      h.type = M_PHASE;
      h.albedo = scl(.7f, mk(vh.rgb[0], vh.rgb[1], vh.rgb[2]));
      tc = vh.t;
    }
  }
  if (iso) {
    auto counted = [&](V3 P) {
      ++samples;
      return um_sample(u, P);
    };
    const IHit ih = iso_isect(&u->grid, counted, isoValue, org, dir, 0.f, tMax, tc, seed, fnv3(0, 1, 0));
    if (ih.hit) {
      h.hit = true;
      h.P = ih.P;
      h.N = ih.N;
      h.t = ih.t;
      h.type = M_LAMBERT;
      h.albedo = mk(.85f * isoColor.x * (1.f / kPi), .85f * isoColor.y * (1.f / kPi), .85f * isoColor.z * (1.f / kPi));
      h.metallic = h.roughness = h.ior = 0.f;
    }
  }
  return h;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// calls f(vertex index) for each vertex of element e (a polyhedron: every
// entry of its face stream)
template <class F> void um_each_vertex(const ref_umesh_s *u, int e, F f) {
  if (!u->cells) {
    for (int q = 0; q < 4; ++q) f(u->tets[4 * (size_t)e + q]);
    return;
  }
  const int *ix = &u->indices[(size_t)u->offsets[e]];
  const int k = u->kinds[e];
  if (k == 4) {
    int pos = 0;
    const int nf = ix[pos++];
    for (int fc = 0; fc < nf; ++fc) {
      const int np = ix[pos++];
      for (int q = 0; q < np; ++q) f(ix[pos + q]);
      pos += np;
    }
    return;
  }
  const int nv = k == 0 ? 4 : k == 1 ? 5 : k == 2 ? 6 : 8;
  for (int q = 0; q < nv; ++q) f(ix[q]);
}

bool um_is_tet(const ref_umesh_s *u, int e) { return !u->cells || u->kinds[e] == 0; }

// rasterBox(grid, worldBounds, box4f) (MCGrid.cuh:67-91): cells other than
// tets; only an inverted box is skipped
void raster_box_cells(ref_volume_s *g, int e, const ref_umesh_s *u) {
  float lo[4], hi[4];
  for (int k = 0; k < 4; ++k) {
    lo[k] = kInf;
    hi[k] = -kInf;
  }
  um_each_vertex(u, e, [&](int vi) {
    const float *c = &u->v4[4 * (size_t)vi];
    for (int k = 0; k < 4; ++k) {
      lo[k] = std::fmin(lo[k], c[k]);
      hi[k] = std::fmax(hi[k], c[k]);
    }
  });
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
  const float org[3] = {g->mcOrigin.x, g->mcOrigin.y, g->mcOrigin.z};
  const float sp[3] = {g->mcSpacing.x, g->mcSpacing.y, g->mcSpacing.z};
  int a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = std::min(std::max((int)((lo[k] - org[k]) * (1.f / sp[k])), 0), g->m[k] - 1);
    b[k] = std::min(std::max((int)((hi[k] - org[k]) * (1.f / sp[k])), 0), g->m[k] - 1);
  }
  for (int z = a[2]; z <= b[2]; ++z)
    for (int y = a[1]; y <= b[1]; ++y)
      for (int x = a[0]; x <= b[0]; ++x) {
        const size_t id = (size_t)x + (size_t)g->m[0] * ((size_t)y + (size_t)g->m[1] * z);
        atomic_fmin(&g->rlo[id], lo[3]);
        atomic_fmax(&g->rhi[id], hi[3]);
      }
}

// element boxes (UMeshField cellBounds), BVH, and the macrocell grid of
// buildInitialMacroCells (UMeshField.cu:110-200): tets by rasterTet<5>,
// other cells by their box
ref_umesh_s *um_build(ref_umesh_s *u, int numElems) {
  V3 wlo = mk(kInf, kInf, kInf), whi = mk(-kInf, -kInf, -kInf);
  for (int e = 0; e < numElems; ++e) {
    Prim pr;
    pr.geom = 0;
    pr.prim = e;
    pr.key = (uint32_t)e;
    for (int k = 0; k < 3; ++k) {
      pr.lo[k] = kInf;
      pr.hi[k] = -kInf;
    }
    um_each_vertex(u, e, [&](int vi) {
      const float *p = &u->v4[4 * (size_t)vi];
      for (int k = 0; k < 3; ++k) {
        pr.lo[k] = std::fmin(pr.lo[k], p[k]);
        pr.hi[k] = std::fmax(pr.hi[k], p[k]);
      }
    });
    wlo = vmin(wlo, mk(pr.lo[0], pr.lo[1], pr.lo[2]));
    whi = vmax(whi, mk(pr.hi[0], pr.hi[1], pr.hi[2]));
    u->bvh.prims.push_back(pr);
  }
  build_bvh(&u->bvh);
  ref_volume_s &g = u->grid;
  g.wlo = wlo;
  g.whi = whi;
  const V3 size = sub(whi, wlo);
  const float maxWidth = maxc(size);
  const int MC = 200 + int(std::sqrt(numElems / 1000.f));
  const float f = (MC - 1) / maxWidth;
  g.m[0] = 1 + (int)(size.x * f);
  g.m[1] = 1 + (int)(size.y * f);
  g.m[2] = 1 + (int)(size.z * f);
  g.mcOrigin = wlo;
  g.mcSpacing = mul(size, inv(mk((float)g.m[0], (float)g.m[1], (float)g.m[2])));
  const size_t nc = (size_t)g.m[0] * g.m[1] * g.m[2];
  g.rlo.assign(nc, kInf);
  g.rhi.assign(nc, -kInf);
  g.maj.assign(nc, 0.f);
  parallel_for(numElems, default_threads(0), [&](int e) {
    if (um_is_tet(u, e)) {
      P4 p[4];
      int q = 0;
      um_each_vertex(u, e, [&](int vi) {
        const float *v = &u->v4[4 * (size_t)vi];
        p[q++] = P4{v[0], v[1], v[2], v[3]};
      });
      raster_tet(&g, 5, p[0], p[1], p[2], p[3]);
    } else {
      raster_box_cells(&g, e, u);
    }
  });
  const float one[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  g.tf.assign(one, one + 8);
  g.numTF = 2;
  return u;
}

void um_vertices(ref_umesh_s *u, const float *vertices, const float *scalars, int numVertices) {
  u->v4.resize(4 * (size_t)numVertices);
  for (int i = 0; i < numVertices; ++i) {
    for (int k = 0; k < 3; ++k) u->v4[4 * (size_t)i + k] = vertices[3 * (size_t)i + k];
    u->v4[4 * (size_t)i + 3] = scalars[i];
  }
}

}  // namespace

extern "C" {

ref_umesh *ref_umesh_create(const float *vertices, const float *scalars, int numVertices, const int *tets, int numTets) {
  auto *u = new ref_umesh_s;
  um_vertices(u, vertices, scalars, numVertices);
  u->tets.assign(tets, tets + 4 * (size_t)numTets);
  return um_build(u, numTets);
}

ref_umesh *ref_umesh_create_cells(const float *vertices, const float *scalars, int numVertices, const int *indices,
                                  int numIndices, const int *offsets, const unsigned char *types, int numCells) {
  auto *u = new ref_umesh_s;
  u->cells = true;
  um_vertices(u, vertices, scalars, numVertices);
  u->indices.assign(indices, indices + numIndices);
  u->offsets.assign(offsets, offsets + numCells);
  u->kinds.resize(numCells);
  for (int c = 0; c < numCells; ++c) {
    int k = -1, nv = 0;
    switch (types[c]) {  // UMeshField's VTK and ANARI codes
      case 10: case 0: k = 0; nv = 4; break;
      case 14: case 3: k = 1; nv = 5; break;
      case 13: case 2: k = 2; nv = 6; break;
      case 12: case 1: k = 3; nv = 8; break;
      case 42: k = 4; nv = 0; break;
    }
    bool ok = k >= 0 && offsets[c] >= 0 && (long long)offsets[c] + nv <= numIndices;
    for (int q = 0; ok && q < nv; ++q) ok = indices[offsets[c] + q] >= 0 && indices[offsets[c] + q] < numVertices;
    if (ok && k == 4) {  // FaceStreamReader's checks (UMeshField.h:179-262)
      long long pos = offsets[c];
      const int nf = pos < numIndices ? indices[pos++] : 0;
      ok = nf > 0 && nf <= 8192;
      for (int f = 0; ok && f < nf; ++f) {
        const int np = pos < numIndices ? indices[pos++] : 0;
        ok = np >= 3 && np <= 1024 && pos + np <= numIndices;
        for (int q = 0; ok && q < np; ++q) ok = indices[pos + q] >= 0 && indices[pos + q] < numVertices;
        pos += np;
      }
    }
    if (!ok) {
      delete u;
      return nullptr;
    }
    u->kinds[c] = k;
  }
  return um_build(u, numCells);
}

void ref_umesh_destroy(ref_umesh *u) { delete u; }

void ref_umesh_set_tf(ref_umesh *u, const float *values, int numValues, float lo, float hi, float base) {
  ref_volume_s &g = u->grid;
  g.tf.assign(values, values + 4 * (size_t)numValues);
  g.numTF = numValues;
  g.dlo = lo;
  g.dhi = hi;
  g.base = base;
  for (size_t i = 0; i < g.maj.size(); ++i) g.maj[i] = tfmaj(&g, g.rlo[i], g.rhi[i]);
}

void ref_umesh_mc_dims(const ref_umesh *u, int *dims) {
  for (int k = 0; k < 3; ++k) dims[k] = u->grid.m[k];
}

void ref_umesh_get_mcs(const ref_umesh *u, float *ranges, float *maj) {
  const ref_volume_s &g = u->grid;
  for (size_t i = 0; i < g.maj.size(); ++i) {
    if (ranges) {
      ranges[2 * i] = g.rlo[i];
      ranges[2 * i + 1] = g.rhi[i];
    }
    if (maj) maj[i] = g.maj[i];
  }
}

// MCIsoSurfaceAccel over a structured field (MCAccelerator.h:338-567 with
// StructuredDataSampler, StructuredData.h:104-115): the same iso walk as the
// tet mesh, over the structured volume's macrocell value ranges.
void ref_volume_iso_rays(const ref_volume *v, const float *rays, float isoValue, float *out, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * (size_t)i;
    Rng seed;
    auto field = [&](V3 P) { return vsample(v, P); };
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    const IHit h = iso_isect(v, field, isoValue, org, dir, r[3], r[7], r[7], seed, fnv3(0, 0, 0));
    float *o = out + 4 * (size_t)i;
    o[0] = h.hit ? h.t : kInf;
    o[1] = h.hit ? h.N.x : 0.f;
    o[2] = h.hit ? h.N.y : 0.f;
    o[3] = h.hit ? h.N.z : 0.f;
  });
}

// MCIsoSurfaceAccel over a NanoVDB field (MCAccelerator.h:338-567 with
// NanoVDBDataSampler, NanoVDB.h:98-127) over its macrocell value ranges.
void ref_nvdb_iso_rays(const ref_nvdb *g, const float *rays, float isoValue, float *out, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * (size_t)i;
    Rng seed;
    auto field = [&](V3 P) { return nv_sample(g, P); };
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    const IHit h = iso_isect(&g->grid, field, isoValue, org, dir, r[3], r[7], r[7], seed, fnv3(0, 0, 0));
    float *o = out + 4 * (size_t)i;
    o[0] = h.hit ? h.t : kInf;
    o[1] = h.hit ? h.N.x : 0.f;
    o[2] = h.hit ? h.N.y : 0.f;
    o[3] = h.hit ? h.N.z : 0.f;
  });
}

void ref_umesh_iso_rays(const ref_umesh *u, const float *rays, float isoValue, float *out, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * (size_t)i;
    Rng seed;
    auto field = [&](V3 P) { return um_sample(u, P); };
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    const IHit h = iso_isect(&u->grid, field, isoValue, org, dir, r[3], r[7], r[7], seed, fnv3(0, 1, 0));
    float *o = out + 4 * (size_t)i;
    o[0] = h.hit ? h.t : kInf;
    o[1] = h.hit ? h.N.x : 0.f;
    o[2] = h.hit ? h.N.y : 0.f;
    o[3] = h.hit ? h.N.z : 0.f;
  });
}

// Brute force: the first element in index order whose inside test passes
// (for mixed meshes, whose own box also holds P: the smallest index that
// um_sample's BVH query must also return)
void ref_umesh_sample_brute(const ref_umesh *u, const float *points, float *out, int n, int threads) {
  const int numElems = u->cells ? (int)u->kinds.size() : (int)u->tets.size() / 4;
  // element boxes by element index (the BVH holds them in its own order)
  std::vector<const Prim *> box(numElems);
  for (const Prim &pr : u->bvh.prims) box[pr.prim] = &pr;
  parallel_for(n, default_threads(threads), [&](int i) {
    const V3 P = ld3(points + 3 * (size_t)i);
    const float p[3] = {P.x, P.y, P.z};
    out[i] = NAN;
    for (int e = 0; e < numElems; ++e) {
      const Prim &pr = *box[e];
      if (u->cells && !(p[0] >= pr.lo[0] && p[1] >= pr.lo[1] && p[2] >= pr.lo[2] && p[0] <= pr.hi[0] &&
                        p[1] <= pr.hi[1] && p[2] <= pr.hi[2]))
        continue;
      float v;
      if (elem_value(u, e, P, v)) {
        out[i] = v;
        break;
      }
    }
  });
}

void ref_umesh_sample(const ref_umesh *u, const float *points, float *out, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) { out[i] = um_sample(u, ld3(points + 3 * (size_t)i)); });
}

long long ref_umesh_render(const ref_umesh *u, const ref_camera *cam, const ref_render_params *p, int volume, int iso,
                           float isoValue, const float *isoColor, float *accum, int y0, int y1, int threads,
                           long long *samples) {
  const Cam c = make_cam(cam);
  const int W = p->width;
  const V3 icol = ld3(isoColor);
  const int maxDepth = p->maxPathDepth > 0 ? p->maxPathDepth : 10;
  const int maxDiff = p->maxDiffuseBounces >= 0 ? p->maxDiffuseBounces : 1;
  std::atomic<long long> nrays(0), ns(0);
  parallel_for((y1 - y0) * W, default_threads(threads), [&](int li) {
    const int ix = li % W, iy = y0 + li / W, pix = ix + iy * W;
    float acc[4] = {0, 0, 0, 0};
    long long nr = 0, cnt = 0;
    for (int pp = 0; pp < p->pathsPerPixel; ++pp) {
      const int a = p->accumID + pp;
      PRay ray = cam_ray(c, p, ix, iy, a);
      const float cmax = 10.f * (1 + a);
      V3 sum = mk(0, 0, 0);
      float alpha = 0.f;
      for (int gen = 0; gen == 0 || ray.tMax > 0.f; ++gen) {
        const V3 d = fix_dir(ray.dir);
        const HitInfo hi = um_trace(u, volume != 0, iso != 0, isoValue, icol, ray.org, d, ray.tMax, ray.rng, cnt);
        ++nr;
        if (gen == 0) alpha = hi.hit ? 1.f : ray.missA;
        PRay sh;
        add_frag(sum, bounce(p, ray, hi, sh, gen, maxDepth, maxDiff), cmax);
        if (sh.tMax > 0.f) {
          const HitInfo hs = um_trace(u, volume != 0, iso != 0, isoValue, icol, sh.org, fix_dir(sh.dir), sh.tMax,
                                      sh.rng, cnt);
          ++nr;
          PRay dummy;
          add_frag(sum, bounce(p, sh, hs, dummy, gen + 1, maxDepth, maxDiff), cmax);
        }
      }
      acc[0] += sum.x; acc[1] += sum.y; acc[2] += sum.z; acc[3] += alpha;
    }
    float *o = accum + 4 * pix;
    if (p->accumID == 0) { o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2]; o[3] = acc[3]; }
    else { o[0] += acc[0]; o[1] += acc[1]; o[2] += acc[2]; o[3] += acc[3]; }
    nrays += nr;
    ns += cnt;
  });
  if (samples) *samples = ns.load();
  return nrays.load();
}

}  // extern "C"

// ============================================================================
// Instanced worlds (rtcore/embree/Group.cpp:181-217 InstanceGroup): each
// instance = a scene + affine transform; rays go to object space through
// rcp(xfm) (computed in double, rounded once), hits come back through xfm.
// ============================================================================
struct ref_world_s {
  std::vector<const ref_scene *> groups;
  std::vector<int> instGroup, instID;
  std::vector<std::array<float, 12>> o2w, w2o;
};

namespace {

V3 xpoint(const float *m, V3 p) {
  return mk(m[0] * p.x + m[3] * p.y + m[6] * p.z + m[9], m[1] * p.x + m[4] * p.y + m[7] * p.z + m[10],
            m[2] * p.x + m[5] * p.y + m[8] * p.z + m[11]);
}
V3 xvector(const float *m, V3 v) {
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

// inverse of the affine map (columns vx, vy, vz, p): cofactor matrix over
// the determinant in double, p' = -(L^-1 p)
void invert12(const float *m, float *out) {
  double L[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) L[r][c] = m[3 * c + r];
  double C[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      C[r][c] = L[r1][c1] * L[r2][c2] - L[r1][c2] * L[r2][c1];
    }
  const double det = L[0][0] * C[0][0] + L[0][1] * C[0][1] + L[0][2] * C[0][2];
  double inv[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) inv[r][c] = C[c][r] / det;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) out[3 * c + r] = (float)inv[r][c];
  for (int r = 0; r < 3; ++r) out[9 + r] = (float)-(inv[r][0] * m[9] + inv[r][1] * m[10] + inv[r][2] * m[11]);
}

}  // namespace

extern "C" {

ref_world *ref_world_create(const ref_scene *const *groups, int numGroups, const int *instGroup, const int *instIDs,
                            const float *xfms, int numInstances) {
  auto *w = new ref_world_s;
  w->groups.assign(groups, groups + numGroups);
  for (int i = 0; i < numInstances; ++i) {
    w->instGroup.push_back(instGroup[i]);
    w->instID.push_back(instIDs[i]);
    std::array<float, 12> a, b;
    for (int k = 0; k < 12; ++k) a[k] = xfms[12 * (size_t)i + k];
    invert12(a.data(), b.data());
    w->o2w.push_back(a);
    w->w2o.push_back(b);
  }
  return w;
}

void ref_world_destroy(ref_world *w) { delete w; }

void ref_world_trace(const ref_world *w, const float *rays, ref_hit *hits, int n, int threads) {
  parallel_for(n, default_threads(threads), [&](int i) {
    const float *r = rays + 8 * (size_t)i;
    const V3 org = ld3(r), dir = fix_dir(ld3(r + 4));
    Best best;
    best.t = r[7];
    best.u = best.v = 0.f;
    best.key = 0xffffffffu;
    best.geom = best.prim = -1;
    int bestInst = -1;
    V3 bo = org, bd = dir;
    for (size_t ii = 0; ii < w->instGroup.size(); ++ii) {
      const ref_scene *sc = w->groups[w->instGroup[ii]];
      const V3 o = xpoint(w->w2o[ii].data(), org);
      const V3 d = fix_dir(xvector(w->w2o[ii].data(), dir));
      const Best b = closest(sc, o, d, r[3], r[7], false, nullptr);
      // closest over instances; ties to the smaller (instance, key)
      if (b.geom >= 0 && (bestInst < 0 || b.t < best.t)) {
        best = b;
        bestInst = (int)ii;
        bo = o;
        bd = d;
      }
    }
    ref_hit &h = hits[i];
    h.t = best.t;
    h.u = best.u;
    h.v = best.v;
    if (bestInst < 0) {
      h.primID = h.geomID = h.instID = h.instIndex = -1;
      h.N[0] = h.N[1] = h.N[2] = 0.f;
    } else {
      V3 P, N;
      hit_point_normal(w->groups[w->instGroup[bestInst]], best, bo, bd, r[3], r[7], P, N);
      N = xvector(w->o2w[bestInst].data(), N);
      h.primID = best.prim;
      h.geomID = best.geom;
      h.instID = w->instID[bestInst];
      h.instIndex = bestInst;
      h.N[0] = N.x; h.N[1] = N.y; h.N[2] = N.z;
    }
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Frame-buffer tiles and readback (barney/fb/TiledFB.cu, FrameBuffer.cu,
// barney/common/math.h). Tiles are 32x32 pixels; pixel i of a tile sits at
// (lower.x + i % 32, lower.y + i / 32).
// ---------------------------------------------------------------------------
namespace {

constexpr int kFbTile = 32;
constexpr int kFbTilePixels = kFbTile * kFbTile;

// make_8bit (barney/common/math.h:94-97): min(255, max(0, int(f * 256)));
// int() as the GPU converts (saturating, NaN -> 0).
uint32_t fb_8bit(float f) {
  const float s = f * 256.f;
  if (s >= 256.f) return 255u;
  if (s > 0.f) return (uint32_t)(int)s;
  return 0u;
}

// linear_to_srgb (barney/common/math.h:52-58); the multiply-add fused as
// nvcc contracts it by default.
float fb_srgb(float x) {
  if (x <= 0.0031308f) return 12.92f * x;
  return std::fma(1.055f, powf(x, 1.f / 2.4f), -0.055f);
}

// saturate (barney/common/math.h:39-46), fminf/fmaxf semantics.
float fb_saturate(float f) { return std::fmax(0.f, std::fmin(f, 1.f)); }

// make_rgba(vec4f) (barney/common/math.h:109-116), sRGB on R, G, B only
// (math.h:63-66).
uint32_t fb_rgba(const float *c, bool srgb) {
  float r = c[0], g = c[1], b = c[2];
  if (srgb) {
    r = fb_srgb(r);
    g = fb_srgb(g);
    b = fb_srgb(b);
  }
  return fb_8bit(r) | (fb_8bit(g) << 8) | (fb_8bit(b) << 16) | (fb_8bit(c[3]) << 24);
}

void fb_store(void *out, size_t idx, const float *c, int format) {
  if (format == 0)
    std::memcpy((float *)out + 4 * idx, c, 16);
  else
    ((uint32_t *)out)[idx] = fb_rgba(c, format == 2);
}

}  // namespace

extern "C" {

uint32_t ref_fb_make_8bit(float f) { return fb_8bit(f); }
float ref_fb_linear_to_srgb(float x) { return fb_srgb(x); }

// TiledFB::resize (barney/fb/TiledFB.cu:243-252) and setTileCoordsKernel
// (TiledFB.cu:226-238).
int ref_fb_tile_descs(int width, int height, int rank, int size, int *lower) {
  const int tx = (width + kFbTile - 1) / kFbTile, ty = (height + kFbTile - 1) / kFbTile;
  const int total = tx * ty;
  const int numActive = (std::max(0, total - rank) + size - 1) / size;
  if (lower)
    for (int t = 0; t < numActive; ++t) {
      const int id = t * size + rank;
      lower[2 * t + 0] = (id % tx) * kFbTile;
      lower[2 * t + 1] = (id / tx) * kFbTile;
    }
  return numActive;
}

// linearizeColorAndNormalKernel (barney/fb/TiledFB.cu:13-48); a tile is
// float accum[1024][4] followed by float normal[1024][3].
void ref_fb_linearize(const float *tiles, const int *lower, int numTiles, int width, int height, float scale,
                      int format, void *out, float *outNormal) {
  for (int t = 0; t < numTiles; ++t) {
    const float *acc = tiles + (size_t)t * kFbTilePixels * 7;
    const float *nrm = acc + kFbTilePixels * 4;
    for (int sub = 0; sub < kFbTilePixels; ++sub) {
      const int ix = lower[2 * t] + sub % kFbTile, iy = lower[2 * t + 1] + sub / kFbTile;
      if (ix >= width || iy >= height) continue;
      const size_t idx = (size_t)ix + (size_t)width * iy;
      const float c[4] = {acc[4 * sub] * scale, acc[4 * sub + 1] * scale, acc[4 * sub + 2] * scale,
                          acc[4 * sub + 3] * scale};
      fb_store(out, idx, c, format);
      if (outNormal) std::memcpy(outNormal + 3 * idx, nrm + 3 * sub, 12);
    }
  }
}

// linearizeAuxTilesKernel (barney/fb/TiledFB.cu:88-105).
void ref_fb_linearize_aux(const uint32_t *aux, const int *lower, int numTiles, int width, int height,
                          uint32_t *out) {
  for (int t = 0; t < numTiles; ++t)
    for (int sub = 0; sub < kFbTilePixels; ++sub) {
      const int ix = lower[2 * t] + sub % kFbTile, iy = lower[2 * t + 1] + sub / kFbTile;
      if (ix >= width || iy >= height) continue;
      out[(size_t)ix + (size_t)width * iy] = aux[(size_t)t * kFbTilePixels + sub];
    }
}

// Linear accumulation -> output format, scaled (the color step of
// linearizeColorAndNormalKernel on an untiled frame).
void ref_fb_resolve(const float *accum, size_t n, float scale, int format, void *out) {
  for (size_t i = 0; i < n; ++i) {
    const float c[4] = {accum[4 * i] * scale, accum[4 * i + 1] * scale, accum[4 * i + 2] * scale,
                        accum[4 * i + 3] * scale};
    fb_store(out, i, c, format);
  }
}

// LinearToFixed8::run (barney/fb/FrameBuffer.cu:116-128).
void ref_fb_to_fixed8(const float *in, size_t n, int srgb, uint32_t *out) {
  for (size_t i = 0; i < n; ++i) {
    const float c[4] = {fb_saturate(in[4 * i]), fb_saturate(in[4 * i + 1]), fb_saturate(in[4 * i + 2]),
                        fb_saturate(in[4 * i + 3])};
    out[i] = fb_rgba(c, srgb != 0);
  }
}

}  // extern "C"
