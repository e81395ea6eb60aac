/* refcpu.h -- CPU oracle for the barney-amd ray-tracing core.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain C++ restatement of barney's
 * embree-backend semantics for the hot path (the reference's embree, OWL and
 * cuBQL dependencies are not available here, so nothing of the reference is
 * linked). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, as the checker; the product never calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the reference repository root).
 */
#ifndef REFCPU_H
#define REFCPU_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- RNG (barney/common/barney-common.h:65-111) --- */
uint64_t ref_rng_seed(uint64_t state, uint32_t a, uint32_t b);
/* advances 'state' n times, writing each 32-bit output */
void ref_rng_sequence(uint64_t *state, uint32_t *out, int n);
/* Random2(a,b) then n floats */
void ref_random2_floats(uint32_t a, uint32_t b, float *out, int n);
uint64_t ref_hash3(uint32_t a, uint32_t b, uint32_t c);
void ref_fmath_eval(const float *x, float *sinOut, float *cosOut, float *pow5Out, int n);
/* atan2(y[i], x[i]) and acos(x[i]) of the shared fmath.h sequences */
void ref_fmath_eval_inv(const float *y, const float *x, float *atan2Out, float *acosOut, int n);
float ref_fast_log(float f);

/* --- geometry input (host arrays) --- */
typedef struct {
  int kind;               /* 0 triangles, 1 spheres, 2 cylinders */
  int numPrims;
  const float *vertices;  /* float3 */
  const int *indices;     /* int3 / int2 */
  const float *radii;     /* or NULL */
  float defaultRadius;
  const float *colors;    /* per-prim float3 or NULL */
  float color[3];         /* material color (matte color / PBR base color) */
  int userID;
  /* material: 0 AnariMatte, 1 AnariPBR (Glass when ior != 1 and
   * transmission >= 1e-3, else NVisii; AnariPBR.h:68-94) */
  int matType;
  float metallic, roughness, transmission, ior;
  /* material "opacity" (AnariMatte / AnariPBR; 1 = opaque): triangles with
   * an alpha below 1 ignore candidate hits at random (Triangles.dev.cu:
   * 121-133) */
  float opacity;
  const float *normals;   /* triangles: float3 per vertex (shading normals) or NULL */
} ref_geom;

typedef struct ref_scene_s ref_scene;

/* Builds the oracle's scene: a binned-SAH binary BVH over all geoms (one
 * identity instance, geomID = index in the array). */
ref_scene *ref_scene_create(const ref_geom *geoms, int numGeoms);
void ref_scene_destroy(ref_scene *s);

typedef struct {
  float t, u, v;
  int primID, geomID, instID;
  float N[3];
  int instIndex;
} ref_hit;

/* closest hit for each ray (rays: 8 floats org,tmin,dir,tmax);
 * bruteForce != 0 tests every primitive. nodeVisits/primTests may be NULL. */
void ref_trace(const ref_scene *s, const float *rays, ref_hit *hits, int n, int bruteForce,
               int threads, long long *primTests);

/* FP64 triangle intersection (no tie rule): returns t or -1 */
double ref_triangle_t_f64(const double *org, const double *dir, const double *v0, const double *v1,
                          const double *v2);

/* float -> IEEE binary16 -> float, round to nearest even */
float ref_half_round(float x);

/* BSDF probes (materials as in ref_geom; color = material color). Sample
 * i uses RNG seed(i, seed); out per sample: dir[3], f_r[3], pdf, scatter
 * type (0 diffuse, 1 glossy, 2 specular, 3 volume), offset direction,
 * changedMedium. */
void ref_bsdf_sample(const ref_geom *mat, const float *color, const float *Ng, const float *wo, int inMedium,
                     uint32_t seed, int n, float *out);
/* out: eval value[3], eval pdf, pdf() */
void ref_bsdf_eval(const ref_geom *mat, const float *color, const float *Ng, const float *wo, const float *wi,
                   float *out);

/* --- camera + surface render (generateRays.cu, shadeRays.cu bounce) --- */
typedef struct {
  float position[3], direction[3], up[3];
  float fovy;
} ref_camera;

typedef struct {
  int width, height, accumID, pathsPerPixel;
  float ambientRadiance;
  float bgColor[4];
  int maxPathDepth, maxDiffuseBounces;
  /* Russian roulette from this path depth on (0 = off; not in barney):
   * survive with p = min(max throughput, 0.95), throughput /= p */
  int russianRoulette;
  /* HDRI environment light (EnvMapLight) or NULL: uniform ambientRadiance */
  const struct ref_envmap_s *envMap;
  /* Renderer "cutPlane" (nx, ny, nz, d): hits with dot(n, P) + d < 0 are
   * ignored; disabled when d <= -1e28 (Renderer.h:59) */
  float cutPlane[4];
  /* quad, directional and point lights (World::DD's lists) or NULL */
  const struct ref_lights_s *lights;
} ref_render_params;

/* --- HDRI environment light (barney/light/EnvMap.cu, EnvMap.h) --- */
typedef struct ref_envmap_s ref_envmap;
/* texels: w*h texels of rtc::DataType format (33 FLOAT4, 30 FLOAT, 1 UCHAR4,
 * 0 UCHAR), row-major; filter 0 point / 1 linear; addr[2] rtc::AddressMode
 * (0 wrap, 1 clamp, 2 border, 3 mirror); direction / up / scale as
 * EnvMapLight's parameters. NULL on a bad format or size. */
ref_envmap *ref_envmap_create(const void *texels, int format, int w, int h, int filter, const int *addr,
                              const float *direction, const float *up, float scale);
void ref_envmap_destroy(ref_envmap *e);

/* --- quad, directional and point lights (barney/light/, World.cpp:80-120) ---
 * quads: n x 12 floats (corner, edge0, edge1, emission); dirs: n x 7
 * (direction, color, radiance); points: n x 8 (position, color, power,
 * intensity; NaN intensity: color * power / (4 pi)) */
typedef struct ref_lights_s ref_lights;
ref_lights *ref_lights_create(const float *quads, int nq, const float *dirs, int nd, const float *points, int np);
void ref_lights_destroy(ref_lights *l);
void ref_envmap_get_cdfs(const ref_envmap *e, float *cdfY, float *cdfX);
/* 3x3 column-major toWorld and toLocal */
void ref_envmap_frame(const ref_envmap *e, float *toWorld, float *toLocal);
/* tex2D at n (u, v) pairs -> float4 */
void ref_envmap_tex2d(const ref_envmap *e, const float *uv, float *out, int n);
/* pdf(dir) and radianceFromEnv(dir) for n directions */
void ref_envmap_eval(const ref_envmap *e, const float *dirs, float *pdf, float *rad, int n);
/* sample i with Random2(i * 0x9E3779B9, seed): dir[3], radiance[3], pdf */
void ref_envmap_sample(const ref_envmap *e, uint32_t seed, float *out, int n);

void ref_trace_filtered(const ref_scene *s, const ref_render_params *p, const float *rays, const uint64_t *seeds,
                        ref_hit *hits, int n, int brute, int shadow, int stochastic, int threads);

/* primary ray for pixel (ix,iy): out = org[3], dir[3], seed (as 2 floats bits) */
void ref_camera_ray(const ref_camera *cam, const ref_render_params *p, int ix, int iy, float *org,
                    float *dir, uint64_t *seed);

/* renders rows [y0, y1) of the frame into accum (float4 per pixel, full
 * frame layout); returns rays traced. */
long long ref_render_surface(const ref_scene *s, const ref_camera *cam, const ref_render_params *p,
                             float *accum, int *ids, int y0, int y1, int threads);

/* --- structured volume (barney/volume) --- */
typedef struct ref_volume_s ref_volume;
ref_volume *ref_volume_create(const float *scalars, const int *dims, const float *origin,
                              const float *spacing, int cellsPerMC);
void ref_volume_destroy(ref_volume *v);
void ref_volume_set_tf(ref_volume *v, const float *values, int numValues, float lo, float hi,
                       float baseDensity);
void ref_volume_mc_dims(const ref_volume *v, int *dims);
void ref_volume_get_mcs(const ref_volume *v, float *ranges, float *majorants);
float ref_volume_sample(const ref_volume *v, const float *P); /* trilinear, NaN outside */
void ref_tf_map(const ref_volume *v, float s, float *out4);
float ref_tf_majorant(const ref_volume *v, float lo, float hi);
/* first accepted Woodcock sample per ray: out = t (inf if none), r, g, b */
void ref_volume_sample_rays(const ref_volume *v, const float *rays, const uint64_t *seeds, float *out,
                            int n, int exactLog, int threads, long long *samples);
/* DDA cell enumeration for one ray (grid space): writes up to maxCells
 * (x,y,z,t0bits,t1bits) records; returns the number of cells visited. */
int ref_dda_cells(const float *org, const float *dir, float tMax, const int *gridSize, int *cells,
                  float *ts, int maxCells);
long long ref_render_volume(const ref_volume *v, const ref_camera *cam, const ref_render_params *p,
                            float *accum, int y0, int y1, int threads, long long *samples);
/* A model holding geometry and a structured volume: the surface path
 * tracer, each ray delta-tracked through the volume up to its closest
 * surface hit; shadow rays not blocked by a surface are tracked through the
 * volume. Returns rays traced; *samples = volume samples. */
long long ref_render_mixed(const ref_scene *s, const ref_volume *v, const ref_camera *cam,
                           const ref_render_params *p, float *accum, int *ids, int y0, int y1, int threads,
                           long long *samples);

/* --- block-structured AMR (barney/amr) --- */
typedef struct ref_amr_s ref_amr;
/* origins/dims int3 per block, levels, offsets (first scalar of each block),
 * refinements per level; NULL for an invalid block */
ref_amr *ref_amr_create(const int *origins, const int *dims, const int *levels, const unsigned long long *offsets,
                        int numBlocks, const int *refinements, int numLevels, const float *scalars,
                        unsigned long long numScalars);
void ref_amr_destroy(ref_amr *a);
void ref_amr_set_tf(ref_amr *a, const float *values, int numValues, float lo, float hi, float baseDensity);
void ref_amr_mc_dims(const ref_amr *a, int *dims);
void ref_amr_get_mcs(const ref_amr *a, float *ranges, float *majorants);
/* field samples (NaN outside); brute != 0 visits every block instead of
 * the lookup grid's lists */
void ref_amr_sample(const ref_amr *a, const float *points, float *out, int n, int brute, int threads);
long long ref_amr_render(const ref_amr *a, const ref_camera *cam, const ref_render_params *p, float *accum, int y0,
                         int y1, int threads, long long *samples);

/* --- NanoVDB fields (barney/volume/NanoVDB.*) --- */
typedef struct ref_nvdb_s ref_nvdb;
/* a NanoVDB 32.x grid buffer of float, Fp4, Fp8, Fp16 or FpN values; NULL
 * when the header is not such a grid or a node lies outside the buffer */
ref_nvdb *ref_nvdb_create(const void *data, unsigned long long bytes);
void ref_nvdb_destroy(ref_nvdb *g);
void ref_nvdb_set_tf(ref_nvdb *g, const float *values, int numValues, float lo, float hi, float baseDensity);
void ref_nvdb_mc_dims(const ref_nvdb *g, int *dims);
void ref_nvdb_get_mcs(const ref_nvdb *g, float *ranges, float *majorants);
/* world box (lo, hi), index box (lo, hi), macrocell origin and spacing */
void ref_nvdb_boxes(const ref_nvdb *g, float *worldLoHi, int *indexLoHi, float *mcOriginSpacing);
void ref_nvdb_sample(const ref_nvdb *g, const float *points, float *out, int n, int threads);
/* ReadAccessor::getValue at integer coordinates (int3 each) */
void ref_nvdb_values(const ref_nvdb *g, const int *coords, float *out, int n);
/* iso-surface rays, as ref_volume_iso_rays */
void ref_nvdb_iso_rays(const ref_nvdb *g, const float *rays, float isoValue, float *out, int n, int threads);
long long ref_nvdb_render(const ref_nvdb *g, const ref_camera *cam, const ref_render_params *p, float *accum, int y0,
                          int y1, int threads, long long *samples);

/* --- unstructured volume (barney/umesh) --- */
typedef struct ref_umesh_s ref_umesh;
/* vertices float3, per-vertex scalars, tets int4 (positively oriented) */
ref_umesh *ref_umesh_create(const float *vertices, const float *scalars, int numVertices, const int *tets,
                            int numTets);
/* general cells in UMeshField's layout: types VTK 10/12/13/14 or ANARI
 * 0-3 (tet, hex, prism, pyramid), vertex indices at indices[offsets[c]..]
 * in VTK order; NULL for an unknown type or an index out of range */
ref_umesh *ref_umesh_create_cells(const float *vertices, const float *scalars, int numVertices, const int *indices,
                                  int numIndices, const int *offsets, const unsigned char *types, int numCells);
void ref_umesh_destroy(ref_umesh *u);
void ref_umesh_set_tf(ref_umesh *u, const float *values, int numValues, float lo, float hi, float baseDensity);
void ref_umesh_mc_dims(const ref_umesh *u, int *dims);
void ref_umesh_get_mcs(const ref_umesh *u, float *ranges, float *majorants);
/* point location + barycentric interpolation; NaN outside */
void ref_umesh_sample(const ref_umesh *u, const float *points, float *out, int n, int threads);
void ref_umesh_sample_brute(const ref_umesh *u, const float *points, float *out, int n, int threads);
/* iso-surface hit per ray (8 floats org,tmin,dir,tmax): out = t (inf if
 * none), normal[3] */
void ref_umesh_iso_rays(const ref_umesh *u, const float *rays, float isoValue, float *out, int n, int threads);
/* iso-surface hit per ray on a structured volume (same layout as
 * ref_umesh_iso_rays) */
void ref_volume_iso_rays(const ref_volume *v, const float *rays, float isoValue, float *out, int n, int threads);
/* volume (delta tracking) and/or iso-surface render, rows [y0, y1) */
long long ref_umesh_render(const ref_umesh *u, const ref_camera *cam, const ref_render_params *p, int volume,
                           int iso, float isoValue, const float *isoColor, float *accum, int y0, int y1,
                           int threads, long long *samples);

/* --- instanced worlds: scenes with affine transforms (columns vx, vy, vz,
 * p; object to world), closest hit over all instances, ties to the
 * smallest (instance, geom, prim) --- */
typedef struct ref_world_s ref_world;
ref_world *ref_world_create(const ref_scene *const *groups, int numGroups, const int *instGroup,
                            const int *instIDs, const float *xfms, int numInstances);
void ref_world_destroy(ref_world *w);
void ref_world_trace(const ref_world *w, const float *rays, ref_hit *hits, int n, int threads);

/* --- frame-buffer tiles and readback (barney/fb/TiledFB.cu:13-130,
 * 226-264, FrameBuffer.cu:105-128, common/math.h:39-116). A tile is
 * float accum[1024][4] then float normal[1024][3]; lower = int2 per tile.
 * format 0 float4, 1 RGBA8, 2 RGBA8 with sRGB on R, G, B. --- */
uint32_t ref_fb_make_8bit(float f);
float ref_fb_linear_to_srgb(float x);
int ref_fb_tile_descs(int width, int height, int rank, int size, int *lower);
void ref_fb_linearize(const float *tiles, const int *lower, int numTiles, int width, int height, float scale,
                      int format, void *out, float *outNormal);
void ref_fb_linearize_aux(const uint32_t *aux, const int *lower, int numTiles, int width, int height,
                          uint32_t *out);
void ref_fb_resolve(const float *accum, size_t n, float scale, int format, void *out);
void ref_fb_to_fixed8(const float *in, size_t n, int srgb, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
