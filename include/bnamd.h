/* bnamd.h -- C ABI of barney-amd, an MI355X-native (gfx950) ray-tracing core
 * that sits behind barney's rtcore TraceInterface / ComputeInterface.
 *
 * Conventions
 *  - Every function returns BNAMD_OK (0) or BNAMD_ERROR (-1); on error
 *    bnamd_last_error() returns a message (the C++ layer rethrows it as
 *    std::runtime_error, as rtcore/cudaCommon/cuda-helper.h:41-60 does).
 *  - Pointers documented as "device" point into GPU memory (from
 *    bnamd_device_malloc, hipMalloc or a torch tensor); everything else is
 *    host memory. Streams are hipStream_t passed as void* (NULL = default).
 *  - Plain C types only: no torch or C++ types cross this boundary.
 *  - Results are complete when the call returns unless a stream is passed,
 *    in which case work is enqueued on it.
 */
#ifndef BNAMD_H
#define BNAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BNAMD_OK 0
#define BNAMD_ERROR (-1)

/* ------------------------------------------------------------------------
 * Runtime / memory.  Replaces rtc::Device (rtcore/cudaCommon/Device.h:34-64:
 * allocMem/freeMem/copy/copyAsync/memsetAsync/sync/setActive).
 * --------------------------------------------------------------------- */
int bnamd_init(int device);
const char *bnamd_last_error(void);
int bnamd_device_malloc(void **dptr, size_t bytes);
int bnamd_device_free(void *dptr);
int bnamd_memcpy_h2d(void *dst, const void *src, size_t bytes);
int bnamd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int bnamd_memcpy_d2d(void *dst, const void *src, size_t bytes);
int bnamd_memset(void *dst, int value, size_t bytes);
int bnamd_synchronize(void);
/* copyAsync / memsetAsync on a stream (NULL = default); copies may be
 * host<->device or device<->device (direction inferred) */
int bnamd_memcpy_async(void *dst, const void *src, size_t bytes, void *stream);
int bnamd_memset_async(void *dst, int value, size_t bytes, void *stream);
/* pinned host memory (allocHost / freeHost) */
int bnamd_host_malloc(void **ptr, size_t bytes);
int bnamd_host_free(void *ptr);
/* active device (setActive / restoreActive) */
int bnamd_get_device(int *device);
int bnamd_set_device(int device);
/* Streaming device copy kernel (bytes % 16 == 0); a traffic reference for
 * hardware byte counters. No counterpart in the reference. */
int bnamd_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);

/* ------------------------------------------------------------------------
 * Geometry and acceleration structures.
 * bnamd_geom mirrors the device data (DD) of barney's Triangles
 * (barney/geometry/Triangles.h), Spheres (Spheres.h) and Cylinders
 * (Cylinders.h) geoms as handed to rtc::Geom::setDD.  A group is the
 * geometry list of rtc::Device::createTrianglesGroup / createUserGeomsGroup
 * (rtcore/cuda/Device.h:36-70); bnamd_world_build is
 * rtc::Device::createInstanceGroup + Group::buildAccel
 * (rtcore/cuda/cudaGroup.cu:144-562) and returns the AccelHandle analogue.
 * --------------------------------------------------------------------- */
#define BNAMD_GEOM_TRIANGLES 0
#define BNAMD_GEOM_SPHERES 1
#define BNAMD_GEOM_CYLINDERS 2
/* user geometry of the programmable path (RTC_EXPORT_USER_GEOM): its
 * primitives are boxes ('vertices' = device float4 pairs lo, hi per
 * primitive, as the bounds program writes them) whose hits its intersect
 * program finds; only rtc::amd::TraceInterface::traceRay traverses worlds
 * holding them (the fused renderers and bnamd_trace reject such worlds) */
#define BNAMD_GEOM_USER 6

typedef struct {
  int kind;              /* BNAMD_GEOM_* */
  int numPrims;          /* triangles / spheres / cylinders */
  int numVertices;       /* entries in 'vertices' */
  int materialID;        /* index into the world's material table */
  const float *vertices; /* device float3[]: tri/cyl vertices, sphere centers */
  const int *indices;    /* device int3[] (triangles) / int2[] (cylinders); unused for spheres */
  const float *radii;    /* device float[numPrims] or NULL (spheres: defaultRadius) */
  const float *colors;   /* device float3[numPrims] per-primitive color or NULL */
  float defaultRadius;
  int userID;            /* 'objID' written to the ID channel */
  const float *normals;  /* triangles: device float3[numVertices] vertex normals, interpolated at
                          * hits as the shading normal (Triangles.dev.cu:70-74), or NULL */
  const void *programs;  /* device program record of the geometry's anyHit / closestHit / intersect
                          * programs and its DD (rtc::amd::ProgramRecord) for the programmable
                          * traversal, or NULL (native programs only) */
} bnamd_geom;

typedef struct {
  const bnamd_geom *geoms; /* host array */
  int numGeoms;
} bnamd_group;

typedef struct {
  int group;      /* index into the group array */
  int instID;     /* user instance ID (InstanceGroup 'instIDs') */
  float xfm[12];  /* affine3f: vx, vy, vz, p (object to world) */
} bnamd_instance;

/* Material table entries: ANARI matte (barney/material/AnariMatte.h) and
 * PBR (AnariPBR.h) reduced to the parameters the BSDFs read. */
#define BNAMD_MAT_MATTE 0
#define BNAMD_MAT_PBR 1
typedef struct {
  int type;
  float color[3];
  float opacity;
  float metallic;
  float roughness;
  float transmission;
  float ior;
  float specular;
} bnamd_material;

typedef struct bnamd_world_s bnamd_world;

typedef struct {
  long long numPrims;
  long long numNodes;
  int depth;
  float buildMs;
  float sceneLower[3];
  float sceneUpper[3];
  int numBlas; /* distinct bottom levels of a two-level world (0: flattened) */
} bnamd_world_info;

int bnamd_world_build(const bnamd_group *groups, int numGroups,
                      const bnamd_instance *instances, int numInstances,
                      const bnamd_material *materials, int numMaterials,
                      void *stream, bnamd_world **out);
int bnamd_world_destroy(bnamd_world *world);
/* Copies the world's BVH8s (both levels) to the host and checks them: child
 * and primitive references in range, every primitive referenced exactly
 * once, every quantized child box containing what it bounds. Builds skip
 * this check unless BNAMD_VALIDATE_BVH=1. No reference counterpart (a
 * test hook). */
int bnamd_world_validate(const bnamd_world *world);
/* The device address of the world's traversal record (the AccelHandle that
 * rtc::amd::getAccelHandle returns and TraceInterface::traceRay reads on the
 * device; rtcore/cuda/Device.h:88 getAccelHandle returns the instance
 * group's device record the same way). Valid until the world is rebuilt,
 * refit or destroyed. */
int bnamd_world_device(const bnamd_world *world, const void **deviceRecord);
int bnamd_world_info_get(const bnamd_world *world, bnamd_world_info *info);

/* Bottom levels built on their own: rtc::GeomGroup::buildAccel of a
 * triangles or user-geometry group (rtcore/cuda/cudaGroup.cu:261-562). A
 * BLAS keeps the group's geometry pointers (the caller's buffers). The
 * material table decides which triangles shadow rays pass (Glass); a world
 * using the BLAS must agree with it on that (checked). */
typedef struct bnamd_blas_s bnamd_blas;
int bnamd_blas_build(const bnamd_group *group, const bnamd_material *materials, int numMaterials,
                     void *stream, bnamd_blas **out);
int bnamd_blas_destroy(bnamd_blas *blas);
/* rtc::InstanceGroup::buildAccel over built groups (cudaGroup.cu:144-258):
 * instances[i].group indexes 'blas'; every instance of one BLAS shares it.
 * The world references the BLASes, which must outlive it. */
int bnamd_world_build_instances(const bnamd_blas *const *blas, int numBlas,
                                const bnamd_instance *instances, int numInstances,
                                const bnamd_material *materials, int numMaterials,
                                void *stream, bnamd_world **out);
/* rtc::InstanceGroup::setTransforms + refitAccel: new object-to-world
 * transforms (12 floats each, bnamd_instance.xfm layout) for every instance
 * of a two-level world; rebuilds the instance level only. */
int bnamd_world_set_transforms(bnamd_world *world, const float *xfms, int numInstances,
                               void *stream);

/* ------------------------------------------------------------------------
 * Trace: closest hit for a batch of rays (rtc::TraceInterface::traceRay
 * semantics, embree oracle rtcore/embree/Device.cpp:107-196).  At equal t the
 * smallest (instance index, geomID, primID) wins.
 * --------------------------------------------------------------------- */
typedef struct {
  float org[3];
  float tmin;
  float dir[3];
  float tmax;
} bnamd_ray; /* 32 B */

typedef struct {
  float t;        /* tmax of the ray if no hit */
  float u, v;     /* triangle barycentrics */
  int primID;     /* -1 if no hit */
  int geomID;
  int instID;     /* user instance ID */
  float N[3];     /* normalized world-space geometric normal */
  int instIndex;  /* flattened instance index */
} bnamd_hit; /* 40 B */

int bnamd_trace(const bnamd_world *world, const bnamd_ray *rays, bnamd_hit *hits, int numRays,
                void *stream);

/* ------------------------------------------------------------------------
 * Surface path tracing (plain per-pixel megakernel): generateRays +
 * traceRays + shadeRays of barney/kernels/ (generateRays.cu, traceRays.cu,
 * shadeRays.cu) fused per pixel.
 * --------------------------------------------------------------------- */
typedef struct {
  float position[3];
  float direction[3];
  float up[3];
  float fovy; /* degrees */
} bnamd_camera;

typedef struct {
  int width, height;
  int accumID;          /* progressive frame index (jitter + seed) */
  int pathsPerPixel;
  float ambientRadiance;
  float bgColor[4];     /* w < 0: barney's vertical gradient */
  int maxPathDepth;     /* barney: 10 */
  int maxDiffuseBounces;/* barney: 1 */
  int russianRoulette;  /* 0: off (barney); >0: start depth for RR */
  /* HDRI environment light (barney's EnvMapLight, light/EnvMap.cu) or
   * NULL: the uniform ambientRadiance (World::DD without an env map) */
  const struct bnamd_envmap_s *envMap;
  /* Renderer "cutPlane" (render/Renderer.h:59, generateRays.cu:171-193,
   * OptixGlobals.h:33-42): (nx, ny, nz, d); surface hits with
   * dot(n, P) + d < 0 are ignored and primary rays are clipped to the
   * visible half-space. Disabled when d <= -1e28 (or all four are 0). */
  float cutPlane[4];
  /* quad, directional and point lights sampled next to the environment
   * (World::DD's light lists, render/World.h:28-35) or NULL: none */
  const struct bnamd_lights_s *lights;
} bnamd_render_params;

/* ------------------------------------------------------------------------
 * 2D textures and the HDRI environment light.
 * bnamd_texture2d replaces rtc::TextureData / rtc::Texture for 2D images
 * (rtcore/cudaCommon/TextureData.h:15-31, Texture.h:12-24); it is sampled
 * as the embree backend's tex2D does (rtcore/embree/Texture.cpp:26-289),
 * texels widened to float4 at creation. bnamd_envmap replaces
 * EnvMapLight::commit / computeCDFs (barney/light/EnvMap.cu:59-150,
 * 178-230) and its DD (EnvMap.h:97-246); renders use it through
 * bnamd_render_params.envMap.
 * --------------------------------------------------------------------- */
typedef struct bnamd_texture2d_s bnamd_texture2d;
typedef struct {
  int width, height;
  int format;          /* rtc::DataType: 33 FLOAT4, 30 FLOAT, 1 UCHAR4, 0 UCHAR */
  const void *texels;  /* host, row-major, width*height texels */
  int filterMode;      /* rtc::FilterMode: 0 point, 1 linear */
  int addressMode[2];  /* rtc::AddressMode per axis: 0 wrap, 1 clamp, 2 border, 3 mirror */
} bnamd_texture2d_desc;
int bnamd_texture2d_create(const bnamd_texture2d_desc *desc, void *stream, bnamd_texture2d **out);
int bnamd_texture2d_destroy(bnamd_texture2d *tex);
/* The device copy tex2D samples: width*height float4 texels (the texels
 * widened as getTexel does, rtcore/embree/Texture.cpp:166-214), row-major.
 * For device-side descriptors (rtc::amd's TextureObject). */
int bnamd_texture2d_device_view(const bnamd_texture2d *tex, const void **texels4, int *width, int *height);
/* tex2D<float4> at n (u, v) pairs: uv device float2[n], out device float4[n] */
int bnamd_texture2d_sample(const bnamd_texture2d *tex, const float *uv, float *out, int n, void *stream);

typedef struct bnamd_envmap_s bnamd_envmap;
typedef struct {
  float direction[3];  /* EnvMapLight "direction" (default 1,0,0) */
  float up[3];         /* "up" (default 0,0,1) */
  float scale;         /* "scale" (default 1) */
} bnamd_envmap_desc;
/* The map references tex's texels: tex must outlive it. */
int bnamd_envmap_create(const bnamd_texture2d *tex, const bnamd_envmap_desc *desc, void *stream,
                        bnamd_envmap **out);
int bnamd_envmap_destroy(bnamd_envmap *env);
/* host copies: cdfY[height], cdfX[height*width]; frames 3x3 column-major */
int bnamd_envmap_get_cdfs(const bnamd_envmap *env, float *cdfY, float *cdfX);
int bnamd_envmap_get_frame(const bnamd_envmap *env, float *toWorld, float *toLocal);
/* EnvMapLight::DD::pdf and radianceFromEnv for n directions: dirs device
 * float3[n], pdf device float[n], radiance device float3[n] */
int bnamd_envmap_eval(const bnamd_envmap *env, const float *dirs, float *pdf, float *radiance, int n,
                      void *stream);
/* EnvMapLight::DD::sample with Random2(i * 0x9E3779B9, seed) for i < n:
 * out device float[n][7] = direction[3], radiance[3], pdf */
int bnamd_envmap_sample(const bnamd_envmap *env, uint32_t seed, float *out, int n, void *stream);

/* ------------------------------------------------------------------------
 * Quad, directional and point lights: World::set of the light lists
 * (barney/render/World.cpp:80-120) with each light's getDD
 * (light/QuadLight.cu:9-18, DirLight.cu:10-20, PointLight.cu:9-17); the
 * instance transform is applied by the caller. Sampled by bounce() as
 * sampleLights does (kernels/shadeRays.cu:46-404): a reservoir of 8 quad
 * candidates, of 2 directional and of 2 point candidates, next to the
 * environment's sample.
 * --------------------------------------------------------------------- */
typedef struct {
  float corner[3], edge0[3], edge1[3]; /* QuadLight "corner", "edge0", "edge1" */
  float emission[3];                   /* "emission" (radiance) */
} bnamd_quad_light;
typedef struct {
  float direction[3]; /* DirLight "direction": the way the light travels */
  float color[3];     /* Light "color" */
  float radiance;     /* "irradiance" if set, else "radiance" (DirLight.cu:15-18) */
} bnamd_dir_light;
typedef struct {
  float position[3];
  float color[3];
  float power;     /* "power" (default 1) */
  float intensity; /* "intensity"; NaN: radiance = color * power / (4 pi) */
} bnamd_point_light;
typedef struct bnamd_lights_s bnamd_lights;
/* Copies the lists to the device (each count >= 0, arrays may be NULL when
 * their count is 0). A quad's edges must span a non-zero area; a
 * directional light's direction must be non-zero. */
int bnamd_lights_create(const bnamd_quad_light *quads, int numQuads, const bnamd_dir_light *dirs, int numDirs,
                        const bnamd_point_light *points, int numPoints, bnamd_lights **out);
int bnamd_lights_destroy(bnamd_lights *lights);

typedef struct {
  long long raysTraced; /* all rays traced: primary, shadow and bounce */
  long long primaryRays;
  float ms;
  long long nodeVisits; /* BVH8 nodes visited (128 B each) */
  long long primBytes;  /* primitive tests x bytes a test reads (triangle
                         * 36, sphere 16, cylinder 28) */
} bnamd_render_stats;

/* accum: device float4[width*height] (RGBA accumulation, as the accum
 * channel of barney/fb/TiledFB.h:21-24 but linear); ids: device int4 per
 * pixel {primID, instID, objID, depth-as-float-bits} or NULL. */
int bnamd_render_surface(const bnamd_world *world, const bnamd_camera *camera,
                         const bnamd_render_params *params, float *accum, int *ids,
                         void *stream, bnamd_render_stats *stats);

/* A model holding geometry and a structured volume (barney's volumes are
 * user geometry of the same world, traced by MCVolumeAccel::isProg,
 * barney/volume/MCAccelerator.h:569-708, next to the surfaces): the surface
 * path tracer of bnamd_render_surface where each ray's closest surface hit
 * is found first and the volume is then delta-tracked up to it (a collision
 * before the surface is a volume hit, shaded with the Phase BSDF); shadow
 * rays that no surface blocks are tracked through the volume. The volume
 * needs its transfer function; the ID channels record surface hits.
 * *volumeSamples (may be NULL) receives the volume samples taken. */
int bnamd_render_mixed(const bnamd_world *world, const struct bnamd_volume_s *volume, const bnamd_camera *camera,
                       const bnamd_render_params *params, float *accum, int *ids, void *stream,
                       bnamd_render_stats *stats, long long *volumeSamples);

/* ------------------------------------------------------------------------
 * Structured volumes: StructuredData + MCGrid + TransferFunction +
 * MCVolumeAccel (barney/volume/).
 * --------------------------------------------------------------------- */
typedef struct {
  /* device float[dims.x*dims.y*dims.z], x fastest. Read during
   * bnamd_volume_create only: the volume keeps its own copy in 8^3-voxel
   * bricks (as StructuredData::commit keeps a texture of the data), so the
   * array may be freed or changed afterwards. */
  const float *scalars;
  int dims[3];          /* number of scalars (cells = dims-1) */
  float gridOrigin[3];
  float gridSpacing[3];
  int cellsPerMC;       /* barney: 8 (StructuredData.cu:25) */
} bnamd_structured_desc;

typedef struct bnamd_volume_s bnamd_volume;

int bnamd_volume_create(const bnamd_structured_desc *desc, bnamd_volume **out);
int bnamd_volume_destroy(bnamd_volume *vol);
/* values: host float4[numValues]; recomputes majorants (MCGrid.cu:94-116). */
int bnamd_volume_set_tf(bnamd_volume *vol, const float *values, int numValues,
                        float domainLo, float domainHi, float baseDensity);
/* mc grid dims, and host copies of ranges (float2 per cell) / majorants */
int bnamd_volume_mc_dims(const bnamd_volume *vol, int dims[3]);
int bnamd_volume_get_mcs(const bnamd_volume *vol, float *ranges, float *majorants);

/* Delta-tracking render with single scattering (SCI_VIS_MODE shade path). */
typedef struct {
  long long samples;  /* sampleAndMap evaluations */
  long long ddaSteps; /* macrocells visited */
  long long raysTraced;
  float ms;
  long long zeroMajorantSteps; /* of ddaSteps: cells with majorant 0 (no draw) */
  /* unstructured meshes rendered with opts->countTraversal (else 0):
   * point location's BVH node visits, element inside tests, and the tests'
   * algorithmic bytes (per vertex a 4-byte index, 12-byte position and
   * 4-byte scalar, e.g. 80 B per tet; a polyhedron's face stream at 4 B per
   * entry plus 16 B per vertex reference; mixed meshes add the 16-byte
   * element-box halves read before a test) */
  long long nodeVisits;
  long long cellTests;
  long long cellBytes;
} bnamd_volume_stats;

int bnamd_volume_render(const bnamd_volume *vol, const bnamd_camera *camera,
                        const bnamd_render_params *params, float *accum, void *stream,
                        bnamd_volume_stats *stats);

/* Per-ray delta-tracking query for parity checks: for each ray returns the
 * first accepted Woodcock sample {t, r, g, b} (t = +inf if none), using the
 * ray seeds given (uint64 per ray). */
int bnamd_volume_sample_rays(const bnamd_volume *vol, const bnamd_ray *rays,
                             const uint64_t *seeds, float *out, int numRays, int exactLog,
                             void *stream);

/* Iso-surface query per ray on the structured field (MCIsoSurfaceAccel over
 * StructuredDataSampler, MCAccelerator.h:338-567): out[i] = {t, N.x, N.y,
 * N.z} of the first crossing of isoValue in [tmin, tmax] found by the
 * macrocell walk with 10 linear sub-steps per cell (t = +inf if none); N is
 * the normalized central-difference gradient. Needs no
 * transfer function. */
int bnamd_volume_iso_rays(const bnamd_volume *vol, const bnamd_ray *rays, float isoValue, float *out,
                          int numRays, void *stream);

/* ------------------------------------------------------------------------
 * Block-structured AMR volumes (barney/amr): BlockStructuredField with its
 * blocks, levels and scalars (BlockStructuredField.h, .cu), sampled as
 * BlockStructuredCuBQLSampler does (BlockStructuredCuBQLSampler.h:66-80:
 * the tent basis functions of every block whose domain contains the point,
 * normalized; NaN where none does), the 256-wide macrocell grid of
 * BlockStructuredField::buildMCs, and MCVolumeAccel delta tracking with the
 * structured-volume render path. Blocks are visited in ascending order (the
 * reference: its BVH's order), which fixes the rounding where domains
 * overlap. A level-L cell is (int)powf(refinements[L], L) units wide; block
 * b covers cells origins[b] .. origins[b] + dims[b] - 1 of its level, its
 * scalars at scalars[offsets[b] + x + dims.x * (y + dims.y * z)].
 * --------------------------------------------------------------------- */
typedef struct {
  int numBlocks;
  const int *origins;                /* host int3 [numBlocks] */
  const int *dims;                   /* host int3 [numBlocks] */
  const int *levels;                 /* host [numBlocks] */
  const unsigned long long *offsets; /* host [numBlocks] */
  int numLevels;
  const int *refinements;            /* host [numLevels] */
  const float *scalars;              /* device [numScalars] */
  unsigned long long numScalars;
} bnamd_amr_desc;

typedef struct bnamd_amr_s bnamd_amr;

typedef struct {
  int numBlocks;
  int mcDims[3];
  int lookupDims[3]; /* the block lookup grid (ascending block lists) */
  float worldLower[3], worldUpper[3];
  float buildMs;
} bnamd_amr_info;

/* Replaces BlockStructuredField::setData/commit/buildMCs and the sampler's
 * BVH build (BlockStructuredField.cu:82-105, 229-300). */
int bnamd_amr_create(const bnamd_amr_desc *desc, void *stream, bnamd_amr **out);
int bnamd_amr_destroy(bnamd_amr *amr);
int bnamd_amr_info_get(const bnamd_amr *amr, bnamd_amr_info *info);
/* values: host float4[numValues]; recomputes the majorants */
int bnamd_amr_set_tf(bnamd_amr *amr, const float *values, int numValues, float domainLo, float domainHi,
                     float baseDensity);
/* host copies of the macrocell ranges (float2 per cell) and majorants */
int bnamd_amr_get_mcs(const bnamd_amr *amr, float *ranges, float *majorants);
/* out[i] = field value at points[i] (device float3), NaN outside */
int bnamd_amr_sample_points(const bnamd_amr *amr, const float *points, float *out, int numPoints, void *stream);
/* delta-tracked render, as bnamd_volume_render */
int bnamd_amr_render(const bnamd_amr *amr, const bnamd_camera *camera, const bnamd_render_params *params,
                     float *accum, void *stream, bnamd_volume_stats *stats);

/* ------------------------------------------------------------------------
 * NanoVDB fields (barney/volume/NanoVDB.h, NanoVDB.cu): one NanoVDB grid
 * (format major version 32: GridData, TreeData, root tiles, 32^3 and 16^3
 * internal nodes, 8^3 leaves) of float, Fp4, Fp8, Fp16 or FpN values, as
 * barney's "NanoVDB" scalar field takes it in its "data" member. Sampled as
 * NanoVDBDataSampler::DD::sample (NanoVDB.h:98-127: worldToIndexF, minus
 * half a voxel, clamped to the index box, SampleFromVoxels<1> trilinear),
 * with the macrocell grid of NanoVDBData::buildMCs / computeMCs
 * (NanoVDB.cu:28-102: divRoundUp(gridSize, 16) cells over the world box),
 * MCVolumeAccel delta tracking and MCIsoSurfaceAccel iso-surfaces.
 * --------------------------------------------------------------------- */
typedef struct {
  const void *data;         /* the grid buffer, 8-byte aligned */
  unsigned long long bytes;
  int dataOnDevice;         /* 0: host memory, 1: device memory */
  int noLeafTable;          /* 0: a per-8^3-block leaf lookup table over the
                             * index box (4 B per block, made when it has at
                             * most 2^24 blocks) spares most samples the tree
                             * walk; 1: every lookup walks the tree. Samples
                             * are the same either way. */
} bnamd_nanovdb_desc;

typedef struct bnamd_nanovdb_s bnamd_nanovdb;

typedef struct {
  int gridType; /* nanovdb::GridType: 1 float, 13 Fp4, 14 Fp8, 15 Fp16, 16 FpN */
  int indexLower[3], indexUpper[3];
  int mcDims[3];
  float worldLower[3], worldUpper[3];
  float mcOrigin[3], mcSpacing[3];
  float voxelSize[3];
  unsigned long long gridBytes;
  float buildMs;
} bnamd_nanovdb_info;

/* Replaces NanoVDBData::setData("data") + commit + buildMCs (NanoVDB.cu:82-102,
 * 190-235). The buffer is read during the call only (the field keeps its own
 * copy in device memory); a buffer that is not such a grid, or whose nodes
 * reach outside it, gives BNAMD_ERROR with a message. */
int bnamd_nanovdb_create(const bnamd_nanovdb_desc *desc, void *stream, bnamd_nanovdb **out);
int bnamd_nanovdb_destroy(bnamd_nanovdb *g);
int bnamd_nanovdb_info_get(const bnamd_nanovdb *g, bnamd_nanovdb_info *info);
/* values: host float4[numValues]; recomputes the majorants */
int bnamd_nanovdb_set_tf(bnamd_nanovdb *g, const float *values, int numValues, float domainLo, float domainHi,
                         float baseDensity);
/* host copies of the macrocell ranges (float2 per cell) and majorants */
int bnamd_nanovdb_get_mcs(const bnamd_nanovdb *g, float *ranges, float *majorants);
/* out[i] = field value at points[i] (device float3) */
int bnamd_nanovdb_sample_points(const bnamd_nanovdb *g, const float *points, float *out, int numPoints,
                                void *stream);
/* delta-tracked render, as bnamd_volume_render */
int bnamd_nanovdb_render(const bnamd_nanovdb *g, const bnamd_camera *camera, const bnamd_render_params *params,
                         float *accum, void *stream, bnamd_volume_stats *stats);
/* iso-surface query per ray, as bnamd_volume_iso_rays */
int bnamd_nanovdb_iso_rays(const bnamd_nanovdb *g, const bnamd_ray *rays, float isoValue, float *out,
                           int numRays, void *stream);

/* ------------------------------------------------------------------------
 * Unstructured volumes (barney/umesh): UMeshField with per-vertex scalars,
 * UMeshCuBQLSampler point location (UMeshCuBQLSampler.h:51-67,
 * UMeshField.h:362-501 eltScalar for tets, pyramids, prisms and hexahedra),
 * the macrocell grid of buildInitialMacroCells (UMeshField.cu:39-200),
 * MCVolumeAccel<UMeshCuBQLSampler> delta tracking and
 * MCIsoSurfaceAccel<UMeshCuBQLSampler> iso-surfaces (MCAccelerator.h:338-708).
 * A mesh is either all tets (tets != NULL) or general elements in
 * UMeshField's layout (tets == NULL): cell c has type cellTypes[c] and its
 * vertex indices at indices[cellOffsets[c] ...] in VTK order. Types are
 * UMeshField's codes: VTK 10 tet, 12 hex, 13 prism (wedge), 14 pyramid, or
 * ANARI 0 tet, 1 hex, 2 prism, 3 pyramid, and VTK 42 polyhedron, whose
 * indices are a face stream (UMeshField.h:179-262 FaceStreamReader).
 * --------------------------------------------------------------------- */
#define BNAMD_CELL_TET 10
#define BNAMD_CELL_HEX 12
#define BNAMD_CELL_PRISM 13
#define BNAMD_CELL_PYR 14
#define BNAMD_CELL_POLYHEDRON 42 /* face stream: numFaces, then numPts + indices per face */

typedef struct {
  const float *vertices; /* device float3 [numVertices] */
  const float *scalars;  /* device float [numVertices] (per-vertex data) */
  int numVertices;
  const int *tets;       /* device int4 [numTets]; positively oriented
                            (det(v1-v0, v2-v0, v3-v0) > 0), as tetScalar's
                            inside test requires; NULL for general cells */
  int numTets;
  const int *indices;            /* device [numIndices] (general cells) */
  int numIndices;
  const int *cellOffsets;        /* device [numCells] */
  const unsigned char *cellTypes;/* device [numCells] */
  int numCells;
} bnamd_umesh_desc;

typedef struct bnamd_umesh_s bnamd_umesh;

typedef struct {
  int numTets, numNodes, bvhDepth; /* numTets: elements (tets or cells) */
  int mcDims[3];
  float worldLower[3], worldUpper[3];
  float buildMs;   /* BVH build + tet packing */
  float mcMs;      /* macrocell rasterization */
} bnamd_umesh_info;

int bnamd_umesh_create(const bnamd_umesh_desc *desc, void *stream, bnamd_umesh **out);
int bnamd_umesh_destroy(bnamd_umesh *mesh);
int bnamd_umesh_info_get(const bnamd_umesh *mesh, bnamd_umesh_info *info);
/* values: host float4[numValues]; recomputes the majorants */
int bnamd_umesh_set_tf(bnamd_umesh *mesh, const float *values, int numValues, float domainLo,
                       float domainHi, float baseDensity);
/* host copies of the macrocell ranges (float2 per cell) and majorants */
int bnamd_umesh_get_mcs(const bnamd_umesh *mesh, float *ranges, float *majorants);
/* point queries: out[i] = interpolated scalar at points[i] (device float3),
 * NaN outside the mesh */
int bnamd_umesh_sample_points(const bnamd_umesh *mesh, const float *points, float *out, int numPoints,
                              void *stream);

/* iso-surface query per ray: out[i] = {t, N.x, N.y, N.z} of the first
 * crossing of isoValue in [tmin, tmax] (t = +inf if none); N is the
 * normalized central-difference gradient (MCAccelerator.h:477-507) */
int bnamd_umesh_iso_rays(const bnamd_umesh *mesh, const bnamd_ray *rays, float isoValue, float *out,
                         int numRays, void *stream);

typedef struct {
  int volume;        /* delta-tracked volume on/off */
  int iso;           /* iso-surface on/off */
  float isoValue;
  float isoColor[3]; /* matte (AnariMatte) color of the iso-surface */
  int countTraversal; /* nonzero: run the instrumented kernel that also fills
                       * stats->nodeVisits / cellTests / cellBytes (the
                       * same image; slower, for measurement) */
} bnamd_umesh_render_opts;

/* Volume and/or iso-surface render with the env light (same bounce as the
 * surface path; volume scattering ends the path as in SCI_VIS_MODE). Both
 * are intersected per ray, the volume first. stats->samples counts field
 * evaluations (Woodcock samples and iso sub-steps). */
int bnamd_umesh_render(const bnamd_umesh *mesh, const bnamd_camera *camera,
                       const bnamd_render_params *params, const bnamd_umesh_render_opts *opts,
                       float *accum, void *stream, bnamd_volume_stats *stats);

/* bnamd_umesh_render without any allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its counters and its iso-surface mask live in one fixed per-GPU array the
 * library holds until the process ends (4 MiB + 256 B): meshes whose
 * macrocell grid exceeds 32 M cells are rejected when opts->iso is set
 * (use bnamd_umesh_render). Every call on a GPU shares that array, so calls
 * that overlap on one GPU race. The image equals bnamd_umesh_render's.
 * stats may be NULL (enqueue only, as under capture); when given, the call
 * waits for the stream to read the counters back, and stats->ms is -1 (not
 * timed). */
int bnamd_umesh_render_noalloc(const bnamd_umesh *mesh, const bnamd_camera *camera,
                               const bnamd_render_params *params, const bnamd_umesh_render_opts *opts,
                               float *accum, void *stream, bnamd_volume_stats *stats);

/* ------------------------------------------------------------------------
 * Emission-absorption volume rendering and its adjoint (differentiable
 * form; not in the reference). Fixed-step march over a structured density
 * grid sampled like StructuredDataSampler (barney/volume/StructuredData.h:
 * 104-115) and mapped through a TransferFunction (TransferFunction.h:67-86):
 *   C = sum_i T_i a_i c_i,  a_i = 1 - exp(-s_i dt),  T_{i+1} = T_i (1 - a_i)
 * with (c_i, s_i) = TF(density(x_i)), x_i = o + (t0 + (i + 1/2) dt) d over
 * the ray's overlap [t0, t1] with the grid box, i < floor((t1 - t0) / dt).
 * The forward pass stores transmittance every 'checkpointEvery' samples;
 * the adjoint replays each segment from its checkpoint in reverse order.
 * --------------------------------------------------------------------- */
typedef struct {
  const float *density; /* device float[dims.x*dims.y*dims.z], x fastest */
  int dims[3];
  float gridOrigin[3];
  float gridSpacing[3];
  const float *tf;      /* device float4[numTF] (rgb, density) */
  int numTF;
  float domainLo, domainHi;
  float baseDensity;
  float step;           /* march step in world units */
} bnamd_ea_volume;

typedef struct {
  int width, height, numViews;
  const bnamd_camera *cameras; /* host array [numViews] */
  int checkpointEvery;         /* samples per checkpoint segment (e.g. 16) */
} bnamd_ea_views;

typedef struct {
  long long samples;
  float ms;
} bnamd_ea_stats;

/* checkpoints per ray the forward pass writes (buffer = rays * count floats) */
int bnamd_ea_checkpoint_count(const bnamd_ea_volume *vol, const bnamd_ea_views *views, int *count);
/* image: device float3[numViews*height*width]; checkpoints: device or NULL */
int bnamd_ea_forward(const bnamd_ea_volume *vol, const bnamd_ea_views *views, float *image,
                     float *checkpoints, void *stream, bnamd_ea_stats *stats);
/* gradImage: device float3 per pixel (dL/dC); gradDensity (dims) and
 * gradTF (float4[numTF]) are accumulated into (caller zeroes them). */
int bnamd_ea_backward(const bnamd_ea_volume *vol, const bnamd_ea_views *views, const float *gradImage,
                      const float *checkpoints, float *gradDensity, float *gradTF, void *stream,
                      bnamd_ea_stats *stats);

/* ------------------------------------------------------------------------
 * Batched multi-view path tracing (not in the reference: barney renders one
 * camera per frame). All views are rendered in one launch sequence; view v's
 * pixels occupy accum[offset_v .. offset_v + w_v*h_v) with offsets the
 * prefix sum of w*h in view order. The per-path computation is barney's
 * generateRays + trace + bounce() (shadeRays.cu:476-878); shadow rays use an
 * occlusion query, which decides the same hit/miss as barney's closest-hit
 * shadow trace.
 *   mode BNAMD_BATCH_MEGAKERNEL: one kernel, a few lanes per pixel.
 *   mode BNAMD_BATCH_WAVEFRONT:  per-bounce queues (trace, sort by BSDF,
 *        shade with compaction, shadow), persistent trace kernels, the bounce
 *        loop captured once in a hipGraph and replayed; queue and launch
 *        sizes from the autotuner (bnamd_batch_autotune) when one ran.
 * --------------------------------------------------------------------- */
#define BNAMD_BATCH_MEGAKERNEL 0
#define BNAMD_BATCH_WAVEFRONT 1
#define BNAMD_BATCH_AUTO 2 /* the renderer bnamd_batch_autotune found faster (wavefront untuned) */

typedef struct {
  int numViews;
  const bnamd_camera *cameras; /* host [numViews] */
  const int *widths;           /* host [numViews] */
  const int *heights;          /* host [numViews] */
} bnamd_views;

typedef struct {
  long long raysTraced;   /* primary + bounce + shadow rays */
  long long primaryRays;
  long long shadowRays;
  float ms;
  int graphReplays;       /* wavefront: graph launches used */
  long long nodeVisits;   /* as bnamd_render_stats */
  long long primBytes;
} bnamd_batch_stats;

/* params->width/height are ignored (per-view sizes); pathsPerPixel paths
 * per pixel starting at params->accumID. accum is overwritten when
 * params->accumID == 0, accumulated into otherwise. ids: NULL, or device
 * int4 per pixel (same offsets as accum) that receives the ID record of
 * bnamd_render_surface ({primID, instID, objID, depth bits} of the primary
 * hit of the path with accumID 0, the hitIDs buffer barney's anyHit
 * programs fill: Triangles.dev.cu:45-60, Spheres.dev.cu:178-198); left
 * untouched when params->accumID != 0. */
int bnamd_render_batch(const bnamd_world *world, const bnamd_views *views,
                       const bnamd_render_params *params, float *accum, int *ids, int mode,
                       void *stream, bnamd_batch_stats *stats);

/* Times launch configurations on this world and views and keeps the fastest
 * with the world for later bnamd_render_batch calls: the wavefront trace
 * grid and work-fetch chunk (one path per pixel), the wavefront queue size
 * (paths per pixel per batch) and the megakernel tile (lanes per pixel) at
 * params->pathsPerPixel, and which of the two renderers is faster
 * (BNAMD_BATCH_AUTO). If path is non-NULL the choice is also written there
 * as JSON. */
int bnamd_batch_autotune(const bnamd_world *world, const bnamd_views *views,
                         const bnamd_render_params *params, const char *path);

/* ------------------------------------------------------------------------
 * Frame-buffer tiles and readback.
 * The accumulation tiles, tile descriptors and aux tiles have the layout of
 * barney's TiledFB (fb/TiledFB.h:16-36): 32x32-pixel tiles, pixel i of a
 * tile at (lower.x + i % 32, lower.y + i / 32).  Device g of G owns global
 * tiles g, g+G, g+2G, ... in row-major tile order.
 * Color formats: BNAMD_FB_FLOAT4 writes float4 per pixel; the UFIXED8
 * formats write one uint32 per pixel, R in the low byte, each channel
 * min(255, max(0, int(c * 256))), the _SRGB one after the sRGB transfer
 * function on R, G and B (alpha stays linear) (common/math.h:52-66,
 * 94-116).  Pixels outside width x height are not written.
 * --------------------------------------------------------------------- */
#define BNAMD_TILE_SIZE 32
#define BNAMD_PIXELS_PER_TILE (BNAMD_TILE_SIZE * BNAMD_TILE_SIZE)
#define BNAMD_FB_FLOAT4 0
#define BNAMD_FB_UFIXED8_RGBA 1
#define BNAMD_FB_UFIXED8_RGBA_SRGB 2

typedef struct {
  int lower[2]; /* pixel coordinates of the tile's lower-left corner */
} bnamd_tile_desc;

typedef struct {
  float accum[BNAMD_PIXELS_PER_TILE][4];  /* RGBA sums over accumID frames */
  float normal[BNAMD_PIXELS_PER_TILE][3];
} bnamd_accum_tile;

/* Number of tiles device globalRank of globalSize owns for a width x height
 * frame, and (if descs != NULL, a device array of that many) their
 * descriptors.  Replaces TiledFB::resize's tile count and
 * setTileCoordsKernel (fb/TiledFB.cu:226-264). */
int bnamd_fb_tile_descs(int width, int height, int globalRank, int globalSize,
                        bnamd_tile_desc *descs, int *numActiveTiles, void *stream);

/* Writes numTiles accumulation tiles into a linear width x height frame:
 * color = accum * accumScale (barney passes 1/accumID) in the given format,
 * and the normal unchanged into outNormal (float3 per pixel) when it is not
 * NULL.  Replaces TiledFB::linearizeColorAndNormal and its kernel
 * (fb/TiledFB.cu:13-83). */
int bnamd_fb_linearize(const bnamd_accum_tile *tiles, const bnamd_tile_desc *descs, int numTiles,
                       int width, int height, float accumScale, int format, void *outColor,
                       float *outNormal, void *stream);

/* Writes numTiles aux tiles (uint32 or float per pixel: depth, primID,
 * instID, objID) into a linear width x height array.  Replaces
 * TiledFB::linearizeAuxTiles (fb/TiledFB.cu:88-130). */
int bnamd_fb_linearize_aux(const uint32_t *auxTiles, const bnamd_tile_desc *descs, int numTiles,
                           int width, int height, uint32_t *out, void *stream);

/* The same color conversion for the linear float4 accumulation buffer that
 * bnamd_render_surface, bnamd_volume_render, bnamd_umesh_render and
 * bnamd_render_batch write (one device owns every pixel). */
int bnamd_fb_resolve(const float *accum, int width, int height, float accumScale, int format,
                     void *out, void *stream);

/* float4 -> RGBA8 after clamping every channel to [0,1], with the sRGB
 * transfer on R, G, B when srgb != 0.  Replaces the LinearToFixed8 compute
 * kernel of the denoised read path (fb/FrameBuffer.cu:105-128, 268-285). */
int bnamd_fb_to_fixed8(const float *in, int width, int height, int srgb, uint32_t *out,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BNAMD_H */
