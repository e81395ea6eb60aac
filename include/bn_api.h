/* The bn* host API over this backend: barney's public C API
 * (barney/include/barney.h) for surface and volume scenes -- context, data
 * arrays (including object lists), triangle / sphere / cylinder geometry
 * (vertex normals, per-primitive colors), matte and physically-based
 * materials (opacity), 2D textures, the environment-map light, 3D float
 * texture data, the "structured", "unstructured", "BlockStructuredAMR" and
 * "NanoVDB" scalar fields, volumes with a transfer function, groups with lights,
 * instances, perspective camera, renderer parameters (including the cut
 * plane), frame buffer with color / depth / ID channels, render, and every
 * bnSet* setter. Names, types, parameter names and enum values are
 * barney's, so an application written against barney.h for such scenes
 * compiles and links against libbn_api.so. One device, one data slot
 * (whichSlot must be 0); a model holds either geometry or one volume (this
 * backend's renderers are separate kernels); of the light types only
 * "envmap" is rendered (bnBuild rejects the others).
 * Objects belong to their context: bnRelease drops the application's
 * reference, storage is freed by bnContextDestroy. Errors throw
 * std::runtime_error from C++ callers' point of view, as barney does. */
#ifndef BN_API_H
#define BN_API_H

#include <cstddef>
#include <cstdint>

#define BN_API extern "C" __attribute__((visibility("default")))

typedef struct _BNContext *BNContext;
typedef struct _BNObject {} *BNObject;
typedef struct _BNData : public _BNObject {} *BNData;
typedef struct _BNGeom : public _BNObject {} *BNGeom;
typedef struct _BNVolume : public _BNObject {} *BNVolume;
typedef struct _BNGroup : public _BNObject {} *BNGroup;
typedef struct _BNModel : public _BNObject {} *BNModel;
typedef struct _BNRenderer : public _BNObject {} *BNRenderer;
typedef struct _BNFrameBuffer : public _BNObject {} *BNFrameBuffer;
typedef struct _BNCamera : public _BNObject {} *BNCamera;
typedef struct _BNMaterial : public _BNObject {} *BNMaterial;
typedef struct _BNTextureData : public _BNObject {} *BNTextureData;
typedef struct _BNScalarField : public _BNObject {} *BNScalarField;
typedef struct _BNLight : public _BNObject {} *BNLight;
typedef struct _BNTexture2D : public _BNObject {} *BNTexture2D;

struct bn_float2 { float x, y; };
struct bn_float3 { float x, y, z; };
struct bn_float4 { float x, y, z, w; };

/* barney.h:85-93 */
struct BNTransform {
  struct {
    bn_float3 vx, vy, vz;
  } l;
  bn_float3 p;
};

/* barney.h:95-102 */
typedef enum {
  BN_FB_COLOR = (1 << 0),
  BN_FB_DEPTH = (1 << 1),
  BN_FB_PRIMID = (1 << 2),
  BN_FB_INSTID = (1 << 3),
  BN_FB_OBJID = (1 << 4),
  BN_FB_NORMAL = (1 << 5),
} BNFrameBufferChannel;

/* the values of barney.h:104-182 that this API accepts */
typedef enum {
  BN_DATA_UNDEFINED = 0,
  BN_OBJECT = 2,
  BN_UINT8 = 110,
  BN_UINT8_VEC4 = 113,
  BN_INT32 = 140,
  BN_INT32_VEC2 = 141,
  BN_INT32_VEC3 = 142,
  BN_INT32_VEC4 = 143,
  BN_UINT32 = 150,
  BN_INT64 = 160,
  BN_UINT64 = 170,
  BN_FLOAT32 = 180,
  BN_FLOAT32_VEC2 = 181,
  BN_FLOAT32_VEC3 = 182,
  BN_FLOAT32_VEC4 = 183,
  BN_FLOAT4 = BN_FLOAT32_VEC4,
  BN_UFIXED8 = 300,
  BN_UFIXED8_RGBA = 301,
  BN_UFIXED8_RGBA_SRGB = 302,
} BNDataType;

/* barney.h:185-213 */
typedef enum {
  BN_UNSTRUCTURED_TET = 10,
  BN_UNSTRUCTURED_HEX = 12,
  BN_UNSTRUCTURED_PRISM = 13,
  BN_UNSTRUCTURED_PYRAMID = 14,
  BN_UNSTRUCTURED_POLYHEDRON = 42
} BNUnstructuredElementType;
typedef enum { BN_TEXTURE_NEAREST, BN_TEXTURE_LINEAR } BNTextureFilterMode;
typedef enum { BN_TEXTURE_WRAP, BN_TEXTURE_CLAMP, BN_TEXTURE_BORDER, BN_TEXTURE_MIRROR } BNTextureAddressMode;
typedef enum { BN_COLOR_SPACE_LINEAR, BN_COLOR_SPACE_SRGB } BNTextureColorSpace;

/* barney.h:308-323 */
BN_API BNContext bnContextCreate(const int *dataRanksOnThisContext = 0, int numDataRanksOnThisContext = 1,
                                 const int *gpuIDs = nullptr, int numGPUs = -1);
BN_API void bnContextDestroy(BNContext context);
/* barney.h:333, 343 */
BN_API void bnRelease(BNObject object);
BN_API void bnAddReference(BNObject object);

/* barney.h:230-244 */
BN_API BNCamera bnCameraCreate(BNContext context, const char *type);
BN_API BNFrameBuffer bnFrameBufferCreate(BNContext context, int deprecated = 0);
BN_API BNModel bnModelCreate(BNContext ctx);
BN_API BNRenderer bnRendererCreate(BNContext ctx, const char *type = "default");

/* barney.h:251-299: parameters and commit */
BN_API void bnCommit(BNObject target);
BN_API void bnSetData(BNObject target, const char *paramName, BNData value);
BN_API void bnSetObject(BNObject target, const char *paramName, const BNObject value);
BN_API void bnSetString(BNObject target, const char *paramName, const char *value);
BN_API void bnSetLight(BNObject target, const char *paramName, BNLight value);
BN_API void bnSet1i(BNObject target, const char *paramName, int value);
BN_API void bnSet2i(BNObject target, const char *paramName, int x, int y);
BN_API void bnSet3i(BNObject target, const char *paramName, int x, int y, int z);
BN_API void bnSet4i(BNObject target, const char *paramName, int x, int y, int z, int w);
BN_API void bnSet1f(BNObject target, const char *paramName, float value);
BN_API void bnSet2f(BNObject target, const char *paramName, float x, float y);
BN_API void bnSet3f(BNObject target, const char *paramName, float x, float y, float z);
BN_API void bnSet4f(BNObject target, const char *paramName, float x, float y, float z, float w);
BN_API void bnSet4x3fv(BNObject target, const char *paramName, const BNTransform *affineMatrix);
BN_API void bnSet4x4fv(BNObject target, const char *paramName, const bn_float4 *xfm);

/* barney.h:346-400 */
BN_API void bnBuild(BNModel model, int whichDataSlot);
BN_API void bnAccumReset(BNFrameBuffer fb);
BN_API void bnFrameBufferResize(BNFrameBuffer fb, BNDataType colorFormat, int sizeX, int sizeY,
                                uint32_t requiredChannels = BN_FB_COLOR);
BN_API void bnFrameBufferRead(BNFrameBuffer fb, BNFrameBufferChannel channelToRead, void *pointerToReadDataInto,
                              BNDataType requiredFormat);
BN_API void bnFrameBufferGetSize(BNFrameBuffer fb, int *sizeX, int *sizeY);
BN_API void bnRender(BNRenderer renderer, BNModel model, BNCamera camera, BNFrameBuffer fb);
BN_API void bnSetInstances(BNModel model, int whichSlot, BNGroup *groupsToInstantiate,
                           BNTransform *instanceTransforms, int numInstances);

/* barney.h:421-425, 454-459, 499-519 */
BN_API BNData bnDataCreate(BNContext context, int whichSlot, BNDataType dataType, size_t numItems, const void *items);
BN_API BNGroup bnGroupCreate(BNContext context, int whichSlot, BNGeom *geoms, int numGeoms, BNVolume *volumes,
                             int numVolumes);
BN_API void bnGroupBuild(BNGroup group);
BN_API BNGeom bnGeometryCreate(BNContext context, int whichSlot, const char *type);
BN_API BNMaterial bnMaterialCreate(BNContext context, int whichSlot, const char *type);

/* barney.h:449-475: lights ("envmap" with "texture", "direction", "up",
 * "scale"; attached to a group as bnSetData(group, "lights", BN_OBJECT
 * data)) and 2D textures (texelFormat BN_FLOAT32_VEC4, BN_FLOAT32,
 * BN_UFIXED8_RGBA or BN_UFIXED8; sampled as barney's embree backend does) */
BN_API BNLight bnLightCreate(BNContext context, int whichSlot, const char *type);
BN_API BNTexture2D bnTexture2DCreate(BNContext context, int whichSlot, BNDataType texelFormat, uint32_t size_x,
                                     uint32_t size_y, const void *texels,
                                     BNTextureFilterMode filterMode = BN_TEXTURE_LINEAR,
                                     BNTextureAddressMode addressMode_x = BN_TEXTURE_WRAP,
                                     BNTextureAddressMode addressMode_y = BN_TEXTURE_WRAP,
                                     BNTextureColorSpace colorSpace = BN_COLOR_SPACE_LINEAR);

/* barney.h:440-446, 507-519: volumes. texelFormat BN_FLOAT32; the
 * "structured" field takes "textureData", "gridOrigin", "gridSpacing"
 * (StructuredData.cu:130-160); "unstructured" takes "vertex.position",
 * "vertex.data", "index", "cell.index", "cell.type" (UMeshField.cu:204-236;
 * per-vertex scalars); "BlockStructuredAMR" takes "grid.origins",
 * "grid.dims", "grid.levels", "grid.offsets", "scalars",
 * "level.refinements" (BlockStructuredField.cu:225-256); "NanoVDB" takes
 * "data", the grid buffer as BN_UINT8 data (NanoVDB.cu:190-235) */
BN_API BNTextureData bnTextureData3DCreate(BNContext context, int whichSlot, BNDataType texelFormat, int width,
                                           int height, int depth, const void *items);
BN_API BNScalarField bnScalarFieldCreate(BNContext context, int whichSlot, const char *type);
BN_API BNVolume bnVolumeCreate(BNContext context, int whichSlot, BNScalarField sf);
BN_API void bnVolumeSetXF(BNVolume volume, bn_float2 domain, const bn_float4 *colorMap, int numColorMapValues,
                          float densityAt1);

#endif /* BN_API_H */
