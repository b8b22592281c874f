/*
 * Host-side state of one VKRT instance. Mirrors the reference's split between
 * scene settings (VKRT_SceneSettingsSnapshot), the SceneData uniform and the
 * per-mesh / per-material records (src/core/internal/vkrt_internal.h), with the
 * Vulkan runtime replaced by one mrt_context on a HIP device.
 */
#ifndef VKRT_HOST_INTERNAL_H
#define VKRT_HOST_INTERNAL_H

#include "../mrt.h"
#include "vkrt.h"

#include <stdint.h>
#include <stdio.h>

#define VKRT_LOG_ERROR(...)                   \
    do {                                      \
        fprintf(stderr, "[vkrt error] ");     \
        fprintf(stderr, __VA_ARGS__);         \
        fprintf(stderr, "\n");                \
    } while (0)

typedef struct HostMesh {
    Vertex* vertices;
    uint32_t* indices;
    MeshInfo info;
    float worldTransform[4][4]; /* column-major, as the reference's Mesh.worldTransform */
    uint32_t materialIndex;
    uint8_t hasMaterialAssignment;
    char name[VKRT_NAME_LEN];
} HostMesh;

typedef struct HostMaterial {
    Material material;
    char name[VKRT_NAME_LEN];
} HostMaterial;

typedef struct HostTexture {
    float* rgba; /* converted to linear RGBA32F */
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t colorSpace;
    char name[VKRT_NAME_LEN];
} HostTexture;

/* Packed buffers as the kernels consume them (rebuilt on scene changes). */
typedef struct PackedScene {
    ShaderVertex* vertices;
    uint32_t vertexCount;
    uint32_t* indices;
    uint32_t indexCount;
    MeshInfo* meshInfos;
    Material* materials;
    uint32_t materialCount;
    EmissiveMesh* emissiveMeshes;
    uint32_t emissiveMeshCount;
    EmissiveTriangle* emissiveTriangles;
    uint32_t emissiveTriangleCount;
    float* meshAliasQ;
    uint32_t* meshAliasIdx;
    float* triAliasQ;
    uint32_t* triAliasIdx;
    uint32_t* emissiveMeshMaterial; /* material of each emissive mesh (for emission gradients) */
    float* worldTriangles;
    uint32_t* triangleMesh;
    uint32_t* meshTriangleBase;
    uint32_t triangleCount;
    float* envAliasQ;
    uint32_t* envAliasIdx;
    float* envPdf;
    uint32_t hasTransmissive;
    /* per global triangle, only when some mesh may reject hits (alpha test):
     * bit 31 = the triangle's mesh may reject a hit, bits 0-23 = the alpha bytes
     * of its three vertex colours; NULL when every mesh is opaque */
    uint32_t* triangleAlpha;
    /* material texture pool: linear RGBA32F texels of every texture a material
     * references, and per texture {first texel, width, height, 0} (width 0 for
     * textures no material uses) */
    float* textureTexels;
    uint64_t textureTexelCount;
    uint32_t* textureInfo;
    uint32_t textureCount;
} PackedScene;

struct VKRT {
    VKRT_SceneSettingsSnapshot settings;
    /* auto-SPP controller (src/core/scene/timing.c:44-178): smoothed trace
     * time per sample, and the frame budget set through the target FPS */
    float autoSPPControlMs;
    float autoSPPTargetFrameMs;
    SceneData sceneData;
    uint32_t renderWidth;
    uint32_t renderHeight;
    uint32_t displayViewport[4];
    uint8_t headless;
    uint8_t shouldDeinit;

    HostMesh* meshes;
    uint32_t meshCount;
    uint32_t meshCapacity;
    HostMaterial* materials;
    uint32_t materialCount;
    uint32_t materialCapacity;
    HostTexture* textures;
    uint32_t textureCount;
    uint32_t textureCapacity;

    PackedScene packed;
    float* rgb2spec;
    uint32_t rgb2specFloatCount;

    /* revision counters, compared with the uploaded revision like the
     * reference's *ResourceRevision pairs (src/core/api/frame.c:152-163) */
    uint32_t geometryRevision, geometryUploaded;
    uint32_t materialRevision, materialUploaded;
    uint32_t environmentRevision, environmentUploaded;
    uint8_t accumulationNeedsReset;

    uint32_t rngPolicy;
    uint64_t rngSeed;
    uint8_t envImportance;
    uint8_t traversalStatistics;

    VKRT_RenderStatusSnapshot renderStatus;
    uint8_t frameTraced;
    uint8_t finalImageDenoiseEnabled; /* src/core/api/lifecycle.c:338 (default on) */

    mrt_context* device;
    VKRT_SystemInfo systemInfo;
    VKRT_TraceStats stats;
    VKRT_NeuralCacheSettings nrc;
    VKRT_NeuralCacheStats nrcStats;
    float sceneMin[3], sceneMax[3]; /* box of the uploaded geometry */
    uint8_t materialSort;
    uint32_t fusedTailRays; /* VKRT_setFusedTail */
    uint8_t denoiserFeatures; /* accumulate the albedo / normal feature images */
};

/* scene preparation (scene_prep.c) */
VKRT_Result vkrtPrepareScene(VKRT* vkrt);
void vkrtFreePackedScene(PackedScene* packed);
int vkrtBuildAliasTable(const float* pmf, uint32_t count, float* outQ, uint32_t* outIdx);
ShaderVertex vkrtPackShaderVertex(const Vertex* vertex);
uint32_t vkrtPackOctNormal32(const float normal[3]);
uint32_t vkrtPackTangent32(const float tangent[4]);
uint32_t vkrtPackColorRGBA8(const float color[4]);

/* camera / transforms (camera_math.c) */
void vkrtSyncCameraMatrices(VKRT* vkrt);
void vkrtMeshWorldTransformFromInfo(const MeshInfo* info, float out[4][4]);

/* settings (api_settings.c) */
void vkrtResetSceneData(VKRT* vkrt);
void vkrtResetAutoSPPState(VKRT* vkrt, int resetSamplesPerPixel);
void vkrtUpdateAutoSPP(VKRT* vkrt);
void vkrtWriteSceneStateUniform(VKRT* vkrt);
void vkrtInitDefaultSettings(VKRT* vkrt, uint32_t width, uint32_t height);

/* rgb2spec table (rgb2spec_table.c) */
VKRT_Result vkrtLoadRGB2SpecTable(VKRT* vkrt);
int vkrtParseRGB2Spec(const uint8_t* data, size_t size, RGB2SpecTableInfo* info, const float** payload,
                      size_t* floatCount);

#endif
