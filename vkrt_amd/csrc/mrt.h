/*
 * Internal C ABI between the C host layer (csrc/host) and the gfx950 kernels
 * (csrc/hip). Plain pointers and sizes only. Every function returns 0 on
 * success and a negative VKRT_Result value on failure; a message describing
 * the last failure is available from mrt_last_error().
 *
 * This seam replaces the reference's Vulkan command recording: the BLAS/TLAS
 * builds (src/core/render/accel/blas.c:222-262, tlas.c:535-559) become
 * mrt_upload_geometry + mrt_build_bvh, and the main trace pass
 * (src/core/runtime/command/record.c:448-486) becomes mrt_render_frame.
 */
#ifndef VKRT_MRT_H
#define VKRT_MRT_H

#include "vkrt_shared.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrt_context mrt_context;

const char* mrt_last_error(void);
int mrt_device_count(void);
int mrt_create(int deviceIndex, mrt_context** out);
void mrt_destroy(mrt_context* ctx);
int mrt_device_synchronize(mrt_context* ctx);
int mrt_device_name(const mrt_context* ctx, char* out, size_t len, uint32_t* vendorId, uint32_t* driverVersion);

typedef struct mrt_geometry_desc {
    const ShaderVertex* vertices;
    uint32_t vertexCount;
    const uint32_t* indices;
    uint32_t indexCount;
    const float* worldTriangles; /* 9 floats per global triangle */
    const uint32_t* triangleMesh;
    uint32_t triangleCount;
    const uint32_t* meshTriangleBase; /* first global triangle of each mesh */
    uint32_t meshCount;
} mrt_geometry_desc;
int mrt_upload_geometry(mrt_context* ctx, const mrt_geometry_desc* desc);

typedef struct mrt_bvh_stats {
    float buildMs;
    uint32_t nodeCount;
    uint32_t leafTriangles;
    uint32_t maxDepth;
} mrt_bvh_stats;
int mrt_build_bvh(mrt_context* ctx, mrt_bvh_stats* stats);

int mrt_upload_meshes(mrt_context* ctx, const MeshInfo* meshInfos, uint32_t meshCount);
int mrt_upload_materials(mrt_context* ctx, const Material* materials, uint32_t materialCount);
/* Per-triangle alpha words (bit 31: the mesh runs the alpha test; bits 0-23:
 * vertex colour alphas); NULL turns the alpha test off. */
int mrt_upload_alpha(mrt_context* ctx, const uint32_t* triangleAlpha, uint32_t triangleCount);
/* Material texture pool: texelCount linear RGBA32F texels and 4 words per
 * texture (first texel, width, height, 0); no texels turns texturing off. */
int mrt_upload_textures(mrt_context* ctx, const float* texels, uint64_t texelCount, const uint32_t* info,
                        uint32_t textureCount);

typedef struct mrt_light_desc {
    const EmissiveMesh* emissiveMeshes;
    uint32_t emissiveMeshCount;
    const EmissiveTriangle* emissiveTriangles;
    uint32_t emissiveTriangleCount;
    const float* meshAliasQ;
    const uint32_t* meshAliasIdx;
    const float* triAliasQ;
    const uint32_t* triAliasIdx;
    const uint32_t* emissiveMeshMaterial; /* material index of each emissive mesh */
} mrt_light_desc;
int mrt_upload_lights(mrt_context* ctx, const mrt_light_desc* desc);

typedef struct mrt_environment_desc {
    const float* pixels; /* RGBA32F, width*height*4, NULL clears */
    uint32_t width;
    uint32_t height;
    const float* aliasQ;   /* width*height entries: importance alias table over texels */
    const uint32_t* aliasIdx;
    const float* pdf;      /* width*height texel selection probabilities */
} mrt_environment_desc;
int mrt_upload_environment(mrt_context* ctx, const mrt_environment_desc* desc);

int mrt_upload_rgb2spec(mrt_context* ctx, const float* table, uint32_t floatCount);

typedef struct mrt_render_params {
    uint32_t width;  /* accumulation extent */
    uint32_t height;
    uint32_t rngPolicy;
    uint64_t rngSeed;
    uint32_t envImportance;
    uint32_t hasTransmissive; /* any material with transmission > 0 */
    uint32_t resetAccumulation;
    uint32_t traversalStatistics;
    /* neural radiance cache (RGB mode) */
    uint32_t nrcEnabled;
    uint32_t nrcPrefix;
    uint32_t nrcTrainingPaths;
    uint32_t nrcTrainingSteps;
    uint32_t nrcBatch;
    uint32_t nrcTrain;
    float nrcLearningRate;
    float sceneMin[3]; /* scene box, for the cache's position input */
    float sceneMax[3];
    uint32_t materialSort; /* shade each depth's hits grouped by material class */
    uint32_t features;     /* accumulate the denoiser albedo / normal feature images */
    uint32_t tailRays;     /* fused tail from the first depth whose queue held at most this many rays
                              in the previous frame (0: every depth is its own pair of launches) */
} mrt_render_params;

typedef struct mrt_render_stats {
    uint64_t extensionRays;
    uint64_t shadowRays;
    uint64_t paths;
    float traceMs;
    float extendMs;
    float shadowMs;
    float shadeMs;
    uint32_t extendLaunches;
    uint32_t shadeLaunches;
    uint32_t shadowLaunches;
    uint64_t nodeVisits;
    uint64_t triangleTests;
    uint32_t nrcQueries;
    uint32_t nrcRecords;
    float nrcLoss;
    float nrcInferMs;
    float nrcTrainMs;
    float extendMsDepth0;  /* the part of extendMs / shadeMs spent on the paths' first rays */
    float shadeMsDepth0;
    uint32_t tailStartDepth;   /* depth from which the fused tail kernel carried the paths (0: none) */
    uint32_t tailLaunches;
    float tailMs;
    uint64_t tailRays;         /* extension rays the fused tail traced */
    uint64_t raysBeforeTail;   /* extension rays of the depths before tailStartDepth (all without a tail) */
    uint64_t nodeVisitsBeforeTail;    /* their node visits and triangle tests, with statistics enabled */
    uint64_t triangleTestsBeforeTail; /* (a statistics frame records them at the depth a tail would start) */
} mrt_render_stats;

/* One frame: samplesPerPixel paths for every viewport pixel, accumulated into
 * the running mean exactly as src/shaders/integrator/path/writeback.slang:72-123. */
int mrt_render_frame(mrt_context* ctx, const SceneData* scene, const mrt_render_params* params,
                     mrt_render_stats* stats);
int mrt_resize_accumulation(mrt_context* ctx, uint32_t width, uint32_t height);
int mrt_clear_accumulation(mrt_context* ctx);
int mrt_read_accumulation(mrt_context* ctx, float* out, size_t floatCount);
int mrt_read_output(mrt_context* ctx, uint16_t* out, size_t count);
/* Feature images of the accumulation extent, RGBA32F each: albedo and shading
 * normal running means with the feature weight in .w
 * (src/shaders/integrator/path/writeback.slang:87-114). Fails before the
 * first frame that computed them. */
int mrt_read_features(mrt_context* ctx, float* albedo, float* normal, size_t floatCount);

/* Neural radiance cache network (nrc.hip). Inputs are 11 floats per sample in
 * [0, 1] (position 3, direction theta/pi phi/2pi, normal likewise, roughness,
 * albedo 3); outputs 3 floats. Weights are the natural fp32 parameters:
 * W0 [64][192], W1..W4 [64][64], W5 [16][64] (rows 3..15 zero). */
int mrt_nrc_parameter_count(void);
int mrt_nrc_reset(mrt_context* ctx, uint32_t seed);
int mrt_nrc_get_weights(mrt_context* ctx, float* out, size_t count);
int mrt_nrc_set_weights(mrt_context* ctx, const float* weights, size_t count);
int mrt_nrc_infer(mrt_context* ctx, const float* inputs, uint32_t n, float* outputs);
int mrt_nrc_train(mrt_context* ctx, const float* inputs, const float* targets, uint32_t n, uint32_t steps,
                  uint32_t batch, float learningRate, float* lossOut);
int mrt_nrc_gradient(mrt_context* ctx, const float* inputs, const float* targets, uint32_t n, float* grad,
                     float* lossOut);

/* Path-replay backpropagation of one frame (RGB mode): gradients of
 * sum_px dImage[px] . I[px], where I is the frame the same parameters and
 * frame number render, with respect to material base colour, roughness and
 * emission and to environment texels. dImage holds width*height*3 floats of
 * the accumulation extent; gradient arrays (materialCount*3, materialCount,
 * materialCount*3, envWidth*envHeight*3) may be NULL. */
typedef struct mrt_gradients {
    float* baseColor;
    float* roughness;
    float* emission;
    uint32_t materialCount;
    float* environment;
    uint32_t environmentWidth;
    uint32_t environmentHeight;
} mrt_gradients;
int mrt_render_backward(mrt_context* ctx, const SceneData* scene, const mrt_render_params* params,
                        const float* dImage, mrt_gradients* grads, mrt_render_stats* stats);

int mrt_trace_rays(mrt_context* ctx, const float* rays, size_t count, uint32_t* outMesh, uint32_t* outPrim,
                   float* outT, float* outBary);
int mrt_trace_shadow_rays(mrt_context* ctx, const float* rays, size_t count, uint32_t* outVisibility);
int mrt_primary_hits(mrt_context* ctx, const SceneData* scene, const mrt_render_params* params, uint32_t sampleIndex,
                     uint32_t* outMesh, uint32_t* outPrim);

/* mrt_trace_rays on device memory with no allocation: rays and outputs are
 * device pointers; the traversal stacks and the overflow counter live in a
 * fixed device array of the library. Calls that overlap on one device race
 * on that array. At most MRT_NOALLOC_MAX_RAYS rays per call. */
#define MRT_NOALLOC_MAX_RAYS (1ull << 30)
int mrt_trace_rays_noalloc(mrt_context* ctx, const float* deviceRays, size_t count, uint32_t* deviceMesh,
                           uint32_t* devicePrim, float* deviceT, float* deviceBary);

#ifdef __cplusplus
}
#endif

#endif
