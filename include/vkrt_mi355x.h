/*
 * Entry points that the MI355X build adds beside the reference API.
 * None of these has a counterpart in src/core/api/vkrt.h; each comment names
 * the reference code whose behaviour it exposes or extends.
 */
#ifndef VKRT_MI355X_H
#define VKRT_MI355X_H

#include "vkrt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Random-number policy for the per-path streams.
 * VKRT_RNG_HASH reproduces src/shaders/sampling/random.slang:7-27 (the default).
 * VKRT_RNG_PCG32 uses a PCG32 stream per path, seeded by `seed` and selected by
 * (pixel, frame, sample); each draw keeps the top 24 bits like random.slang:16-19. */
#define VKRT_RNG_HASH 0u
#define VKRT_RNG_PCG32 1u
VKRT_Result VKRT_setRandomPolicy(VKRT* vkrt, uint32_t policy, uint64_t seed);

/* Environment-map next-event estimation with an alias-table sampler and MIS
 * (power heuristic) against BSDF sampling. The reference only gathers the
 * environment when a BSDF ray misses (src/shaders/light/environment.slang:16-27);
 * enabling this keeps the estimator unbiased but changes the random sequence,
 * so only converged images compare with the reference. Default off. */
VKRT_Result VKRT_setEnvironmentImportanceSampling(VKRT* vkrt, uint8_t enabled);

/* Selects an already-added linear texture as the environment map (the
 * reference's internal replaceEnvironmentTexture, src/core/scene/environment.c:11-41,
 * reachable there only through a file path). */
VKRT_Result VKRT_setEnvironmentTexture(VKRT* vkrt, uint32_t textureIndex);

/* Reads back the accumulation buffer (RGBA32F, .w = sample count, as
 * writeback.slang:110): `outRGBA` holds width*height*4 floats. In spectral
 * modes .xyz is CIE XYZ (equal-energy white), as the reference stores it. */
VKRT_Result VKRT_readAccumulation(const VKRT* vkrt, float* outRGBA, size_t floatCount);

/* Reads back the display image (RGBA16 UNORM, writeback.slang:113). */
VKRT_Result VKRT_readOutputImage(const VKRT* vkrt, uint16_t* outRGBA, size_t count);

typedef struct VKRT_TraceStats {
    uint64_t extensionRays; /* one per path per depth iteration = "bounces" */
    uint64_t shadowRays;
    uint64_t paths;
    float traceMs;          /* wall time of the last VKRT_trace, device-synchronised */
    float extendMs;         /* summed closest-hit traversal kernel time of the last trace */
    float shadowMs;         /* summed shadow traversal kernel time */
    float shadeMs;          /* summed shading kernel time */
    uint32_t extendLaunches;
    float bvhBuildMs;       /* last BVH build, device time */
    uint32_t bvhNodeCount;
    uint32_t bvhLeafTriangles;
    uint32_t bvhMaxDepth;
    uint32_t triangleCount;
    uint32_t shadeLaunches;
    uint32_t shadowLaunches;
    uint64_t nodeVisits;    /* closest-hit traversal only, when statistics are enabled */
    uint64_t triangleTests; /* closest-hit traversal only, when statistics are enabled */
    float backwardMs;       /* device time of the last VKRT_renderBackward (primal + replay) */
    float extendMsDepth0;   /* the part of extendMs spent on the primary rays (depth 0) */
    float shadeMsDepth0;    /* the part of shadeMs spent on the primary hits and misses (depth 0) */
    uint32_t tailStartDepth; /* depth from which the fused tail carried the paths (0: none) */
    uint32_t tailLaunches;
    float tailMs;            /* fused tail kernel time; not part of extendMs or shadeMs */
    uint64_t tailRays;       /* extension rays the fused tail traced (part of extensionRays) */
    uint64_t raysBeforeTail; /* extension rays traced before tailStartDepth (all of them without a tail) */
    uint64_t nodeVisitsBeforeTail;    /* nodeVisits and triangleTests of those rays, with statistics */
    uint64_t triangleTestsBeforeTail; /* enabled (a statistics frame runs no tail but records them) */
} VKRT_TraceStats;
VKRT_Result VKRT_getTraceStats(const VKRT* vkrt, VKRT_TraceStats* outStats);

/* Blocks until every kernel and copy the instance has queued on its device has
 * finished (hipDeviceSynchronize); the reference waits on its frame fences
 * instead (src/core/api/frame.c:74-105). */
VKRT_Result VKRT_deviceSynchronize(VKRT* vkrt);

/* Counts BVH nodes visited and triangles tested by every closest-hit query of
 * the following frames (an instrumented traversal; slower, for measurement). */
VKRT_Result VKRT_setTraversalStatistics(VKRT* vkrt, uint8_t enabled);

/* Gradients returned by VKRT_renderBackward. Arrays are caller-owned; a NULL
 * array is not computed. materialCount must equal VKRT_getMaterialCount. */
typedef struct VKRT_Gradients {
    float* baseColor;          /* materialCount x 3: d loss / d Material.baseColor */
    float* roughness;          /* materialCount: d loss / d Material.roughness */
    float* emission;           /* materialCount x 3: d loss / d (Material.emissionColor * emissionLuminance) */
    uint32_t materialCount;
    float* environment;        /* environmentWidth x environmentHeight x 3: d loss / d environment texel RGB */
    uint32_t environmentWidth; /* must match the environment texture, else environment is skipped */
    uint32_t environmentHeight;
} VKRT_Gradients;

/* Path-replay backpropagation (render_backward): gradients of
 * sum_px dLossdImage[px] . I[px] with respect to material base colour,
 * roughness and emission and to environment texels, where I is the frame of
 * samplesPerPixel samples that the next VKRT_draw would add with the current
 * settings (same random streams, so finite differences of VKRT_draw images
 * share random numbers with it). dLossdImage holds renderWidth x renderHeight
 * x 3 floats. RGB render mode only. Paths are replayed, not stored; visibility
 * edges are not differentiated. Does not advance the frame. No counterpart in
 * the reference (src/core/api/vkrt.h has no differentiable rendering). */
VKRT_Result VKRT_renderBackward(VKRT* vkrt, const float* dLossdImage, VKRT_Gradients* outGradients);

/* Shading order (an addition): when enabled, the hits of every depth are
 * partitioned by material class -- transmissive materials after the others,
 * each class in queue order -- before the shading kernel runs, so a wavefront
 * runs one family of BSDF lobes. Results are unchanged (every path keeps its
 * random stream); only the order of the shading work changes. Default off:
 * on the 1080p spectral workload the partition and the indirect, strided
 * reads of the sorted order cost more than the divergence they remove. */
VKRT_Result VKRT_setMaterialSorting(VKRT* vkrt, uint8_t enabled);

/* Fused tail (an addition): once the queues of a frame have become short, one
 * launch carries every remaining path to its end, tracing and shading vertex
 * after vertex per lane, instead of a traversal and a shading launch per depth
 * that each wait for their slowest ray. It takes over at the first depth whose
 * queue held at most maxRays rays in the previous frame of the same size and
 * depth, in frames of one chunk whose shading needs none of shadow rays, the
 * neural cache, feature images, material sorting, the alpha test, material
 * textures or sheen; other frames keep one launch pair per depth. Results are
 * unchanged (every path keeps its random stream and arithmetic). 0 turns it
 * off; default 2^20 rays (about two per lane of the traversal kernel's grid). */
VKRT_Result VKRT_setFusedTail(VKRT* vkrt, uint32_t maxRays);

/* Denoiser feature images. The reference always writes its albedo and normal
 * outputs for the denoiser (src/shaders/integrator/path/loop.slang:83-111,
 * writeback.slang:87-114); here they cost a pass per depth, so they are off
 * until turned on (the denoiser debug views turn the pass on by themselves).
 * Turning them on restarts accumulation. The images are RGBA32F of the render
 * extent: the running mean of the first non-specular-follow surface's albedo
 * and shading normal, with the accumulated feature weight in .w. */
VKRT_Result VKRT_setDenoiserFeatures(VKRT* vkrt, uint8_t enabled);
VKRT_Result VKRT_readDenoiserFeatures(const VKRT* vkrt, float* outAlbedo, float* outNormal, size_t floatCount);

/* Neural radiance cache (an addition; the reference has no neural cache).
 * RGB render mode and spectral mode with hero wavelength sampling; in the
 * latter the cache's three outputs are CIE XYZ radiance, which an image path
 * uplifts to its four wavelengths through the rgb2spec table (a path whose
 * wavelengths collapsed at a dispersive surface is traced on without the
 * cache). Single-wavelength spectral mode ignores the cache. When enabled,
 * VKRT_draw shades the first
 * prefixBounces bounces of every path as usual and then replaces the rest of
 * the path with the cache's prediction of the radiance scattered at the next
 * surface (emission there is still added); trainingPaths extra paths per frame
 * through random pixels are traced in full, without the cache, and every
 * surface vertex on them becomes a training record (the radiance the rest of
 * that path carried back, divided by its throughput). After the frame the cache
 * takes trainingSteps Adam steps of batchSize records each (relative L2 loss).
 * The network: 12-frequency encoding of position, direction and normal plus
 * roughness and albedo (192 inputs), 5 hidden layers of 64 ReLU units, 3
 * outputs; fp16 weights on MFMA with fp32 accumulation. */
typedef struct VKRT_NeuralCacheSettings {
    uint32_t enabled;
    uint32_t prefixBounces;   /* bounces traced before the cache is queried (1) */
    uint32_t trainingPaths;   /* full paths per frame that produce training records (65536) */
    uint32_t trainingSteps;   /* optimiser steps per frame (1) */
    uint32_t batchSize;       /* records per step (16384) */
    float learningRate;       /* Adam step size (1e-2) */
    uint32_t train;           /* 0 keeps the weights fixed */
} VKRT_NeuralCacheSettings;

typedef struct VKRT_NeuralCacheStats {
    float lastLoss;            /* mean relative L2 loss of the last frame's training steps */
    uint32_t queries;          /* cache queries of the last frame */
    uint32_t trainingRecords;  /* training records of the last frame */
    float inferMs;             /* device time of the last frame's inference */
    float trainMs;             /* device time of the last frame's training */
} VKRT_NeuralCacheStats;

void VKRT_defaultNeuralCacheSettings(VKRT_NeuralCacheSettings* outSettings);
VKRT_Result VKRT_setNeuralCache(VKRT* vkrt, const VKRT_NeuralCacheSettings* settings);
VKRT_Result VKRT_getNeuralCacheStats(const VKRT* vkrt, VKRT_NeuralCacheStats* outStats);

/* Direct access to the cache network (tests and tools). Inputs: 11 floats per
 * sample in [0, 1] -- position in the scene box (3), outgoing direction as
 * theta/pi and phi/(2 pi) (2), shading normal likewise (2), roughness, albedo
 * (3); outputs: 3 floats. Weights: the fp32 master parameters in the order
 * W0 [64][192], W1..W4 [64][64], W5 [16][64] (rows 3..15 are zero). */
uint32_t VKRT_neuralCacheParameterCount(void);
VKRT_Result VKRT_neuralCacheReset(VKRT* vkrt, uint32_t seed);
VKRT_Result VKRT_neuralCacheGetWeights(VKRT* vkrt, float* outWeights, size_t count);
VKRT_Result VKRT_neuralCacheSetWeights(VKRT* vkrt, const float* weights, size_t count);
VKRT_Result VKRT_neuralCacheInfer(VKRT* vkrt, const float* inputs, uint32_t count, float* outputs);
VKRT_Result VKRT_neuralCacheTrain(VKRT* vkrt, const float* inputs, const float* targets, uint32_t count,
                                  uint32_t steps, uint32_t batchSize, float learningRate, float* outLoss);
VKRT_Result VKRT_neuralCacheGradient(VKRT* vkrt, const float* inputs, const float* targets, uint32_t count,
                                     float* outGradient, float* outLoss);

/* Host copies of the packed scene buffers exactly as the kernels consume them
 * (the reference's descriptor bindings 10-24, src/shaders/scene/resources.slang:7-59).
 * Pointers stay valid until the next scene change or VKRT_destroy. */
typedef struct VKRT_SceneBuffers {
    const SceneData* sceneData;
    const ShaderVertex* vertices;
    uint32_t vertexCount;
    const uint32_t* indices;
    uint32_t indexCount;
    const MeshInfo* meshInfos;
    uint32_t meshCount;
    const Material* materials;
    uint32_t materialCount;
    const EmissiveMesh* emissiveMeshes;
    const EmissiveTriangle* emissiveTriangles;
    const float* meshAliasQ;
    const uint32_t* meshAliasIdx;
    const float* triAliasQ;
    const uint32_t* triAliasIdx;
    const float* rgb2specTable; /* res scale floats then 3*res^3*3 coefficients */
    uint32_t rgb2specFloatCount;
    const float* environmentPixels; /* RGBA32F, NULL if none */
    uint32_t environmentWidth;
    uint32_t environmentHeight;
    const float* worldTriangles;    /* 9 floats per triangle: world-space v0,v1,v2 in global order */
    uint32_t triangleCount;
    const uint32_t* triangleMesh;   /* global triangle -> mesh index */
    uint32_t rngPolicy;
    uint64_t rngSeed;
    uint32_t environmentImportanceSampling;
    const float* environmentAliasQ;      /* width*height entries, NULL without an environment texture */
    const uint32_t* environmentAliasIdx;
    const float* environmentPdf;         /* texel selection probabilities */
    const uint32_t* triangleAlpha;       /* per triangle: bit 31 = mesh runs the alpha test, bits 0-23 = the
                                            three vertex colour alphas (8 bits each), bit 30 = the base-colour
                                            texture's alpha takes part; NULL when all opaque */
    const float* textureTexels;          /* material texture pool, linear RGBA32F; NULL without material textures */
    uint64_t textureTexelCount;
    const uint32_t* textureInfo;         /* per texture: first texel, width, height, 0 (width 0: unused) */
    uint32_t textureCount;
} VKRT_SceneBuffers;
VKRT_Result VKRT_exportSceneBuffers(VKRT* vkrt, VKRT_SceneBuffers* out);

/* Closest-hit queries through the GPU traversal for caller-supplied rays.
 * rays: count * 8 floats (ox,oy,oz,tmin, dx,dy,dz,tmax).
 * Outputs per ray: mesh (instance) index, primitive index within the mesh,
 * hit distance and barycentrics (u,v) -- the payload of the reference's
 * closest-hit shader (src/shaders/entry/path/closest_hit.slang:3-5);
 * a miss reports VKRT_INVALID_INDEX. The closest hit is the smallest distance
 * and, among triangles at exactly the same distance (a ray through a shared
 * edge), the one with the lowest global triangle index, whatever the tree's
 * visiting order; the reference's driver traversal leaves that choice open. */
VKRT_Result VKRT_traceRays(VKRT* vkrt, const float* rays, size_t count, uint32_t* outMesh, uint32_t* outPrimitive,
                           float* outT, float* outBarycentrics);

/* Primary-hit query for every pixel of the current viewport at a given sample
 * index: the first TraceRay of the reference's path loop
 * (src/shaders/integrator/path/state.slang:140-146, loop.slang:8-17), camera
 * jitter drawn from the configured RNG policy. */
VKRT_Result VKRT_tracePrimaryHits(VKRT* vkrt, uint32_t sampleIndex, uint32_t* outMesh, uint32_t* outPrimitive);

/* Occlusion queries (any hit) for caller-supplied rays, classified like the
 * reference's shadow closest-hit / miss shaders
 * (src/shaders/entry/shadow/closest_hit.slang:4-8): 0 occluded, 1 visible,
 * 2 occluded by a transmissive material. */
VKRT_Result VKRT_traceShadowRays(VKRT* vkrt, const float* rays, size_t count, uint32_t* outVisibility);

/* VKRT_traceRays on device memory, allocating nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). deviceRays, deviceMesh, devicePrimitive and the optional
 * deviceT / deviceBarycentrics are device pointers on the context's GPU,
 * laid out as in VKRT_traceRays; results are bit-identical to it. The
 * traversal stacks and the overflow counter live in one fixed device array of
 * the library (about 61 MB per GPU, held until the process ends) shared by
 * every call on that GPU: calls that overlap on one GPU race. At most 2^30
 * rays per call (VKRT_ERROR_INVALID_ARGUMENT beyond). The scene must already
 * be on the device (VKRT_ERROR_OPERATION_FAILED when it changed since the
 * last upload, as uploading allocates). Returns after the results are
 * written. */
VKRT_Result VKRT_traceRaysDeviceNoAlloc(VKRT* vkrt, const float* deviceRays, size_t count, uint32_t* deviceMesh,
                                        uint32_t* devicePrimitive, float* deviceT, float* deviceBarycentrics);

#ifdef __cplusplus
}
#endif

#endif
