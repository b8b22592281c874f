/*
 * Public C API of the MI355X renderer: the drop-in boundary for the reference's
 * trace step. Every function here keeps the name, argument meaning and error
 * behaviour of the reference's public API (src/core/api/vkrt.h:10-117); the
 * line cited beside each declaration is the reference declaration it replaces.
 * cglm's vec3 / mat4 parameters are spelled as the plain arrays they decay to,
 * so callers that pass cglm values compile unchanged.
 *
 * The reference records vkCmdTraceRaysKHR against a driver-built BVH
 * (src/core/runtime/command/record.c:448-486); here VKRT_trace launches the
 * gfx950 wavefront pipeline and VKRT_updateScene builds the BVH on the GPU.
 */
#ifndef VKRT_H
#define VKRT_H

#ifdef __cplusplus
extern "C" {
#endif

#include "vkrt_types.h"

#include <stddef.h>

/* lifecycle: vkrt.h:10-17 */
void VKRT_defaultCreateInfo(VKRT_CreateInfo* createInfo);                         /* vkrt.h:10 */
VKRT_Result VKRT_create(VKRT** outVkrt);                                           /* vkrt.h:11 */
void VKRT_destroy(VKRT* vkrt);                                                     /* vkrt.h:12 */
VKRT_Result VKRT_initWithCreateInfo(VKRT* vkrt, const VKRT_CreateInfo* createInfo); /* vkrt.h:13 */
VKRT_Result VKRT_init(VKRT* vkrt);                                                 /* vkrt.h:14 */
void VKRT_deinit(VKRT* vkrt);                                                      /* vkrt.h:15 */
int VKRT_shouldDeinit(VKRT* vkrt);                                                 /* vkrt.h:16 */
void VKRT_poll(VKRT* vkrt);                                                        /* vkrt.h:17 */

/* frame loop: vkrt.h:19-25 */
VKRT_Result VKRT_beginFrame(VKRT* vkrt);  /* vkrt.h:19 */
VKRT_Result VKRT_updateScene(VKRT* vkrt); /* vkrt.h:20 */
VKRT_Result VKRT_trace(VKRT* vkrt);       /* vkrt.h:21 */
VKRT_Result VKRT_present(VKRT* vkrt);     /* vkrt.h:22 */
VKRT_Result VKRT_endFrame(VKRT* vkrt);    /* vkrt.h:23 */
VKRT_Result VKRT_draw(VKRT* vkrt);        /* vkrt.h:25 */

/* geometry: vkrt.h:27-35 */
VKRT_Result VKRT_uploadMeshData(VKRT* vkrt, const Vertex* vertices, size_t vertexCount, const uint32_t* indices,
                                size_t indexCount);                                                    /* vkrt.h:27 */
VKRT_Result VKRT_uploadMeshDataBatch(VKRT* vkrt, const VKRT_MeshUpload* uploads, size_t uploadCount); /* vkrt.h:34 */
VKRT_Result VKRT_removeMesh(VKRT* vkrt, uint32_t meshIndex);                                          /* vkrt.h:35 */

/* Orbit / pan / zoom of the viewport camera from one frame of mouse input;
 * src/core/scene/camera.c:90-120. */
VKRT_Result VKRT_applyCameraInput(VKRT* vkrt, const VKRT_CameraInput* input);          /* vkrt.h:36 */

/* settings: vkrt.h:37-53 */
VKRT_Result VKRT_invalidateAccumulation(VKRT* vkrt);                                  /* vkrt.h:37 */
VKRT_Result VKRT_setSamplesPerPixel(VKRT* vkrt, uint32_t samplesPerPixel);            /* vkrt.h:38 */
VKRT_Result VKRT_setPathDepth(VKRT* vkrt, uint32_t rrMinDepth, uint32_t rrMaxDepth);  /* vkrt.h:39 */
VKRT_Result VKRT_setAutoSPPEnabled(VKRT* vkrt, uint8_t enabled);                      /* vkrt.h:40 */
VKRT_Result VKRT_setAutoSPPTargetFPS(VKRT* vkrt, uint32_t targetFPS);                 /* vkrt.h:41 */
VKRT_Result VKRT_setToneMappingMode(VKRT* vkrt, VKRT_ToneMappingMode toneMappingMode); /* vkrt.h:42 */
VKRT_Result VKRT_setRenderMode(VKRT* vkrt, VKRT_RenderMode renderMode);               /* vkrt.h:43 */
VKRT_Result VKRT_setSpectralSamplingMode(VKRT* vkrt, VKRT_SpectralSamplingMode mode); /* vkrt.h:44 */
VKRT_Result VKRT_setExposure(VKRT* vkrt, float exposure);                             /* vkrt.h:45 */
VKRT_Result VKRT_setAutoExposureEnabled(VKRT* vkrt, uint8_t enabled);                 /* vkrt.h:46 */
VKRT_Result VKRT_setEnvironmentLight(VKRT* vkrt, const float color[3], float strength); /* vkrt.h:47 */
VKRT_Result VKRT_setEnvironmentRotation(VKRT* vkrt, float rotationDegrees);           /* vkrt.h:48 */
/* Loads a PNG, JPEG or EXR file as a linear texture and makes it the
 * environment map (src/core/scene/environment.c:42-60). */
VKRT_Result VKRT_setEnvironmentTextureFromFile(VKRT* vkrt, const char* path);           /* vkrt.h:49 */
VKRT_Result VKRT_clearEnvironmentTexture(VKRT* vkrt);                                 /* vkrt.h:50 */
VKRT_Result VKRT_setDebugMode(VKRT* vkrt, VKRT_DebugMode mode);                       /* vkrt.h:51 */
VKRT_Result VKRT_setMisNeeEnabled(VKRT* vkrt, uint8_t enabled);                       /* vkrt.h:52 */
VKRT_Result VKRT_setTimeRange(VKRT* vkrt, float timeBase, float timeStep);            /* vkrt.h:53 */

/* image export: vkrt.h:54-58. The path's extension picks PNG (16-bit), JPEG
 * or EXR (linear RGBA32F); no extension means PNG. Denoising is not available
 * in this build: a request for it exports the raw render. */
void VKRT_defaultRenderExportSettings(VKRT_RenderExportSettings* settings);                    /* vkrt.h:54 */
/* The flag is kept and reported as in the reference (src/core/api/render.c:27-37).
 * This build carries no denoiser, so a finished render always reaches
 * VKRT_RENDER_PHASE_COMPLETE_RAW: the outcome the reference reports when its
 * denoise step fails (src/core/utility/export/api.c:160-167). */
VKRT_Result VKRT_setRenderDenoiseEnabled(VKRT* vkrt, uint8_t enabled);                         /* vkrt.h:55 */
VKRT_Result VKRT_denoiseRenderToViewport(VKRT* vkrt);                                          /* vkrt.h:56 */
VKRT_Result VKRT_saveRenderImageEx(VKRT* vkrt, const char* path, const VKRT_RenderExportSettings* settings); /* vkrt.h:57 */
VKRT_Result VKRT_saveRenderImage(VKRT* vkrt, const char* path);                               /* vkrt.h:58 */

/* render sessions and queries: vkrt.h:59-66 */
VKRT_Result VKRT_startRender(VKRT* vkrt, uint32_t width, uint32_t height, uint32_t targetSamples); /* vkrt.h:59 */
VKRT_Result VKRT_continueRender(VKRT* vkrt, uint32_t targetSamples);                            /* vkrt.h:60 */
VKRT_Result VKRT_stopRenderSampling(VKRT* vkrt);                                                /* vkrt.h:61 */
VKRT_Result VKRT_stopRender(VKRT* vkrt);                                                        /* vkrt.h:62 */
VKRT_Result VKRT_getSceneSettings(const VKRT* vkrt, VKRT_SceneSettingsSnapshot* outSettings);   /* vkrt.h:63 */
VKRT_Result VKRT_getRenderStatus(const VKRT* vkrt, VKRT_RenderStatusSnapshot* outStatus);       /* vkrt.h:64 */
VKRT_Result VKRT_getRuntimeSnapshot(const VKRT* vkrt, VKRT_RuntimeSnapshot* outRuntime);       /* vkrt.h:65 */
VKRT_Result VKRT_getSystemInfo(const VKRT* vkrt, VKRT_SystemInfo* outSystemInfo);               /* vkrt.h:66 */

/* scene objects: vkrt.h:73-112 */
VKRT_Result VKRT_getMeshCount(const VKRT* vkrt, uint32_t* outMeshCount);                              /* vkrt.h:73 */
VKRT_Result VKRT_getMeshSnapshot(const VKRT* vkrt, uint32_t meshIndex, VKRT_MeshSnapshot* outMesh);   /* vkrt.h:74 */
VKRT_Result VKRT_getMaterialCount(const VKRT* vkrt, uint32_t* outMaterialCount);                      /* vkrt.h:75 */
VKRT_Result VKRT_getMaterialSnapshot(const VKRT* vkrt, uint32_t materialIndex,
                                     VKRT_MaterialSnapshot* outMaterial);                             /* vkrt.h:76 */
VKRT_Result VKRT_getTextureCount(const VKRT* vkrt, uint32_t* outTextureCount);                        /* vkrt.h:77 */
VKRT_Result VKRT_getTextureSnapshot(const VKRT* vkrt, uint32_t textureIndex,
                                    VKRT_TextureSnapshot* outTexture);                               /* vkrt.h:78 */
VKRT_Result VKRT_addTextureFromPixels(VKRT* vkrt, const VKRT_TextureUpload* upload,
                                      uint32_t* outTextureIndex);                                     /* vkrt.h:79 */
/* PNG (libspng's RGBA8 / RGBA16 output rules), baseline JPEG (RGBA8) or EXR
 * (RGBA16F when every channel is half, else RGBA32F); src/core/utility/image.c:123-399. */
VKRT_Result VKRT_addTextureFromFile(VKRT* vkrt, const char* path, const char* name, uint32_t colorSpace,
                                    uint32_t* outTextureIndex);                                      /* vkrt.h:80 */
VKRT_Result VKRT_removeTexture(VKRT* vkrt, uint32_t textureIndex);                                  /* vkrt.h:87 */
VKRT_Result VKRT_addTexturesBatch(VKRT* vkrt, const VKRT_TextureUpload* uploads, size_t uploadCount,
                                  uint32_t* outTextureIndices);                                      /* vkrt.h:88 */
VKRT_Result VKRT_setMaterialTexture(VKRT* vkrt, uint32_t materialIndex, uint32_t textureSlot,
                                    uint32_t textureIndex);                                          /* vkrt.h:94 */
VKRT_Result VKRT_addMaterial(VKRT* vkrt, const Material* material, const char* name,
                             uint32_t* outMaterialIndex);                                             /* vkrt.h:95 */
VKRT_Result VKRT_removeMaterial(VKRT* vkrt, uint32_t materialIndex);                                 /* vkrt.h:96 */
VKRT_Result VKRT_setMaterialName(VKRT* vkrt, uint32_t materialIndex, const char* name);              /* vkrt.h:97 */
VKRT_Result VKRT_setMaterial(VKRT* vkrt, uint32_t materialIndex, const Material* material);           /* vkrt.h:98 */
VKRT_Result VKRT_setMeshMaterialIndex(VKRT* vkrt, uint32_t meshIndex, uint32_t materialIndex);       /* vkrt.h:99 */
VKRT_Result VKRT_clearMeshMaterialAssignment(VKRT* vkrt, uint32_t meshIndex);                        /* vkrt.h:100 */
VKRT_Result VKRT_setMeshOpacity(VKRT* vkrt, uint32_t meshIndex, float opacity);                      /* vkrt.h:101 */
VKRT_Result VKRT_setMeshName(VKRT* vkrt, uint32_t meshIndex, const char* name);                      /* vkrt.h:102 */
VKRT_Result VKRT_setMeshTransform(VKRT* vkrt, uint32_t meshIndex, const float position[3], const float rotation[3],
                                  const float scale[3]);                                             /* vkrt.h:103 */
VKRT_Result VKRT_setMeshTransformMatrix(VKRT* vkrt, uint32_t meshIndex, float worldTransform[4][4]); /* vkrt.h:104 */
/* Stored and reported as in the reference; the reference traces with no culling
 * ray flag, so the setting does not change which triangles rays hit. */
VKRT_Result VKRT_setMeshRenderBackfaces(VKRT* vkrt, uint32_t meshIndex, uint32_t enabled);           /* vkrt.h:105 */
VKRT_Result VKRT_setRenderViewport(VKRT* vkrt, uint32_t x, uint32_t y, uint32_t width, uint32_t height); /* vkrt.h:110 */
VKRT_Result VKRT_cameraSetPose(VKRT* vkrt, const float position[3], const float target[3], const float upVector[3],
                               float vfov);                                                          /* vkrt.h:111 */
VKRT_Result VKRT_cameraGetPose(const VKRT* vkrt, float position[3], float target[3], float upVector[3],
                               float* vfov);                                                         /* vkrt.h:112 */

/* transform helpers: vkrt.h:114-117 */
void VKRT_buildMeshTransformMatrix(const float position[3], const float rotationDegrees[3], const float scale[3],
                                   float outMatrix[4][4]);                                           /* vkrt.h:114 */
void VKRT_buildImportedNodeTransform(float worldTransform[4][4], float outEngineTransform[4][4]);   /* vkrt.h:115 */
void VKRT_decomposeMeshTransform(float worldTransform[4][4], float outPosition[3], float outRotation[3],
                                 float outScale[3]);                                                 /* vkrt.h:116 */
void VKRT_decomposeMeshNodeTransform(float worldTransform[4][4], float outPosition[3], float outRotation[3],
                                     float outScale[3]);                                             /* vkrt.h:117 */

#ifdef __cplusplus
}
#endif

#include "vkrt_mi355x.h"

#endif
