/*
 * Public API types of the MI355X renderer. Names, values and layouts follow the
 * reference's public types (src/core/api/vkrt_types.h:1-240, src/core/api/config.h:3-8)
 * so a caller of the reference API compiles against this header unchanged.
 */
#ifndef VKRT_TYPES_H
#define VKRT_TYPES_H

#include "vkrt_shared.h"

#include <stddef.h>
#include <stdint.h>

enum {
    VKRT_DEFAULT_WIDTH = 1600u,
    VKRT_DEFAULT_HEIGHT = 900u,
    VKRT_MAX_FRAMES_IN_FLIGHT = 2u,
    VKRT_FRAMETIME_HISTORY_SIZE = 128u,
};

enum {
    VKRT_DEVICE_NAME_LEN = 256,
    VKRT_NAME_LEN = 256,
};

/* src/core/api/vkrt_types.h:16-26 */
typedef enum VKRT_Result {
    VKRT_SUCCESS = 0,
    VKRT_ERROR_INVALID_ARGUMENT = -1,
    VKRT_ERROR_OPERATION_FAILED = -2,
    VKRT_ERROR_OUT_OF_MEMORY = -3,
    VKRT_ERROR_DEVICE_LOST = -4,
    VKRT_ERROR_INITIALIZATION_FAILED = -5,
    VKRT_ERROR_SWAPCHAIN_OUT_OF_DATE = -6,
    VKRT_ERROR_PIPELINE_CREATION_FAILED = -7,
    VKRT_ERROR_SHADER_COMPILATION_FAILED = -8,
} VKRT_Result;

typedef uint32_t VKRT_ToneMappingMode;
typedef uint32_t VKRT_RenderMode;
typedef uint32_t VKRT_SpectralSamplingMode;
typedef uint32_t VKRT_DebugMode;
typedef uint32_t VKRT_MaterialTextureSlot;
typedef uint32_t VKRT_TextureColorSpace;
typedef uint32_t VKRT_TextureFormat;

typedef struct Camera {
    vkrt_f3 pos, target, up;
    float nearZ, farZ, vfov;
} Camera;

/* src/core/api/vkrt_types.h:37-46: one frame of viewport mouse input */
typedef struct VKRT_CameraInput {
    float orbitDx;
    float orbitDy;
    float panDx;
    float panDy;
    float scroll;
    uint8_t orbiting;
    uint8_t panning;
    uint8_t captureMouse;
} VKRT_CameraInput;

typedef struct VKRT_CreateInfo {
    uint32_t width;
    uint32_t height;
    const char* title;
    uint8_t startMaximized;
    uint8_t startFullscreen;
    uint8_t headless;
    uint8_t disableSER;
    int32_t preferredDeviceIndex;
    const char* preferredDeviceName;
} VKRT_CreateInfo;

typedef struct VKRT_MeshUpload {
    const Vertex* vertices;
    size_t vertexCount;
    const uint32_t* indices;
    size_t indexCount;
} VKRT_MeshUpload;

typedef struct VKRT_TextureUpload {
    const char* name;
    const void* pixels;
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t colorSpace;
} VKRT_TextureUpload;

typedef struct VKRT VKRT;

typedef struct VKRT_SceneSettingsSnapshot {
    Camera camera;
    uint32_t samplesPerPixel;
    uint32_t rrMaxDepth;
    uint32_t rrMinDepth;
    VKRT_ToneMappingMode toneMappingMode;
    VKRT_RenderMode renderMode;
    uint32_t spectralSamplingMode;
    float exposure;
    uint8_t autoExposureEnabled;
    uint8_t autoSPPEnabled;
    uint32_t autoSPPTargetFPS;
    vkrt_f3 environmentColor;
    float environmentStrength;
    float environmentRotation;
    uint32_t environmentTextureIndex;
    float timeBase;
    float timeStep;
    uint32_t debugMode;
    uint32_t misNeeEnabled;
    uint32_t selectionEnabled;
    uint32_t selectedMeshIndex;
} VKRT_SceneSettingsSnapshot;

typedef enum VKRT_RenderPhase {
    VKRT_RENDER_PHASE_INACTIVE = 0,
    VKRT_RENDER_PHASE_SAMPLING,
    VKRT_RENDER_PHASE_DENOISING,
    VKRT_RENDER_PHASE_COMPLETE_RAW,
    VKRT_RENDER_PHASE_COMPLETE_DENOISED,
} VKRT_RenderPhase;

typedef struct VKRT_RenderStatusSnapshot {
    uint32_t framesPerSecond;
    float averageFrametime;
    float frametimes[VKRT_FRAMETIME_HISTORY_SIZE];
    float displayTimeMs;
    float renderTimeMs;
    uint32_t accumulationFrame;
    uint64_t totalSamples;
    VKRT_RenderPhase renderPhase;
    uint8_t renderDenoiseEnabled;
    uint32_t renderTargetSamples;
    float displayRenderTimeMs;
    float displayFrameTimeMs;
} VKRT_RenderStatusSnapshot;

/* Phase and status predicates of src/core/api/vkrt_types.h:172-227 (the
 * offline benchmark and the render controller poll these). */
static inline uint8_t VKRT_renderPhaseIsActive(VKRT_RenderPhase phase) { return phase != VKRT_RENDER_PHASE_INACTIVE; }
static inline uint8_t VKRT_renderPhaseIsSampling(VKRT_RenderPhase phase) { return phase == VKRT_RENDER_PHASE_SAMPLING; }
static inline uint8_t VKRT_renderPhaseIsDenoising(VKRT_RenderPhase phase) { return phase == VKRT_RENDER_PHASE_DENOISING; }
static inline uint8_t VKRT_renderPhaseIsBusy(VKRT_RenderPhase phase) { return VKRT_renderPhaseIsDenoising(phase); }
static inline uint8_t VKRT_renderPhaseIsComplete(VKRT_RenderPhase phase) {
    return phase == VKRT_RENDER_PHASE_COMPLETE_RAW || phase == VKRT_RENDER_PHASE_COMPLETE_DENOISED;
}
static inline uint8_t VKRT_renderPhaseSamplingFinished(VKRT_RenderPhase phase) {
    return VKRT_renderPhaseIsBusy(phase) || VKRT_renderPhaseIsComplete(phase);
}
static inline uint8_t VKRT_renderPhaseIsDenoised(VKRT_RenderPhase phase) {
    return phase == VKRT_RENDER_PHASE_COMPLETE_DENOISED;
}
static inline uint8_t VKRT_renderStatusIsActive(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsActive(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusIsSampling(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsSampling(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusIsDenoising(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsDenoising(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusIsBusy(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsBusy(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusIsComplete(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsComplete(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusSamplingFinished(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseSamplingFinished(s->renderPhase);
}
static inline uint8_t VKRT_renderStatusIsDenoised(const VKRT_RenderStatusSnapshot* s) {
    return s && VKRT_renderPhaseIsDenoised(s->renderPhase);
}

/* src/core/api/vkrt_types.h:229-238. Headless: the display and swapchain
 * extents are the render extent and presentMode is 0. */
typedef struct VKRT_RuntimeSnapshot {
    uint32_t displayWidth;
    uint32_t displayHeight;
    uint32_t swapchainWidth;
    uint32_t swapchainHeight;
    uint32_t renderWidth;
    uint32_t renderHeight;
    uint32_t displayViewportRect[4];
    uint32_t presentMode;
    float displayRefreshHz;
} VKRT_RuntimeSnapshot;

typedef struct VKRT_RenderExportSettings {
    uint8_t denoiseEnabled;
} VKRT_RenderExportSettings;

typedef struct VKRT_SystemInfo {
    char deviceName[VKRT_DEVICE_NAME_LEN];
    uint32_t vendorID;
    uint32_t driverVersion;
} VKRT_SystemInfo;

typedef struct VKRT_MeshSnapshot {
    MeshInfo info;
    Material material;
    uint32_t materialIndex;
    uint32_t geometrySource;
    uint8_t hasMaterialAssignment;
    uint8_t ownsGeometry;
    char name[VKRT_NAME_LEN];
} VKRT_MeshSnapshot;

typedef struct VKRT_MaterialSnapshot {
    Material material;
    uint32_t useCount;
    char name[VKRT_NAME_LEN];
} VKRT_MaterialSnapshot;

/* src/core/api/vkrt_types.h:265-272 */
typedef struct VKRT_TextureSnapshot {
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t colorSpace;
    uint32_t useCount;
    char name[VKRT_NAME_LEN];
} VKRT_TextureSnapshot;

/* Default material of the reference (src/core/api/vkrt_types.h:200-242). */
static inline Material VKRT_materialDefault(void) {
    Material m = {0};
    m.baseColor[0] = m.baseColor[1] = m.baseColor[2] = 0.8f;
    m.roughness = 0.5f;
    m.diffuseRoughness = 0.0f;
    m.emissionColor[0] = m.emissionColor[1] = m.emissionColor[2] = 1.0f;
    m.emissionLuminance = 0.0f;
    m.metallic = 0.0f;
    m.specular = 0.5f;
    m.specularTint = 0.0f;
    m.anisotropic = 0.0f;
    m.sheenTintWeight[0] = m.sheenTintWeight[1] = m.sheenTintWeight[2] = 1.0f;
    m.sheenTintWeight[3] = 0.0f;
    m.clearcoat = 0.0f;
    m.clearcoatGloss = 1.0f;
    m.ior = 1.5f;
    m.abbeNumber = 0.0f;
    m.transmission = 0.0f;
    m.subsurface = 0.0f;
    m.sheenRoughness = 0.5f;
    m.absorptionCoefficient = 0.0f;
    m.attenuationColor[0] = m.attenuationColor[1] = m.attenuationColor[2] = 1.0f;
    m.normalTextureScale = 1.0f;
    m.baseColorTextureIndex = VKRT_INVALID_INDEX;
    m.metallicRoughnessTextureIndex = VKRT_INVALID_INDEX;
    m.normalTextureIndex = VKRT_INVALID_INDEX;
    m.emissiveTextureIndex = VKRT_INVALID_INDEX;
    m.opacity = 1.0f;
    m.alphaCutoff = 0.5f;
    m.alphaMode = VKRT_MATERIAL_ALPHA_MODE_OPAQUE;
    m.textureTexcoordSets = 0u;
    m.baseColorTextureTransform[0] = m.baseColorTextureTransform[1] = 1.0f;
    m.metallicRoughnessTextureTransform[0] = m.metallicRoughnessTextureTransform[1] = 1.0f;
    m.normalTextureTransform[0] = m.normalTextureTransform[1] = 1.0f;
    m.emissiveTextureTransform[0] = m.emissiveTextureTransform[1] = 1.0f;
    return m;
}

#endif
