/*
 * Render settings, camera and the SceneData uniform. Defaults and the packing
 * of settings into SceneData follow src/core/scene/uniform.c:93-245; setter
 * validation and reset-on-change follow src/core/api/settings.c.
 */
#include "internal.h"

#include <math.h>
#include <string.h>

void vkrtInitDefaultSettings(VKRT* vkrt, uint32_t width, uint32_t height) {
    VKRT_SceneSettingsSnapshot* s = &vkrt->settings;
    memset(s, 0, sizeof(*s));
    s->samplesPerPixel = 8;
    s->rrMaxDepth = 8;
    s->rrMinDepth = 4;
    s->toneMappingMode = VKRT_TONE_MAPPING_MODE_ACES;
    s->renderMode = VKRT_RENDER_MODE_RGB;
    s->spectralSamplingMode = VKRT_SPECTRAL_SAMPLING_MODE_HERO;
    s->exposure = 1.0f;
    s->autoExposureEnabled = 0u;
    s->environmentColor[0] = s->environmentColor[1] = s->environmentColor[2] = 0.25f;
    s->environmentStrength = 1.0f;
    s->environmentRotation = 0.0f;
    s->environmentTextureIndex = VKRT_INVALID_INDEX;
    s->autoSPPEnabled = 1;
    s->autoSPPTargetFPS = 60;
    s->camera.nearZ = 0.001f;
    s->camera.farZ = 10000.0f;
    s->camera.vfov = 40.0f;
    s->camera.pos[0] = -0.5f;
    s->camera.pos[1] = 0.2f;
    s->camera.pos[2] = -0.2f;
    s->camera.up[2] = 1.0f;
    s->timeBase = -1.0f;
    s->timeStep = 0.5f;
    s->debugMode = VKRT_DEBUG_MODE_NONE;
    s->misNeeEnabled = 1u;
    s->selectionEnabled = 0;
    s->selectedMeshIndex = VKRT_INVALID_INDEX;

    vkrt->renderWidth = width;
    vkrt->renderHeight = height;
    vkrt->displayViewport[0] = 0;
    vkrt->displayViewport[1] = 0;
    vkrt->displayViewport[2] = width;
    vkrt->displayViewport[3] = height;
    vkrt->sceneData.viewportRect[0] = 0;
    vkrt->sceneData.viewportRect[1] = 0;
    vkrt->sceneData.viewportRect[2] = width;
    vkrt->sceneData.viewportRect[3] = height;
}

/* src/core/scene/uniform.c:153-174 */
void vkrtWriteSceneStateUniform(VKRT* vkrt) {
    const VKRT_SceneSettingsSnapshot* s = &vkrt->settings;
    SceneData* d = &vkrt->sceneData;
    d->samplesPerPixel = s->samplesPerPixel > 0u ? s->samplesPerPixel : 1u;
    d->rrMaxDepth = s->rrMaxDepth;
    d->rrMinDepth = s->rrMinDepth;
    d->packedRenderSettings = VKRT_SCENE_PACK_RENDER_SETTINGS(s->toneMappingMode, s->renderMode, s->spectralSamplingMode);
    d->exposure = s->exposure;
    d->timeBase = s->timeBase;
    d->timeStep = s->timeStep;
    d->environmentLight[0] = s->environmentColor[0] * s->environmentStrength;
    d->environmentLight[1] = s->environmentColor[1] * s->environmentStrength;
    d->environmentLight[2] = s->environmentColor[2] * s->environmentStrength;
    d->environmentLight[3] = s->environmentStrength;
    d->environmentTextureIndex = s->environmentTextureIndex;
    d->environmentRotation = s->environmentRotation;
    d->debugMode = s->debugMode;
    d->misNeeEnabled = s->misNeeEnabled ? 1u : 0u;
    d->selectionEnabled = s->selectionEnabled ? 1u : 0u;
    d->selectedMeshIndex = s->selectedMeshIndex;
}

/* src/core/scene/uniform.c:227-245: any change restarts accumulation. */
/* src/core/scene/timing.c:62-75: a scene change restarts the controller at
 * one sample per pixel; a settings change only forgets its measurements. */
void vkrtResetAutoSPPState(VKRT* vkrt, int resetSamplesPerPixel) {
    if (resetSamplesPerPixel && vkrt->settings.autoSPPEnabled) vkrt->settings.samplesPerPixel = 1u;
    vkrt->autoSPPControlMs = 0.0f;
    vkrt->renderStatus.displayTimeMs = 0.0f;
    vkrt->renderStatus.renderTimeMs = 0.0f;
    vkrt->renderStatus.displayRenderTimeMs = 0.0f;
    vkrt->renderStatus.displayFrameTimeMs = 0.0f;
    vkrt->renderStatus.framesPerSecond = 0u;
}

/* src/core/scene/timing.c:121-178, run once per frame that added samples:
 * the trace time per sample is smoothed, the sample count aimed at 90 % of
 * the frame budget, moves inside a dead band are ignored, and one step is
 * limited to x1.25 up and x0.6 down, within [1, 2048]. */
void vkrtUpdateAutoSPP(VKRT* vkrt) {
    if (!vkrt->settings.autoSPPEnabled) return;
    const float smoothing = 0.35f, budgetScale = 0.90f, upDeadband = 0.18f, downDeadband = 0.08f;
    const float maxUp = 1.25f, maxDown = 0.60f;
    float targetMs = vkrt->autoSPPTargetFrameMs;
    if (!(targetMs > 0.0f)) {
        uint32_t fps = vkrt->settings.autoSPPTargetFPS ? vkrt->settings.autoSPPTargetFPS : 60u;
        targetMs = 1000.0f / (float)fps;
    }
    float controlMs = vkrt->renderStatus.renderTimeMs;
    if (!(controlMs > 0.0f)) controlMs = vkrt->renderStatus.displayRenderTimeMs;
    if (!(controlMs > 0.0f)) controlMs = vkrt->renderStatus.displayFrameTimeMs;
    if (!(controlMs > 0.0f)) return;
    uint32_t spp = vkrt->settings.samplesPerPixel ? vkrt->settings.samplesPerPixel : 1u;
    float sppf = (float)spp;
    float perSample = controlMs / sppf;
    if (!(perSample > 0.0f)) return;
    vkrt->autoSPPControlMs = vkrt->autoSPPControlMs <= 0.0f
                                 ? perSample
                                 : vkrt->autoSPPControlMs * (1.0f - smoothing) + perSample * smoothing;
    float desired = (targetMs * budgetScale) / vkrt->autoSPPControlMs;
    desired = fminf(fmaxf(desired, 1.0f), 2048.0f);
    float delta = desired - sppf;
    float deadband = fmaxf(sppf * (delta > 0.0f ? upDeadband : downDeadband), 1.0f);
    if (fabsf(delta) <= deadband) return;
    uint32_t next;
    if (delta > 0.0f) {
        next = (uint32_t)floorf(fminf(desired, ceilf(sppf * maxUp)));
        if (next <= spp && spp < 2048u) next = spp + 1u;
    } else {
        next = (uint32_t)ceilf(fmaxf(desired, floorf(sppf * maxDown)));
        if (next >= spp && spp > 1u) next = spp - 1u;
    }
    if (next < 1u) next = 1u;
    if (next > 2048u) next = 2048u;
    if (next != spp) {
        vkrt->settings.samplesPerPixel = next;
        vkrtWriteSceneStateUniform(vkrt); /* the accumulation continues */
    }
}

void vkrtResetSceneData(VKRT* vkrt) {
    vkrtResetAutoSPPState(vkrt, 1);
    vkrt->sceneData.frameNumber = 0;
    if (vkrt->renderStatus.renderPhase != VKRT_RENDER_PHASE_INACTIVE) {
        vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_SAMPLING;
    }
    vkrt->renderStatus.accumulationFrame = 0;
    vkrt->renderStatus.totalSamples = 0;
    vkrt->renderStatus.averageFrametime = 0.0f;
    memset(vkrt->renderStatus.frametimes, 0, sizeof(vkrt->renderStatus.frametimes));
    vkrt->accumulationNeedsReset = 1;
    vkrtWriteSceneStateUniform(vkrt);
}

VKRT_Result VKRT_invalidateAccumulation(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setSamplesPerPixel(VKRT* vkrt, uint32_t samplesPerPixel) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (samplesPerPixel == 0) samplesPerPixel = 1;
    if (vkrt->settings.samplesPerPixel == samplesPerPixel) return VKRT_SUCCESS;
    vkrt->settings.samplesPerPixel = samplesPerPixel;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setPathDepth(VKRT* vkrt, uint32_t rrMinDepth, uint32_t rrMaxDepth) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (rrMaxDepth < 1u) rrMaxDepth = 1u;
    if (rrMaxDepth > 64u) rrMaxDepth = 64u;
    if (rrMinDepth > rrMaxDepth) rrMinDepth = rrMaxDepth;
    if (vkrt->settings.rrMinDepth == rrMinDepth && vkrt->settings.rrMaxDepth == rrMaxDepth) return VKRT_SUCCESS;
    vkrt->settings.rrMinDepth = rrMinDepth;
    vkrt->settings.rrMaxDepth = rrMaxDepth;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setAutoSPPEnabled(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.autoSPPEnabled = enabled ? 1u : 0u;
    vkrtResetAutoSPPState(vkrt, 0);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setAutoSPPTargetFPS(VKRT* vkrt, uint32_t targetFPS) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (targetFPS == 0) targetFPS = 60; /* the display refresh rate; no display here */
    if (targetFPS < 30) targetFPS = 30;
    if (targetFPS > 360) targetFPS = 360;
    vkrt->settings.autoSPPTargetFPS = targetFPS;
    vkrt->autoSPPTargetFrameMs = 1000.0f / (float)targetFPS;
    vkrtResetAutoSPPState(vkrt, 0);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setToneMappingMode(VKRT* vkrt, VKRT_ToneMappingMode mode) {
    if (!vkrt || mode >= VKRT_TONE_MAPPING_MODE_COUNT) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.toneMappingMode = mode;
    vkrtWriteSceneStateUniform(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setRenderMode(VKRT* vkrt, VKRT_RenderMode mode) {
    if (!vkrt || mode >= VKRT_RENDER_MODE_COUNT) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->settings.renderMode == mode) return VKRT_SUCCESS;
    vkrt->settings.renderMode = mode;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setSpectralSamplingMode(VKRT* vkrt, VKRT_SpectralSamplingMode mode) {
    if (!vkrt || mode >= VKRT_SPECTRAL_SAMPLING_MODE_COUNT) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->settings.spectralSamplingMode == mode) return VKRT_SUCCESS;
    vkrt->settings.spectralSamplingMode = mode;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setExposure(VKRT* vkrt, float exposure) {
    if (!vkrt || !isfinite(exposure) || exposure < 0.0f) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.exposure = exposure;
    vkrtWriteSceneStateUniform(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setAutoExposureEnabled(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.autoExposureEnabled = enabled ? 1u : 0u;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setEnvironmentLight(VKRT* vkrt, const float color[3], float strength) {
    if (!vkrt || !color || !isfinite(strength) || strength < 0.0f) return VKRT_ERROR_INVALID_ARGUMENT;
    for (int i = 0; i < 3; i++) {
        if (!isfinite(color[i]) || color[i] < 0.0f) return VKRT_ERROR_INVALID_ARGUMENT;
    }
    memcpy(vkrt->settings.environmentColor, color, sizeof(float) * 3);
    vkrt->settings.environmentStrength = strength;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setEnvironmentRotation(VKRT* vkrt, float rotationDegrees) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!isfinite(rotationDegrees)) rotationDegrees = 0.0f;
    rotationDegrees = fmodf(rotationDegrees, 360.0f);
    if (rotationDegrees < -180.0f) rotationDegrees += 360.0f;
    if (rotationDegrees >= 180.0f) rotationDegrees -= 360.0f;
    if (vkrt->settings.environmentRotation == rotationDegrees) return VKRT_SUCCESS;
    vkrt->settings.environmentRotation = rotationDegrees;
    vkrt->environmentRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setDebugMode(VKRT* vkrt, VKRT_DebugMode mode) {
    if (!vkrt || mode >= VKRT_DEBUG_MODE_COUNT) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.debugMode = mode;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setMisNeeEnabled(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.misNeeEnabled = enabled ? 1u : 0u;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setTimeRange(VKRT* vkrt, float timeBase, float timeStep) {
    if (!vkrt || !isfinite(timeBase) || !isfinite(timeStep)) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->settings.timeBase = timeBase;
    vkrt->settings.timeStep = timeStep;
    vkrtWriteSceneStateUniform(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setRandomPolicy(VKRT* vkrt, uint32_t policy, uint64_t seed) {
    if (!vkrt || (policy != VKRT_RNG_HASH && policy != VKRT_RNG_PCG32)) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->rngPolicy = policy;
    vkrt->rngSeed = seed;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setEnvironmentImportanceSampling(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->envImportance = enabled ? 1u : 0u;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

static int finite3(const float v[3]) { return v && isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

/* src/core/api/render.c:402-427: non-finite poses and degenerate views are rejected. */
VKRT_Result VKRT_cameraSetPose(VKRT* vkrt, const float position[3], const float target[3], const float upVector[3],
                               float vfov) {
    if (!vkrt || !finite3(position) || !finite3(target) || !finite3(upVector) || !isfinite(vfov)) {
        return VKRT_ERROR_INVALID_ARGUMENT;
    }
    if (vfov <= 0.0f || vfov >= 180.0f) return VKRT_ERROR_INVALID_ARGUMENT;
    float d[3] = {target[0] - position[0], target[1] - position[1], target[2] - position[2]};
    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= 0.0f) return VKRT_ERROR_INVALID_ARGUMENT;
    if (upVector[0] * upVector[0] + upVector[1] * upVector[1] + upVector[2] * upVector[2] <= 0.0f)
        return VKRT_ERROR_INVALID_ARGUMENT;
    memcpy(vkrt->settings.camera.pos, position, sizeof(float) * 3);
    memcpy(vkrt->settings.camera.target, target, sizeof(float) * 3);
    memcpy(vkrt->settings.camera.up, upVector, sizeof(float) * 3);
    vkrt->settings.camera.vfov = vfov;
    vkrtSyncCameraMatrices(vkrt);
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_cameraGetPose(const VKRT* vkrt, float position[3], float target[3], float upVector[3], float* vfov) {
    if (!vkrt || !position || !target || !upVector || !vfov) return VKRT_ERROR_INVALID_ARGUMENT;
    memcpy(position, vkrt->settings.camera.pos, sizeof(float) * 3);
    memcpy(target, vkrt->settings.camera.target, sizeof(float) * 3);
    memcpy(upVector, vkrt->settings.camera.up, sizeof(float) * 3);
    *vfov = vkrt->settings.camera.vfov;
    return VKRT_SUCCESS;
}

/* src/core/api/settings.c:15-21 and src/core/scene/camera.c:12-120: pan moves
 * position and target along the view's right/up axes by 0.001 x distance per
 * unit; orbit turns the position about the target by 0.004 rad per unit with
 * the pitch held 0.001 rad short of the poles; scroll moves the position along
 * the (unnormalised) view vector by 0.075 per unit. Each step recomputes the
 * camera and restarts accumulation; nothing moves while the mouse is captured. */
VKRT_Result VKRT_applyCameraInput(VKRT* vkrt, const VKRT_CameraInput* input) {
    if (!vkrt || !input) return VKRT_ERROR_INVALID_ARGUMENT;
    if (input->captureMouse) return VKRT_SUCCESS;
    float* pos = vkrt->settings.camera.pos;
    float* tgt = vkrt->settings.camera.target;
    const float view[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
    /* right = normalize(view x +z), or +y when looking straight up or down; up = normalize(right x view) */
    float right[3] = {view[1], -view[0], 0.0f};
    float rl = sqrtf(right[0] * right[0] + right[1] * right[1]);
    if (rl < 1e-6f) {
        right[0] = 0.0f;
        right[1] = 1.0f;
    } else {
        right[0] /= rl;
        right[1] /= rl;
    }
    float up[3] = {right[1] * view[2] - right[2] * view[1], right[2] * view[0] - right[0] * view[2],
                   right[0] * view[1] - right[1] * view[0]};
    float ul = sqrtf(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    if (ul > 0.0f)
        for (int a = 0; a < 3; a++) up[a] /= ul;
    const float dist = sqrtf(view[0] * view[0] + view[1] * view[1] + view[2] * view[2]);
    int moved = 0;
    if (input->panning && (input->panDx != 0.0f || input->panDy != 0.0f)) {
        const float scale = 0.001f * dist, mx = -input->panDx * scale, my = input->panDy * scale;
        for (int a = 0; a < 3; a++) {
            const float delta = right[a] * mx + up[a] * my;
            pos[a] += delta;
            tgt[a] += delta;
        }
        moved = 1;
    }
    if (input->orbiting && (input->orbitDx != 0.0f || input->orbitDy != 0.0f)) {
        const float halfPi = 1.5707963267948966f, eps = 0.001f;
        float yaw = atan2f(view[1], view[0]) - input->orbitDx * 0.004f;
        float pitch = atan2f(view[2], sqrtf(view[0] * view[0] + view[1] * view[1])) - input->orbitDy * 0.004f;
        if (pitch < -halfPi + eps) pitch = -halfPi + eps;
        if (pitch > halfPi - eps) pitch = halfPi - eps;
        const float forward[3] = {cosf(pitch) * cosf(yaw), cosf(pitch) * sinf(yaw), sinf(pitch)};
        for (int a = 0; a < 3; a++) pos[a] = tgt[a] - forward[a] * dist;
        moved = 1;
    }
    if (input->scroll != 0.0f) {
        const float step = input->scroll * -0.075f;
        for (int a = 0; a < 3; a++) pos[a] -= view[a] * step;
        moved = 1;
    }
    if (moved) {
        vkrtSyncCameraMatrices(vkrt);
        vkrtResetSceneData(vkrt);
    }
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setRenderViewport(VKRT* vkrt, uint32_t x, uint32_t y, uint32_t width, uint32_t height) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (x >= vkrt->renderWidth) x = vkrt->renderWidth ? vkrt->renderWidth - 1 : 0;
    if (y >= vkrt->renderHeight) y = vkrt->renderHeight ? vkrt->renderHeight - 1 : 0;
    if (width > vkrt->renderWidth - x) width = vkrt->renderWidth - x;
    if (height > vkrt->renderHeight - y) height = vkrt->renderHeight - y;
    vkrt->displayViewport[0] = x;
    vkrt->displayViewport[1] = y;
    vkrt->displayViewport[2] = width;
    vkrt->displayViewport[3] = height;
    vkrt->sceneData.viewportRect[0] = x;
    vkrt->sceneData.viewportRect[1] = y;
    vkrt->sceneData.viewportRect[2] = width;
    vkrt->sceneData.viewportRect[3] = height;
    vkrtSyncCameraMatrices(vkrt);
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getSceneSettings(const VKRT* vkrt, VKRT_SceneSettingsSnapshot* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    *out = vkrt->settings;
    return VKRT_SUCCESS;
}
