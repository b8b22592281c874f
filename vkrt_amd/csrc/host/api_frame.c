/*
 * Lifecycle, frame loop and render sessions.
 *
 * The frame loop keeps the reference's contract (src/core/api/frame.c:74-420):
 * VKRT_draw = beginFrame, updateScene, trace, present, endFrame. Without a
 * window every frame is "offscreen" and is traced only while a render session
 * is sampling (frame.c:107-113); endFrame advances frameNumber and
 * totalSamples only when the trace contributed (frame.c:370-402) and stops
 * sampling once the target is reached.
 *
 * VKRT_updateScene re-packs dirty host data and uploads it; a geometry change
 * rebuilds the BVH on the GPU (the reference's BLAS/TLAS builds,
 * src/core/render/accel/blas.c:222-262, tlas.c:535-559). VKRT_trace launches
 * the wavefront path tracer in place of vkCmdTraceRaysKHR
 * (src/core/runtime/command/record.c:448-486).
 */
#include "internal.h"

#include <stdlib.h>
#include <string.h>

void VKRT_defaultCreateInfo(VKRT_CreateInfo* createInfo) {
    if (!createInfo) return;
    memset(createInfo, 0, sizeof(*createInfo));
    createInfo->width = VKRT_DEFAULT_WIDTH;
    createInfo->height = VKRT_DEFAULT_HEIGHT;
    createInfo->title = "vkrt";
    createInfo->headless = 1;
    createInfo->preferredDeviceIndex = -1;
}

VKRT_Result VKRT_create(VKRT** outVkrt) {
    if (!outVkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT* vkrt = (VKRT*)calloc(1, sizeof(VKRT));
    if (!vkrt) return VKRT_ERROR_OUT_OF_MEMORY;
    vkrtInitDefaultSettings(vkrt, VKRT_DEFAULT_WIDTH, VKRT_DEFAULT_HEIGHT);
    vkrt->sceneData.exposure = 1.0f;
    vkrtWriteSceneStateUniform(vkrt);
    vkrtSyncCameraMatrices(vkrt);
    vkrt->accumulationNeedsReset = 1;
    vkrt->rngPolicy = VKRT_RNG_HASH;
    VKRT_defaultNeuralCacheSettings(&vkrt->nrc);
    vkrt->materialSort = 0;
    vkrt->fusedTailRays = 1u << 20;
    vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_INACTIVE;
    vkrt->finalImageDenoiseEnabled = 1u;
    *outVkrt = vkrt;
    return VKRT_SUCCESS;
}

void VKRT_deinit(VKRT* vkrt) {
    if (!vkrt) return;
    if (vkrt->device) {
        mrt_destroy(vkrt->device);
        vkrt->device = NULL;
    }
}

void VKRT_destroy(VKRT* vkrt) {
    if (!vkrt) return;
    VKRT_deinit(vkrt);
    for (uint32_t i = 0; i < vkrt->meshCount; i++) {
        free(vkrt->meshes[i].vertices);
        free(vkrt->meshes[i].indices);
    }
    free(vkrt->meshes);
    free(vkrt->materials);
    for (uint32_t i = 0; i < vkrt->textureCount; i++) free(vkrt->textures[i].rgba);
    free(vkrt->textures);
    vkrtFreePackedScene(&vkrt->packed);
    free(vkrt->rgb2spec);
    free(vkrt);
}

VKRT_Result VKRT_initWithCreateInfo(VKRT* vkrt, const VKRT_CreateInfo* createInfo) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_CreateInfo info;
    if (createInfo)
        info = *createInfo;
    else
        VKRT_defaultCreateInfo(&info);
    uint32_t w = info.width ? info.width : VKRT_DEFAULT_WIDTH;
    uint32_t h = info.height ? info.height : VKRT_DEFAULT_HEIGHT;
    vkrt->headless = 1;
    vkrt->renderWidth = w;
    vkrt->renderHeight = h;
    VKRT_setRenderViewport(vkrt, 0, 0, w, h);
    VKRT_Result r = vkrtLoadRGB2SpecTable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (vkrt->device) return VKRT_SUCCESS;
    int device = info.preferredDeviceIndex >= 0 ? info.preferredDeviceIndex : 0;
    if (mrt_create(device, &vkrt->device) != 0) {
        VKRT_LOG_ERROR("GPU initialisation failed: %s", mrt_last_error());
        vkrt->device = NULL;
        return VKRT_ERROR_INITIALIZATION_FAILED;
    }
    uint32_t vendor = 0, driver = 0;
    mrt_device_name(vkrt->device, vkrt->systemInfo.deviceName, sizeof(vkrt->systemInfo.deviceName), &vendor, &driver);
    vkrt->systemInfo.vendorID = vendor;
    vkrt->systemInfo.driverVersion = driver;
    if (mrt_upload_rgb2spec(vkrt->device, vkrt->rgb2spec, vkrt->rgb2specFloatCount) != 0) {
        VKRT_LOG_ERROR("%s", mrt_last_error());
        return VKRT_ERROR_INITIALIZATION_FAILED;
    }
    /* force a full upload on the first frame */
    vkrt->geometryUploaded = vkrt->geometryRevision - 1u;
    vkrt->materialUploaded = vkrt->materialRevision - 1u;
    vkrt->environmentUploaded = vkrt->environmentRevision - 1u;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_init(VKRT* vkrt) { return VKRT_initWithCreateInfo(vkrt, NULL); }

int VKRT_shouldDeinit(VKRT* vkrt) { return vkrt ? vkrt->shouldDeinit : 1; }

void VKRT_poll(VKRT* vkrt) { (void)vkrt; }

VKRT_Result VKRT_getSystemInfo(const VKRT* vkrt, VKRT_SystemInfo* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    *out = vkrt->systemInfo;
    return VKRT_SUCCESS;
}

static VKRT_Result deviceError(void) {
    VKRT_LOG_ERROR("%s", mrt_last_error());
    return VKRT_ERROR_OPERATION_FAILED;
}

/* Re-packs and uploads whatever changed since the last upload. */
static VKRT_Result syncScene(VKRT* vkrt, int requireDevice) {
    int geometryDirty = vkrt->geometryUploaded != vkrt->geometryRevision;
    int materialDirty = vkrt->materialUploaded != vkrt->materialRevision;
    int envDirty = vkrt->environmentUploaded != vkrt->environmentRevision;
    if (!geometryDirty && !materialDirty && !envDirty && vkrt->packed.meshInfos) return VKRT_SUCCESS;
    VKRT_Result r = vkrtPrepareScene(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (!vkrt->device) {
        if (requireDevice) return VKRT_ERROR_INITIALIZATION_FAILED;
        /* packed buffers stay valid until the next change; device initialisation
         * forces a full upload of whatever is prepared by then */
        vkrt->geometryUploaded = vkrt->geometryRevision;
        vkrt->materialUploaded = vkrt->materialRevision;
        vkrt->environmentUploaded = vkrt->environmentRevision;
        return VKRT_SUCCESS;
    }
    PackedScene* p = &vkrt->packed;
    if (geometryDirty) {
        mrt_geometry_desc g = {p->vertices, p->vertexCount, p->indices, p->indexCount, p->worldTriangles,
                               p->triangleMesh, p->triangleCount, p->meshTriangleBase, vkrt->meshCount};
        if (mrt_upload_geometry(vkrt->device, &g) != 0) return deviceError();
        mrt_bvh_stats bs;
        memset(&bs, 0, sizeof(bs));
        if (mrt_build_bvh(vkrt->device, &bs) != 0) return deviceError();
        vkrt->stats.bvhBuildMs = bs.buildMs;
        vkrt->stats.bvhNodeCount = bs.nodeCount;
        vkrt->stats.bvhLeafTriangles = bs.leafTriangles;
        vkrt->stats.bvhMaxDepth = bs.maxDepth;
        vkrt->stats.triangleCount = p->triangleCount;
        /* scene box (the neural cache's position input is normalised by it) */
        for (int a = 0; a < 3; a++) {
            vkrt->sceneMin[a] = 0.0f;
            vkrt->sceneMax[a] = 1.0f;
        }
        for (uint32_t t = 0; t < p->triangleCount; t++) {
            for (int v = 0; v < 3; v++) {
                const float* q = p->worldTriangles + 9u * t + 3u * v;
                for (int a = 0; a < 3; a++) {
                    if (t == 0 && v == 0) vkrt->sceneMin[a] = vkrt->sceneMax[a] = q[a];
                    if (q[a] < vkrt->sceneMin[a]) vkrt->sceneMin[a] = q[a];
                    if (q[a] > vkrt->sceneMax[a]) vkrt->sceneMax[a] = q[a];
                }
            }
        }
    }
    if (mrt_upload_materials(vkrt->device, p->materials, p->materialCount) != 0) return deviceError();
    if ((geometryDirty || materialDirty) && mrt_upload_alpha(vkrt->device, p->triangleAlpha, p->triangleCount) != 0)
        return deviceError();
    if ((geometryDirty || materialDirty) &&
        mrt_upload_textures(vkrt->device, p->textureTexels, p->textureTexelCount, p->textureInfo, p->textureCount) != 0)
        return deviceError();
    if (mrt_upload_meshes(vkrt->device, p->meshInfos, vkrt->meshCount) != 0) return deviceError();
    mrt_light_desc l = {p->emissiveMeshes, p->emissiveMeshCount, p->emissiveTriangles, p->emissiveTriangleCount,
                        p->meshAliasQ,     p->meshAliasIdx,      p->triAliasQ,         p->triAliasIdx,
                        p->emissiveMeshMaterial};
    if (mrt_upload_lights(vkrt->device, &l) != 0) return deviceError();
    if (envDirty) {
        mrt_environment_desc e;
        memset(&e, 0, sizeof(e));
        uint32_t ti = vkrt->settings.environmentTextureIndex;
        if (ti != VKRT_INVALID_INDEX && ti < vkrt->textureCount) {
            e.pixels = vkrt->textures[ti].rgba;
            e.width = vkrt->textures[ti].width;
            e.height = vkrt->textures[ti].height;
            e.aliasQ = p->envAliasQ;
            e.aliasIdx = p->envAliasIdx;
            e.pdf = p->envPdf;
        }
        if (mrt_upload_environment(vkrt->device, &e) != 0) return deviceError();
    }
    vkrt->geometryUploaded = vkrt->geometryRevision;
    vkrt->materialUploaded = vkrt->materialRevision;
    vkrt->environmentUploaded = vkrt->environmentRevision;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_beginFrame(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->frameTraced = 0;
    return VKRT_SUCCESS;
}

static int sessionSampling(const VKRT* vkrt) { return vkrt->renderStatus.renderPhase == VKRT_RENDER_PHASE_SAMPLING; }

VKRT_Result VKRT_updateScene(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!sessionSampling(vkrt)) return VKRT_SUCCESS;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    return syncScene(vkrt, 1);
}

static void fillRenderParams(const VKRT* vkrt, mrt_render_params* rp) {
    memset(rp, 0, sizeof(*rp));
    rp->width = vkrt->renderWidth;
    rp->height = vkrt->renderHeight;
    rp->rngPolicy = vkrt->rngPolicy;
    rp->rngSeed = vkrt->rngSeed;
    rp->envImportance = vkrt->envImportance;
    rp->hasTransmissive = vkrt->packed.hasTransmissive;
    rp->traversalStatistics = vkrt->traversalStatistics;
    rp->materialSort = vkrt->materialSort;
    rp->tailRays = vkrt->fusedTailRays;
    rp->features = vkrt->denoiserFeatures;
    /* the cache runs in RGB and in hero spectral mode (in CIE XYZ there) */
    const int cacheMode = vkrt->settings.renderMode == VKRT_RENDER_MODE_RGB ||
                          (vkrt->settings.renderMode == VKRT_RENDER_MODE_SPECTRAL &&
                           vkrt->settings.spectralSamplingMode == VKRT_SPECTRAL_SAMPLING_MODE_HERO);
    if (vkrt->nrc.enabled && cacheMode) {
        rp->nrcEnabled = 1u;
        rp->nrcPrefix = vkrt->nrc.prefixBounces;
        rp->nrcTrainingPaths = vkrt->nrc.trainingPaths;
        rp->nrcTrainingSteps = vkrt->nrc.trainingSteps;
        rp->nrcBatch = vkrt->nrc.batchSize;
        rp->nrcTrain = vkrt->nrc.train;
        rp->nrcLearningRate = vkrt->nrc.learningRate;
        memcpy(rp->sceneMin, vkrt->sceneMin, sizeof(rp->sceneMin));
        memcpy(rp->sceneMax, vkrt->sceneMax, sizeof(rp->sceneMax));
    }
}

VKRT_Result VKRT_trace(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!sessionSampling(vkrt)) return VKRT_SUCCESS;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    mrt_render_params rp;
    fillRenderParams(vkrt, &rp);
    if (mrt_resize_accumulation(vkrt->device, vkrt->renderWidth, vkrt->renderHeight) != 0) return deviceError();
    if (vkrt->accumulationNeedsReset) {
        vkrt->renderStatus.accumulationFrame = 0;
        vkrt->renderStatus.totalSamples = 0;
        rp.resetAccumulation = 1;
        vkrt->accumulationNeedsReset = 0;
    }
    mrt_render_stats rs;
    memset(&rs, 0, sizeof(rs));
    if (mrt_render_frame(vkrt->device, &vkrt->sceneData, &rp, &rs) != 0) return deviceError();
    vkrt->stats.extensionRays = rs.extensionRays;
    vkrt->stats.shadowRays = rs.shadowRays;
    vkrt->stats.paths = rs.paths;
    vkrt->stats.traceMs = rs.traceMs;
    vkrt->stats.extendMs = rs.extendMs;
    vkrt->stats.shadowMs = rs.shadowMs;
    vkrt->stats.shadeMs = rs.shadeMs;
    vkrt->stats.extendMsDepth0 = rs.extendMsDepth0;
    vkrt->stats.shadeMsDepth0 = rs.shadeMsDepth0;
    vkrt->stats.extendLaunches = rs.extendLaunches;
    vkrt->stats.shadeLaunches = rs.shadeLaunches;
    vkrt->stats.shadowLaunches = rs.shadowLaunches;
    vkrt->stats.nodeVisits = rs.nodeVisits;
    vkrt->stats.triangleTests = rs.triangleTests;
    vkrt->stats.tailStartDepth = rs.tailStartDepth;
    vkrt->stats.tailLaunches = rs.tailLaunches;
    vkrt->stats.tailMs = rs.tailMs;
    vkrt->stats.tailRays = rs.tailRays;
    vkrt->stats.raysBeforeTail = rs.raysBeforeTail;
    vkrt->stats.nodeVisitsBeforeTail = rs.nodeVisitsBeforeTail;
    vkrt->stats.triangleTestsBeforeTail = rs.triangleTestsBeforeTail;
    vkrt->nrcStats.lastLoss = rs.nrcLoss;
    vkrt->nrcStats.queries = rs.nrcQueries;
    vkrt->nrcStats.trainingRecords = rs.nrcRecords;
    vkrt->nrcStats.inferMs = rs.nrcInferMs;
    vkrt->nrcStats.trainMs = rs.nrcTrainMs;
    vkrt->renderStatus.renderTimeMs = rs.traceMs;
    vkrt->frameTraced = 1;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_present(VKRT* vkrt) { return vkrt ? VKRT_SUCCESS : VKRT_ERROR_INVALID_ARGUMENT; }

VKRT_Result VKRT_endFrame(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->frameTraced && sessionSampling(vkrt)) {
        uint32_t spp = vkrt->sceneData.samplesPerPixel ? vkrt->sceneData.samplesPerPixel : 1u;
        vkrtUpdateAutoSPP(vkrt); /* src/core/api/frame.c:380-381 */
        vkrt->renderStatus.accumulationFrame++;
        vkrt->renderStatus.totalSamples += spp;
        vkrt->sceneData.frameNumber++;
        uint32_t slot = vkrt->renderStatus.accumulationFrame % VKRT_FRAMETIME_HISTORY_SIZE;
        vkrt->renderStatus.frametimes[slot] = vkrt->renderStatus.renderTimeMs;
        if (vkrt->renderStatus.renderTargetSamples > 0 &&
            vkrt->renderStatus.totalSamples >= vkrt->renderStatus.renderTargetSamples) {
            vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_COMPLETE_RAW;
        }
    }
    vkrt->frameTraced = 0;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_draw(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = VKRT_beginFrame(vkrt);
    if (r == VKRT_SUCCESS) r = VKRT_updateScene(vkrt);
    if (r == VKRT_SUCCESS) r = VKRT_trace(vkrt);
    if (r == VKRT_SUCCESS) r = VKRT_present(vkrt);
    if (r == VKRT_SUCCESS) r = VKRT_endFrame(vkrt);
    return r;
}

/* src/core/api/render.c:27-37 */
VKRT_Result VKRT_setRenderDenoiseEnabled(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->finalImageDenoiseEnabled = enabled ? 1u : 0u;
    if (vkrt->renderStatus.renderPhase == VKRT_RENDER_PHASE_SAMPLING)
        vkrt->renderStatus.renderDenoiseEnabled = vkrt->finalImageDenoiseEnabled;
    return VKRT_SUCCESS;
}

/* src/core/api/render.c:46-81. Valid once sampling has finished and with no
 * debug view. The reference then denoises on a worker thread and reports
 * COMPLETE_RAW when that fails (src/core/utility/export/api.c:160-167); this
 * build has no denoiser, so the request completes at once as COMPLETE_RAW. */
VKRT_Result VKRT_denoiseRenderToViewport(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    const VKRT_RenderPhase phase = vkrt->renderStatus.renderPhase;
    if (!VKRT_renderPhaseIsActive(phase) || !VKRT_renderPhaseSamplingFinished(phase)) return VKRT_ERROR_OPERATION_FAILED;
    if (VKRT_renderPhaseIsBusy(phase)) return VKRT_SUCCESS;
    if (vkrt->settings.debugMode != VKRT_DEBUG_MODE_NONE) return VKRT_ERROR_OPERATION_FAILED;
    vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_COMPLETE_RAW;
    return VKRT_SUCCESS;
}

/* src/core/api/query.c:25-38. Headless: the display and swapchain are the
 * render target, with no present mode or refresh rate. */
VKRT_Result VKRT_getRuntimeSnapshot(const VKRT* vkrt, VKRT_RuntimeSnapshot* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    memset(out, 0, sizeof(*out));
    out->displayWidth = out->swapchainWidth = out->renderWidth = vkrt->renderWidth;
    out->displayHeight = out->swapchainHeight = out->renderHeight = vkrt->renderHeight;
    memcpy(out->displayViewportRect, vkrt->displayViewport, sizeof(out->displayViewportRect));
    return VKRT_SUCCESS;
}

/* src/core/api/render.c:248-265 */
VKRT_Result VKRT_startRender(VKRT* vkrt, uint32_t width, uint32_t height, uint32_t targetSamples) {
    if (!vkrt || width == 0 || height == 0) return VKRT_ERROR_INVALID_ARGUMENT;
    if (width > 16384) width = 16384;
    if (height > 16384) height = 16384;
    vkrt->renderWidth = width;
    vkrt->renderHeight = height;
    vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_SAMPLING;
    vkrt->renderStatus.renderDenoiseEnabled = vkrt->finalImageDenoiseEnabled; /* render.c:214 */
    vkrt->renderStatus.renderTargetSamples = targetSamples;
    vkrtResetAutoSPPState(vkrt, 1); /* src/core/api/render.c:275 */
    vkrtWriteSceneStateUniform(vkrt);
    return VKRT_setRenderViewport(vkrt, 0, 0, width, height);
}

VKRT_Result VKRT_continueRender(VKRT* vkrt, uint32_t targetSamples) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->renderStatus.renderPhase != VKRT_RENDER_PHASE_COMPLETE_RAW &&
        vkrt->renderStatus.renderPhase != VKRT_RENDER_PHASE_COMPLETE_DENOISED)
        return VKRT_ERROR_OPERATION_FAILED;
    vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_SAMPLING;
    vkrt->renderStatus.renderDenoiseEnabled = vkrt->finalImageDenoiseEnabled;
    vkrt->renderStatus.renderTargetSamples = targetSamples;
    vkrtWriteSceneStateUniform(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_stopRenderSampling(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (sessionSampling(vkrt)) vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_COMPLETE_RAW;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_stopRender(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->renderStatus.renderPhase == VKRT_RENDER_PHASE_INACTIVE) return VKRT_SUCCESS;
    vkrt->renderStatus.renderPhase = VKRT_RENDER_PHASE_INACTIVE;
    vkrt->renderStatus.renderTargetSamples = 0;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getRenderStatus(const VKRT* vkrt, VKRT_RenderStatusSnapshot* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    *out = vkrt->renderStatus;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getTraceStats(const VKRT* vkrt, VKRT_TraceStats* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    *out = vkrt->stats;
    return VKRT_SUCCESS;
}

static VKRT_Result requireTraceable(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    return syncScene(vkrt, 1);
}

VKRT_Result VKRT_renderBackward(VKRT* vkrt, const float* dLossdImage, VKRT_Gradients* out) {
    if (!vkrt || !dLossdImage || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = requireTraceable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (vkrt->settings.renderMode != VKRT_RENDER_MODE_RGB) {
        VKRT_LOG_ERROR("VKRT_renderBackward supports the RGB render mode");
        return VKRT_ERROR_INVALID_ARGUMENT;
    }
    if (out->materialCount != vkrt->packed.materialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->packed.textureTexelCount) {
        VKRT_LOG_ERROR("VKRT_renderBackward does not differentiate through material textures");
        return VKRT_ERROR_INVALID_ARGUMENT;
    }
    mrt_render_params rp;
    fillRenderParams(vkrt, &rp);
    if (mrt_resize_accumulation(vkrt->device, vkrt->renderWidth, vkrt->renderHeight) != 0) return deviceError();
    /* the next frame starts from an empty accumulation when a reset is pending:
     * replay the samples of that frame */
    if (vkrt->accumulationNeedsReset && mrt_clear_accumulation(vkrt->device) != 0) return deviceError();
    mrt_gradients g = {out->baseColor,   out->roughness,        out->emission,         out->materialCount,
                       out->environment, out->environmentWidth, out->environmentHeight};
    mrt_render_stats rs;
    memset(&rs, 0, sizeof(rs));
    if (mrt_render_backward(vkrt->device, &vkrt->sceneData, &rp, dLossdImage, &g, &rs) != 0) return deviceError();
    vkrt->stats.backwardMs = rs.traceMs;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_deviceSynchronize(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_INITIALIZATION_FAILED;
    return mrt_device_synchronize(vkrt->device) == 0 ? VKRT_SUCCESS : deviceError();
}

VKRT_Result VKRT_setMaterialSorting(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->materialSort = enabled ? 1u : 0u;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setFusedTail(VKRT* vkrt, uint32_t maxRays) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->fusedTailRays = maxRays;
    return VKRT_SUCCESS;
}

/* The feature images restart with the accumulation when they are turned on. */
VKRT_Result VKRT_setDenoiserFeatures(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    enabled = enabled ? 1u : 0u;
    if (vkrt->denoiserFeatures == enabled) return VKRT_SUCCESS;
    vkrt->denoiserFeatures = enabled;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_readDenoiserFeatures(const VKRT* vkrt, float* outAlbedo, float* outNormal, size_t floatCount) {
    if (!vkrt || !outAlbedo || !outNormal) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    if (floatCount != (size_t)vkrt->renderWidth * vkrt->renderHeight * 4u) return VKRT_ERROR_INVALID_ARGUMENT;
    return mrt_read_features(vkrt->device, outAlbedo, outNormal, floatCount) == 0 ? VKRT_SUCCESS
                                                                                    : VKRT_ERROR_OPERATION_FAILED;
}

VKRT_Result VKRT_setTraversalStatistics(VKRT* vkrt, uint8_t enabled) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->traversalStatistics = enabled ? 1u : 0u;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_readAccumulation(const VKRT* vkrt, float* out, size_t floatCount) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    if (floatCount != (size_t)vkrt->renderWidth * vkrt->renderHeight * 4u) return VKRT_ERROR_INVALID_ARGUMENT;
    return mrt_read_accumulation(vkrt->device, out, floatCount) == 0 ? VKRT_SUCCESS : deviceError();
}

VKRT_Result VKRT_readOutputImage(const VKRT* vkrt, uint16_t* out, size_t count) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    if (count != (size_t)vkrt->renderWidth * vkrt->renderHeight * 4u) return VKRT_ERROR_INVALID_ARGUMENT;
    return mrt_read_output(vkrt->device, out, count) == 0 ? VKRT_SUCCESS : deviceError();
}

VKRT_Result VKRT_exportSceneBuffers(VKRT* vkrt, VKRT_SceneBuffers* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = vkrtLoadRGB2SpecTable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    r = syncScene(vkrt, 0);
    if (r != VKRT_SUCCESS) return r;
    const PackedScene* p = &vkrt->packed;
    memset(out, 0, sizeof(*out));
    out->sceneData = &vkrt->sceneData;
    out->vertices = p->vertices;
    out->vertexCount = p->vertexCount;
    out->indices = p->indices;
    out->indexCount = p->indexCount;
    out->meshInfos = p->meshInfos;
    out->meshCount = vkrt->meshCount;
    out->materials = p->materials;
    out->materialCount = p->materialCount;
    out->emissiveMeshes = p->emissiveMeshes;
    out->emissiveTriangles = p->emissiveTriangles;
    out->meshAliasQ = p->meshAliasQ;
    out->meshAliasIdx = p->meshAliasIdx;
    out->triAliasQ = p->triAliasQ;
    out->triAliasIdx = p->triAliasIdx;
    out->rgb2specTable = vkrt->rgb2spec;
    out->rgb2specFloatCount = vkrt->rgb2specFloatCount;
    uint32_t ti = vkrt->settings.environmentTextureIndex;
    if (ti != VKRT_INVALID_INDEX && ti < vkrt->textureCount) {
        out->environmentPixels = vkrt->textures[ti].rgba;
        out->environmentWidth = vkrt->textures[ti].width;
        out->environmentHeight = vkrt->textures[ti].height;
    }
    out->worldTriangles = p->worldTriangles;
    out->triangleCount = p->triangleCount;
    out->triangleMesh = p->triangleMesh;
    out->rngPolicy = vkrt->rngPolicy;
    out->rngSeed = vkrt->rngSeed;
    out->environmentImportanceSampling = vkrt->envImportance;
    out->triangleAlpha = p->triangleAlpha;
    out->textureTexels = p->textureTexels;
    out->textureTexelCount = p->textureTexelCount;
    out->textureInfo = p->textureInfo;
    out->textureCount = p->textureCount;
    if (out->environmentPixels) {
        out->environmentAliasQ = p->envAliasQ;
        out->environmentAliasIdx = p->envAliasIdx;
        out->environmentPdf = p->envPdf;
    }
    return VKRT_SUCCESS;
}


VKRT_Result VKRT_traceRays(VKRT* vkrt, const float* rays, size_t count, uint32_t* outMesh, uint32_t* outPrimitive,
                           float* outT, float* outBarycentrics) {
    if (!rays || !outMesh || !outPrimitive) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = requireTraceable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (count == 0) return VKRT_SUCCESS;
    return mrt_trace_rays(vkrt->device, rays, count, outMesh, outPrimitive, outT, outBarycentrics) == 0 ? VKRT_SUCCESS
                                                                                                        : deviceError();
}

VKRT_Result VKRT_traceRaysDeviceNoAlloc(VKRT* vkrt, const float* deviceRays, size_t count, uint32_t* deviceMesh,
                                        uint32_t* devicePrimitive, float* deviceT, float* deviceBarycentrics) {
    if (!vkrt || !deviceRays || !deviceMesh || !devicePrimitive) return VKRT_ERROR_INVALID_ARGUMENT;
    if (!vkrt->device) return VKRT_ERROR_DEVICE_LOST;
    if (count > MRT_NOALLOC_MAX_RAYS) {
        VKRT_LOG_ERROR("VKRT_traceRaysDeviceNoAlloc: %zu rays exceed the maximum of 2^30 per call", count);
        return VKRT_ERROR_INVALID_ARGUMENT;
    }
    /* uploading a changed scene allocates: the scene must already be on the device */
    if (vkrt->geometryUploaded != vkrt->geometryRevision || vkrt->materialUploaded != vkrt->materialRevision ||
        vkrt->environmentUploaded != vkrt->environmentRevision || !vkrt->packed.meshInfos) {
        VKRT_LOG_ERROR("VKRT_traceRaysDeviceNoAlloc: the scene changed since its last upload; call VKRT_traceRays "
                       "or draw a frame first");
        return VKRT_ERROR_OPERATION_FAILED;
    }
    if (count == 0) return VKRT_SUCCESS;
    return mrt_trace_rays_noalloc(vkrt->device, deviceRays, count, deviceMesh, devicePrimitive, deviceT,
                                  deviceBarycentrics) == 0
               ? VKRT_SUCCESS
               : deviceError();
}

VKRT_Result VKRT_traceShadowRays(VKRT* vkrt, const float* rays, size_t count, uint32_t* outVisibility) {
    if (!rays || !outVisibility) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = requireTraceable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (count == 0) return VKRT_SUCCESS;
    return mrt_trace_shadow_rays(vkrt->device, rays, count, outVisibility) == 0 ? VKRT_SUCCESS : deviceError();
}

VKRT_Result VKRT_tracePrimaryHits(VKRT* vkrt, uint32_t sampleIndex, uint32_t* outMesh, uint32_t* outPrimitive) {
    if (!outMesh || !outPrimitive) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = requireTraceable(vkrt);
    if (r != VKRT_SUCCESS) return r;
    mrt_render_params rp;
    fillRenderParams(vkrt, &rp);
    return mrt_primary_hits(vkrt->device, &vkrt->sceneData, &rp, sampleIndex, outMesh, outPrimitive) == 0
               ? VKRT_SUCCESS
               : deviceError();
}
