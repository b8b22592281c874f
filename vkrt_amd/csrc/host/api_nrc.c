/*
 * Neural radiance cache settings and direct network access. An addition of
 * this build: the reference (src/core/api/vkrt.h) has no neural cache.
 */
#include "internal.h"

#include <string.h>

void VKRT_defaultNeuralCacheSettings(VKRT_NeuralCacheSettings* out) {
    if (!out) return;
    memset(out, 0, sizeof(*out));
    /* measured best on the 1080p Cornell orbit: querying after one bounce with
     * one step of 16384 records at 1e-2 per frame (two bounces, more steps and
     * smaller steps all reached a given error later) */
    out->prefixBounces = 1u;
    out->trainingPaths = 65536u;
    out->trainingSteps = 1u;
    out->batchSize = 16384u;
    out->learningRate = 1e-2f;
    out->train = 1u;
}

VKRT_Result VKRT_setNeuralCache(VKRT* vkrt, const VKRT_NeuralCacheSettings* s) {
    if (!vkrt || !s) return VKRT_ERROR_INVALID_ARGUMENT;
    if (s->enabled && (s->prefixBounces == 0u || s->batchSize == 0u || !(s->learningRate >= 0.0f)))
        return VKRT_ERROR_INVALID_ARGUMENT;
    uint8_t changed = vkrt->nrc.enabled != s->enabled || vkrt->nrc.prefixBounces != s->prefixBounces ||
                      vkrt->nrc.trainingPaths != s->trainingPaths;
    vkrt->nrc = *s;
    if (changed) vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getNeuralCacheStats(const VKRT* vkrt, VKRT_NeuralCacheStats* out) {
    if (!vkrt || !out) return VKRT_ERROR_INVALID_ARGUMENT;
    *out = vkrt->nrcStats;
    return VKRT_SUCCESS;
}

uint32_t VKRT_neuralCacheParameterCount(void) { return (uint32_t)mrt_nrc_parameter_count(); }

static VKRT_Result nrcDevice(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    return vkrt->device ? VKRT_SUCCESS : VKRT_ERROR_DEVICE_LOST;
}

static VKRT_Result nrcResult(int code) {
    if (code == 0) return VKRT_SUCCESS;
    if (code == -1) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_LOG_ERROR("neural cache: %s", mrt_last_error());
    return VKRT_ERROR_OPERATION_FAILED;
}

VKRT_Result VKRT_neuralCacheReset(VKRT* vkrt, uint32_t seed) {
    VKRT_Result r = nrcDevice(vkrt);
    return r != VKRT_SUCCESS ? r : nrcResult(mrt_nrc_reset(vkrt->device, seed));
}

VKRT_Result VKRT_neuralCacheGetWeights(VKRT* vkrt, float* out, size_t count) {
    VKRT_Result r = nrcDevice(vkrt);
    return r != VKRT_SUCCESS ? r : nrcResult(mrt_nrc_get_weights(vkrt->device, out, count));
}

VKRT_Result VKRT_neuralCacheSetWeights(VKRT* vkrt, const float* w, size_t count) {
    VKRT_Result r = nrcDevice(vkrt);
    return r != VKRT_SUCCESS ? r : nrcResult(mrt_nrc_set_weights(vkrt->device, w, count));
}

VKRT_Result VKRT_neuralCacheInfer(VKRT* vkrt, const float* inputs, uint32_t count, float* outputs) {
    VKRT_Result r = nrcDevice(vkrt);
    return r != VKRT_SUCCESS ? r : nrcResult(mrt_nrc_infer(vkrt->device, inputs, count, outputs));
}

VKRT_Result VKRT_neuralCacheTrain(VKRT* vkrt, const float* inputs, const float* targets, uint32_t count,
                                  uint32_t steps, uint32_t batchSize, float learningRate, float* outLoss) {
    VKRT_Result r = nrcDevice(vkrt);
    if (r != VKRT_SUCCESS) return r;
    return nrcResult(mrt_nrc_train(vkrt->device, inputs, targets, count, steps, batchSize, learningRate, outLoss));
}

VKRT_Result VKRT_neuralCacheGradient(VKRT* vkrt, const float* inputs, const float* targets, uint32_t count,
                                     float* outGradient, float* outLoss) {
    VKRT_Result r = nrcDevice(vkrt);
    return r != VKRT_SUCCESS ? r : nrcResult(mrt_nrc_gradient(vkrt->device, inputs, targets, count, outGradient, outLoss));
}
