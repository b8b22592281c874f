/*
 * Loads the sRGB -> spectrum coefficient table. The header check follows
 * src/core/scene/rgb2spec.c:17-59 ("SPEC" magic, res >= 2, exact payload size);
 * the table itself is produced by csrc/tools/rgb2spec_opt.c at build time.
 * Search order: $VKRT_RGB2SPEC_PATH, then data/srgb.coeff beside the library.
 */
#define _GNU_SOURCE
#include "internal.h"

#include <dlfcn.h>
#include <stdlib.h>
#include <string.h>

int vkrtParseRGB2Spec(const uint8_t* data, size_t size, RGB2SpecTableInfo* info, const float** payload,
                      size_t* floatCount) {
    if (!data || !info || !payload || !floatCount) return 0;
    if (size < 8u || memcmp(data, "SPEC", 4u) != 0) return 0;
    uint32_t res = 0;
    memcpy(&res, data + 4, 4);
    if (res < 2u) return 0;
    uint64_t count = (uint64_t)res + (uint64_t)3u * res * res * res * 3u;
    if (size != 8u + count * sizeof(float)) return 0;
    info->res = res;
    info->scaleOffset = 0u;
    info->dataOffset = res;
    *payload = (const float*)(data + 8);
    *floatCount = (size_t)count;
    return 1;
}

static uint8_t* readWholeFile(const char* path, size_t* outSize) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return NULL;
    }
    long size = ftell(f);
    if (size <= 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return NULL;
    }
    uint8_t* buf = (uint8_t*)malloc((size_t)size);
    if (buf && fread(buf, 1, (size_t)size, f) != (size_t)size) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    if (buf) *outSize = (size_t)size;
    return buf;
}

static int defaultTablePath(char* out, size_t len) {
    Dl_info dl;
    if (!dladdr((void*)&defaultTablePath, &dl) || !dl.dli_fname) return 0;
    const char* slash = strrchr(dl.dli_fname, '/');
    size_t dirLen = slash ? (size_t)(slash - dl.dli_fname) : 0u;
    int n = snprintf(out, len, "%.*s%sdata/srgb.coeff", (int)dirLen, dl.dli_fname, dirLen ? "/" : "");
    return n > 0 && (size_t)n < len;
}

VKRT_Result vkrtLoadRGB2SpecTable(VKRT* vkrt) {
    if (vkrt->rgb2spec) return VKRT_SUCCESS;
    char path[4096];
    const char* env = getenv("VKRT_RGB2SPEC_PATH");
    if (env && env[0]) {
        snprintf(path, sizeof(path), "%s", env);
    } else if (!defaultTablePath(path, sizeof(path))) {
        return VKRT_ERROR_INITIALIZATION_FAILED;
    }
    size_t size = 0;
    uint8_t* data = readWholeFile(path, &size);
    if (!data) {
        VKRT_LOG_ERROR("rgb2spec table not found at %s", path);
        return VKRT_ERROR_INITIALIZATION_FAILED;
    }
    RGB2SpecTableInfo info;
    const float* payload = NULL;
    size_t count = 0;
    if (!vkrtParseRGB2Spec(data, size, &info, &payload, &count)) {
        free(data);
        VKRT_LOG_ERROR("rgb2spec table at %s is malformed", path);
        return VKRT_ERROR_INITIALIZATION_FAILED;
    }
    vkrt->rgb2spec = (float*)malloc(count * sizeof(float));
    if (!vkrt->rgb2spec) {
        free(data);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    memcpy(vkrt->rgb2spec, payload, count * sizeof(float));
    vkrt->rgb2specFloatCount = (uint32_t)count;
    vkrt->sceneData.rgb2specSRGB = info;
    free(data);
    return VKRT_SUCCESS;
}
