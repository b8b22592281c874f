/*
 * Image files for textures and environment maps, with the reference's codec
 * choice and output rules (src/core/utility/image.c:86-399):
 *   - the codec comes from the MIME type ("image/png", "image/jpeg",
 *     "image/exr" or "image/x-exr"), else from the file signature;
 *   - PNG: RGBA8 in the requested colour space, or RGBA16 for 16-bit files,
 *     whose sRGB samples are converted to linear 16-bit values at load
 *     (rounded to nearest) and stored as linear;
 *   - JPEG: RGBA8 in the requested colour space;
 *   - EXR: RGBA16F or RGBA32F, always linear.
 */
#include "image_io.h"

#include "exr_decode.h"
#include "internal.h"
#include "jpeg_decode.h"
#include "png_decode.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { CODEC_UNKNOWN, CODEC_PNG, CODEC_JPEG, CODEC_EXR };

static int startsWith(const char* s, const char* prefix) { return strncmp(s, prefix, strlen(prefix)) == 0; }

static int codecOf(const uint8_t* data, size_t size, const char* mimeType) {
    if (mimeType && mimeType[0]) {
        if (startsWith(mimeType, "image/png")) return CODEC_PNG;
        if (startsWith(mimeType, "image/jpeg")) return CODEC_JPEG;
        if (startsWith(mimeType, "image/exr") || startsWith(mimeType, "image/x-exr")) return CODEC_EXR;
    }
    static const uint8_t png[8] = {137u, 80u, 78u, 71u, 13u, 10u, 26u, 10u};
    static const uint8_t jpeg[3] = {0xffu, 0xd8u, 0xffu};
    static const uint8_t exr[4] = {0x76u, 0x2fu, 0x31u, 0x01u};
    if (size >= 8 && memcmp(data, png, 8) == 0) return CODEC_PNG;
    if (size >= 3 && memcmp(data, jpeg, 3) == 0) return CODEC_JPEG;
    if (size >= 4 && memcmp(data, exr, 4) == 0) return CODEC_EXR;
    return CODEC_UNKNOWN;
}

static float srgbDecode(float v) { return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f); }

int vkrtLoadImageFromMemory(const void* data, size_t size, const char* mimeType, uint32_t preferredColorSpace,
                            VkrtLoadedImage* out, const char* label) {
    memset(out, 0, sizeof(*out));
    if (!data || size == 0) return -1;
    const char* what = label ? label : "image";
    char err[256] = "";
    switch (codecOf((const uint8_t*)data, size, mimeType)) {
        case CODEC_PNG: {
            PngImage png;
            if (vkrtDecodePng((const uint8_t*)data, size, &png) != 0) {
                VKRT_LOG_ERROR("PNG decode from %s failed", what);
                return -1;
            }
            out->pixels = png.pixels;
            out->width = png.width;
            out->height = png.height;
            out->colorSpace = preferredColorSpace;
            out->format = VKRT_TEXTURE_FORMAT_RGBA8_UNORM;
            if (png.bits == 16) {
                out->format = VKRT_TEXTURE_FORMAT_RGBA16_UNORM;
                if (preferredColorSpace == VKRT_TEXTURE_COLOR_SPACE_SRGB) {
                    uint16_t* p = (uint16_t*)png.pixels;
                    const size_t n = (size_t)png.width * png.height;
                    for (size_t i = 0; i < n; i++) {
                        for (int c = 0; c < 3; c++) {
                            long u = lrintf(srgbDecode((float)p[i * 4 + c] / 65535.0f) * 65535.0f);
                            p[i * 4 + c] = (uint16_t)(u < 0 ? 0 : (u > 65535 ? 65535 : u));
                        }
                    }
                    out->colorSpace = VKRT_TEXTURE_COLOR_SPACE_LINEAR;
                }
            }
            return 0;
        }
        case CODEC_JPEG: {
            JpegImage jpg;
            if (vkrtDecodeJpeg((const uint8_t*)data, size, &jpg, err, sizeof(err)) != 0) {
                VKRT_LOG_ERROR("JPEG decode from %s failed (%s)", what, err);
                return -1;
            }
            out->pixels = jpg.pixels;
            out->width = jpg.width;
            out->height = jpg.height;
            out->format = VKRT_TEXTURE_FORMAT_RGBA8_UNORM;
            out->colorSpace = preferredColorSpace;
            return 0;
        }
        case CODEC_EXR: {
            ExrImage exr;
            if (vkrtDecodeExr((const uint8_t*)data, size, &exr, err, sizeof(err)) != 0) {
                VKRT_LOG_ERROR("EXR decode from %s failed (%s)", what, err);
                return -1;
            }
            out->pixels = exr.pixels;
            out->width = exr.width;
            out->height = exr.height;
            out->format = exr.half ? VKRT_TEXTURE_FORMAT_RGBA16_SFLOAT : VKRT_TEXTURE_FORMAT_RGBA32_SFLOAT;
            out->colorSpace = VKRT_TEXTURE_COLOR_SPACE_LINEAR;
            return 0;
        }
        default:
            if (mimeType && mimeType[0]) VKRT_LOG_ERROR("Unsupported image format for %s (%s)", what, mimeType);
            else VKRT_LOG_ERROR("Unsupported image format for %s", what);
            return -1;
    }
}

int vkrtLoadImageFromFile(const char* path, uint32_t preferredColorSpace, VkrtLoadedImage* out) {
    memset(out, 0, sizeof(*out));
    if (!path || !path[0]) return -1;
    FILE* f = fopen(path, "rb");
    if (!f) {
        VKRT_LOG_ERROR("Failed to open image file: %s", path);
        return -1;
    }
    uint8_t* bytes = NULL;
    long length = -1;
    if (fseek(f, 0, SEEK_END) == 0) length = ftell(f);
    if (length > 0 && fseek(f, 0, SEEK_SET) == 0) {
        bytes = (uint8_t*)malloc((size_t)length);
        if (bytes && fread(bytes, 1, (size_t)length, f) != (size_t)length) {
            free(bytes);
            bytes = NULL;
        }
    }
    fclose(f);
    if (!bytes) {
        VKRT_LOG_ERROR("Failed to read image file: %s", path);
        return -1;
    }
    int rc = vkrtLoadImageFromMemory(bytes, (size_t)length, NULL, preferredColorSpace, out, path);
    free(bytes);
    return rc;
}
