#pragma once
#include <stddef.h>
#include <stdint.h>

/* A decoded texture in one of the VKRT_TEXTURE_FORMAT_* layouts. */
typedef struct VkrtLoadedImage {
    void* pixels;
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t colorSpace;
} VkrtLoadedImage;

/* Decode PNG, JPEG or EXR bytes (codec from the MIME type, else from the
 * signature). Returns 0 on success; the caller frees pixels. */
int vkrtLoadImageFromMemory(const void* data, size_t size, const char* mimeType, uint32_t preferredColorSpace,
                            VkrtLoadedImage* out, const char* label);
int vkrtLoadImageFromFile(const char* path, uint32_t preferredColorSpace, VkrtLoadedImage* out);
