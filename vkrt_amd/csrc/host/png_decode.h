#pragma once
#include <stddef.h>
#include <stdint.h>

typedef struct PngImage {
    void* pixels; /* width * height * 4 channels of uint8 (bits 8) or uint16 (bits 16), row 0 first */
    uint32_t width;
    uint32_t height;
    uint32_t bits;
} PngImage;

/* Returns 0 on success; the caller frees pixels. */
int vkrtDecodePng(const uint8_t* data, size_t size, PngImage* out);
