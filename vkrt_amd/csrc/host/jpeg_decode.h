#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

typedef struct JpegImage {
    uint8_t* pixels; /* width * height RGBA8, row 0 first, alpha 255 */
    uint32_t width;
    uint32_t height;
} JpegImage;

/* Returns 0 on success (the caller frees pixels); otherwise writes a reason to err. */
int vkrtDecodeJpeg(const uint8_t* data, size_t size, JpegImage* out, char* err, size_t errSize);
