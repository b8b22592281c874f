#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

typedef struct ExrImage {
    void* pixels; /* width * height RGBA: uint16 half bits when half != 0, else float */
    uint32_t width;
    uint32_t height;
    int half;
} ExrImage;

/* Returns 0 on success (the caller frees pixels); otherwise writes a reason to err. */
int vkrtDecodeExr(const uint8_t* data, size_t size, ExrImage* out, char* err, size_t errSize);
