/* PNG decoding for texture and environment files. The reference decodes PNG
 * with libspng into RGBA8, or RGBA16 for 16-bit files, applying palette alpha
 * but no colour-key transparency (spng_decode_image with flags 0,
 * src/core/utility/image.c:123-192); this decoder gives the same samples.
 * Every colour type and bit depth of the PNG specification (1, 2, 4, 8 and
 * 16 bits), the five scanline filters and Adam7 interlacing are supported.
 * Greyscale below 8 bits scales to 8 bits by v * 255 / (2^depth - 1), which
 * is exact for depths 1, 2 and 4. */
#include "png_decode.h"

#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

static uint8_t paeth(uint8_t a, uint8_t b, uint8_t c) {
    int p = (int)a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

/* Undoes the scanline filters of one (sub)image in place; rows of `row` bytes
 * each preceded by the filter byte. */
static int unfilter(uint8_t* raw, uint32_t h, size_t row, uint32_t bpp) {
    for (uint32_t y = 0; y < h; y++) {
        uint8_t* cur = raw + y * (row + 1u) + 1u;
        const uint8_t* prev = y ? cur - (row + 1u) : NULL;
        const uint8_t filter = cur[-1];
        for (size_t x = 0; x < row; x++) {
            uint8_t a = x >= bpp ? cur[x - bpp] : 0, b = prev ? prev[x] : 0, c = (prev && x >= bpp) ? prev[x - bpp] : 0;
            switch (filter) {
                case 0: break;
                case 1: cur[x] = (uint8_t)(cur[x] + a); break;
                case 2: cur[x] = (uint8_t)(cur[x] + b); break;
                case 3: cur[x] = (uint8_t)(cur[x] + ((a + b) >> 1)); break;
                case 4: cur[x] = (uint8_t)(cur[x] + paeth(a, b, c)); break;
                default: return -1;
            }
        }
    }
    return 0;
}

/* Sample c of pixel x in an unfiltered row, at any bit depth. */
static uint32_t sampleAt(const uint8_t* src, uint32_t x, uint32_t ch, uint32_t c, uint32_t depth) {
    if (depth == 16) return ((uint32_t)src[(x * ch + c) * 2u] << 8) | src[(x * ch + c) * 2u + 1u];
    if (depth == 8) return src[x * ch + c];
    const uint32_t bit = (x * ch + c) * depth; /* sub-byte depths only occur with one channel */
    return (src[bit >> 3] >> (8u - depth - (bit & 7u))) & ((1u << depth) - 1u);
}

int vkrtDecodePng(const uint8_t* data, size_t size, PngImage* out) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    memset(out, 0, sizeof(*out));
    if (!data || size < 8 || memcmp(data, sig, 8) != 0) return -1;
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0, paletteCount = 0;
    uint8_t palette[256][4];
    memset(palette, 255, sizeof(palette));
    uint8_t* idat = NULL;
    size_t idatSize = 0, pos = 8;
    int ok = 0;
    while (pos + 12 <= size) {
        uint32_t n = be32(data + pos);
        const uint8_t* type = data + pos + 4;
        const uint8_t* body = data + pos + 8;
        if (n > size - pos - 12) break;
        if (!memcmp(type, "IHDR", 4) && n >= 13) {
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!memcmp(type, "PLTE", 4)) {
            paletteCount = n / 3u > 256u ? 256u : n / 3u;
            for (uint32_t i = 0; i < paletteCount; i++)
                for (int c = 0; c < 3; c++) palette[i][c] = body[3 * i + c];
        } else if (!memcmp(type, "tRNS", 4) && ctype == 3) {
            for (uint32_t i = 0; i < n && i < 256; i++) palette[i][3] = body[i];
        } else if (!memcmp(type, "IDAT", 4)) {
            uint8_t* grown = (uint8_t*)realloc(idat, idatSize + n + 1);
            if (!grown) break;
            idat = grown;
            memcpy(idat + idatSize, body, n);
            idatSize += n;
        } else if (!memcmp(type, "IEND", 4)) {
            ok = 1;
            break;
        }
        pos += 12u + n;
    }
    static const uint32_t channelsOf[7] = {1, 0, 3, 1, 2, 0, 4};
    int depthOk = ctype <= 6 && channelsOf[ctype] != 0 &&
                  ((ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                   (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                   ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16)));
    if (!ok || !idat || w == 0 || h == 0 || w > 32768 || h > 32768 || interlace > 1 || !depthOk ||
        (ctype == 3 && paletteCount == 0)) {
        free(idat);
        return -1;
    }
    const uint32_t ch = channelsOf[ctype];
    const uint32_t bitsPerPixel = ch * depth, bpp = bitsPerPixel < 8 ? 1u : bitsPerPixel / 8u;
    /* Adam7 passes: start and step in x and y; a non-interlaced image is one pass */
    static const uint32_t x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const uint32_t dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    const int passes = interlace ? 7 : 1;
    size_t total = 0;
    for (int p = 0; p < passes; p++) {
        uint32_t pw = interlace ? (w + dx[p] - 1 - x0[p]) / dx[p] : w, ph = interlace ? (h + dy[p] - 1 - y0[p]) / dy[p] : h;
        if (w <= x0[p]) pw = 0;
        if (h <= y0[p]) ph = 0;
        if (pw && ph) total += ((size_t)pw * bitsPerPixel + 7u) / 8u * ph + ph;
    }
    uLongf rawSize = (uLongf)total;
    uint8_t* raw = (uint8_t*)malloc(total ? total : 1u);
    if (!raw || uncompress(raw, &rawSize, idat, (uLong)idatSize) != Z_OK || rawSize != total) {
        free(raw);
        free(idat);
        return -1;
    }
    free(idat);
    const int wide = depth == 16;
    void* px = malloc((size_t)w * h * 4u * (wide ? 2u : 1u));
    if (!px) {
        free(raw);
        return -1;
    }
    const uint32_t greyScale = depth < 8 ? 255u / ((1u << depth) - 1u) : 1u;
    uint8_t* passData = raw;
    for (int p = 0; p < passes; p++) {
        uint32_t pw = interlace ? (w + dx[p] - 1 - x0[p]) / dx[p] : w, ph = interlace ? (h + dy[p] - 1 - y0[p]) / dy[p] : h;
        if (w <= x0[p]) pw = 0;
        if (h <= y0[p]) ph = 0;
        if (!pw || !ph) continue;
        const size_t row = ((size_t)pw * bitsPerPixel + 7u) / 8u;
        if (unfilter(passData, ph, row, bpp) != 0) {
            free(raw);
            free(px);
            return -1;
        }
        for (uint32_t py = 0; py < ph; py++) {
            const uint8_t* src = passData + py * (row + 1u) + 1u;
            const uint32_t y = interlace ? y0[p] + py * dy[p] : py;
            for (uint32_t qx = 0; qx < pw; qx++) {
                const uint32_t x = interlace ? x0[p] + qx * dx[p] : qx;
                uint32_t v[4] = {0, 0, 0, wide ? 65535u : 255u};
                uint32_t s[4];
                for (uint32_t c = 0; c < ch; c++) s[c] = sampleAt(src, qx, ch, c, depth);
                if (ctype == 0 || ctype == 4) {
                    v[0] = v[1] = v[2] = s[0] * greyScale;
                    if (ctype == 4) v[3] = s[1];
                } else if (ctype == 3) {
                    const uint8_t* pe = palette[s[0]];
                    for (int c = 0; c < 4; c++) v[c] = pe[c];
                } else {
                    for (int c = 0; c < 3; c++) v[c] = s[c];
                    if (ctype == 6) v[3] = s[3];
                }
                size_t o = ((size_t)y * w + x) * 4u;
                for (int c = 0; c < 4; c++) {
                    if (wide)
                        ((uint16_t*)px)[o + c] = (uint16_t)v[c];
                    else
                        ((uint8_t*)px)[o + c] = (uint8_t)v[c];
                }
            }
        }
        passData += (row + 1u) * ph;
    }
    free(raw);
    out->pixels = px;
    out->width = w;
    out->height = h;
    out->bits = wide ? 16u : 8u;
    return 0;
}
