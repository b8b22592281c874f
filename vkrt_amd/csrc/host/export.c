/*
 * Render image export: VKRT_saveRenderImage / VKRT_saveRenderImageEx
 * (src/core/api/vkrt.h:54-58, src/core/utility/export/api.c:170-242,
 * src/core/utility/export/image.c).
 *
 * Behaviour follows the reference:
 *  - the path's extension picks the format (.png, .jpg/.jpeg, .exr, any case);
 *    a path without one gets ".png"; any other extension is an error;
 *  - PNG and JPEG without denoising write the display image (exposure, tone
 *    map, sRGB; RGBA16), JPEG as 8 bits (v * 255 + 32767) / 65535, 4:4:4,
 *    quality 95;
 *  - EXR, or any format with denoising requested, starts from the linear
 *    accumulation: spectral modes convert XYZ (equal-energy white) to linear
 *    sRGB through a Bradford adaptation to D65 (image.c:388-423), alpha is 1,
 *    non-finite values become 0 (image.c:425-447); PNG/JPEG then tone-map on
 *    the host exactly as image.c:641-699 does.
 * The reference denoises with OIDN, which this library does not carry: a
 * request to denoise logs a note once and exports the raw image, as the
 * reference's fallback does when its denoiser fails.
 *
 * Encoders are this library's own: PNG through zlib (16-bit RGBA, filter 0),
 * EXR as uncompressed scanlines of four FLOAT channels (A, B, G, R, the order
 * OpenEXR sorts channel names in), baseline JPEG with the ITU-T T.81 Annex K
 * tables. The reference encodes with libspng, tinyexr and libjpeg-turbo; files
 * decode to the same pixels except for JPEG's lossy coding.
 */
#include "internal.h"

#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

typedef enum { kFormatPng, kFormatJpeg, kFormatExr } ImageFormat;

/* ------------------------------------------------------------------ paths */

static int extensionIs(const char* ext, const char* want) {
    while (*ext && *want) {
        if (tolower((unsigned char)*ext) != *want) return 0;
        ext++;
        want++;
    }
    return *ext == '\0' && *want == '\0';
}

/* image.c:74-148 */
static char* resolvePath(const char* path, ImageFormat* format) {
    const char* base = strrchr(path, '/');
    base = base ? base + 1 : path;
    const char* dot = strrchr(base, '.');
    size_t n = strlen(path);
    if (!dot || dot == base || dot[1] == '\0') {
        char* out = (char*)malloc(n + 5);
        if (!out) return NULL;
        snprintf(out, n + 5, "%s.png", path);
        *format = kFormatPng;
        return out;
    }
    if (extensionIs(dot, ".png"))
        *format = kFormatPng;
    else if (extensionIs(dot, ".jpg") || extensionIs(dot, ".jpeg"))
        *format = kFormatJpeg;
    else if (extensionIs(dot, ".exr"))
        *format = kFormatExr;
    else {
        VKRT_LOG_ERROR("Unsupported render export format: %s", path);
        return NULL;
    }
    char* out = (char*)malloc(n + 1);
    if (out) memcpy(out, path, n + 1);
    return out;
}

/* ------------------------------------------------------------------- PNG */

static void be32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

static int writeChunk(FILE* f, const char* type, const uint8_t* data, uint32_t len) {
    uint8_t head[8];
    be32(head, len);
    memcpy(head + 4, type, 4);
    uLong crc = crc32(0L, (const Bytef*)type, 4);
    if (len) crc = crc32(crc, data, len);
    uint8_t tail[4];
    be32(tail, (uint32_t)crc);
    return fwrite(head, 1, 8, f) == 8 && (len == 0 || fwrite(data, 1, len, f) == len) && fwrite(tail, 1, 4, f) == 4;
}

static int writePng(const char* path, const uint16_t* rgba16, uint32_t w, uint32_t h) {
    size_t row = (size_t)w * 8u + 1u;
    size_t rawSize = row * h;
    uint8_t* raw = (uint8_t*)malloc(rawSize);
    uLongf zSize = compressBound((uLong)rawSize);
    uint8_t* z = (uint8_t*)malloc(zSize);
    if (!raw || !z) {
        free(raw);
        free(z);
        return 0;
    }
    for (uint32_t y = 0; y < h; y++) {
        uint8_t* r = raw + row * y;
        r[0] = 0; /* filter: none */
        const uint16_t* src = rgba16 + (size_t)y * w * 4u;
        for (size_t i = 0; i < (size_t)w * 4u; i++) {
            r[1 + 2 * i] = (uint8_t)(src[i] >> 8);
            r[2 + 2 * i] = (uint8_t)src[i];
        }
    }
    int ok = compress2(z, &zSize, raw, (uLong)rawSize, 6) == Z_OK && zSize <= 0xFFFFFFFFu;
    free(raw);
    FILE* f = ok ? fopen(path, "wb") : NULL;
    if (f) {
        static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
        uint8_t ihdr[13];
        be32(ihdr, w);
        be32(ihdr + 4, h);
        ihdr[8] = 16; /* bit depth */
        ihdr[9] = 6;  /* RGBA */
        ihdr[10] = ihdr[11] = ihdr[12] = 0;
        ok = fwrite(sig, 1, 8, f) == 8 && writeChunk(f, "IHDR", ihdr, 13) && writeChunk(f, "IDAT", z, (uint32_t)zSize) &&
             writeChunk(f, "IEND", NULL, 0);
        ok = (fclose(f) == 0) && ok;
    } else {
        ok = 0;
    }
    free(z);
    return ok;
}

/* ------------------------------------------------------------------- EXR */

static void put(uint8_t** p, const void* src, size_t n) {
    memcpy(*p, src, n);
    *p += n;
}
static void putStr(uint8_t** p, const char* s) { put(p, s, strlen(s) + 1); }
static void putI32(uint8_t** p, int32_t v) { put(p, &v, 4); } /* little-endian host */
static void putAttr(uint8_t** p, const char* name, const char* type, int32_t size) {
    putStr(p, name);
    putStr(p, type);
    putI32(p, size);
}

/* OpenEXR 2.0 single-part scanline file, NO_COMPRESSION, FLOAT channels. */
static int writeExr(const char* path, const float* rgba, uint32_t w, uint32_t h) {
    uint8_t header[512];
    uint8_t* p = header;
    const int32_t magic = 20000630, version = 2;
    putI32(&p, magic);
    putI32(&p, version);
    static const char* names[4] = {"A", "B", "G", "R"};
    putAttr(&p, "channels", "chlist", 4 * (2 + 16) + 1);
    for (int c = 0; c < 4; c++) {
        putStr(&p, names[c]);
        putI32(&p, 2); /* FLOAT */
        uint8_t linear[4] = {0, 0, 0, 0};
        put(&p, linear, 4);
        putI32(&p, 1); /* x sampling */
        putI32(&p, 1); /* y sampling */
    }
    *p++ = 0;
    putAttr(&p, "compression", "compression", 1);
    *p++ = 0; /* NO_COMPRESSION */
    int32_t box[4] = {0, 0, (int32_t)w - 1, (int32_t)h - 1};
    putAttr(&p, "dataWindow", "box2i", 16);
    put(&p, box, 16);
    putAttr(&p, "displayWindow", "box2i", 16);
    put(&p, box, 16);
    putAttr(&p, "lineOrder", "lineOrder", 1);
    *p++ = 0; /* INCREASING_Y */
    float one = 1.0f, zero2[2] = {0.0f, 0.0f};
    putAttr(&p, "pixelAspectRatio", "float", 4);
    put(&p, &one, 4);
    putAttr(&p, "screenWindowCenter", "v2f", 8);
    put(&p, zero2, 8);
    putAttr(&p, "screenWindowWidth", "float", 4);
    put(&p, &one, 4);
    *p++ = 0; /* end of header */
    size_t headerSize = (size_t)(p - header);
    size_t lineBytes = (size_t)w * 16u;
    size_t blockSize = 8u + lineBytes;
    FILE* f = fopen(path, "wb");
    if (!f) return 0;
    int ok = fwrite(header, 1, headerSize, f) == headerSize;
    uint64_t first = headerSize + 8u * (uint64_t)h;
    for (uint32_t y = 0; ok && y < h; y++) {
        uint64_t off = first + (uint64_t)y * blockSize;
        ok = fwrite(&off, 8, 1, f) == 1;
    }
    uint8_t* block = (uint8_t*)malloc(blockSize);
    if (!block) ok = 0;
    for (uint32_t y = 0; ok && y < h; y++) {
        uint8_t* q = block;
        putI32(&q, (int32_t)y);
        putI32(&q, (int32_t)lineBytes);
        static const int src[4] = {3, 2, 1, 0}; /* A, B, G, R */
        for (int c = 0; c < 4; c++)
            for (uint32_t x = 0; x < w; x++) put(&q, &rgba[((size_t)y * w + x) * 4u + (size_t)src[c]], 4);
        ok = fwrite(block, 1, blockSize, f) == blockSize;
    }
    free(block);
    ok = (fclose(f) == 0) && ok;
    return ok;
}

/* ------------------------------------------------------------------ JPEG */
/* Baseline sequential DCT, Huffman coded (ITU-T T.81), JFIF, 4:4:4. */

static const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
/* T.81 Annex K.1, natural (row-major) order */
static const uint8_t kLumaQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                                   14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                                   18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                                   49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
static const uint8_t kChromaQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                     24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                     99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                     99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
/* T.81 Annex K.3: code counts per length 1..16, then symbols */
static const uint8_t kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
static const uint8_t kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
static const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
static const uint8_t kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
static const uint8_t kAcLumaVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
static const uint8_t kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
static const uint8_t kAcChromaVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

typedef struct HuffTable {
    uint16_t code[256];
    uint8_t size[256];
} HuffTable;

/* Canonical codes from the per-length counts (T.81 Annex C). */
static void buildHuff(const uint8_t bits[16], const uint8_t* vals, HuffTable* t) {
    memset(t, 0, sizeof(*t));
    uint16_t code = 0;
    int k = 0;
    for (int len = 1; len <= 16; len++) {
        for (int i = 0; i < bits[len - 1]; i++, k++) {
            t->code[vals[k]] = code++;
            t->size[vals[k]] = (uint8_t)len;
        }
        code <<= 1;
    }
}

typedef struct BitWriter {
    uint8_t* buf;
    size_t len, cap;
    uint32_t acc;
    int nbits;
    int failed;
} BitWriter;

static void emitByte(BitWriter* b, uint8_t v) {
    if (b->len + 2 > b->cap) {
        size_t cap = b->cap ? b->cap * 2 : 65536;
        uint8_t* g = (uint8_t*)realloc(b->buf, cap);
        if (!g) {
            b->failed = 1;
            return;
        }
        b->buf = g;
        b->cap = cap;
    }
    b->buf[b->len++] = v;
}
static void emitBits(BitWriter* b, uint32_t value, int n) {
    b->acc = (b->acc << n) | (value & ((1u << n) - 1u));
    b->nbits += n;
    while (b->nbits >= 8) {
        uint8_t byte = (uint8_t)(b->acc >> (b->nbits - 8));
        emitByte(b, byte);
        if (byte == 0xFF) emitByte(b, 0); /* byte stuffing */
        b->nbits -= 8;
    }
}
static void flushBits(BitWriter* b) {
    if (b->nbits > 0) emitBits(b, 0x7F, 8 - b->nbits); /* pad with ones */
}

static int category(int v) {
    int a = v < 0 ? -v : v, n = 0;
    while (a) {
        n++;
        a >>= 1;
    }
    return n;
}

/* 8x8 forward DCT-II (orthonormal scaling of T.81 A.3.3) and quantisation. */
static void encodeBlock(BitWriter* bw, const float in[64], const float qdiv[64], int* prevDc, const HuffTable* dc,
                        const HuffTable* ac) {
    static float cosTable[8][8];
    static int ready = 0;
    if (!ready) {
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++) cosTable[x][u] = (float)cos((2.0 * x + 1.0) * u * M_PI / 16.0);
        ready = 1;
    }
    float tmp[64], coef[64];
    for (int y = 0; y < 8; y++)
        for (int u = 0; u < 8; u++) {
            float s = 0.0f;
            for (int x = 0; x < 8; x++) s += in[y * 8 + x] * cosTable[x][u];
            tmp[y * 8 + u] = s * (u == 0 ? (float)M_SQRT1_2 : 1.0f) * 0.5f;
        }
    for (int u = 0; u < 8; u++)
        for (int v = 0; v < 8; v++) {
            float s = 0.0f;
            for (int y = 0; y < 8; y++) s += tmp[y * 8 + u] * cosTable[y][v];
            coef[v * 8 + u] = s * (v == 0 ? (float)M_SQRT1_2 : 1.0f) * 0.5f;
        }
    int q[64];
    for (int i = 0; i < 64; i++) {
        int n = kZigzag[i];
        q[i] = (int)lroundf(coef[n] / qdiv[n]);
    }
    int diff = q[0] - *prevDc;
    *prevDc = q[0];
    int c = category(diff);
    emitBits(bw, dc->code[c], dc->size[c]);
    if (c) emitBits(bw, (uint32_t)(diff < 0 ? diff - 1 : diff), c);
    int run = 0;
    for (int i = 1; i < 64; i++) {
        if (q[i] == 0) {
            run++;
            continue;
        }
        while (run > 15) {
            emitBits(bw, ac->code[0xF0], ac->size[0xF0]);
            run -= 16;
        }
        int cat = category(q[i]);
        int sym = (run << 4) | cat;
        emitBits(bw, ac->code[sym], ac->size[sym]);
        emitBits(bw, (uint32_t)(q[i] < 0 ? q[i] - 1 : q[i]), cat);
        run = 0;
    }
    if (run > 0) emitBits(bw, ac->code[0x00], ac->size[0x00]);
}

static void putMarker(BitWriter* b, uint8_t m) {
    emitByte(b, 0xFF);
    emitByte(b, m);
}
static void put16(BitWriter* b, uint32_t v) {
    emitByte(b, (uint8_t)(v >> 8));
    emitByte(b, (uint8_t)v);
}

static int writeJpeg(const char* path, const uint8_t* rgba8, uint32_t w, uint32_t h, int quality) {
    if (w > 65535u || h > 65535u) {
        VKRT_LOG_ERROR("JPEG export dimensions exceed codec limits: %ux%u", w, h);
        return 0;
    }
    int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    uint8_t qt[2][64];
    float qdiv[2][64];
    for (int i = 0; i < 64; i++) {
        int a = (kLumaQ[i] * scale + 50) / 100, c = (kChromaQ[i] * scale + 50) / 100;
        qt[0][i] = (uint8_t)(a < 1 ? 1 : (a > 255 ? 255 : a));
        qt[1][i] = (uint8_t)(c < 1 ? 1 : (c > 255 ? 255 : c));
        qdiv[0][i] = (float)qt[0][i];
        qdiv[1][i] = (float)qt[1][i];
    }
    HuffTable dcL, acL, dcC, acC;
    buildHuff(kDcLumaBits, kDcVals, &dcL);
    buildHuff(kAcLumaBits, kAcLumaVals, &acL);
    buildHuff(kDcChromaBits, kDcVals, &dcC);
    buildHuff(kAcChromaBits, kAcChromaVals, &acC);

    BitWriter bw = {0};
    putMarker(&bw, 0xD8); /* SOI */
    putMarker(&bw, 0xE0); /* APP0 JFIF 1.01, no density */
    put16(&bw, 16);
    static const uint8_t jfif[14] = {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0};
    for (int i = 0; i < 14; i++) emitByte(&bw, jfif[i]);
    for (int t = 0; t < 2; t++) { /* DQT */
        putMarker(&bw, 0xDB);
        put16(&bw, 67);
        emitByte(&bw, (uint8_t)t);
        for (int i = 0; i < 64; i++) emitByte(&bw, qt[t][kZigzag[i]]);
    }
    putMarker(&bw, 0xC0); /* SOF0 */
    put16(&bw, 17);
    emitByte(&bw, 8);
    put16(&bw, h);
    put16(&bw, w);
    emitByte(&bw, 3);
    for (int c = 0; c < 3; c++) {
        emitByte(&bw, (uint8_t)(c + 1));
        emitByte(&bw, 0x11);
        emitByte(&bw, (uint8_t)(c ? 1 : 0));
    }
    const uint8_t* bitsList[4] = {kDcLumaBits, kAcLumaBits, kDcChromaBits, kAcChromaBits};
    const uint8_t* valsList[4] = {kDcVals, kAcLumaVals, kDcVals, kAcChromaVals};
    const uint8_t classId[4] = {0x00, 0x10, 0x01, 0x11};
    for (int t = 0; t < 4; t++) { /* DHT */
        int nv = 0;
        for (int i = 0; i < 16; i++) nv += bitsList[t][i];
        putMarker(&bw, 0xC4);
        put16(&bw, (uint32_t)(2 + 1 + 16 + nv));
        emitByte(&bw, classId[t]);
        for (int i = 0; i < 16; i++) emitByte(&bw, bitsList[t][i]);
        for (int i = 0; i < nv; i++) emitByte(&bw, valsList[t][i]);
    }
    putMarker(&bw, 0xDA); /* SOS */
    put16(&bw, 12);
    emitByte(&bw, 3);
    for (int c = 0; c < 3; c++) {
        emitByte(&bw, (uint8_t)(c + 1));
        emitByte(&bw, (uint8_t)(c ? 0x11 : 0x00));
    }
    emitByte(&bw, 0);
    emitByte(&bw, 63);
    emitByte(&bw, 0);
    int dcPrev[3] = {0, 0, 0};
    for (uint32_t by = 0; by < h && !bw.failed; by += 8) {
        for (uint32_t bx = 0; bx < w; bx += 8) {
            float blk[3][64];
            for (int y = 0; y < 8; y++) {
                uint32_t sy = by + (uint32_t)y < h ? by + (uint32_t)y : h - 1u; /* replicate the edge */
                for (int x = 0; x < 8; x++) {
                    uint32_t sx = bx + (uint32_t)x < w ? bx + (uint32_t)x : w - 1u;
                    const uint8_t* px = rgba8 + ((size_t)sy * w + sx) * 4u;
                    float r = px[0], g = px[1], b = px[2];
                    blk[0][y * 8 + x] = 0.299f * r + 0.587f * g + 0.114f * b - 128.0f;
                    blk[1][y * 8 + x] = -0.168736f * r - 0.331264f * g + 0.5f * b;
                    blk[2][y * 8 + x] = 0.5f * r - 0.418688f * g - 0.081312f * b;
                }
            }
            encodeBlock(&bw, blk[0], qdiv[0], &dcPrev[0], &dcL, &acL);
            encodeBlock(&bw, blk[1], qdiv[1], &dcPrev[1], &dcC, &acC);
            encodeBlock(&bw, blk[2], qdiv[1], &dcPrev[2], &dcC, &acC);
        }
    }
    flushBits(&bw);
    putMarker(&bw, 0xD9); /* EOI */
    int ok = !bw.failed;
    FILE* f = ok ? fopen(path, "wb") : NULL;
    if (f) {
        ok = fwrite(bw.buf, 1, bw.len, f) == bw.len;
        ok = (fclose(f) == 0) && ok;
    } else {
        ok = 0;
    }
    free(bw.buf);
    return ok;
}

/* ------------------------------------------------------- pixel transforms */

/* image.c:381-423 */
static void mul3(const float m[9], const float in[3], float out[3]) {
    out[0] = (m[0] * in[0]) + (m[1] * in[1]) + (m[2] * in[2]);
    out[1] = (m[3] * in[0]) + (m[4] * in[1]) + (m[5] * in[2]);
    out[2] = (m[6] * in[0]) + (m[7] * in[1]) + (m[8] * in[2]);
}
static void xyzToLinearSrgb(float px[3]) {
    static const float bradford[9] = {0.8951f, 0.2664f, -0.1614f, -0.7502f, 1.7135f, 0.0367f, 0.0389f, -0.0685f, 1.0296f};
    static const float bradfordInv[9] = {0.9869929f, -0.1470543f, 0.1599627f, 0.4323053f, 0.5183603f,
                                         0.0492912f, -0.0085287f, 0.0400428f, 0.9684867f};
    static const float eToD65[3] = {0.9413344f, 1.0404175f, 1.0895327f};
    static const float toSrgb[9] = {3.2404542f, -1.5371385f, -0.4985314f, -0.9692660f, 1.8760108f,
                                    0.0415560f, 0.0556434f,  -0.2040259f, 1.0572252f};
    float lms[3], a[3];
    mul3(bradford, px, lms);
    for (int c = 0; c < 3; c++) lms[c] *= eToD65[c];
    mul3(bradfordInv, lms, a);
    mul3(toSrgb, a, px);
}

static float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
static float srgbEncode(float v) {
    v = clamp01(v);
    return v <= 0.0031308f ? 12.92f * v : (1.055f * powf(v, 1.0f / 2.4f)) - 0.055f;
}

/* image.c:641-699 */
static void linearToDisplay(const VKRT_SceneSettingsSnapshot* s, const float* lin, uint16_t* out, size_t n) {
    int debug = s->debugMode != VKRT_DEBUG_MODE_NONE;
    float exposure = s->exposure;
    if (!isfinite(exposure) || exposure < 0.0f) exposure = 1.0f;
    for (size_t i = 0; i < n; i++) {
        float m[3] = {lin[i * 4], lin[i * 4 + 1], lin[i * 4 + 2]};
        if (!debug) {
            for (int c = 0; c < 3; c++) m[c] *= exposure;
            if (s->toneMappingMode == VKRT_TONE_MAPPING_MODE_ACES)
                for (int c = 0; c < 3; c++) m[c] = (m[c] * ((2.51f * m[c]) + 0.03f)) / ((m[c] * ((2.43f * m[c]) + 0.59f)) + 0.14f);
        }
        for (int c = 0; c < 3; c++) out[i * 4 + (size_t)c] = (uint16_t)((clamp01(srgbEncode(m[c])) * 65535.0f) + 0.5f);
        out[i * 4 + 3] = 65535u;
    }
}

/* ------------------------------------------------------------------- API */

void VKRT_defaultRenderExportSettings(VKRT_RenderExportSettings* settings) {
    if (settings) settings->denoiseEnabled = 1u;
}

VKRT_Result VKRT_saveRenderImageEx(VKRT* vkrt, const char* path, const VKRT_RenderExportSettings* settings) {
    if (!vkrt || !path || !path[0]) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_RenderExportSettings es;
    VKRT_defaultRenderExportSettings(&es);
    if (settings) es = *settings;
    ImageFormat format = kFormatPng;
    char* resolved = resolvePath(path, &format);
    if (!resolved) return VKRT_ERROR_OPERATION_FAILED;
    uint32_t w = vkrt->renderWidth, h = vkrt->renderHeight;
    if (!vkrt->device || w == 0 || h == 0) {
        VKRT_LOG_ERROR("Cannot save render image before the accumulation image is initialized");
        free(resolved);
        return VKRT_ERROR_OPERATION_FAILED;
    }
    size_t n = (size_t)w * h;
    static int denoiseNoted = 0;
    if (es.denoiseEnabled && !denoiseNoted) {
        denoiseNoted = 1;
        VKRT_LOG_ERROR("denoising is not available in this build; exporting the raw render");
    }
    int ok = 0;
    uint16_t* display = (uint16_t*)malloc(sizeof(uint16_t) * 4u * n);
    float* linear = NULL;
    if (!display) goto done;
    if (format != kFormatExr && !es.denoiseEnabled) {
        /* the display image the frame loop already produced (api.c:208-218) */
        if (VKRT_readOutputImage(vkrt, display, 4u * n) != VKRT_SUCCESS) goto done;
    } else {
        linear = (float*)malloc(sizeof(float) * 4u * n);
        if (!linear || VKRT_readAccumulation(vkrt, linear, 4u * n) != VKRT_SUCCESS) goto done;
        int spectral = vkrt->settings.renderMode == VKRT_RENDER_MODE_SPECTRAL;
        for (size_t i = 0; i < n; i++) {
            float* px = &linear[i * 4];
            if (spectral) xyzToLinearSrgb(px);
            for (int c = 0; c < 3; c++)
                if (!isfinite(px[c])) px[c] = 0.0f;
            px[3] = 1.0f;
        }
        if (format != kFormatExr) linearToDisplay(&vkrt->settings, linear, display, n);
    }
    if (format == kFormatExr) {
        ok = writeExr(resolved, linear, w, h);
    } else if (format == kFormatPng) {
        ok = writePng(resolved, display, w, h);
    } else {
        uint8_t* rgba8 = (uint8_t*)malloc(4u * n);
        if (rgba8) {
            for (size_t i = 0; i < 4u * n; i++) rgba8[i] = (uint8_t)((((uint32_t)display[i] * 255u) + 32767u) / 65535u);
            ok = writeJpeg(resolved, rgba8, w, h, 95);
            free(rgba8);
        }
    }
done:
    if (!ok) {
        remove(resolved);
        VKRT_LOG_ERROR("Render export failed for '%s'", resolved);
    }
    free(display);
    free(linear);
    free(resolved);
    return ok ? VKRT_SUCCESS : VKRT_ERROR_OPERATION_FAILED;
}

VKRT_Result VKRT_saveRenderImage(VKRT* vkrt, const char* path) { return VKRT_saveRenderImageEx(vkrt, path, NULL); }
