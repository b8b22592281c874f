/*
 * OpenEXR reader for environment maps and linear textures: single-part
 * scanline files with NONE, RLE, ZIPS or ZIP compression and UINT, HALF or
 * FLOAT channels (the OpenEXR 2 file layout: magic, version, attribute
 * header, line offset table, chunks of `y, size, data`).
 *
 * The channel rules follow the reference's reader (src/core/utility/exr.cpp):
 *   - R, G, B and A are found by name, or by the part after the last '.'
 *     ("diffuse.R"); a file with no R/G/B but a Y channel is greyscale;
 *   - the result is RGBA16F when every channel of the file is HALF, else
 *     RGBA32F; a missing alpha is 1;
 *   - the pixels are the data window, row 0 first, colour space linear.
 * Tiled, deep and multi-part files, sub-sampled channels and the PIZ, PXR24,
 * B44 and DWA codecs are rejected with an error.
 */
#include "exr_decode.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

typedef struct {
    char name[256];
    int type; /* 0 UINT, 1 HALF, 2 FLOAT */
    int xs, ys;
} ExrChannel;

static uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
static uint64_t le64(const uint8_t* p) { return (uint64_t)le32(p) | (uint64_t)le32(p + 4) << 32; }

static float halfBitsToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, exp = (h >> 10) & 0x1fu, man = h & 0x3ffu, bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            int e = -1;
            do {
                e++;
                man <<= 1;
            } while ((man & 0x400u) == 0);
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &bits, sizeof(f));
    return f;
}

static int nameMatches(const char* channel, const char* component) {
    if (strcmp(channel, component) == 0) return 1;
    const char* dot = strrchr(channel, '.');
    return dot && strcmp(dot + 1, component) == 0;
}

/* OpenEXR ZIP/RLE post-processing: undo the byte delta predictor, then
 * re-interleave the two half-buffers (first half = even bytes). */
static void unpredictAndInterleave(uint8_t* tmp, size_t n, uint8_t* out) {
    for (size_t i = 1; i < n; i++) tmp[i] = (uint8_t)(tmp[i - 1] + tmp[i] - 128);
    const uint8_t* a = tmp;
    const uint8_t* b = tmp + (n + 1) / 2;
    for (size_t i = 0; i < n; i++) out[i] = (i & 1) ? *b++ : *a++;
}

static int rleDecode(const uint8_t* in, size_t inSize, uint8_t* out, size_t outSize) {
    size_t o = 0, i = 0;
    while (i < inSize) {
        int count = (int8_t)in[i++];
        if (count < 0) {
            size_t n = (size_t)(-count);
            if (i + n > inSize || o + n > outSize) return -1;
            memcpy(out + o, in + i, n);
            i += n;
            o += n;
        } else {
            size_t n = (size_t)count + 1u;
            if (i >= inSize || o + n > outSize) return -1;
            memset(out + o, in[i++], n);
            o += n;
        }
    }
    return o == outSize ? 0 : -1;
}

#define EXR_FAIL(...)                                     \
    do {                                                  \
        if (err) snprintf(err, errSize, __VA_ARGS__);     \
        goto fail;                                        \
    } while (0)

int vkrtDecodeExr(const uint8_t* data, size_t size, ExrImage* out, char* err, size_t errSize) {
    memset(out, 0, sizeof(*out));
    ExrChannel channels[64];
    int channelCount = 0, compression = -1, haveWindow = 0;
    int32_t xMin = 0, yMin = 0, xMax = -1, yMax = -1;
    uint8_t* block = NULL;
    uint8_t* tmp = NULL;
    void* pixels = NULL;
    if (!data || size < 8 || le32(data) != 20000630u) {
        if (err) snprintf(err, errSize, "not an OpenEXR file");
        return -1;
    }
    {
        uint32_t version = le32(data + 4);
        if ((version & 0xffu) != 2u) EXR_FAIL("OpenEXR version %u is not supported", version & 0xffu);
        if (version & 0x200u) EXR_FAIL("tiled OpenEXR files are not supported");
        if (version & 0x1800u) EXR_FAIL("deep or multi-part OpenEXR files are not supported");
    }
    size_t pos = 8;
    /* attribute header: name\0 type\0 size value, ended by an empty name */
    for (;;) {
        if (pos >= size) EXR_FAIL("truncated header");
        if (data[pos] == 0) {
            pos++;
            break;
        }
        const char* name = (const char*)data + pos;
        size_t nl = strnlen(name, size - pos);
        if (pos + nl + 1 >= size) EXR_FAIL("truncated header");
        const char* type = name + nl + 1;
        size_t tl = strnlen(type, size - pos - nl - 1);
        size_t at = pos + nl + 1 + tl + 1;
        if (at + 4 > size) EXR_FAIL("truncated header");
        uint32_t len = le32(data + at);
        at += 4;
        if (len > size - at) EXR_FAIL("truncated attribute '%s'", name);
        const uint8_t* v = data + at;
        if (strcmp(name, "channels") == 0 && strcmp(type, "chlist") == 0) {
            size_t c = 0;
            while (c < len && v[c] != 0) {
                if (channelCount == 64) EXR_FAIL("too many channels");
                ExrChannel* ch = &channels[channelCount++];
                size_t cl = strnlen((const char*)v + c, len - c);
                if (cl >= sizeof(ch->name) || c + cl + 1 + 16 > len) EXR_FAIL("bad channel list");
                memcpy(ch->name, v + c, cl);
                ch->name[cl] = 0;
                c += cl + 1;
                ch->type = (int)le32(v + c);
                ch->xs = (int)le32(v + c + 8);
                ch->ys = (int)le32(v + c + 12);
                c += 16;
                if (ch->type < 0 || ch->type > 2) EXR_FAIL("bad pixel type in channel '%s'", ch->name);
                if (ch->xs != 1 || ch->ys != 1) EXR_FAIL("sub-sampled channel '%s' is not supported", ch->name);
            }
        } else if (strcmp(name, "compression") == 0 && len >= 1) {
            compression = v[0];
        } else if (strcmp(name, "dataWindow") == 0 && len >= 16) {
            xMin = (int32_t)le32(v);
            yMin = (int32_t)le32(v + 4);
            xMax = (int32_t)le32(v + 8);
            yMax = (int32_t)le32(v + 12);
            haveWindow = 1;
        }
        pos = at + len;
    }
    if (!haveWindow || channelCount == 0 || compression < 0) EXR_FAIL("header lacks channels, compression or dataWindow");
    if (xMax < xMin || yMax < yMin) EXR_FAIL("empty data window");
    {
        const int64_t width = (int64_t)xMax - xMin + 1, height = (int64_t)yMax - yMin + 1;
        if (width > 65536 || height > 65536) EXR_FAIL("image too large");
        int linesPerBlock;
        switch (compression) {
            case 0: case 1: case 2: linesPerBlock = 1; break;
            case 3: linesPerBlock = 16; break;
            default: EXR_FAIL("OpenEXR compression %d (PIZ, PXR24, B44 or DWA) is not supported", compression);
        }
        int idx[4] = {-1, -1, -1, -1}, yIdx = -1, allHalf = 1;
        static const char* comps[4] = {"R", "G", "B", "A"};
        for (int c = 0; c < channelCount; c++) {
            for (int k = 0; k < 4; k++)
                if (idx[k] < 0 && nameMatches(channels[c].name, comps[k])) idx[k] = c;
            if (yIdx < 0 && nameMatches(channels[c].name, "Y")) yIdx = c;
            if (channels[c].type != 1) allHalf = 0;
        }
        if ((idx[0] < 0 || idx[1] < 0 || idx[2] < 0) && yIdx < 0) EXR_FAIL("no RGB(A) or Y channels");
        size_t lineBytes = 0;
        size_t offsetInLine[64];
        for (int c = 0; c < channelCount; c++) {
            offsetInLine[c] = lineBytes;
            lineBytes += (size_t)width * (channels[c].type == 1 ? 2u : 4u);
        }
        const size_t chunks = (size_t)((height + linesPerBlock - 1) / linesPerBlock);
        if (pos + chunks * 8u > size) EXR_FAIL("truncated offset table");
        const size_t n = (size_t)width * (size_t)height;
        pixels = malloc(n * 4u * (allHalf ? 2u : 4u));
        block = (uint8_t*)malloc(lineBytes * (size_t)linesPerBlock);
        tmp = (uint8_t*)malloc(lineBytes * (size_t)linesPerBlock);
        if (!pixels || !block || !tmp) EXR_FAIL("out of memory");
        for (size_t k = 0; k < chunks; k++) {
            uint64_t off = le64(data + pos + k * 8u);
            if (off + 8u > size) EXR_FAIL("chunk offset out of range");
            int32_t y = (int32_t)le32(data + off);
            uint32_t dataSize = le32(data + off + 4);
            if (dataSize > size - off - 8u) EXR_FAIL("truncated chunk");
            if (y < yMin || y > yMax) EXR_FAIL("chunk line %d outside the data window", y);
            const int lines = (int)(((int64_t)yMax - y + 1) < linesPerBlock ? ((int64_t)yMax - y + 1) : linesPerBlock);
            const size_t expected = lineBytes * (size_t)lines;
            const uint8_t* src = data + off + 8u;
            if (dataSize == expected || compression == 0) {
                if (dataSize != expected) EXR_FAIL("uncompressed chunk of the wrong size");
                memcpy(block, src, expected); /* stored raw when compressing did not help */
            } else if (compression == 1) {
                if (rleDecode(src, dataSize, tmp, expected) != 0) EXR_FAIL("corrupt RLE chunk");
                unpredictAndInterleave(tmp, expected, block);
            } else {
                uLongf got = (uLongf)expected;
                if (uncompress(tmp, &got, src, dataSize) != Z_OK || got != expected) EXR_FAIL("corrupt ZIP chunk");
                unpredictAndInterleave(tmp, expected, block);
            }
            for (int l = 0; l < lines; l++) {
                const uint8_t* line = block + (size_t)l * lineBytes;
                const size_t row = (size_t)(y - yMin + l);
                for (int comp = 0; comp < 4; comp++) {
                    const int c = idx[comp] >= 0 ? idx[comp] : (comp < 3 ? yIdx : -1);
                    for (int64_t x = 0; x < width; x++) {
                        const size_t o = (row * (size_t)width + (size_t)x) * 4u + (size_t)comp;
                        if (allHalf) {
                            uint16_t h = 0x3c00u; /* 1.0 */
                            if (c >= 0) {
                                const uint8_t* p = line + offsetInLine[c] + (size_t)x * 2u;
                                h = (uint16_t)(p[0] | p[1] << 8);
                            }
                            ((uint16_t*)pixels)[o] = h;
                        } else {
                            float f = 1.0f;
                            if (c >= 0) {
                                const int t = channels[c].type;
                                const uint8_t* p = line + offsetInLine[c] + (size_t)x * (t == 1 ? 2u : 4u);
                                if (t == 1) f = halfBitsToFloat((uint16_t)(p[0] | p[1] << 8));
                                else if (t == 2) memcpy(&f, p, 4);
                                else f = (float)le32(p);
                            }
                            ((float*)pixels)[o] = f;
                        }
                    }
                }
            }
        }
        out->pixels = pixels;
        out->width = (uint32_t)width;
        out->height = (uint32_t)height;
        out->half = allHalf;
        pixels = NULL;
    }
    free(block);
    free(tmp);
    return 0;
fail:
    free(block);
    free(tmp);
    free(pixels);
    return -1;
}
