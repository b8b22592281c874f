/*
 * JPEG decoder for texture and environment files: baseline and extended
 * sequential (SOF0/SOF1) and progressive (SOF2) Huffman-coded, 8-bit samples,
 * one (greyscale) or three (YCbCr, or RGB with an Adobe transform flag of 0)
 * components, any sampling factors up to 4, restart intervals.
 *
 * The reference decodes JPEG with turbojpeg's tjDecompress2(TJPF_RGBA,
 * TJFLAG_ACCURATEDCT) (src/core/utility/image.c:194-279). To give the same
 * pixels this decoder follows the same published steps that call performs:
 *   - ITU-T T.81 Huffman decoding (Annex F sequential, Annex G progressive);
 *   - the slow-but-accurate integer IDCT of Loeffler, Ligtenberg and Moschytz
 *     (ICASSP 1989), 13-bit fixed-point constants, 2 extra bits between the
 *     column and the row pass, output clamped to 0..255 around 128;
 *   - triangle-filter ("fancy") chroma upsampling for 2:1 horizontal and/or
 *     vertical factors, edges replicated, pixel replication for other ratios;
 *   - the JFIF YCbCr -> RGB conversion of ITU-T T.871 in 16-bit fixed point.
 * Arithmetic coding, 12-bit samples, lossless and CMYK files are rejected.
 */
#include "jpeg_decode.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint8_t lookupBits[512]; /* 9-bit fast path: code length (0 = slow path) */
    uint8_t lookupValue[512];
    int32_t maxCode[18];
    int32_t valOffset[17];
    uint8_t values[256];
    int present;
} Huffman;

typedef struct {
    int id, h, v, tq;
    int blocksW, blocksH;   /* blocks covering the component, padded to whole MCUs */
    int compW, compH;       /* samples of the component before padding (downsampled size) */
    int16_t* coef;          /* blocksW * blocksH * 64, natural order */
    uint8_t* plane;         /* blocksW*8 x blocksH*8 samples after the IDCT */
    int dcPred;
    int dcTable, acTable;
} Component;

typedef struct {
    const uint8_t* data;
    size_t size, pos;
    uint32_t bitBuffer;
    int bitCount;
    int hitMarker;

    uint16_t quant[4][64];
    Huffman dc[4], ac[4];
    Component comp[4];
    int compCount;
    int width, height;
    int hMax, vMax, mcuX, mcuY;
    int progressive;
    int restartInterval;
    int adobeTransform; /* -1 when absent */
    int eobRun;
    char* err;
    size_t errSize;
} Decoder;

static const uint8_t kZigzag[64 + 16] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
    /* a corrupt run past coefficient 63 lands here and is ignored */
    63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

#define FAIL(...)                                               \
    do {                                                        \
        if (d->err) snprintf(d->err, d->errSize, __VA_ARGS__);  \
        return -1;                                              \
    } while (0)

/* ---------------------------------------------------------------- bits -- */
static void fillBits(Decoder* d) {
    while (d->bitCount <= 24) {
        uint32_t byte = 0;
        if (!d->hitMarker && d->pos < d->size) {
            byte = d->data[d->pos];
            if (byte == 0xFF) {
                uint8_t next = d->pos + 1 < d->size ? d->data[d->pos + 1] : 0;
                if (next == 0x00) {
                    d->pos += 2;
                } else {
                    d->hitMarker = 1; /* a marker ends the entropy-coded segment: feed zeros */
                    byte = 0;
                }
            } else {
                d->pos++;
            }
        }
        d->bitBuffer |= byte << (24 - d->bitCount);
        d->bitCount += 8;
    }
}

static int getBits(Decoder* d, int n) {
    if (n == 0) return 0;
    fillBits(d);
    int v = (int)(d->bitBuffer >> (32 - n));
    d->bitBuffer <<= n;
    d->bitCount -= n;
    return v;
}

static int getBit(Decoder* d) { return getBits(d, 1); }

/* T.81 F.2.2.1 EXTEND */
static int extendValue(int v, int n) { return n == 0 ? 0 : (v < (1 << (n - 1)) ? v - (1 << n) + 1 : v); }

static int decodeHuffman(Decoder* d, const Huffman* h) {
    fillBits(d);
    uint32_t peek = d->bitBuffer >> 23;
    int len = h->lookupBits[peek];
    if (len) {
        d->bitBuffer <<= len;
        d->bitCount -= len;
        return h->lookupValue[peek];
    }
    for (len = 10; len <= 16; len++) {
        int32_t code = (int32_t)(d->bitBuffer >> (32 - len));
        if (code <= h->maxCode[len]) {
            d->bitBuffer <<= len;
            d->bitCount -= len;
            int idx = h->valOffset[len] + code;
            return (idx >= 0 && idx < 256) ? h->values[idx] : 0;
        }
    }
    return -1; /* corrupt code */
}

/* T.81 C.2 / F.2.2.3: canonical code assignment from the 16 length counts */
static int buildHuffman(Huffman* h, const uint8_t counts[16], const uint8_t* values, int total) {
    memset(h, 0, sizeof(*h));
    if (total > 256) return -1;
    memcpy(h->values, values, (size_t)total);
    int code = 0, k = 0;
    for (int len = 1; len <= 16; len++) {
        h->valOffset[len] = k - code;
        if (counts[len - 1]) {
            for (int i = 0; i < counts[len - 1]; i++, k++, code++) {
                if (len <= 9) {
                    int shift = 9 - len;
                    for (int fill = 0; fill < (1 << shift); fill++) {
                        int idx = (code << shift) | fill;
                        if (idx < 512) {
                            h->lookupBits[idx] = (uint8_t)len;
                            h->lookupValue[idx] = values[k];
                        }
                    }
                }
            }
            h->maxCode[len] = code - 1;
        } else {
            h->maxCode[len] = -1;
        }
        if (code > (1 << len)) return -1;
        code <<= 1;
    }
    h->maxCode[17] = 0x7fffffff;
    h->present = 1;
    return 0;
}

/* ------------------------------------------------------------ markers -- */
static int readU16(Decoder* d, int* out) {
    if (d->pos + 2 > d->size) return -1;
    *out = (d->data[d->pos] << 8) | d->data[d->pos + 1];
    d->pos += 2;
    return 0;
}

static int parseDQT(Decoder* d, size_t end) {
    while (d->pos < end) {
        int pq = d->data[d->pos] >> 4, tq = d->data[d->pos] & 15;
        d->pos++;
        if (tq > 3 || pq > 1) FAIL("bad quantisation table");
        if (d->pos + (size_t)(pq ? 128 : 64) > end) FAIL("truncated quantisation table");
        for (int i = 0; i < 64; i++) {
            int v = pq ? (d->data[d->pos] << 8 | d->data[d->pos + 1]) : d->data[d->pos];
            d->pos += pq ? 2 : 1;
            d->quant[tq][kZigzag[i]] = (uint16_t)v;
        }
    }
    return 0;
}

static int parseDHT(Decoder* d, size_t end) {
    while (d->pos < end) {
        if (d->pos + 17 > end) FAIL("truncated Huffman table");
        int tc = d->data[d->pos] >> 4, th = d->data[d->pos] & 15;
        const uint8_t* counts = d->data + d->pos + 1;
        int total = 0;
        for (int i = 0; i < 16; i++) total += counts[i];
        d->pos += 17;
        if (tc > 1 || th > 3 || d->pos + (size_t)total > end) FAIL("bad Huffman table");
        if (buildHuffman(tc ? &d->ac[th] : &d->dc[th], counts, d->data + d->pos, total) != 0)
            FAIL("bad Huffman code lengths");
        d->pos += (size_t)total;
    }
    return 0;
}

static int parseSOF(Decoder* d, size_t end) {
    if (end - d->pos < 6) FAIL("truncated frame header");
    int precision = d->data[d->pos];
    d->height = d->data[d->pos + 1] << 8 | d->data[d->pos + 2];
    d->width = d->data[d->pos + 3] << 8 | d->data[d->pos + 4];
    d->compCount = d->data[d->pos + 5];
    d->pos += 6;
    if (precision != 8) FAIL("%d-bit samples are not supported", precision);
    if (d->width <= 0 || d->height <= 0) FAIL("image without lines (DNL) is not supported");
    if (d->compCount != 1 && d->compCount != 3) FAIL("%d components are not supported", d->compCount);
    if (end - d->pos < (size_t)(3 * d->compCount)) FAIL("truncated frame header");
    d->hMax = d->vMax = 1;
    for (int c = 0; c < d->compCount; c++) {
        Component* k = &d->comp[c];
        k->id = d->data[d->pos];
        k->h = d->data[d->pos + 1] >> 4;
        k->v = d->data[d->pos + 1] & 15;
        k->tq = d->data[d->pos + 2];
        d->pos += 3;
        if (k->h < 1 || k->h > 4 || k->v < 1 || k->v > 4 || k->tq > 3) FAIL("bad sampling factors");
        if (k->h > d->hMax) d->hMax = k->h;
        if (k->v > d->vMax) d->vMax = k->v;
    }
    if (d->compCount == 1) d->comp[0].h = d->comp[0].v = d->hMax = d->vMax = 1; /* T.81 A.2.2 */
    d->mcuX = (d->width + 8 * d->hMax - 1) / (8 * d->hMax);
    d->mcuY = (d->height + 8 * d->vMax - 1) / (8 * d->vMax);
    for (int c = 0; c < d->compCount; c++) {
        Component* k = &d->comp[c];
        k->compW = (d->width * k->h + d->hMax - 1) / d->hMax;
        k->compH = (d->height * k->v + d->vMax - 1) / d->vMax;
        k->blocksW = d->mcuX * k->h;
        k->blocksH = d->mcuY * k->v;
        size_t blocks = (size_t)k->blocksW * (size_t)k->blocksH;
        if (blocks > ((size_t)1 << 26)) FAIL("image too large");
        k->coef = (int16_t*)calloc(blocks * 64u, sizeof(int16_t));
        if (!k->coef) FAIL("out of memory");
    }
    return 0;
}

/* --------------------------------------------------------- block decode -- */
static int decodeBlockSequential(Decoder* d, Component* k, int16_t* blk) {
    int t = decodeHuffman(d, &d->dc[k->dcTable]);
    if (t < 0 || t > 16) FAIL("corrupt DC code");
    int diff = extendValue(getBits(d, t), t);
    k->dcPred += diff;
    blk[0] = (int16_t)k->dcPred;
    for (int i = 1; i < 64;) {
        int rs = decodeHuffman(d, &d->ac[k->acTable]);
        if (rs < 0) FAIL("corrupt AC code");
        int r = rs >> 4, s = rs & 15;
        if (s == 0) {
            if (r != 15) break; /* EOB */
            i += 16;
            continue;
        }
        i += r;
        if (i > 63) FAIL("AC run past the block");
        blk[kZigzag[i]] = (int16_t)extendValue(getBits(d, s), s);
        i++;
    }
    return 0;
}

/* T.81 G.1.2.1: DC first and refinement scans */
static int decodeDCProgressive(Decoder* d, Component* k, int16_t* blk, int ah, int al) {
    if (ah == 0) {
        int t = decodeHuffman(d, &d->dc[k->dcTable]);
        if (t < 0 || t > 16) FAIL("corrupt DC code");
        k->dcPred += extendValue(getBits(d, t), t);
        blk[0] = (int16_t)(k->dcPred * (1 << al));
    } else if (getBit(d)) {
        blk[0] = (int16_t)(blk[0] | (1 << al));
    }
    return 0;
}

/* T.81 G.1.2.2: AC first scan with end-of-band runs */
static int decodeACFirst(Decoder* d, Component* k, int16_t* blk, int ss, int se, int al) {
    if (d->eobRun > 0) {
        d->eobRun--;
        return 0;
    }
    for (int i = ss; i <= se;) {
        int rs = decodeHuffman(d, &d->ac[k->acTable]);
        if (rs < 0) FAIL("corrupt AC code");
        int r = rs >> 4, s = rs & 15;
        if (s == 0) {
            if (r < 15) {
                d->eobRun = (1 << r) - 1;
                if (r) d->eobRun += getBits(d, r);
                break;
            }
            i += 16;
            continue;
        }
        i += r;
        if (i > 63) FAIL("AC run past the block");
        blk[kZigzag[i]] = (int16_t)(extendValue(getBits(d, s), s) * (1 << al));
        i++;
    }
    return 0;
}

/* Adds the next bit of precision to an already non-zero coefficient. */
static void refineCoefficient(Decoder* d, int16_t* c, int bit) {
    if (getBit(d) && (*c & bit) == 0) *c = (int16_t)(*c >= 0 ? *c + bit : *c - bit);
}

/* T.81 G.1.2.3: AC successive-approximation refinement */
static int decodeACRefine(Decoder* d, Component* k, int16_t* blk, int ss, int se, int al) {
    const int bit = 1 << al;
    int i = ss;
    if (d->eobRun == 0) {
        for (; i <= se;) {
            int rs = decodeHuffman(d, &d->ac[k->acTable]);
            if (rs < 0) FAIL("corrupt AC code");
            int r = rs >> 4, s = rs & 15, value = 0;
            if (s == 0) {
                if (r < 15) {
                    d->eobRun = 1 << r;
                    if (r) d->eobRun += getBits(d, r);
                    break; /* the rest of the band is refined below */
                }
                /* ZRL: skip 16 zero coefficients, refining non-zero ones on the way */
            } else {
                if (s != 1) FAIL("bad refinement magnitude");
                value = getBit(d) ? bit : -bit;
            }
            while (i <= se) {
                int16_t* c = &blk[kZigzag[i]];
                if (*c != 0) {
                    refineCoefficient(d, c, bit);
                } else {
                    if (r == 0) {
                        if (value) *c = (int16_t)value;
                        i++;
                        break;
                    }
                    r--;
                }
                i++;
            }
        }
    }
    if (d->eobRun > 0) {
        for (; i <= se; i++) {
            int16_t* c = &blk[kZigzag[i]];
            if (*c != 0) refineCoefficient(d, c, bit);
        }
        d->eobRun--;
    }
    return 0;
}

static int expectRestart(Decoder* d) {
    /* drop the remaining bits, find the RSTn marker, restart predictors */
    d->bitBuffer = 0;
    d->bitCount = 0;
    d->hitMarker = 0;
    while (d->pos + 1 < d->size && !(d->data[d->pos] == 0xFF && d->data[d->pos + 1] >= 0xD0 && d->data[d->pos + 1] <= 0xD7)) {
        if (d->data[d->pos] == 0xFF && d->data[d->pos + 1] != 0x00 && d->data[d->pos + 1] != 0xFF) break;
        d->pos++;
    }
    if (d->pos + 1 < d->size && d->data[d->pos] == 0xFF && d->data[d->pos + 1] >= 0xD0 && d->data[d->pos + 1] <= 0xD7)
        d->pos += 2;
    for (int c = 0; c < d->compCount; c++) d->comp[c].dcPred = 0;
    d->eobRun = 0;
    return 0;
}

static int parseSOS(Decoder* d, size_t end) {
    if (d->compCount == 0) FAIL("scan before frame header");
    int ns = d->data[d->pos++];
    if (ns < 1 || ns > d->compCount || end - d->pos < (size_t)(2 * ns + 3)) FAIL("bad scan header");
    Component* sc[4];
    for (int i = 0; i < ns; i++) {
        int id = d->data[d->pos], tables = d->data[d->pos + 1];
        d->pos += 2;
        sc[i] = NULL;
        for (int c = 0; c < d->compCount; c++)
            if (d->comp[c].id == id) sc[i] = &d->comp[c];
        if (!sc[i]) FAIL("scan names an unknown component");
        sc[i]->dcTable = tables >> 4;
        sc[i]->acTable = tables & 15;
        if (sc[i]->dcTable > 3 || sc[i]->acTable > 3) FAIL("bad table selector");
    }
    int ss = d->data[d->pos], se = d->data[d->pos + 1], ah = d->data[d->pos + 2] >> 4, al = d->data[d->pos + 2] & 15;
    d->pos += 3;
    if (!d->progressive) {
        ss = 0;
        se = 63;
        ah = al = 0;
    } else if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13) {
        FAIL("bad progressive scan parameters");
    }
    for (int i = 0; i < ns; i++) {
        if ((ss == 0 && !d->dc[sc[i]->dcTable].present) ||
            ((!d->progressive || ss > 0) && !d->ac[sc[i]->acTable].present))
            FAIL("scan uses a missing Huffman table");
        sc[i]->dcPred = 0;
    }
    d->bitBuffer = 0;
    d->bitCount = 0;
    d->hitMarker = 0;
    d->eobRun = 0;

    int restartsLeft = d->restartInterval;
    if (ns == 1) {
        /* non-interleaved: one block per MCU over the component's own block grid */
        Component* k = sc[0];
        int bw = (k->compW + 7) / 8, bh = (k->compH + 7) / 8;
        for (int by = 0; by < bh; by++) {
            for (int bx = 0; bx < bw; bx++) {
                if (d->restartInterval && restartsLeft == 0) {
                    expectRestart(d);
                    restartsLeft = d->restartInterval;
                }
                int16_t* blk = k->coef + ((size_t)by * k->blocksW + bx) * 64u;
                int rc;
                if (!d->progressive) rc = decodeBlockSequential(d, k, blk);
                else if (ss == 0) rc = decodeDCProgressive(d, k, blk, ah, al);
                else if (ah == 0) rc = decodeACFirst(d, k, blk, ss, se, al);
                else rc = decodeACRefine(d, k, blk, ss, se, al);
                if (rc) return rc;
                restartsLeft--;
            }
        }
    } else {
        for (int my = 0; my < d->mcuY; my++) {
            for (int mx = 0; mx < d->mcuX; mx++) {
                if (d->restartInterval && restartsLeft == 0) {
                    expectRestart(d);
                    restartsLeft = d->restartInterval;
                }
                for (int i = 0; i < ns; i++) {
                    Component* k = sc[i];
                    for (int v = 0; v < k->v; v++) {
                        for (int h = 0; h < k->h; h++) {
                            int bx = mx * k->h + h, by = my * k->v + v;
                            int16_t* blk = k->coef + ((size_t)by * k->blocksW + bx) * 64u;
                            int rc = d->progressive ? decodeDCProgressive(d, k, blk, ah, al)
                                                    : decodeBlockSequential(d, k, blk);
                            if (rc) return rc;
                        }
                    }
                }
                restartsLeft--;
            }
        }
    }
    /* skip to the next marker */
    d->bitBuffer = 0;
    d->bitCount = 0;
    d->hitMarker = 0;
    while (d->pos + 1 < d->size) {
        if (d->data[d->pos] == 0xFF && d->data[d->pos + 1] != 0x00 &&
            !(d->data[d->pos + 1] >= 0xD0 && d->data[d->pos + 1] <= 0xD7) && d->data[d->pos + 1] != 0xFF)
            break;
        d->pos++;
    }
    return 0;
}

/* ------------------------------------------------------------------ IDCT -- */
/* LL&M 8-point IDCT in 13-bit fixed point: FIXK(x) = round(x * 2^13). */
#define CONST_BITS 13
#define PASS1_BITS 2
#define FIXK(x) ((int32_t)((x) * 8192.0 + 0.5))

typedef struct {
    int32_t c0_298631336, c0_390180644, c0_541196100, c0_765366865, c0_899976223, c1_175875602;
    int32_t c1_501321110, c1_847759065, c1_961570560, c2_053119869, c2_562915447, c3_072711026;
} IdctConstants;

/* The twelve rotation products of the LL&M flow graph, from c(k) = cos(k*pi/16):
 * sqrt2*(c6), sqrt2*(c2-c6), sqrt2*(c2+c6), sqrt2*(-c1+c3+c5-c7), ... */
static IdctConstants idctConstants(void) {
    const double s2 = sqrt(2.0);
    double c[8];
    for (int k = 0; k < 8; k++) c[k] = cos(k * M_PI / 16.0);
    IdctConstants t;
    t.c0_541196100 = FIXK(s2 * c[6]);
    t.c0_765366865 = FIXK(s2 * (c[2] - c[6]));
    t.c1_847759065 = FIXK(s2 * (c[2] + c[6]));
    t.c1_175875602 = FIXK(s2 * c[3]);
    t.c0_298631336 = FIXK(s2 * (-c[1] + c[3] + c[5] - c[7]));
    t.c2_053119869 = FIXK(s2 * (c[1] + c[3] - c[5] + c[7]));
    t.c3_072711026 = FIXK(s2 * (c[1] + c[3] + c[5] - c[7]));
    t.c1_501321110 = FIXK(s2 * (c[1] + c[3] - c[5] - c[7]));
    t.c0_899976223 = FIXK(s2 * (c[7] - c[3]) * -1.0);
    t.c2_562915447 = FIXK(s2 * (c[1] + c[3]));
    t.c1_961570560 = FIXK(s2 * (c[3] + c[5]));
    t.c0_390180644 = FIXK(s2 * (c[3] - c[5]));
    return t;
}

static inline int32_t descale(int64_t x, int n) { return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n); }

/* One 8-point pass; in[] stride `is`, results before the final descale. */
static inline void idct8(const IdctConstants* K, const int32_t* in, int is, int64_t out[8]) {
    int64_t z2 = in[2 * is], z3 = in[6 * is];
    int64_t z1 = (z2 + z3) * K->c0_541196100;
    int64_t tmp2 = z1 - z3 * K->c1_847759065;
    int64_t tmp3 = z1 + z2 * K->c0_765366865;
    z2 = in[0];
    z3 = in[4 * is];
    int64_t tmp0 = (z2 + z3) * (1 << CONST_BITS);
    int64_t tmp1 = (z2 - z3) * (1 << CONST_BITS);
    int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;

    int64_t o0 = in[7 * is], o1 = in[5 * is], o2 = in[3 * is], o3 = in[1 * is];
    int64_t a1 = o0 + o3, a2 = o1 + o2, a3 = o0 + o2, a4 = o1 + o3;
    int64_t a5 = (a3 + a4) * K->c1_175875602;
    o0 *= K->c0_298631336;
    o1 *= K->c2_053119869;
    o2 *= K->c3_072711026;
    o3 *= K->c1_501321110;
    a1 *= -K->c0_899976223;
    a2 *= -K->c2_562915447;
    a3 = a3 * -K->c1_961570560 + a5;
    a4 = a4 * -K->c0_390180644 + a5;
    o0 += a1 + a3;
    o1 += a2 + a4;
    o2 += a2 + a3;
    o3 += a1 + a4;

    out[0] = tmp10 + o3;
    out[7] = tmp10 - o3;
    out[1] = tmp11 + o2;
    out[6] = tmp11 - o2;
    out[2] = tmp12 + o1;
    out[5] = tmp12 - o1;
    out[3] = tmp13 + o0;
    out[4] = tmp13 - o0;
}

static inline uint8_t clampSample(int32_t x) {
    /* 10-bit wrap then clamp around the 128 level shift, as the reference decoder's range table */
    int32_t v = ((x + 512) & 1023) - 512 + 128;
    return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
}

static void idctBlock(const IdctConstants* K, const int16_t* coef, const uint16_t* q, uint8_t* out, int stride) {
    int32_t in[64], ws[64];
    for (int i = 0; i < 64; i++) in[i] = (int32_t)coef[i] * (int32_t)q[i];
    int64_t t[8];
    for (int col = 0; col < 8; col++) {
        idct8(K, in + col, 8, t);
        for (int r = 0; r < 8; r++) ws[r * 8 + col] = descale(t[r], CONST_BITS - PASS1_BITS);
    }
    for (int row = 0; row < 8; row++) {
        idct8(K, ws + row * 8, 1, t);
        for (int c = 0; c < 8; c++) out[row * stride + c] = clampSample(descale(t[c], CONST_BITS + PASS1_BITS + 3));
    }
}

/* ------------------------------------------------------ upsample / colour -- */
static uint8_t samplePlane(const Component* k, int x, int y) {
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    if (x >= k->compW) x = k->compW - 1;
    if (y >= k->compH) y = k->compH - 1;
    return k->plane[(size_t)y * (size_t)(k->blocksW * 8) + (size_t)x];
}

/* Full-resolution samples of one component into out (width x height). */
static int upsampleComponent(const Decoder* d, const Component* k, uint8_t* out) {
    const int W = d->width, H = d->height;
    const int fx = d->hMax / k->h, fy = d->vMax / k->v;
    const int hRatio2 = (d->hMax == 2 * k->h), vRatio2 = (d->vMax == 2 * k->v);
    const int fancyH2V1 = hRatio2 && d->vMax == k->v && k->compW > 2;
    const int fancyH2V2 = hRatio2 && vRatio2 && k->compW > 2;
    const int fancyH1V2 = d->hMax == k->h && vRatio2;
    if (d->hMax % k->h || d->vMax % k->v) return -1;
    for (int y = 0; y < H; y++) {
        uint8_t* row = out + (size_t)y * W;
        if (fx == 1 && fy == 1) {
            for (int x = 0; x < W; x++) row[x] = samplePlane(k, x, y);
        } else if (fancyH2V1) {
            for (int x = 0; x < W; x++) {
                int sx = x >> 1;
                int cur = samplePlane(k, sx, y);
                if (sx == 0 && (x & 1) == 0) row[x] = (uint8_t)cur;
                else if (sx == k->compW - 1 && (x & 1)) row[x] = (uint8_t)cur;
                else if (x & 1) row[x] = (uint8_t)((cur * 3 + samplePlane(k, sx + 1, y) + 2) >> 2);
                else row[x] = (uint8_t)((cur * 3 + samplePlane(k, sx - 1, y) + 1) >> 2);
            }
        } else if (fancyH2V2) {
            int sy = y >> 1, ny = (y & 1) ? sy + 1 : sy - 1; /* neighbour row, replicated at the edges */
            for (int x = 0; x < W; x++) {
                int sx = x >> 1;
                int this_ = samplePlane(k, sx, sy) * 3 + samplePlane(k, sx, ny);
                if ((x & 1) == 0) {
                    if (sx == 0) row[x] = (uint8_t)((this_ * 4 + 8) >> 4);
                    else row[x] = (uint8_t)((this_ * 3 + samplePlane(k, sx - 1, sy) * 3 + samplePlane(k, sx - 1, ny) + 8) >> 4);
                } else {
                    if (sx == k->compW - 1) row[x] = (uint8_t)((this_ * 4 + 7) >> 4);
                    else row[x] = (uint8_t)((this_ * 3 + samplePlane(k, sx + 1, sy) * 3 + samplePlane(k, sx + 1, ny) + 7) >> 4);
                }
            }
        } else if (fancyH1V2) {
            int sy = y >> 1, ny = (y & 1) ? sy + 1 : sy - 1, bias = (y & 1) ? 2 : 1;
            for (int x = 0; x < W; x++)
                row[x] = (uint8_t)((samplePlane(k, x, sy) * 3 + samplePlane(k, x, ny) + bias) >> 2);
        } else {
            for (int x = 0; x < W; x++) row[x] = samplePlane(k, x / fx, y / fy);
        }
    }
    return 0;
}

int vkrtDecodeJpeg(const uint8_t* data, size_t size, JpegImage* out, char* err, size_t errSize) {
    Decoder dec;
    Decoder* d = &dec;
    memset(d, 0, sizeof(*d));
    d->data = data;
    d->size = size;
    d->err = err;
    d->errSize = errSize;
    d->adobeTransform = -1;
    memset(out, 0, sizeof(*out));
    int rc = -1, frames = 0;
    uint8_t* planes = NULL;
    uint8_t* pixels = NULL;
    if (!data || size < 4 || data[0] != 0xFF || data[1] != 0xD8) {
        if (err) snprintf(err, errSize, "not a JPEG stream");
        return -1;
    }
    d->pos = 2;
    for (;;) {
        while (d->pos < d->size && d->data[d->pos] != 0xFF) d->pos++;
        while (d->pos < d->size && d->data[d->pos] == 0xFF) d->pos++;
        if (d->pos >= d->size) {
            if (err) snprintf(err, errSize, "missing end of image");
            goto done;
        }
        int marker = d->data[d->pos++];
        if (marker == 0xD9) break;                          /* EOI */
        if (marker == 0x01 || (marker >= 0xD0 && marker <= 0xD7)) continue;
        int len = 0;
        if (readU16(d, &len) != 0 || len < 2 || d->pos + (size_t)len - 2 > d->size) {
            if (err) snprintf(err, errSize, "truncated marker segment");
            goto done;
        }
        size_t end = d->pos + (size_t)len - 2;
        int r = 0;
        switch (marker) {
            case 0xDB: r = parseDQT(d, end); break;
            case 0xC4: r = parseDHT(d, end); break;
            case 0xC0:
            case 0xC1:
            case 0xC2:
                if (frames++) {
                    if (err) snprintf(err, errSize, "more than one frame");
                    goto done;
                }
                d->progressive = marker == 0xC2;
                r = parseSOF(d, end);
                break;
            case 0xDD:
                if (len < 4) r = -1;
                else d->restartInterval = d->data[d->pos] << 8 | d->data[d->pos + 1];
                break;
            case 0xEE: /* Adobe: transform flag at byte 11 of the segment payload */
                if (len >= 14 && memcmp(d->data + d->pos, "Adobe", 5) == 0) d->adobeTransform = d->data[d->pos + 11];
                break;
            case 0xDA:
                r = parseSOS(d, end);
                end = d->pos; /* parseSOS moved past the entropy-coded data */
                break;
            default:
                if ((marker >= 0xC3 && marker <= 0xCF && marker != 0xC4 && marker != 0xC8 && marker != 0xCC) ||
                    marker == 0xC8) {
                    if (err) snprintf(err, errSize, "JPEG process 0x%02X (arithmetic, lossless or hierarchical) is not supported", marker);
                    goto done;
                }
                break; /* APPn, COM, ... */
        }
        if (r != 0) goto done;
        d->pos = end;
    }
    if (!frames) {
        if (err) snprintf(err, errSize, "no frame header");
        goto done;
    }
    {
        const IdctConstants K = idctConstants();
        for (int c = 0; c < d->compCount; c++) {
            Component* k = &d->comp[c];
            size_t pw = (size_t)k->blocksW * 8u;
            k->plane = (uint8_t*)malloc(pw * (size_t)k->blocksH * 8u);
            if (!k->plane) goto done;
            for (int by = 0; by < k->blocksH; by++)
                for (int bx = 0; bx < k->blocksW; bx++)
                    idctBlock(&K, k->coef + ((size_t)by * k->blocksW + bx) * 64u, d->quant[k->tq],
                              k->plane + (size_t)by * 8u * pw + (size_t)bx * 8u, (int)pw);
        }
        const size_t n = (size_t)d->width * (size_t)d->height;
        planes = (uint8_t*)malloc(n * (size_t)d->compCount);
        pixels = (uint8_t*)malloc(n * 4u);
        if (!planes || !pixels) goto done;
        for (int c = 0; c < d->compCount; c++) {
            if (upsampleComponent(d, &d->comp[c], planes + n * (size_t)c) != 0) {
                if (err) snprintf(err, errSize, "unsupported sampling factors");
                goto done;
            }
        }
        const int rgbInput = d->compCount == 3 &&
                             (d->adobeTransform == 0 ||
                              (d->adobeTransform < 0 && d->comp[0].id == 'R' && d->comp[1].id == 'G' && d->comp[2].id == 'B'));
        /* T.871 YCbCr -> RGB in 16-bit fixed point, rounded like the reference decoder */
        const int32_t ONE_HALF = 1 << 15;
        const int32_t kCrR = (int32_t)(1.40200 * 65536.0 + 0.5), kCbB = (int32_t)(1.77200 * 65536.0 + 0.5);
        const int32_t kCrG = (int32_t)(0.71414 * 65536.0 + 0.5), kCbG = (int32_t)(0.34414 * 65536.0 + 0.5);
        for (size_t i = 0; i < n; i++) {
            uint8_t* p = pixels + i * 4u;
            if (d->compCount == 1) {
                p[0] = p[1] = p[2] = planes[i];
            } else if (rgbInput) {
                p[0] = planes[i];
                p[1] = planes[n + i];
                p[2] = planes[2 * n + i];
            } else {
                int y = planes[i], cb = planes[n + i] - 128, cr = planes[2 * n + i] - 128;
                int r = y + ((kCrR * cr + ONE_HALF) >> 16);
                int g = y + ((-kCbG * cb + ONE_HALF - kCrG * cr) >> 16);
                int b = y + ((kCbB * cb + ONE_HALF) >> 16);
                p[0] = (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
                p[1] = (uint8_t)(g < 0 ? 0 : g > 255 ? 255 : g);
                p[2] = (uint8_t)(b < 0 ? 0 : b > 255 ? 255 : b);
            }
            p[3] = 255;
        }
        out->pixels = pixels;
        out->width = (uint32_t)d->width;
        out->height = (uint32_t)d->height;
        pixels = NULL;
        rc = 0;
    }
done:
    for (int c = 0; c < 4; c++) {
        free(d->comp[c].coef);
        free(d->comp[c].plane);
    }
    free(planes);
    free(pixels);
    return rc;
}
