/*
 * glTF 2.0 / GLB reader producing MeshImportData (src/app/mesh/loader.h).
 *
 * The reference parses with cgltf and converts in src/app/mesh/loader.c; this
 * reader parses the container and JSON itself (json.c) and then applies the
 * same conversions:
 *  - positions, normals and tangents go from glTF's +Y-up frame to the
 *    engine's +Z-up frame, (x, y, z) -> (x, -z, y) (loader.c:1816-1820,
 *    1756-1776); vertex colour defaults to white, texcoords to 0;
 *  - a primitive without normals gets area-weighted vertex normals
 *    (loader.c:1599-1644); with normals, each triangle whose winding disagrees
 *    with its vertex normals is flipped (loader.c:1894-1907) and tangents are
 *    imported, generated from texcoords or built from the normal
 *    (loader.c:1871-1892);
 *  - materials map metallic-roughness / specular-glossiness and the
 *    KHR_materials_{specular,ior,transmission,volume,clearcoat,sheen,
 *    emissive_strength} extensions onto Material (loader.c:942-1078,
 *    1472-1555);
 *  - nodes are walked depth-first from the scene's roots, each local transform
 *    conjugated into engine space (loader.c:1557-1597, 2027-2107).
 * Accessor reads follow the glTF 2.0 specification (strided buffer views,
 * sparse substitution, missing buffer view = zeros); normalized integers are
 * divided by the type's maximum, as the reference's loader library does.
 */
#include "internal.h"
#include "json.h"
#include "image_io.h"
#include "vkrt_import.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GLTF_LOG(...) VKRT_LOG_ERROR(__VA_ARGS__)

enum {
    kGlbMagic = 0x46546C67u,     /* "glTF" */
    kGlbChunkJson = 0x4E4F534Au, /* "JSON" */
    kGlbChunkBin = 0x004E4942u,  /* "BIN\0" */
    kModeTriangles = 4,
};

/* loader.c:27: host staging allowed for one primitive */
static const size_t kMaxPrimitiveBytes = (size_t)1024u * 1024u * 1024u;

typedef struct GltfBuffer {
    const uint8_t* data;
    size_t size;
    uint8_t* owned;
} GltfBuffer;

typedef struct GltfDoc {
    JsonValue root;
    uint8_t* file;
    size_t fileSize;
    const uint8_t* bin; /* GLB binary chunk */
    size_t binSize;
    GltfBuffer* buffers;
    uint32_t bufferCount;
    char directory[4096];
    const JsonValue* accessors;
    const JsonValue* bufferViews;
    const JsonValue* materials;
    const JsonValue* meshes;
    const JsonValue* nodes;
    int textureFailed; /* an image that does not decode fails the import, as in the reference */
} GltfDoc;

/* ---------------------------------------------------------------- files */

static uint8_t* readWholeFile(const char* path, size_t* outSize) {
    *outSize = 0;
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return NULL;
    }
    long end = ftell(f);
    if (end < 0 || fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return NULL;
    }
    uint8_t* data = (uint8_t*)malloc((size_t)end + 1u);
    if (!data) {
        fclose(f);
        return NULL;
    }
    size_t got = fread(data, 1, (size_t)end, f);
    fclose(f);
    if (got != (size_t)end) {
        free(data);
        return NULL;
    }
    data[end] = 0;
    *outSize = (size_t)end;
    return data;
}

static void directoryOf(const char* path, char* out, size_t outSize) {
    const char* slash = strrchr(path, '/');
    if (!slash) {
        snprintf(out, outSize, ".");
        return;
    }
    size_t n = (size_t)(slash - path);
    if (n == 0) n = 1; /* "/file" */
    if (n >= outSize) n = outSize - 1;
    memcpy(out, path, n);
    out[n] = '\0';
}

static int hexValue(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

/* URIs in glTF are percent-encoded relative references. */
static int decodeUriPath(const char* uri, char* out, size_t outSize) {
    size_t n = 0;
    for (size_t i = 0; uri[i]; i++) {
        char c = uri[i];
        if (c == '%' && hexValue(uri[i + 1]) >= 0 && hexValue(uri[i + 2]) >= 0) {
            c = (char)(hexValue(uri[i + 1]) * 16 + hexValue(uri[i + 2]));
            i += 2;
        }
        if (n + 1 >= outSize) return 0;
        out[n++] = c;
    }
    out[n] = '\0';
    return 1;
}

static int base64Value(char c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
}

static uint8_t* decodeBase64(const char* text, size_t* outSize) {
    size_t len = strlen(text);
    uint8_t* out = (uint8_t*)malloc(len / 4u * 3u + 3u);
    if (!out) return NULL;
    uint32_t acc = 0;
    int bits = 0;
    size_t n = 0;
    for (size_t i = 0; i < len; i++) {
        char c = text[i];
        if (c == '=') break;
        int v = base64Value(c);
        if (v < 0) {
            free(out);
            return NULL;
        }
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out[n++] = (uint8_t)(acc >> bits);
        }
    }
    *outSize = n;
    return out;
}

/* -------------------------------------------------------------- buffers */

static int loadBuffers(GltfDoc* doc) {
    const JsonValue* buffers = vkrtJsonGet(&doc->root, "buffers");
    if (!buffers) return 0;
    if (buffers->type != JSON_ARRAY) return -1;
    doc->bufferCount = buffers->count;
    doc->buffers = (GltfBuffer*)calloc(buffers->count ? buffers->count : 1u, sizeof(GltfBuffer));
    if (!doc->buffers) return -1;
    for (uint32_t i = 0; i < buffers->count; i++) {
        const JsonValue* b = &buffers->items[i];
        uint32_t byteLength = 0;
        if (!vkrtJsonToU32(vkrtJsonGet(b, "byteLength"), &byteLength)) {
            GLTF_LOG("glTF buffer %u has no valid byteLength", i);
            return -1;
        }
        const JsonValue* uri = vkrtJsonGet(b, "uri");
        GltfBuffer* gb = &doc->buffers[i];
        if (!uri) {
            if (i != 0 || !doc->bin) {
                GLTF_LOG("glTF buffer %u has no uri and no GLB binary chunk", i);
                return -1;
            }
            gb->data = doc->bin;
            gb->size = doc->binSize;
        } else if (uri->type == JSON_STRING && strncmp(uri->string, "data:", 5) == 0) {
            const char* comma = strstr(uri->string, ";base64,");
            if (!comma) {
                GLTF_LOG("glTF buffer %u: only base64 data URIs are supported", i);
                return -1;
            }
            gb->owned = decodeBase64(comma + 8, &gb->size);
            if (!gb->owned) {
                GLTF_LOG("glTF buffer %u: malformed base64", i);
                return -1;
            }
            gb->data = gb->owned;
        } else if (uri->type == JSON_STRING) {
            char rel[4096];
            char full[8192];
            if (!decodeUriPath(uri->string, rel, sizeof(rel))) return -1;
            if (rel[0] == '/')
                snprintf(full, sizeof(full), "%s", rel);
            else
                snprintf(full, sizeof(full), "%s/%s", doc->directory, rel);
            gb->owned = readWholeFile(full, &gb->size);
            if (!gb->owned) {
                GLTF_LOG("glTF buffer file not found: %s", full);
                return -1;
            }
            gb->data = gb->owned;
        } else {
            return -1;
        }
        if (gb->size < byteLength) {
            GLTF_LOG("glTF buffer %u holds %zu bytes, byteLength says %u", i, gb->size, byteLength);
            return -1;
        }
        gb->size = byteLength;
    }
    return 0;
}

/* ------------------------------------------------------------ accessors */

static uint32_t componentCount(const char* type) {
    if (!type) return 0;
    if (strcmp(type, "SCALAR") == 0) return 1;
    if (strcmp(type, "VEC2") == 0) return 2;
    if (strcmp(type, "VEC3") == 0) return 3;
    if (strcmp(type, "VEC4") == 0) return 4;
    if (strcmp(type, "MAT2") == 0) return 4;
    if (strcmp(type, "MAT3") == 0) return 9;
    if (strcmp(type, "MAT4") == 0) return 16;
    return 0;
}

static uint32_t componentSize(uint32_t componentType) {
    switch (componentType) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
        default: return 0;
    }
}

static float readComponentFloat(const uint8_t* p, uint32_t componentType, int normalized) {
    switch (componentType) {
        case 5126: {
            float f;
            memcpy(&f, p, 4);
            return f;
        }
        case 5120: {
            int8_t v = (int8_t)p[0];
            return normalized ? (float)v / 127.0f : (float)v;
        }
        case 5121: return normalized ? (float)p[0] / 255.0f : (float)p[0];
        case 5122: {
            int16_t v;
            memcpy(&v, p, 2);
            return normalized ? (float)v / 32767.0f : (float)v;
        }
        case 5123: {
            uint16_t v;
            memcpy(&v, p, 2);
            return normalized ? (float)v / 65535.0f : (float)v;
        }
        case 5125: {
            uint32_t v;
            memcpy(&v, p, 4);
            return normalized ? 0.0f : (float)v;
        }
        default: return 0.0f;
    }
}

static uint32_t readComponentIndex(const uint8_t* p, uint32_t componentType) {
    switch (componentType) {
        case 5121: return p[0];
        case 5123: {
            uint16_t v;
            memcpy(&v, p, 2);
            return v;
        }
        case 5125: {
            uint32_t v;
            memcpy(&v, p, 4);
            return v;
        }
        default: return UINT32_MAX;
    }
}

typedef struct AccessorView {
    const uint8_t* base; /* NULL: all zeros */
    size_t stride;
    uint32_t componentType;
    uint32_t components;
    uint8_t normalized;
    size_t count;
} AccessorView;

/* Resolves a buffer-view span [offset, offset + stride*(count-1) + elemSize)
 * and checks it against the view and its buffer. */
static int resolveView(const GltfDoc* doc, uint32_t viewIndex, size_t offset, size_t elemSize, size_t count,
                       int allowStride, const uint8_t** outBase, size_t* outStride) {
    const JsonValue* view = vkrtJsonAt(doc->bufferViews, viewIndex);
    uint32_t bufferIndex = 0, viewOffset = 0, viewLength = 0, stride = 0;
    if (!view || !vkrtJsonToU32(vkrtJsonGet(view, "buffer"), &bufferIndex) ||
        !vkrtJsonToU32(vkrtJsonGet(view, "byteLength"), &viewLength) || bufferIndex >= doc->bufferCount)
        return -1;
    const JsonValue* ov = vkrtJsonGet(view, "byteOffset");
    if (ov && !vkrtJsonToU32(ov, &viewOffset)) return -1;
    const JsonValue* sv = vkrtJsonGet(view, "byteStride");
    if (sv && !vkrtJsonToU32(sv, &stride)) return -1;
    if (!allowStride || stride == 0) stride = (uint32_t)elemSize;
    if (stride < elemSize) return -1;
    const GltfBuffer* b = &doc->buffers[bufferIndex];
    if ((size_t)viewOffset + viewLength > b->size) return -1;
    if (count > 0) {
        size_t span = offset;
        if ((count - 1u) > (SIZE_MAX - elemSize - span) / stride) return -1;
        span += (count - 1u) * (size_t)stride + elemSize;
        if (span > viewLength) return -1;
    }
    *outBase = b->data + viewOffset + offset;
    *outStride = stride;
    return 0;
}

static int viewAccessor(const GltfDoc* doc, uint32_t accessorIndex, AccessorView* out) {
    memset(out, 0, sizeof(*out));
    const JsonValue* a = vkrtJsonAt(doc->accessors, accessorIndex);
    uint32_t componentType = 0, count = 0, offset = 0, viewIndex = 0;
    const JsonValue* type = vkrtJsonGet(a, "type");
    if (!a || !vkrtJsonToU32(vkrtJsonGet(a, "componentType"), &componentType) ||
        !vkrtJsonToU32(vkrtJsonGet(a, "count"), &count) || !type || type->type != JSON_STRING)
        return -1;
    out->componentType = componentType;
    out->components = componentCount(type->string);
    out->count = count;
    const JsonValue* norm = vkrtJsonGet(a, "normalized");
    out->normalized = norm && norm->type == JSON_BOOL && norm->boolean;
    uint32_t csize = componentSize(componentType);
    if (!out->components || !csize) return -1;
    const JsonValue* ov = vkrtJsonGet(a, "byteOffset");
    if (ov && !vkrtJsonToU32(ov, &offset)) return -1;
    const JsonValue* bv = vkrtJsonGet(a, "bufferView");
    if (!bv) return 0; /* zeros, possibly overridden by sparse values */
    if (!vkrtJsonToU32(bv, &viewIndex)) return -1;
    return resolveView(doc, viewIndex, offset, (size_t)csize * out->components, count, 1, &out->base, &out->stride);
}

/* Dense float copy of an accessor with exactly `components` components per
 * element (sparse substitution applied). */
static float* readAccessorFloats(const GltfDoc* doc, uint32_t accessorIndex, uint32_t components, size_t* outCount) {
    AccessorView v;
    if (viewAccessor(doc, accessorIndex, &v) != 0 || v.components != components) return NULL;
    float* out = (float*)calloc(v.count ? v.count * components : 1u, sizeof(float));
    if (!out) return NULL;
    uint32_t csize = componentSize(v.componentType);
    if (v.base) {
        for (size_t i = 0; i < v.count; i++) {
            const uint8_t* e = v.base + i * v.stride;
            for (uint32_t c = 0; c < components; c++)
                out[i * components + c] = readComponentFloat(e + (size_t)c * csize, v.componentType, v.normalized);
        }
    }
    const JsonValue* sparse = vkrtJsonGet(vkrtJsonAt(doc->accessors, accessorIndex), "sparse");
    if (sparse) {
        uint32_t sc = 0, iview = 0, vview = 0, ioff = 0, voff = 0, itype = 0;
        const JsonValue* si = vkrtJsonGet(sparse, "indices");
        const JsonValue* sv = vkrtJsonGet(sparse, "values");
        const uint8_t *ib = NULL, *vb = NULL;
        size_t istride = 0, vstride = 0;
        const JsonValue* o1 = vkrtJsonGet(si, "byteOffset");
        const JsonValue* o2 = vkrtJsonGet(sv, "byteOffset");
        if (!vkrtJsonToU32(vkrtJsonGet(sparse, "count"), &sc) || !vkrtJsonToU32(vkrtJsonGet(si, "bufferView"), &iview) ||
            !vkrtJsonToU32(vkrtJsonGet(si, "componentType"), &itype) ||
            !vkrtJsonToU32(vkrtJsonGet(sv, "bufferView"), &vview) || (o1 && !vkrtJsonToU32(o1, &ioff)) ||
            (o2 && !vkrtJsonToU32(o2, &voff)) || componentSize(itype) == 0 || itype == 5126 ||
            resolveView(doc, iview, ioff, componentSize(itype), sc, 0, &ib, &istride) != 0 ||
            resolveView(doc, vview, voff, (size_t)csize * components, sc, 0, &vb, &vstride) != 0) {
            free(out);
            return NULL;
        }
        for (uint32_t k = 0; k < sc; k++) {
            uint32_t target = readComponentIndex(ib + (size_t)k * istride, itype);
            if (target >= v.count) {
                free(out);
                return NULL;
            }
            for (uint32_t c = 0; c < components; c++)
                out[(size_t)target * components + c] =
                    readComponentFloat(vb + (size_t)k * vstride + (size_t)c * csize, v.componentType, v.normalized);
        }
    }
    *outCount = v.count;
    return out;
}

static uint32_t* readAccessorIndices(const GltfDoc* doc, uint32_t accessorIndex, size_t* outCount) {
    AccessorView v;
    if (viewAccessor(doc, accessorIndex, &v) != 0 || v.components != 1 ||
        (v.componentType != 5121 && v.componentType != 5123 && v.componentType != 5125))
        return NULL;
    if (vkrtJsonGet(vkrtJsonAt(doc->accessors, accessorIndex), "sparse")) return NULL;
    uint32_t* out = (uint32_t*)calloc(v.count ? v.count : 1u, sizeof(uint32_t));
    if (!out) return NULL;
    if (v.base) {
        for (size_t i = 0; i < v.count; i++) out[i] = readComponentIndex(v.base + i * v.stride, v.componentType);
    }
    *outCount = v.count;
    return out;
}

static size_t accessorCount(const GltfDoc* doc, uint32_t accessorIndex) {
    uint32_t count = 0;
    vkrtJsonToU32(vkrtJsonGet(vkrtJsonAt(doc->accessors, accessorIndex), "count"), &count);
    return count;
}

static int accessorComponents(const GltfDoc* doc, uint32_t accessorIndex) {
    const JsonValue* type = vkrtJsonGet(vkrtJsonAt(doc->accessors, accessorIndex), "type");
    return type && type->type == JSON_STRING ? (int)componentCount(type->string) : 0;
}

/* ------------------------------------------------------------ materials */

static float readFloatOr(const JsonValue* object, const char* key, float fallback) {
    float v;
    return vkrtJsonToFloat(vkrtJsonGet(object, key), &v) ? v : fallback;
}

static void readFloatsOr(const JsonValue* object, const char* key, float* out, uint32_t n, const float* fallback) {
    if (!vkrtJsonToFloats(vkrtJsonGet(object, key), out, n)) memcpy(out, fallback, sizeof(float) * n);
}

static float max3f(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

/* loader.c:1472-1555 with the factor mappings of loader.c:942-1078 */
static Material buildMaterial(const JsonValue* src, uint32_t* outTextureRefs) {
    static const float one4[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    static const float zero3[3] = {0.0f, 0.0f, 0.0f};
    Material m = VKRT_materialDefault();
    if (!src) return m;
    const JsonValue* ext = vkrtJsonGet(src, "extensions");
    const JsonValue* sg = vkrtJsonGet(ext, "KHR_materials_pbrSpecularGlossiness");
    const JsonValue* pbr = vkrtJsonGet(src, "pbrMetallicRoughness");
    uint32_t refs = 0;
    if (sg) {
        float diffuse[4];
        readFloatsOr(sg, "diffuseFactor", diffuse, 4, one4);
        for (int c = 0; c < 3; c++) m.baseColor[c] = diffuse[c];
        m.opacity = diffuse[3];
        m.roughness = 1.0f - readFloatOr(sg, "glossinessFactor", 1.0f);
        refs += vkrtJsonGet(sg, "diffuseTexture") != NULL;
    } else if (pbr) {
        float base[4];
        readFloatsOr(pbr, "baseColorFactor", base, 4, one4);
        for (int c = 0; c < 3; c++) m.baseColor[c] = base[c];
        m.opacity = base[3];
        m.metallic = readFloatOr(pbr, "metallicFactor", 1.0f);
        m.roughness = readFloatOr(pbr, "roughnessFactor", 1.0f);
        refs += vkrtJsonGet(pbr, "baseColorTexture") != NULL;
        refs += vkrtJsonGet(pbr, "metallicRoughnessTexture") != NULL;
    }
    const JsonValue* spec = vkrtJsonGet(ext, "KHR_materials_specular");
    if (spec) m.specular = readFloatOr(spec, "specularFactor", 1.0f);
    const JsonValue* ior = vkrtJsonGet(ext, "KHR_materials_ior");
    if (ior) m.ior = readFloatOr(ior, "ior", 1.5f);
    const JsonValue* tr = vkrtJsonGet(ext, "KHR_materials_transmission");
    if (tr) m.transmission = readFloatOr(tr, "transmissionFactor", 0.0f);
    const JsonValue* vol = vkrtJsonGet(ext, "KHR_materials_volume");
    if (vol) {
        float att[3];
        readFloatsOr(vol, "attenuationColor", att, 3, one4);
        for (int c = 0; c < 3; c++) m.attenuationColor[c] = att[c];
        /* an absent attenuationDistance means infinity; the reference's library
         * stores FLT_MAX for it, which is finite, giving 1 / FLT_MAX */
        float dist = readFloatOr(vol, "attenuationDistance", FLT_MAX);
        m.absorptionCoefficient = dist > 0.0f && isfinite(dist) ? 1.0f / dist : 0.0f;
    }
    const JsonValue* cc = vkrtJsonGet(ext, "KHR_materials_clearcoat");
    if (cc) {
        m.clearcoat = readFloatOr(cc, "clearcoatFactor", 0.0f);
        m.clearcoatGloss = 1.0f - readFloatOr(cc, "clearcoatRoughnessFactor", 0.0f);
    }
    const JsonValue* sheen = vkrtJsonGet(ext, "KHR_materials_sheen");
    if (sheen) {
        float col[3];
        readFloatsOr(sheen, "sheenColorFactor", col, 3, zero3);
        float w = max3f(col[0], col[1], col[2]);
        for (int c = 0; c < 3; c++) m.sheenTintWeight[c] = w > 0.0f ? col[c] / w : 1.0f;
        m.sheenTintWeight[3] = w > 0.0f ? w : 0.0f;
        m.sheenRoughness = readFloatOr(sheen, "sheenRoughnessFactor", 0.0f);
    }
    /* loader.c:1056-1078 */
    float emissive[3];
    readFloatsOr(src, "emissiveFactor", emissive, 3, zero3);
    const JsonValue* es = vkrtJsonGet(ext, "KHR_materials_emissive_strength");
    float strength = es ? readFloatOr(es, "emissiveStrength", 1.0f) : 1.0f;
    for (int c = 0; c < 3; c++) emissive[c] *= strength;
    float emax = max3f(emissive[0], emissive[1], emissive[2]);
    for (int c = 0; c < 3; c++) m.emissionColor[c] = emax > 0.0f ? emissive[c] / emax : 1.0f;
    m.emissionLuminance = emax > 0.0f ? emax : 0.0f;

    refs += vkrtJsonGet(src, "normalTexture") != NULL;
    refs += vkrtJsonGet(src, "emissiveTexture") != NULL;
    const JsonValue* nt = vkrtJsonGet(src, "normalTexture");
    if (nt) {
        float s = readFloatOr(nt, "scale", 1.0f);
        m.normalTextureScale = s > 0.0f ? s : 1.0f;
    }
    const JsonValue* mode = vkrtJsonGet(src, "alphaMode");
    if (mode && mode->type == JSON_STRING) {
        if (strcmp(mode->string, "MASK") == 0) {
            m.alphaMode = VKRT_MATERIAL_ALPHA_MODE_MASK;
            m.alphaCutoff = readFloatOr(src, "alphaCutoff", 0.5f);
        } else if (strcmp(mode->string, "BLEND") == 0) {
            m.alphaMode = VKRT_MATERIAL_ALPHA_MODE_BLEND;
            m.alphaCutoff = 1.0f / 255.0f;
        }
    }
    if (outTextureRefs) *outTextureRefs += refs;
    return m;
}

static char* duplicateString(const char* s);
static const char* jsonName(const JsonValue* object);

/* ------------------------------------------------------------- textures */

/* One image of the file as a texture entry (src/app/mesh/loader.c:432-499),
 * shared by every view of the same image in the same colour space. The image
 * is decoded from its file, data URI or buffer view by the library's image
 * reader (PNG, JPEG, EXR: the reference's formats and output rules); an image
 * that does not decode fails the whole import, as the reference's does. */
static uint32_t importTexture(GltfDoc* doc, MeshImportData* out, uint64_t* keys, uint32_t textureIndex,
                              uint32_t colorSpace, uint32_t* outWrap) {
    const JsonValue* tex = vkrtJsonAt(vkrtJsonGet(&doc->root, "textures"), textureIndex);
    uint32_t source = 0, samplerIndex = 0;
    *outWrap = VKRT_TEXTURE_WRAP_DEFAULT;
    if (!tex || !vkrtJsonToU32(vkrtJsonGet(tex, "source"), &source)) return VKRT_INVALID_INDEX;
    const JsonValue* image = vkrtJsonAt(vkrtJsonGet(&doc->root, "images"), source);
    if (!image) return VKRT_INVALID_INDEX;
    if (vkrtJsonToU32(vkrtJsonGet(tex, "sampler"), &samplerIndex)) {
        const JsonValue* smp = vkrtJsonAt(vkrtJsonGet(&doc->root, "samplers"), samplerIndex);
        uint32_t ws = VKRT_TEXTURE_WRAP_REPEAT, wt = VKRT_TEXTURE_WRAP_REPEAT; /* loader.c:362-366 */
        uint32_t v = 0;
        if (smp && vkrtJsonToU32(vkrtJsonGet(smp, "wrapS"), &v)) ws = v;
        if (smp && vkrtJsonToU32(vkrtJsonGet(smp, "wrapT"), &v)) wt = v;
        *outWrap = ws | (wt << 16);
    }
    const uint64_t key = ((uint64_t)source << 1) | (colorSpace == VKRT_TEXTURE_COLOR_SPACE_SRGB);
    for (uint32_t i = 0; i < out->textureCount; i++)
        if (keys[i] == key) return i;
    const JsonValue* uri = vkrtJsonGet(image, "uri");
    const JsonValue* mime = vkrtJsonGet(image, "mimeType");
    char mimeType[64] = "";
    if (mime && mime->type == JSON_STRING) snprintf(mimeType, sizeof(mimeType), "%s", mime->string);
    uint8_t* owned = NULL;
    const uint8_t* bytes = NULL;
    size_t size = 0;
    uint32_t view = 0;
    if (uri && uri->type == JSON_STRING && strncmp(uri->string, "data:", 5) == 0) {
        const char* comma = strchr(uri->string, ',');
        const char* semi = strchr(uri->string, ';');
        if (semi && comma && semi < comma && (size_t)(semi - uri->string - 5) < sizeof(mimeType))
            snprintf(mimeType, sizeof(mimeType), "%.*s", (int)(semi - uri->string - 5), uri->string + 5);
        if (comma) owned = decodeBase64(comma + 1, &size);
        bytes = owned;
    } else if (uri && uri->type == JSON_STRING) {
        char rel[4096], full[8192];
        if (decodeUriPath(uri->string, rel, sizeof(rel))) {
            snprintf(full, sizeof(full), "%s/%s", doc->directory, rel);
            owned = readWholeFile(full, &size);
        }
        bytes = owned;
        mimeType[0] = 0; /* loader.c:443-449: files are decoded by their contents */
    } else if (vkrtJsonToU32(vkrtJsonGet(image, "bufferView"), &view)) {
        uint32_t length = 0;
        size_t stride = 0;
        const JsonValue* bv = vkrtJsonAt(doc->bufferViews, view);
        if (bv && vkrtJsonToU32(vkrtJsonGet(bv, "byteLength"), &length) && length &&
            resolveView(doc, view, 0, 1, length, 0, &bytes, &stride) == 0)
            size = length;
    }
    VkrtLoadedImage decoded;
    char label[64];
    snprintf(label, sizeof(label), "glTF image %u", source);
    int ok = bytes && size && vkrtLoadImageFromMemory(bytes, size, mimeType, colorSpace, &decoded, label) == 0;
    free(owned);
    if (!ok) {
        GLTF_LOG("glTF image %u could not be decoded", source);
        doc->textureFailed = 1;
        return VKRT_INVALID_INDEX;
    }
    TextureImportEntry e;
    memset(&e, 0, sizeof(e));
    e.width = decoded.width;
    e.height = decoded.height;
    e.pixels = decoded.pixels;
    e.format = decoded.format;
    e.colorSpace = decoded.colorSpace;
    const char* name = jsonName(image) ? jsonName(image) : (jsonName(tex) ? jsonName(tex) : "Texture");
    e.name = duplicateString(name);
    if (out->textureCount == out->textureCapacity) {
        uint32_t cap = out->textureCapacity ? out->textureCapacity * 2u : 8u;
        TextureImportEntry* grown = (TextureImportEntry*)realloc(out->textures, cap * sizeof(TextureImportEntry));
        if (!grown || !e.name) {
            if (grown) out->textures = grown;
            free(e.pixels);
            free(e.name);
            return VKRT_INVALID_INDEX;
        }
        out->textures = grown;
        out->textureCapacity = cap;
    }
    keys[out->textureCount] = key;
    out->textures[out->textureCount] = e;
    return out->textureCount++;
}

/* A material texture view (loader.c:372-404, 907-1010): texture, sampler
 * wrap, texture-coordinate set and KHR_texture_transform (its texCoord
 * overrides the view's). The set is stored in 8 bits as the reference does;
 * shading reads TEXCOORD_1 for set 1 and TEXCOORD_0 for any other set
 * (textures.slang). */
static void attachTexture(GltfDoc* doc, MeshImportData* out, uint64_t* keys, Material* m, const JsonValue* view,
                          uint32_t slot, uint32_t colorSpace) {
    uint32_t index = 0, set = 0, wrap = 0;
    if (!view || !vkrtJsonToU32(vkrtJsonGet(view, "index"), &index)) return;
    const JsonValue* tt = vkrtJsonGet(vkrtJsonGet(view, "extensions"), "KHR_texture_transform");
    if (!(tt && vkrtJsonToU32(vkrtJsonGet(tt, "texCoord"), &set)) &&
        !vkrtJsonToU32(vkrtJsonGet(view, "texCoord"), &set))
        set = 0;
    const uint32_t t = importTexture(doc, out, keys, index, colorSpace, &wrap);
    if (t == VKRT_INVALID_INDEX) return;
    float xf[4] = {1.0f, 1.0f, 0.0f, 0.0f};
    float rotation = 0.0f;
    if (tt) {
        static const float one2[2] = {1.0f, 1.0f}, zero2[2] = {0.0f, 0.0f};
        readFloatsOr(tt, "scale", xf, 2, one2);
        readFloatsOr(tt, "offset", xf + 2, 2, zero2);
        rotation = readFloatOr(tt, "rotation", 0.0f);
    }
    uint32_t* idx[4] = {&m->baseColorTextureIndex, &m->metallicRoughnessTextureIndex, &m->normalTextureIndex,
                        &m->emissiveTextureIndex};
    uint32_t* wraps[4] = {&m->baseColorTextureWrap, &m->metallicRoughnessTextureWrap, &m->normalTextureWrap,
                          &m->emissiveTextureWrap};
    float* xfs[4] = {m->baseColorTextureTransform, m->metallicRoughnessTextureTransform, m->normalTextureTransform,
                     m->emissiveTextureTransform};
    *idx[slot] = t;
    *wraps[slot] = wrap;
    memcpy(xfs[slot], xf, sizeof(xf));
    m->textureRotations[slot] = rotation;
    m->textureTexcoordSets = (m->textureTexcoordSets & ~(0xffu << (8u * slot))) | ((set & 0xffu) << (8u * slot));
}

static void attachMaterialTextures(GltfDoc* doc, MeshImportData* out, uint64_t* keys, Material* m,
                                   const JsonValue* src) {
    const JsonValue* pbr = vkrtJsonGet(src, "pbrMetallicRoughness");
    attachTexture(doc, out, keys, m, vkrtJsonGet(pbr, "baseColorTexture"), VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR,
                  VKRT_TEXTURE_COLOR_SPACE_SRGB);
    attachTexture(doc, out, keys, m, vkrtJsonGet(pbr, "metallicRoughnessTexture"),
                  VKRT_MATERIAL_TEXTURE_SLOT_METALLIC_ROUGHNESS, VKRT_TEXTURE_COLOR_SPACE_LINEAR);
    attachTexture(doc, out, keys, m, vkrtJsonGet(src, "normalTexture"), VKRT_MATERIAL_TEXTURE_SLOT_NORMAL,
                  VKRT_TEXTURE_COLOR_SPACE_LINEAR);
    attachTexture(doc, out, keys, m, vkrtJsonGet(src, "emissiveTexture"), VKRT_MATERIAL_TEXTURE_SLOT_EMISSIVE,
                  VKRT_TEXTURE_COLOR_SPACE_SRGB);
}

/* ------------------------------------------------------------- geometry */

static void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static int normalize3(float v[3]) {
    float n2 = dot3(v, v);
    if (n2 <= 1e-12f) return 0;
    float inv = 1.0f / sqrtf(n2);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
    return 1;
}

static int triangleAt(const uint32_t* idx, size_t vertexCount, size_t o, uint32_t t[3]) {
    t[0] = idx[o];
    t[1] = idx[o + 1];
    t[2] = idx[o + 2];
    return t[0] < vertexCount && t[1] < vertexCount && t[2] < vertexCount;
}

static void faceNormal(const Vertex* v, const uint32_t t[3], float out[3]) {
    float e1[3], e2[3];
    for (int c = 0; c < 3; c++) {
        e1[c] = v[t[1]].position[c] - v[t[0]].position[c];
        e2[c] = v[t[2]].position[c] - v[t[0]].position[c];
    }
    cross3(e1, e2, out);
}

/* loader.c:1599-1644 */
static void generateNormals(Vertex* v, size_t vertexCount, const uint32_t* idx, size_t indexCount) {
    for (size_t i = 0; i < vertexCount; i++) memset(v[i].normal, 0, sizeof(v[i].normal));
    for (size_t o = 0; o + 2 < indexCount; o += 3) {
        uint32_t t[3];
        float n[3];
        if (!triangleAt(idx, vertexCount, o, t)) continue;
        faceNormal(v, t, n);
        if (dot3(n, n) <= 1e-12f) continue;
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) v[t[k]].normal[c] += n[c];
    }
    for (size_t i = 0; i < vertexCount; i++) {
        float n[3] = {v[i].normal[0], v[i].normal[1], v[i].normal[2]};
        if (!normalize3(n)) n[2] = 1.0f;
        v[i].normal[0] = n[0];
        v[i].normal[1] = n[1];
        v[i].normal[2] = n[2];
        v[i].normal[3] = 0.0f;
    }
}

/* loader.c:1894-1907 with the alignment test of loader.c:1343-1384 */
static void alignWindingToNormals(const Vertex* v, size_t vertexCount, uint32_t* idx, size_t indexCount) {
    for (size_t o = 0; o + 2 < indexCount; o += 3) {
        uint32_t t[3];
        float n[3], avg[3];
        if (!triangleAt(idx, vertexCount, o, t)) continue;
        faceNormal(v, t, n);
        if (dot3(n, n) <= 1e-12f) continue;
        for (int c = 0; c < 3; c++) avg[c] = v[t[0]].normal[c] + v[t[1]].normal[c] + v[t[2]].normal[c];
        if (dot3(avg, avg) <= 1e-12f) continue;
        if (dot3(n, avg) >= 0.0f) continue;
        uint32_t tmp = idx[o + 1];
        idx[o + 1] = idx[o + 2];
        idx[o + 2] = tmp;
    }
}

/* loader.c:1386-1446 */
static int orthonormalTangent(const float normal[3], const float tangentIn[3], float handedness, float out[4]) {
    float n[3] = {normal[0], normal[1], normal[2]};
    float t[3] = {tangentIn[0], tangentIn[1], tangentIn[2]};
    if (dot3(n, n) <= 1e-12f || dot3(t, t) <= 1e-12f) return 0;
    normalize3(n);
    float d = dot3(t, n);
    for (int c = 0; c < 3; c++) t[c] -= n[c] * d;
    if (!normalize3(t)) return 0;
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
    out[3] = handedness < 0.0f ? -1.0f : 1.0f;
    return 1;
}

static void fallbackTangent(const float normal[3], float out[4]) {
    float n[3] = {normal[0], normal[1], normal[2]};
    if (dot3(n, n) <= 1e-12f)
        n[2] = 1.0f;
    else
        normalize3(n);
    float up[3] = {0.0f, 0.0f, 1.0f};
    if (fabsf(n[2]) > 0.999f) {
        up[0] = 1.0f;
        up[2] = 0.0f;
    }
    float t[3];
    cross3(up, n, t);
    if (!normalize3(t)) {
        t[0] = 1.0f;
        t[1] = 0.0f;
        t[2] = 0.0f;
    }
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
    out[3] = 1.0f;
}

/* loader.c:1646-1754: per-vertex sums of the UV-space tangent and bitangent */
static void generateTangents(Vertex* v, size_t vertexCount, const uint32_t* idx, size_t indexCount,
                             const float* uv) {
    float* t1 = (float*)calloc(vertexCount * 3u, sizeof(float));
    float* t2 = (float*)calloc(vertexCount * 3u, sizeof(float));
    if (!t1 || !t2) {
        free(t1);
        free(t2);
        for (size_t i = 0; i < vertexCount; i++) fallbackTangent(v[i].normal, v[i].tangent);
        return;
    }
    for (size_t o = 0; o + 2 < indexCount; o += 3) {
        uint32_t t[3];
        if (!triangleAt(idx, vertexCount, o, t)) continue;
        float e1[3], e2[3];
        for (int c = 0; c < 3; c++) {
            e1[c] = v[t[1]].position[c] - v[t[0]].position[c];
            e2[c] = v[t[2]].position[c] - v[t[0]].position[c];
        }
        float du1 = uv[t[1] * 2u] - uv[t[0] * 2u], dv1 = uv[t[1] * 2u + 1u] - uv[t[0] * 2u + 1u];
        float du2 = uv[t[2] * 2u] - uv[t[0] * 2u], dv2 = uv[t[2] * 2u + 1u] - uv[t[0] * 2u + 1u];
        float det = du1 * dv2 - dv1 * du2;
        if (fabsf(det) <= 1e-12f) continue;
        float inv = 1.0f / det;
        for (int k = 0; k < 3; k++) {
            for (int c = 0; c < 3; c++) {
                t1[t[k] * 3u + (uint32_t)c] += (dv2 * e1[c] - dv1 * e2[c]) * inv;
                t2[t[k] * 3u + (uint32_t)c] += (du1 * e2[c] - du2 * e1[c]) * inv;
            }
        }
    }
    for (size_t i = 0; i < vertexCount; i++) {
        float* tan = &t1[i * 3u];
        float n[3] = {v[i].normal[0], v[i].normal[1], v[i].normal[2]};
        if (dot3(n, n) <= 1e-12f || dot3(tan, tan) <= 1e-12f) {
            fallbackTangent(v[i].normal, v[i].tangent);
            continue;
        }
        normalize3(n);
        float nxt[3];
        cross3(n, tan, nxt);
        float handed = dot3(nxt, &t2[i * 3u]) < 0.0f ? -1.0f : 1.0f;
        if (!orthonormalTangent(v[i].normal, tan, handed, v[i].tangent)) fallbackTangent(v[i].normal, v[i].tangent);
    }
    free(t1);
    free(t2);
}

/* ----------------------------------------------------------- primitives */

typedef struct PrimitiveAttributes {
    int position, normal, tangent, texcoord0, texcoord1, color; /* accessor indices, -1 if absent */
} PrimitiveAttributes;

static int attributeIndex(const JsonValue* attributes, const char* name) {
    uint32_t v;
    return vkrtJsonToU32(vkrtJsonGet(attributes, name), &v) ? (int)v : -1;
}

static char* duplicateString(const char* s) {
    size_t n = strlen(s);
    char* d = (char*)malloc(n + 1);
    if (d) memcpy(d, s, n + 1);
    return d;
}

static const char* jsonName(const JsonValue* object) {
    const JsonValue* n = vkrtJsonGet(object, "name");
    return n && n->type == JSON_STRING && n->string[0] ? n->string : NULL;
}

static void releaseEntry(MeshImportEntry* e) {
    free(e->name);
    free(e->vertices);
    free(e->indices);
    memset(e, 0, sizeof(*e));
}

/* loader.c:1909-1996. Returns 1 when an entry was built, 0 when the primitive
 * is skipped (not triangles, no positions), -1 on error. */
static int buildPrimitive(const GltfDoc* doc, const JsonValue* node, const JsonValue* mesh, uint32_t primitiveIndex,
                          MeshImportEntry* out) {
    memset(out, 0, sizeof(*out));
    const JsonValue* prims = vkrtJsonGet(mesh, "primitives");
    const JsonValue* prim = vkrtJsonAt(prims, primitiveIndex);
    if (!prim) return -1;
    uint32_t mode = kModeTriangles;
    const JsonValue* mv = vkrtJsonGet(prim, "mode");
    if (mv && !vkrtJsonToU32(mv, &mode)) return -1;
    if (mode != kModeTriangles) return 0;
    const JsonValue* attrs = vkrtJsonGet(prim, "attributes");
    PrimitiveAttributes a = {attributeIndex(attrs, "POSITION"), attributeIndex(attrs, "NORMAL"),
                             attributeIndex(attrs, "TANGENT"),  attributeIndex(attrs, "TEXCOORD_0"),
                             attributeIndex(attrs, "TEXCOORD_1"), attributeIndex(attrs, "COLOR_0")};
    int useImportedTangents = a.normal >= 0 && a.tangent >= 0;
    if (a.position < 0 || accessorCount(doc, (uint32_t)a.position) == 0) return 0;

    uint32_t materialIndex = VKRT_INVALID_INDEX;
    const JsonValue* matv = vkrtJsonGet(prim, "material");
    const JsonValue* material = NULL;
    if (matv) {
        if (!vkrtJsonToU32(matv, &materialIndex) || !(material = vkrtJsonAt(doc->materials, materialIndex))) return -1;
    }
    const JsonValue* ds = vkrtJsonGet(material, "doubleSided");
    out->renderBackfaces = ds && ds->type == JSON_BOOL && ds->boolean;
    out->materialIndex = materialIndex;
    uint32_t tangentSet = 0;
    const JsonValue* nt = vkrtJsonGet(material, "normalTexture");
    if (nt) {
        const JsonValue* tt = vkrtJsonGet(vkrtJsonGet(nt, "extensions"), "KHR_texture_transform");
        uint32_t set = 0;
        if (!(tt && vkrtJsonToU32(vkrtJsonGet(tt, "texCoord"), &set))) vkrtJsonToU32(vkrtJsonGet(nt, "texCoord"), &set);
        tangentSet = set > 1u ? 0u : set;
    }
    int tangentUv = tangentSet == 1u ? a.texcoord1 : a.texcoord0;

    size_t vertexCount = 0;
    float* pos = readAccessorFloats(doc, (uint32_t)a.position, 3, &vertexCount);
    if (!pos) {
        GLTF_LOG("glTF primitive has an unreadable POSITION accessor");
        return -1;
    }
    size_t indexCount = vertexCount;
    uint32_t* idx = NULL;
    const JsonValue* iv = vkrtJsonGet(prim, "indices");
    if (iv) {
        uint32_t ia = 0;
        if (!vkrtJsonToU32(iv, &ia) || !(idx = readAccessorIndices(doc, ia, &indexCount))) {
            free(pos);
            GLTF_LOG("glTF primitive has an unreadable index accessor");
            return -1;
        }
    }
    if (vertexCount > VKRT_INVALID_INDEX || indexCount == 0 || indexCount > VKRT_INVALID_INDEX ||
        vertexCount * sizeof(Vertex) + indexCount * sizeof(uint32_t) + vertexCount * 8u > kMaxPrimitiveBytes) {
        free(pos);
        free(idx);
        GLTF_LOG("glTF primitive with %zu vertices / %zu indices exceeds the import limits", vertexCount, indexCount);
        return -1;
    }

    float *nrm = NULL, *tan = NULL, *uv0 = NULL, *uv1 = NULL, *col = NULL;
    size_t n = 0;
    int colorComponents = a.color >= 0 ? accessorComponents(doc, (uint32_t)a.color) : 0;
    int ok = (a.normal < 0 || ((nrm = readAccessorFloats(doc, (uint32_t)a.normal, 3, &n)) && n >= vertexCount)) &&
             (!useImportedTangents ||
              ((tan = readAccessorFloats(doc, (uint32_t)a.tangent, 4, &n)) && n >= vertexCount)) &&
             (a.texcoord0 < 0 || ((uv0 = readAccessorFloats(doc, (uint32_t)a.texcoord0, 2, &n)) && n >= vertexCount)) &&
             (a.texcoord1 < 0 || ((uv1 = readAccessorFloats(doc, (uint32_t)a.texcoord1, 2, &n)) && n >= vertexCount)) &&
             (a.color < 0 || ((colorComponents == 3 || colorComponents == 4) &&
                              (col = readAccessorFloats(doc, (uint32_t)a.color, (uint32_t)colorComponents, &n)) &&
                              n >= vertexCount));
    Vertex* verts = ok ? (Vertex*)calloc(vertexCount, sizeof(Vertex)) : NULL;
    if (!idx && ok) {
        idx = (uint32_t*)malloc(sizeof(uint32_t) * indexCount);
        if (idx)
            for (size_t i = 0; i < indexCount; i++) idx[i] = (uint32_t)i;
    }
    if (!ok || !verts || !idx) {
        free(pos);
        free(idx);
        free(nrm);
        free(tan);
        free(uv0);
        free(uv1);
        free(col);
        free(verts);
        GLTF_LOG("glTF primitive has an unreadable vertex attribute");
        return -1;
    }
    for (size_t i = 0; i < indexCount; i++) {
        if (idx[i] >= vertexCount) {
            free(pos);
            free(idx);
            free(nrm);
            free(tan);
            free(uv0);
            free(uv1);
            free(col);
            free(verts);
            GLTF_LOG("glTF index %u out of range (%zu vertices)", idx[i], vertexCount);
            return -1;
        }
    }
    /* loader.c:1796-1832 */
    for (size_t i = 0; i < vertexCount; i++) {
        Vertex* v = &verts[i];
        const float* p = &pos[i * 3u];
        v->position[0] = p[0];
        v->position[1] = -p[2];
        v->position[2] = p[1];
        v->position[3] = 1.0f;
        if (nrm) {
            const float* q = &nrm[i * 3u];
            v->normal[0] = q[0];
            v->normal[1] = -q[2];
            v->normal[2] = q[1];
        }
        if (tan) {
            const float* q = &tan[i * 4u];
            v->tangent[0] = q[0];
            v->tangent[1] = -q[2];
            v->tangent[2] = q[1];
            v->tangent[3] = q[3];
        }
        v->color[0] = v->color[1] = v->color[2] = v->color[3] = 1.0f;
        if (col)
            for (int c = 0; c < colorComponents; c++) v->color[c] = col[i * (size_t)colorComponents + (size_t)c];
        if (uv0) {
            v->texcoord0[0] = uv0[i * 2u];
            v->texcoord0[1] = uv0[i * 2u + 1u];
        }
        if (uv1) {
            v->texcoord1[0] = uv1[i * 2u];
            v->texcoord1[1] = uv1[i * 2u + 1u];
        }
    }
    /* loader.c:1871-1892 */
    if (!nrm) {
        generateNormals(verts, vertexCount, idx, indexCount);
    } else {
        alignWindingToNormals(verts, vertexCount, idx, indexCount);
        if (useImportedTangents) {
            for (size_t i = 0; i < vertexCount; i++) {
                float handed = verts[i].tangent[3] < 0.0f ? -1.0f : 1.0f;
                float t[3] = {verts[i].tangent[0], verts[i].tangent[1], verts[i].tangent[2]};
                if (!orthonormalTangent(verts[i].normal, t, handed, verts[i].tangent))
                    fallbackTangent(verts[i].normal, verts[i].tangent);
            }
        } else if (tangentUv >= 0) {
            float* uv = (float*)malloc(sizeof(float) * 2u * vertexCount);
            if (uv) {
                for (size_t i = 0; i < vertexCount; i++) {
                    const float* s = tangentSet == 1u ? verts[i].texcoord1 : verts[i].texcoord0;
                    uv[i * 2u] = s[0];
                    uv[i * 2u + 1u] = s[1];
                }
                generateTangents(verts, vertexCount, idx, indexCount, uv);
                free(uv);
            } else {
                for (size_t i = 0; i < vertexCount; i++) fallbackTangent(verts[i].normal, verts[i].tangent);
            }
        } else {
            for (size_t i = 0; i < vertexCount; i++) fallbackTangent(verts[i].normal, verts[i].tangent);
        }
    }
    free(pos);
    free(nrm);
    free(tan);
    free(uv0);
    free(uv1);
    free(col);

    /* loader.c:1097-1118 */
    char name[256];
    const char* base = jsonName(node) ? jsonName(node) : (jsonName(mesh) ? jsonName(mesh) : "mesh");
    if (prims->count > 1)
        snprintf(name, sizeof(name), "%s_%u", base, primitiveIndex);
    else
        snprintf(name, sizeof(name), "%s", base);
    out->name = duplicateString(name);
    out->vertices = verts;
    out->indices = idx;
    out->vertexCount = vertexCount;
    out->indexCount = indexCount;
    out->nodeIndex = VKRT_INVALID_INDEX;
    out->scale[0] = out->scale[1] = out->scale[2] = 1.0f;
    if (!out->name) {
        releaseEntry(out);
        return -1;
    }
    return 1;
}

/* ---------------------------------------------------------------- nodes */

/* glTF 2.0 node.matrix, or T * R * S from translation / rotation (x, y, z, w) /
 * scale; column-major. */
static int nodeLocalMatrix(const JsonValue* node, float out[4][4]) {
    float m[16];
    if (vkrtJsonToFloats(vkrtJsonGet(node, "matrix"), m, 16)) {
        memcpy(out, m, sizeof(m));
        return 1;
    }
    static const float zero3[3] = {0.0f, 0.0f, 0.0f};
    static const float one3[3] = {1.0f, 1.0f, 1.0f};
    static const float unitQ[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    float t[3], q[4], s[3];
    readFloatsOr(node, "translation", t, 3, zero3);
    readFloatsOr(node, "rotation", q, 4, unitQ);
    readFloatsOr(node, "scale", s, 3, one3);
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float r[3][3] = {
        {1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w)},
        {2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w)},
        {2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y)},
    }; /* r[column][row] */
    for (int c = 0; c < 3; c++) {
        for (int rr = 0; rr < 3; rr++) out[c][rr] = r[c][rr] * s[c];
        out[c][3] = 0.0f;
    }
    out[3][0] = t[0];
    out[3][1] = t[1];
    out[3][2] = t[2];
    out[3][3] = 1.0f;
    return 1;
}

static void releaseNode(NodeImportEntry* n) { free(n->name); }

static int appendEntry(MeshImportData* d, const MeshImportEntry* e) {
    if (d->count == d->entryCapacity) {
        uint32_t cap = d->entryCapacity ? d->entryCapacity * 2u : 8u;
        MeshImportEntry* grown = (MeshImportEntry*)realloc(d->entries, sizeof(MeshImportEntry) * cap);
        if (!grown) return -1;
        d->entries = grown;
        d->entryCapacity = cap;
    }
    d->entries[d->count++] = *e;
    return 0;
}

static int appendNode(MeshImportData* d, const NodeImportEntry* n, uint32_t* outIndex) {
    if (d->nodeCount == d->nodeCapacity) {
        uint32_t cap = d->nodeCapacity ? d->nodeCapacity * 2u : 8u;
        NodeImportEntry* grown = (NodeImportEntry*)realloc(d->nodes, sizeof(NodeImportEntry) * cap);
        if (!grown) return -1;
        d->nodes = grown;
        d->nodeCapacity = cap;
    }
    *outIndex = d->nodeCount;
    d->nodes[d->nodeCount++] = *n;
    return 0;
}

/* loader.c:2027-2071: depth-first, roots and children in file order. */
static int collectNodes(const GltfDoc* doc, const uint32_t* roots, uint32_t rootCount, MeshImportData* d) {
    uint32_t nodeTotal = doc->nodes ? doc->nodes->count : 0u;
    typedef struct StackItem {
        uint32_t node;
        uint32_t parent;
    } StackItem;
    /* each node is visited at most once (checked below), so the stack never
     * holds more than every node once plus the pending roots */
    StackItem* stack = (StackItem*)malloc(sizeof(StackItem) * ((size_t)nodeTotal + rootCount + 1u));
    uint8_t* visited = (uint8_t*)calloc(nodeTotal + 1u, 1);
    if (!stack || !visited) {
        free(stack);
        free(visited);
        return -1;
    }
    size_t top = 0;
    for (uint32_t r = rootCount; r > 0; r--) stack[top++] = (StackItem){roots[r - 1u], VKRT_INVALID_INDEX};
    int result = 0;
    while (top > 0 && result == 0) {
        StackItem it = stack[--top];
        const JsonValue* node = vkrtJsonAt(doc->nodes, it.node);
        if (!node || visited[it.node]) {
            GLTF_LOG("glTF node %u is missing or reached twice (cycle or shared child)", it.node);
            result = -1;
            break;
        }
        visited[it.node] = 1;
        NodeImportEntry entry;
        memset(&entry, 0, sizeof(entry));
        entry.parentIndex = it.parent;
        float raw[4][4];
        nodeLocalMatrix(node, raw);
        VKRT_buildImportedNodeTransform(raw, entry.localTransform);
        VKRT_decomposeMeshTransform(entry.localTransform, entry.position, entry.rotation, entry.scale);
        if (jsonName(node) && !(entry.name = duplicateString(jsonName(node)))) {
            result = -1;
            break;
        }
        uint32_t nodeIndex = 0;
        if (appendNode(d, &entry, &nodeIndex) != 0) {
            releaseNode(&entry);
            result = -1;
            break;
        }
        uint32_t meshIndex = 0;
        const JsonValue* mv = vkrtJsonGet(node, "mesh");
        if (mv) {
            const JsonValue* mesh = NULL;
            if (!vkrtJsonToU32(mv, &meshIndex) || !(mesh = vkrtJsonAt(doc->meshes, meshIndex))) {
                result = -1;
                break;
            }
            const JsonValue* prims = vkrtJsonGet(mesh, "primitives");
            uint32_t primCount = prims && prims->type == JSON_ARRAY ? prims->count : 0u;
            for (uint32_t p = 0; p < primCount && result == 0; p++) {
                MeshImportEntry e;
                int built = buildPrimitive(doc, node, mesh, p, &e);
                if (built < 0) {
                    result = -1;
                } else if (built > 0) {
                    e.nodeIndex = nodeIndex;
                    if (appendEntry(d, &e) != 0) {
                        releaseEntry(&e);
                        result = -1;
                    } else {
                        d->nodes[nodeIndex].meshEntryCount++;
                    }
                }
            }
        }
        const JsonValue* children = vkrtJsonGet(node, "children");
        if (result == 0 && children) {
            if (children->type != JSON_ARRAY || top + children->count > (size_t)nodeTotal + rootCount) {
                result = -1;
                break;
            }
            for (uint32_t c = children->count; c > 0; c--) {
                uint32_t child = 0;
                if (!vkrtJsonToU32(&children->items[c - 1u], &child) || child >= nodeTotal) {
                    result = -1;
                    break;
                }
                stack[top++] = (StackItem){child, nodeIndex};
            }
        }
    }
    free(stack);
    free(visited);
    return result;
}

/* loader.c:2087-2107: the default scene's roots, else every parentless node. */
static int collectRoots(const GltfDoc* doc, MeshImportData* d) {
    uint32_t nodeTotal = doc->nodes ? doc->nodes->count : 0u;
    const JsonValue* scenes = vkrtJsonGet(&doc->root, "scenes");
    uint32_t sceneIndex = 0;
    const JsonValue* sv = vkrtJsonGet(&doc->root, "scene");
    const JsonValue* scene = NULL;
    if (sv) {
        if (!vkrtJsonToU32(sv, &sceneIndex) || !(scene = vkrtJsonAt(scenes, sceneIndex))) return -1;
    }
    const JsonValue* sceneNodes = vkrtJsonGet(scene, "nodes");
    uint32_t* roots = (uint32_t*)malloc(sizeof(uint32_t) * (nodeTotal + 1u));
    if (!roots) return -1;
    uint32_t rootCount = 0;
    if (sceneNodes && sceneNodes->type == JSON_ARRAY && sceneNodes->count > 0) {
        if (sceneNodes->count > nodeTotal) {
            free(roots);
            return -1;
        }
        for (uint32_t i = 0; i < sceneNodes->count; i++) {
            if (!vkrtJsonToU32(&sceneNodes->items[i], &roots[rootCount]) || roots[rootCount] >= nodeTotal) {
                free(roots);
                return -1;
            }
            rootCount++;
        }
    } else {
        uint8_t* isChild = (uint8_t*)calloc(nodeTotal + 1u, 1);
        if (!isChild) {
            free(roots);
            return -1;
        }
        for (uint32_t i = 0; i < nodeTotal; i++) {
            const JsonValue* children = vkrtJsonGet(&doc->nodes->items[i], "children");
            for (uint32_t c = 0; children && children->type == JSON_ARRAY && c < children->count; c++) {
                uint32_t child = 0;
                if (vkrtJsonToU32(&children->items[c], &child) && child < nodeTotal) isChild[child] = 1;
            }
        }
        for (uint32_t i = 0; i < nodeTotal; i++)
            if (!isChild[i]) roots[rootCount++] = i;
        free(isChild);
    }
    int r = collectNodes(doc, roots, rootCount, d);
    free(roots);
    return r;
}

/* ------------------------------------------------------------ container */

static uint32_t readU32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

static int parseContainer(GltfDoc* doc) {
    const char* jsonText = (const char*)doc->file;
    size_t jsonLength = doc->fileSize;
    if (doc->fileSize >= 12 && readU32(doc->file) == kGlbMagic) {
        uint32_t version = readU32(doc->file + 4);
        uint32_t total = readU32(doc->file + 8);
        if (version != 2 || total > doc->fileSize || total < 20) {
            GLTF_LOG("unsupported GLB header (version %u, length %u)", version, total);
            return -1;
        }
        size_t off = 12;
        uint32_t chunk = 0;
        jsonText = NULL;
        while (off + 8 <= total) {
            uint32_t len = readU32(doc->file + off);
            uint32_t type = readU32(doc->file + off + 4);
            off += 8;
            if (len > total - off) {
                GLTF_LOG("GLB chunk %u overruns the file", chunk);
                return -1;
            }
            if (chunk == 0) {
                if (type != kGlbChunkJson) return -1;
                jsonText = (const char*)doc->file + off;
                jsonLength = len;
            } else if (chunk == 1 && type == kGlbChunkBin) {
                doc->bin = doc->file + off;
                doc->binSize = len;
            }
            off += len;
            chunk++;
        }
        if (!jsonText) return -1;
    }
    char err[256];
    if (vkrtJsonParse(jsonText, jsonLength, &doc->root, err, sizeof(err)) != 0) {
        GLTF_LOG("glTF JSON: %s", err);
        return -1;
    }
    const JsonValue* asset = vkrtJsonGet(&doc->root, "asset");
    const JsonValue* version = vkrtJsonGet(asset, "version");
    if (!version || version->type != JSON_STRING || version->string[0] != '2') {
        GLTF_LOG("glTF asset.version must be 2.x");
        return -1;
    }
    doc->accessors = vkrtJsonGet(&doc->root, "accessors");
    doc->bufferViews = vkrtJsonGet(&doc->root, "bufferViews");
    doc->materials = vkrtJsonGet(&doc->root, "materials");
    doc->meshes = vkrtJsonGet(&doc->root, "meshes");
    doc->nodes = vkrtJsonGet(&doc->root, "nodes");
    return loadBuffers(doc);
}

static void releaseDoc(GltfDoc* doc) {
    for (uint32_t i = 0; i < doc->bufferCount; i++) free(doc->buffers[i].owned);
    free(doc->buffers);
    vkrtJsonFree(&doc->root);
    free(doc->file);
}

void meshReleaseImportData(MeshImportData* d) {
    if (!d) return;
    for (uint32_t i = 0; i < d->count; i++) releaseEntry(&d->entries[i]);
    for (uint32_t i = 0; i < d->nodeCount; i++) releaseNode(&d->nodes[i]);
    for (uint32_t i = 0; i < d->materialCount; i++) free(d->materials[i].name);
    for (uint32_t i = 0; i < d->textureCount; i++) {
        free(d->textures[i].name);
        free(d->textures[i].pixels);
    }
    free(d->entries);
    free(d->nodes);
    free(d->materials);
    free(d->textures);
    memset(d, 0, sizeof(*d));
}

int meshLoadFromFile(const char* filePath, MeshImportData* out) {
    if (!filePath || !filePath[0] || !out) return -1;
    memset(out, 0, sizeof(*out));
    GltfDoc* doc = (GltfDoc*)calloc(1, sizeof(GltfDoc));
    if (!doc) return -1;
    doc->file = readWholeFile(filePath, &doc->fileSize);
    if (!doc->file) {
        GLTF_LOG("Mesh file not found: %s", filePath);
        free(doc);
        return -1;
    }
    directoryOf(filePath, doc->directory, sizeof(doc->directory));
    int result = parseContainer(doc);
    uint32_t textureRefs = 0;
    /* (image, colour space) of each imported texture; at most four per material */
    uint64_t* keys = NULL;
    if (result == 0 && doc->materials && doc->materials->type == JSON_ARRAY && doc->materials->count > 0) {
        keys = (uint64_t*)calloc((size_t)doc->materials->count * 4u, sizeof(uint64_t));
        out->materials = (MaterialImportEntry*)calloc(doc->materials->count, sizeof(MaterialImportEntry));
        if (!out->materials) result = -1;
        for (uint32_t i = 0; result == 0 && i < doc->materials->count; i++) {
            const JsonValue* src = &doc->materials->items[i];
            char name[VKRT_NAME_LEN];
            if (jsonName(src))
                snprintf(name, sizeof(name), "%s", jsonName(src));
            else
                snprintf(name, sizeof(name), "Material %u", i);
            out->materials[i].material = buildMaterial(src, &textureRefs);
            if (keys) attachMaterialTextures(doc, out, keys, &out->materials[i].material, src);
            out->materials[i].name = duplicateString(name);
            out->materialCount = i + 1u;
            if (!out->materials[i].name) result = -1;
        }
    }
    if (result == 0 && doc->textureFailed) result = -1;
    if (result == 0) result = collectRoots(doc, out);
    if (result == 0 && out->count == 0) {
        GLTF_LOG("No triangle mesh primitives were found in '%s'", filePath);
        result = -1;
    }
    if (result == 0 && textureRefs > out->textureCount)
        GLTF_LOG("'%s': %u material texture reference(s), %u texture(s) imported", filePath, textureRefs,
                 out->textureCount);
    free(keys);
    releaseDoc(doc);
    free(doc);
    if (result != 0) {
        meshReleaseImportData(out);
        GLTF_LOG("Failed to import '%s'", filePath);
        return -1;
    }
    return 0;
}
