/*
 * Scene objects: meshes, materials, textures and transforms.
 * Behaviour follows src/core/api/mesh.c and src/core/scene/geometry.c:
 *  - the first upload or material creates the default material at index 0
 *    (src/core/internal/state.c:42-54);
 *  - vertex / index arrays are copied (geometry.c:411-429);
 *  - material fields are clamped rather than rejected (mesh.c:107-195);
 *  - every change restarts accumulation (resetSceneData).
 */
#include "internal.h"
#include "image_io.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static float finiteClamp(float v, float fallback, float lo, float hi) {
    if (!isfinite(v)) {
        if (isinf(v) && v > 0.0f && isinf(hi)) return v;
        return fallback;
    }
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

static Material sanitizeMaterial(const VKRT* vkrt, Material m) {
    for (int c = 0; c < 3; c++) {
        m.baseColor[c] = finiteClamp(m.baseColor[c], 0.0f, 0.0f, 1.0f);
        m.emissionColor[c] = finiteClamp(m.emissionColor[c], 0.0f, 0.0f, INFINITY);
        m.sheenTintWeight[c] = finiteClamp(m.sheenTintWeight[c], 0.0f, 0.0f, 1.0f);
        m.attenuationColor[c] = finiteClamp(m.attenuationColor[c], 1.0f, 0.0f, 1.0f);
        m.eta[c] = finiteClamp(m.eta[c], 0.0f, 0.0f, INFINITY);
        m.k[c] = finiteClamp(m.k[c], 0.0f, 0.0f, INFINITY);
    }
    m.metallic = finiteClamp(m.metallic, 0.0f, 0.0f, 1.0f);
    m.roughness = finiteClamp(m.roughness, 0.0f, 0.0f, 1.0f);
    m.diffuseRoughness = finiteClamp(m.diffuseRoughness, 0.0f, 0.0f, 1.0f);
    m.specular = finiteClamp(m.specular, 0.0f, 0.0f, 1.0f);
    m.specularTint = finiteClamp(m.specularTint, 0.0f, 0.0f, 1.0f);
    m.anisotropic = finiteClamp(m.anisotropic, 0.0f, 0.0f, 1.0f);
    m.sheenTintWeight[3] = finiteClamp(m.sheenTintWeight[3], 0.0f, 0.0f, 1.0f);
    m.clearcoat = finiteClamp(m.clearcoat, 0.0f, 0.0f, 1.0f);
    m.clearcoatGloss = finiteClamp(m.clearcoatGloss, 0.0f, 0.0f, 1.0f);
    m.ior = finiteClamp(m.ior, 1.0f, 1.0f, 4.0f);
    m.abbeNumber = finiteClamp(m.abbeNumber, 0.0f, 0.0f, 200.0f);
    m.transmission = finiteClamp(m.transmission, 0.0f, 0.0f, 1.0f);
    m.subsurface = finiteClamp(m.subsurface, 0.0f, 0.0f, 1.0f);
    m.sheenRoughness = finiteClamp(m.sheenRoughness, 0.0f, 0.0f, 1.0f);
    m.absorptionCoefficient = finiteClamp(m.absorptionCoefficient, 0.0f, 0.0f, VKRT_MAX_ABSORPTION_COEFFICIENT);
    m.emissionLuminance = finiteClamp(m.emissionLuminance, 0.0f, 0.0f, INFINITY);
    if (!isfinite(m.normalTextureScale)) m.normalTextureScale = 1.0f;
    if (m.normalTextureScale < 0.0f) m.normalTextureScale = 0.0f;
    m.opacity = finiteClamp(m.opacity, 1.0f, 0.0f, 1.0f);
    m.alphaCutoff = finiteClamp(m.alphaCutoff, 0.5f, 0.0f, 1.0f);
    if (m.alphaMode != VKRT_MATERIAL_ALPHA_MODE_MASK && m.alphaMode != VKRT_MATERIAL_ALPHA_MODE_BLEND) {
        m.alphaMode = VKRT_MATERIAL_ALPHA_MODE_OPAQUE;
    }
    uint32_t* slots[4] = {&m.baseColorTextureIndex, &m.metallicRoughnessTextureIndex, &m.normalTextureIndex,
                          &m.emissiveTextureIndex};
    for (int i = 0; i < 4; i++) {
        if (*slots[i] != VKRT_INVALID_INDEX && *slots[i] >= vkrt->textureCount) *slots[i] = VKRT_INVALID_INDEX;
    }
    return m;
}

static VKRT_Result ensureDefaultMaterial(VKRT* vkrt) {
    if (vkrt->materialCount > 0) return VKRT_SUCCESS;
    vkrt->materials = (HostMaterial*)calloc(4, sizeof(HostMaterial));
    if (!vkrt->materials) return VKRT_ERROR_OUT_OF_MEMORY;
    vkrt->materialCapacity = 4;
    vkrt->materials[0].material = VKRT_materialDefault();
    snprintf(vkrt->materials[0].name, VKRT_NAME_LEN, "%s", "Default");
    vkrt->materialCount = 1;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_uploadMeshDataBatch(VKRT* vkrt, const VKRT_MeshUpload* uploads, size_t uploadCount) {
    if (!vkrt || (!uploads && uploadCount)) return VKRT_ERROR_INVALID_ARGUMENT;
    for (size_t i = 0; i < uploadCount; i++) {
        const VKRT_MeshUpload* u = &uploads[i];
        if (!u->vertices || !u->indices || u->vertexCount == 0 || u->indexCount == 0 || u->indexCount % 3u != 0)
            return VKRT_ERROR_INVALID_ARGUMENT;
        if (u->vertexCount > UINT32_MAX || u->indexCount > UINT32_MAX) return VKRT_ERROR_INVALID_ARGUMENT;
        /* every index must address a vertex of its own mesh: the kernels fetch vertices unchecked */
        for (size_t k = 0; k < u->indexCount; k++) {
            if (u->indices[k] >= u->vertexCount) return VKRT_ERROR_INVALID_ARGUMENT;
        }
    }
    VKRT_Result r = ensureDefaultMaterial(vkrt);
    if (r != VKRT_SUCCESS) return r;
    uint32_t needed = vkrt->meshCount + (uint32_t)uploadCount;
    if (needed > vkrt->meshCapacity) {
        uint32_t cap = vkrt->meshCapacity ? vkrt->meshCapacity : 8u;
        while (cap < needed) cap *= 2u;
        HostMesh* grown = (HostMesh*)realloc(vkrt->meshes, sizeof(HostMesh) * cap);
        if (!grown) return VKRT_ERROR_OUT_OF_MEMORY;
        vkrt->meshes = grown;
        vkrt->meshCapacity = cap;
    }
    for (size_t i = 0; i < uploadCount; i++) {
        const VKRT_MeshUpload* u = &uploads[i];
        HostMesh* mesh = &vkrt->meshes[vkrt->meshCount];
        memset(mesh, 0, sizeof(*mesh));
        mesh->vertices = (Vertex*)malloc(sizeof(Vertex) * u->vertexCount);
        mesh->indices = (uint32_t*)malloc(sizeof(uint32_t) * u->indexCount);
        if (!mesh->vertices || !mesh->indices) {
            free(mesh->vertices);
            free(mesh->indices);
            return VKRT_ERROR_OUT_OF_MEMORY;
        }
        memcpy(mesh->vertices, u->vertices, sizeof(Vertex) * u->vertexCount);
        memcpy(mesh->indices, u->indices, sizeof(uint32_t) * u->indexCount);
        for (int c = 0; c < 4; c++)
            for (int rr = 0; rr < 4; rr++) mesh->worldTransform[c][rr] = c == rr ? 1.0f : 0.0f;
        mesh->info.vertexCount = (uint32_t)u->vertexCount;
        mesh->info.indexCount = (uint32_t)u->indexCount;
        mesh->info.materialIndex = 0u;
        mesh->info.opacity = 1.0f;
        mesh->info.scale[0] = mesh->info.scale[1] = mesh->info.scale[2] = 1.0f;
        snprintf(mesh->name, VKRT_NAME_LEN, "Mesh %u", vkrt->meshCount);
        vkrt->meshCount++;
    }
    vkrt->geometryRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_uploadMeshData(VKRT* vkrt, const Vertex* vertices, size_t vertexCount, const uint32_t* indices,
                                size_t indexCount) {
    VKRT_MeshUpload u = {vertices, vertexCount, indices, indexCount};
    return VKRT_uploadMeshDataBatch(vkrt, &u, 1);
}

VKRT_Result VKRT_removeMesh(VKRT* vkrt, uint32_t meshIndex) {
    if (!vkrt || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    free(vkrt->meshes[meshIndex].vertices);
    free(vkrt->meshes[meshIndex].indices);
    memmove(&vkrt->meshes[meshIndex], &vkrt->meshes[meshIndex + 1],
            sizeof(HostMesh) * (vkrt->meshCount - meshIndex - 1u));
    vkrt->meshCount--;
    vkrt->geometryRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getMeshCount(const VKRT* vkrt, uint32_t* outMeshCount) {
    if (!vkrt || !outMeshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    *outMeshCount = vkrt->meshCount;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getMeshSnapshot(const VKRT* vkrt, uint32_t meshIndex, VKRT_MeshSnapshot* out) {
    if (!vkrt || !out || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    memset(out, 0, sizeof(*out));
    const HostMesh* mesh = &vkrt->meshes[meshIndex];
    out->info = mesh->info;
    if (mesh->info.materialIndex < vkrt->materialCount) out->material = vkrt->materials[mesh->info.materialIndex].material;
    out->materialIndex = mesh->info.materialIndex;
    out->geometrySource = meshIndex;
    out->hasMaterialAssignment = mesh->hasMaterialAssignment;
    out->ownsGeometry = 1u;
    snprintf(out->name, VKRT_NAME_LEN, "%s", mesh->name);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setMeshName(VKRT* vkrt, uint32_t meshIndex, const char* name) {
    if (!vkrt || !name || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    snprintf(vkrt->meshes[meshIndex].name, VKRT_NAME_LEN, "%s", name);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getMaterialCount(const VKRT* vkrt, uint32_t* outMaterialCount) {
    if (!vkrt || !outMaterialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    *outMaterialCount = vkrt->materialCount;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getMaterialSnapshot(const VKRT* vkrt, uint32_t materialIndex, VKRT_MaterialSnapshot* out) {
    if (!vkrt || !out || materialIndex >= vkrt->materialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    memset(out, 0, sizeof(*out));
    out->material = vkrt->materials[materialIndex].material;
    for (uint32_t i = 0; i < vkrt->meshCount; i++) out->useCount += vkrt->meshes[i].info.materialIndex == materialIndex;
    snprintf(out->name, VKRT_NAME_LEN, "%s", vkrt->materials[materialIndex].name);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_addMaterial(VKRT* vkrt, const Material* material, const char* name, uint32_t* outMaterialIndex) {
    if (outMaterialIndex) *outMaterialIndex = VKRT_INVALID_INDEX;
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = ensureDefaultMaterial(vkrt);
    if (r != VKRT_SUCCESS) return r;
    if (vkrt->materialCount == vkrt->materialCapacity) {
        uint32_t cap = vkrt->materialCapacity * 2u;
        HostMaterial* grown = (HostMaterial*)realloc(vkrt->materials, sizeof(HostMaterial) * cap);
        if (!grown) return VKRT_ERROR_OUT_OF_MEMORY;
        vkrt->materials = grown;
        vkrt->materialCapacity = cap;
    }
    uint32_t index = vkrt->materialCount++;
    HostMaterial* hm = &vkrt->materials[index];
    memset(hm, 0, sizeof(*hm));
    hm->material = sanitizeMaterial(vkrt, material ? *material : VKRT_materialDefault());
    if (name && name[0])
        snprintf(hm->name, VKRT_NAME_LEN, "%s", name);
    else
        snprintf(hm->name, VKRT_NAME_LEN, "Material %u", index);
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    if (outMaterialIndex) *outMaterialIndex = index;
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setMaterial(VKRT* vkrt, uint32_t materialIndex, const Material* material) {
    if (!vkrt || !material || materialIndex >= vkrt->materialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    vkrt->materials[materialIndex].material = sanitizeMaterial(vkrt, *material);
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setMeshMaterialIndex(VKRT* vkrt, uint32_t meshIndex, uint32_t materialIndex) {
    if (!vkrt || meshIndex >= vkrt->meshCount || materialIndex >= vkrt->materialCount)
        return VKRT_ERROR_INVALID_ARGUMENT;
    HostMesh* mesh = &vkrt->meshes[meshIndex];
    if (mesh->info.materialIndex == materialIndex && mesh->hasMaterialAssignment) return VKRT_SUCCESS;
    mesh->info.materialIndex = materialIndex;
    mesh->hasMaterialAssignment = 1u;
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

static int matrixValid(float m[4][4]) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            if (!isfinite(m[c][r])) return 0;
    return 1;
}

VKRT_Result VKRT_setMeshTransformMatrix(VKRT* vkrt, uint32_t meshIndex, float worldTransform[4][4]) {
    if (!vkrt || !worldTransform || meshIndex >= vkrt->meshCount || !matrixValid(worldTransform))
        return VKRT_ERROR_INVALID_ARGUMENT;
    HostMesh* mesh = &vkrt->meshes[meshIndex];
    if (memcmp(mesh->worldTransform, worldTransform, sizeof(mesh->worldTransform)) == 0) return VKRT_SUCCESS;
    memcpy(mesh->worldTransform, worldTransform, sizeof(mesh->worldTransform));
    VKRT_decomposeMeshTransform(mesh->worldTransform, mesh->info.position, mesh->info.rotation, mesh->info.scale);
    vkrt->geometryRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_setMeshTransform(VKRT* vkrt, uint32_t meshIndex, const float position[3], const float rotation[3],
                                  const float scale[3]) {
    if (!vkrt || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    const MeshInfo* info = &vkrt->meshes[meshIndex].info;
    float p[3], r[3], s[3];
    for (int i = 0; i < 3; i++) {
        p[i] = position && isfinite(position[i]) ? position[i] : info->position[i];
        r[i] = rotation && isfinite(rotation[i]) ? rotation[i] : info->rotation[i];
        s[i] = scale && isfinite(scale[i]) ? scale[i] : info->scale[i];
    }
    float m[4][4];
    VKRT_buildMeshTransformMatrix(p, r, s, m);
    return VKRT_setMeshTransformMatrix(vkrt, meshIndex, m);
}

static float halfToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    uint32_t bits;
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

static float srgbToLinear(float v) { return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f); }

VKRT_Result VKRT_addTextureFromPixels(VKRT* vkrt, const VKRT_TextureUpload* upload, uint32_t* outTextureIndex) {
    if (outTextureIndex) *outTextureIndex = VKRT_INVALID_INDEX;
    if (!vkrt || !upload || !upload->pixels || upload->width == 0 || upload->height == 0 ||
        upload->format >= VKRT_TEXTURE_FORMAT_COUNT || upload->colorSpace > VKRT_TEXTURE_COLOR_SPACE_LINEAR)
        return VKRT_ERROR_INVALID_ARGUMENT;
    if (vkrt->textureCount == vkrt->textureCapacity) {
        uint32_t cap = vkrt->textureCapacity ? vkrt->textureCapacity * 2u : 4u;
        HostTexture* grown = (HostTexture*)realloc(vkrt->textures, sizeof(HostTexture) * cap);
        if (!grown) return VKRT_ERROR_OUT_OF_MEMORY;
        vkrt->textures = grown;
        vkrt->textureCapacity = cap;
    }
    size_t n = (size_t)upload->width * upload->height * 4u;
    float* rgba = (float*)malloc(sizeof(float) * n);
    if (!rgba) return VKRT_ERROR_OUT_OF_MEMORY;
    for (size_t i = 0; i < n; i++) {
        float v;
        switch (upload->format) {
            case VKRT_TEXTURE_FORMAT_RGBA8_UNORM: v = (float)((const uint8_t*)upload->pixels)[i] / 255.0f; break;
            case VKRT_TEXTURE_FORMAT_RGBA16_UNORM: v = (float)((const uint16_t*)upload->pixels)[i] / 65535.0f; break;
            case VKRT_TEXTURE_FORMAT_RGBA16_SFLOAT: v = halfToFloat(((const uint16_t*)upload->pixels)[i]); break;
            default: v = ((const float*)upload->pixels)[i]; break;
        }
        if (upload->colorSpace == VKRT_TEXTURE_COLOR_SPACE_SRGB && (i & 3u) != 3u) v = srgbToLinear(v);
        rgba[i] = v;
    }
    HostTexture* t = &vkrt->textures[vkrt->textureCount];
    memset(t, 0, sizeof(*t));
    t->rgba = rgba;
    t->width = upload->width;
    t->height = upload->height;
    t->format = upload->format;
    t->colorSpace = upload->colorSpace;
    snprintf(t->name, VKRT_NAME_LEN, "%s", upload->name ? upload->name : "Texture");
    if (outTextureIndex) *outTextureIndex = vkrt->textureCount;
    vkrt->textureCount++;
    return VKRT_SUCCESS;
}

/* src/core/scene/textures.c:374-381: material slots that name the texture,
 * plus one when it is the environment map. */
static uint32_t textureUsers(const VKRT* vkrt, uint32_t textureIndex) {
    uint32_t users = vkrt->settings.environmentTextureIndex == textureIndex ? 1u : 0u;
    for (uint32_t m = 0; m < vkrt->materialCount; m++) {
        const Material* mat = &vkrt->materials[m].material;
        users += (mat->baseColorTextureIndex == textureIndex) + (mat->metallicRoughnessTextureIndex == textureIndex) +
                 (mat->normalTextureIndex == textureIndex) + (mat->emissiveTextureIndex == textureIndex);
    }
    return users;
}

/* src/core/api/query.c:110-124 */
VKRT_Result VKRT_getTextureSnapshot(const VKRT* vkrt, uint32_t textureIndex, VKRT_TextureSnapshot* out) {
    if (!vkrt || !out || textureIndex >= vkrt->textureCount) return VKRT_ERROR_INVALID_ARGUMENT;
    const HostTexture* t = &vkrt->textures[textureIndex];
    out->width = t->width;
    out->height = t->height;
    out->format = t->format;
    out->colorSpace = t->colorSpace;
    out->useCount = textureUsers(vkrt, textureIndex);
    snprintf(out->name, sizeof(out->name), "%s", t->name);
    return VKRT_SUCCESS;
}

/* src/core/api/texture.c:16-28 and src/core/scene/textures.c:484-515: the file
 * must exist; the texture is named after the file unless a name is given. */
VKRT_Result VKRT_addTextureFromFile(VKRT* vkrt, const char* path, const char* name, uint32_t colorSpace,
                                    uint32_t* outTextureIndex) {
    if (outTextureIndex) *outTextureIndex = VKRT_INVALID_INDEX;
    if (!vkrt || !path || !path[0] || colorSpace > VKRT_TEXTURE_COLOR_SPACE_LINEAR) return VKRT_ERROR_INVALID_ARGUMENT;
    FILE* probe = fopen(path, "rb");
    if (!probe) return VKRT_ERROR_INVALID_ARGUMENT;
    fclose(probe);
    VkrtLoadedImage image;
    if (vkrtLoadImageFromFile(path, colorSpace, &image) != 0) return VKRT_ERROR_OPERATION_FAILED;
    const char* base = strrchr(path, '/');
    VKRT_TextureUpload upload;
    memset(&upload, 0, sizeof(upload));
    upload.name = name && name[0] ? name : (base ? base + 1 : path);
    upload.pixels = image.pixels;
    upload.width = image.width;
    upload.height = image.height;
    upload.format = image.format;
    upload.colorSpace = image.colorSpace;
    VKRT_Result r = VKRT_addTextureFromPixels(vkrt, &upload, outTextureIndex);
    free(image.pixels);
    return r;
}

/* src/core/api/texture.c:48-53 and src/core/scene/textures.c:580-622: colour
 * slots (base colour, emissive) take sRGB or linear textures, the others
 * linear only; binding resets the slot's wrap, coordinate set, transform and
 * rotation, and a texture left without users is released. */
VKRT_Result VKRT_setMaterialTexture(VKRT* vkrt, uint32_t materialIndex, uint32_t textureSlot, uint32_t textureIndex) {
    if (!vkrt || textureSlot >= VKRT_MATERIAL_TEXTURE_SLOT_COUNT || materialIndex >= vkrt->materialCount)
        return VKRT_ERROR_INVALID_ARGUMENT;
    Material* m = &vkrt->materials[materialIndex].material;
    uint32_t* index[4] = {&m->baseColorTextureIndex, &m->metallicRoughnessTextureIndex, &m->normalTextureIndex,
                          &m->emissiveTextureIndex};
    uint32_t* wrap[4] = {&m->baseColorTextureWrap, &m->metallicRoughnessTextureWrap, &m->normalTextureWrap,
                         &m->emissiveTextureWrap};
    float* xf[4] = {m->baseColorTextureTransform, m->metallicRoughnessTextureTransform, m->normalTextureTransform,
                    m->emissiveTextureTransform};
    if (textureIndex != VKRT_INVALID_INDEX) {
        if (textureIndex >= vkrt->textureCount) return VKRT_ERROR_INVALID_ARGUMENT;
        const int colourSlot =
            textureSlot == VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR || textureSlot == VKRT_MATERIAL_TEXTURE_SLOT_EMISSIVE;
        if (!colourSlot && vkrt->textures[textureIndex].colorSpace != VKRT_TEXTURE_COLOR_SPACE_LINEAR)
            return VKRT_ERROR_INVALID_ARGUMENT;
    }
    if (*index[textureSlot] == textureIndex) return VKRT_SUCCESS;
    const uint32_t previous = *index[textureSlot];
    *index[textureSlot] = textureIndex;
    *wrap[textureSlot] = VKRT_TEXTURE_WRAP_DEFAULT;
    m->textureTexcoordSets &= ~(0xffu << (textureSlot * 8u));
    xf[textureSlot][0] = xf[textureSlot][1] = 1.0f;
    xf[textureSlot][2] = xf[textureSlot][3] = 0.0f;
    m->textureRotations[textureSlot] = 0.0f;
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    if (previous != VKRT_INVALID_INDEX && previous < vkrt->textureCount && textureUsers(vkrt, previous) == 0u)
        return VKRT_removeTexture(vkrt, previous);
    return VKRT_SUCCESS;
}

VKRT_Result VKRT_getTextureCount(const VKRT* vkrt, uint32_t* outTextureCount) {
    if (!vkrt || !outTextureCount) return VKRT_ERROR_INVALID_ARGUMENT;
    *outTextureCount = vkrt->textureCount;
    return VKRT_SUCCESS;
}

/* src/core/scene/textures.c (vkrtSceneRemoveTexture): removing the environment's
 * texture first clears the environment; a texture still bound to a material slot
 * cannot be removed; later textures move down one index, in the material slots
 * and in the environment setting alike. */
VKRT_Result VKRT_removeTexture(VKRT* vkrt, uint32_t textureIndex) {
    if (!vkrt || textureIndex >= vkrt->textureCount) return VKRT_ERROR_INVALID_ARGUMENT;
    for (uint32_t m = 0; m < vkrt->materialCount; m++) {
        const Material* mat = &vkrt->materials[m].material;
        if (mat->baseColorTextureIndex == textureIndex || mat->metallicRoughnessTextureIndex == textureIndex ||
            mat->normalTextureIndex == textureIndex || mat->emissiveTextureIndex == textureIndex)
            return VKRT_ERROR_OPERATION_FAILED;
    }
    if (vkrt->settings.environmentTextureIndex == textureIndex) {
        VKRT_Result r = VKRT_setEnvironmentTexture(vkrt, VKRT_INVALID_INDEX);
        if (r != VKRT_SUCCESS) return r;
    }
    free(vkrt->textures[textureIndex].rgba);
    memmove(&vkrt->textures[textureIndex], &vkrt->textures[textureIndex + 1u],
            sizeof(HostTexture) * (vkrt->textureCount - textureIndex - 1u));
    vkrt->textureCount--;
    uint8_t materialsMoved = 0;
    for (uint32_t m = 0; m < vkrt->materialCount; m++) {
        Material* mat = &vkrt->materials[m].material;
        uint32_t* slots[4] = {&mat->baseColorTextureIndex, &mat->metallicRoughnessTextureIndex, &mat->normalTextureIndex,
                              &mat->emissiveTextureIndex};
        for (int k = 0; k < 4; k++) {
            if (*slots[k] != VKRT_INVALID_INDEX && *slots[k] > textureIndex) {
                (*slots[k])--;
                materialsMoved = 1;
            }
        }
    }
    if (materialsMoved) vkrt->materialRevision++;
    if (vkrt->settings.environmentTextureIndex != VKRT_INVALID_INDEX &&
        vkrt->settings.environmentTextureIndex > textureIndex) {
        vkrt->settings.environmentTextureIndex--;
        vkrt->environmentRevision++;
    }
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

/* src/core/scene/environment.c:11-41: only linear textures may light the scene. */
VKRT_Result VKRT_setEnvironmentTexture(VKRT* vkrt, uint32_t textureIndex) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    if (textureIndex != VKRT_INVALID_INDEX) {
        if (textureIndex >= vkrt->textureCount) return VKRT_ERROR_INVALID_ARGUMENT;
        if (vkrt->textures[textureIndex].colorSpace != VKRT_TEXTURE_COLOR_SPACE_LINEAR)
            return VKRT_ERROR_INVALID_ARGUMENT;
    }
    if (vkrt->settings.environmentTextureIndex == textureIndex) return VKRT_SUCCESS;
    vkrt->settings.environmentTextureIndex = textureIndex;
    vkrt->environmentRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

/* src/core/scene/environment.c:11-38,63-65: the cleared map is released when
 * no material uses it. */
VKRT_Result VKRT_clearEnvironmentTexture(VKRT* vkrt) {
    if (!vkrt) return VKRT_ERROR_INVALID_ARGUMENT;
    const uint32_t previous = vkrt->settings.environmentTextureIndex;
    VKRT_Result r = VKRT_setEnvironmentTexture(vkrt, VKRT_INVALID_INDEX);
    if (r == VKRT_SUCCESS && previous != VKRT_INVALID_INDEX && previous < vkrt->textureCount &&
        textureUsers(vkrt, previous) == 0u)
        r = VKRT_removeTexture(vkrt, previous);
    return r;
}

/* src/core/scene/environment.c:11-60: the file loads as a linear texture and
 * replaces the environment map; the previous map is released when nothing
 * else uses it, and the new texture is released again if it cannot be used. */
VKRT_Result VKRT_setEnvironmentTextureFromFile(VKRT* vkrt, const char* path) {
    if (!vkrt || !path || !path[0]) return VKRT_ERROR_INVALID_ARGUMENT;
    uint32_t index = VKRT_INVALID_INDEX;
    VKRT_Result r = VKRT_addTextureFromFile(vkrt, path, NULL, VKRT_TEXTURE_COLOR_SPACE_LINEAR, &index);
    if (r != VKRT_SUCCESS) return r;
    if (vkrt->textures[index].colorSpace != VKRT_TEXTURE_COLOR_SPACE_LINEAR) {
        VKRT_removeTexture(vkrt, index);
        return VKRT_ERROR_INVALID_ARGUMENT;
    }
    const uint32_t previous = vkrt->settings.environmentTextureIndex;
    r = VKRT_setEnvironmentTexture(vkrt, index);
    if (r != VKRT_SUCCESS) {
        VKRT_removeTexture(vkrt, index);
        return r;
    }
    if (previous != VKRT_INVALID_INDEX && previous < vkrt->textureCount && textureUsers(vkrt, previous) == 0u)
        return VKRT_removeTexture(vkrt, previous);
    return VKRT_SUCCESS;
}

/* src/core/api/texture.c:35-46: an empty batch is invalid; a batch is added
 * whole or not at all. */
VKRT_Result VKRT_addTexturesBatch(VKRT* vkrt, const VKRT_TextureUpload* uploads, size_t uploadCount,
                                  uint32_t* outTextureIndices) {
    if (!vkrt || !uploads || uploadCount == 0u) return VKRT_ERROR_INVALID_ARGUMENT;
    uint32_t before = vkrt->textureCount;
    for (size_t i = 0; i < uploadCount; i++) {
        uint32_t index = VKRT_INVALID_INDEX;
        VKRT_Result r = VKRT_addTextureFromPixels(vkrt, &uploads[i], &index);
        if (r != VKRT_SUCCESS) {
            while (vkrt->textureCount > before) VKRT_removeTexture(vkrt, vkrt->textureCount - 1u);
            return r;
        }
        if (outTextureIndices) outTextureIndices[i] = index;
    }
    return VKRT_SUCCESS;
}

/* src/core/api/mesh.c:296-340,447-457: the default material (index 0) cannot
 * be removed; meshes that used the removed material fall back to it, later
 * materials move down one index, and textures only the removed material
 * referenced are released. */
VKRT_Result VKRT_removeMaterial(VKRT* vkrt, uint32_t materialIndex) {
    if (!vkrt || materialIndex >= vkrt->materialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    if (materialIndex == 0u) return VKRT_ERROR_OPERATION_FAILED;
    Material removed = vkrt->materials[materialIndex].material;
    for (uint32_t i = 0; i < vkrt->meshCount; i++) {
        HostMesh* mesh = &vkrt->meshes[i];
        if (mesh->info.materialIndex == materialIndex) {
            mesh->info.materialIndex = 0u;
            mesh->hasMaterialAssignment = 0u;
        } else if (mesh->info.materialIndex > materialIndex) {
            mesh->info.materialIndex--;
        }
    }
    memmove(&vkrt->materials[materialIndex], &vkrt->materials[materialIndex + 1u],
            sizeof(HostMaterial) * (vkrt->materialCount - materialIndex - 1u));
    vkrt->materialCount--;
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    uint32_t slots[4] = {removed.baseColorTextureIndex, removed.metallicRoughnessTextureIndex,
                         removed.normalTextureIndex, removed.emissiveTextureIndex};
    /* release in descending index order so earlier removals do not shift later ones */
    for (int pass = 0; pass < 4; pass++) {
        uint32_t best = VKRT_INVALID_INDEX;
        int bestSlot = -1;
        for (int k = 0; k < 4; k++) {
            if (slots[k] == VKRT_INVALID_INDEX) continue;
            if (best == VKRT_INVALID_INDEX || slots[k] > best) {
                best = slots[k];
                bestSlot = k;
            }
        }
        if (bestSlot < 0) break;
        for (int k = 0; k < 4; k++)
            if (slots[k] == best) slots[k] = VKRT_INVALID_INDEX;
        if (best >= vkrt->textureCount || best == vkrt->settings.environmentTextureIndex) continue;
        /* VKRT_removeTexture refuses a texture another material still binds */
        VKRT_removeTexture(vkrt, best);
    }
    return VKRT_SUCCESS;
}

/* src/core/api/mesh.c:287-294,459-473 */
VKRT_Result VKRT_setMaterialName(VKRT* vkrt, uint32_t materialIndex, const char* name) {
    if (!vkrt || !name || materialIndex >= vkrt->materialCount) return VKRT_ERROR_INVALID_ARGUMENT;
    if (name[0])
        snprintf(vkrt->materials[materialIndex].name, VKRT_NAME_LEN, "%s", name);
    else
        snprintf(vkrt->materials[materialIndex].name, VKRT_NAME_LEN, "Material %u", materialIndex);
    return VKRT_SUCCESS;
}

/* src/core/api/mesh.c:592-621: the mesh goes back to the default material. */
VKRT_Result VKRT_clearMeshMaterialAssignment(VKRT* vkrt, uint32_t meshIndex) {
    if (!vkrt || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    VKRT_Result r = ensureDefaultMaterial(vkrt);
    if (r != VKRT_SUCCESS) return r;
    HostMesh* mesh = &vkrt->meshes[meshIndex];
    if (!mesh->hasMaterialAssignment && mesh->info.materialIndex == 0u) return VKRT_SUCCESS;
    mesh->info.materialIndex = 0u;
    mesh->hasMaterialAssignment = 0u;
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

/* src/core/api/mesh.c:270-272,623-638: opacity must be finite and in [0, 1]. */
VKRT_Result VKRT_setMeshOpacity(VKRT* vkrt, uint32_t meshIndex, float opacity) {
    if (!vkrt || meshIndex >= vkrt->meshCount || !isfinite(opacity) || opacity < 0.0f || opacity > 1.0f)
        return VKRT_ERROR_INVALID_ARGUMENT;
    MeshInfo* info = &vkrt->meshes[meshIndex].info;
    if (info->opacity == opacity) return VKRT_SUCCESS;
    info->opacity = opacity;
    vkrt->materialRevision++; /* light-sampling eligibility depends on it */
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

/* src/core/api/mesh.c:646-663 */
VKRT_Result VKRT_setMeshRenderBackfaces(VKRT* vkrt, uint32_t meshIndex, uint32_t enabled) {
    if (!vkrt || meshIndex >= vkrt->meshCount) return VKRT_ERROR_INVALID_ARGUMENT;
    MeshInfo* info = &vkrt->meshes[meshIndex].info;
    uint32_t v = enabled ? 1u : 0u;
    if (info->renderBackfaces == v) return VKRT_SUCCESS;
    info->renderBackfaces = v;
    vkrt->materialRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}
