/*
 * Host scene preparation: turns the API-side meshes / materials / textures
 * into the packed buffers the kernels read.
 *
 *  - ShaderVertex quantisation keeps the reference's packed format
 *    (src/core/utility/packing.c:92-156): octahedral normal in two snorm16,
 *    octahedral tangent in two snorm15 plus a sign bit, colour in RGBA8.
 *  - Emissive light tables follow src/core/scene/lighting.c:55-433: per-mesh
 *    triangle alias tables over world-space areas, a mesh alias table over
 *    area * luminance(emission), and MeshInfo.lightPdfArea = pmfMesh / area.
 *  - World-space triangles use the mesh's 3x4 world matrix the way the
 *    reference feeds its TLAS instances (src/core/scene/transform.c:212-223).
 *  - The environment importance table (no counterpart in the reference) is an
 *    alias table over texels weighted by luminance * sin(theta).
 */
#include "internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static float clampUnit(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
static float clampSigned(float v) { return v < -1.0f ? -1.0f : (v > 1.0f ? 1.0f : v); }

static void unitVector(const float in[3], float out[3]) {
    float len2 = in[0] * in[0] + in[1] * in[1] + in[2] * in[2];
    if (len2 <= 1e-20f) {
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = 1.0f;
        return;
    }
    float inv = 1.0f / sqrtf(len2);
    out[0] = in[0] * inv;
    out[1] = in[1] * inv;
    out[2] = in[2] * inv;
}

/* Octahedral projection of a unit vector onto the [-1,1]^2 square. */
static void octahedralProject(const float n[3], float* px, float* py) {
    float invL1 = 1.0f / (fabsf(n[0]) + fabsf(n[1]) + fabsf(n[2]));
    float x = n[0] * invL1;
    float y = n[1] * invL1;
    if (n[2] < 0.0f) {
        float fx = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        float fy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = fx;
        y = fy;
    }
    *px = x;
    *py = y;
}

uint32_t vkrtPackOctNormal32(const float normal[3]) {
    float n[3];
    unitVector(normal, n);
    float x, y;
    octahedralProject(n, &x, &y);
    int32_t qx = (int32_t)lroundf(clampSigned(x) * 32767.0f);
    int32_t qy = (int32_t)lroundf(clampSigned(y) * 32767.0f);
    return ((uint32_t)qx & 0xffffu) | (((uint32_t)qy & 0xffffu) << 16u);
}

static uint32_t quantizeSnorm15(float v) {
    int32_t q = (int32_t)lroundf(clampSigned(v) * 16383.0f);
    return (uint32_t)q & 0x7fffu;
}

uint32_t vkrtPackTangent32(const float tangent[4]) {
    float n[3];
    unitVector(tangent, n);
    float x, y;
    octahedralProject(n, &x, &y);
    uint32_t packed = quantizeSnorm15(x) | (quantizeSnorm15(y) << 15u);
    if (tangent[3] < 0.0f) packed |= 0x80000000u;
    return packed;
}

uint32_t vkrtPackColorRGBA8(const float color[4]) {
    uint32_t out = 0u;
    for (int c = 0; c < 4; c++) {
        uint32_t q = (uint32_t)lroundf(clampUnit(color[c]) * 255.0f);
        out |= q << (8u * (uint32_t)c);
    }
    return out;
}

ShaderVertex vkrtPackShaderVertex(const Vertex* vertex) {
    ShaderVertex sv;
    memset(&sv, 0, sizeof(sv));
    if (!vertex) return sv;
    memcpy(sv.position, vertex->position, sizeof(sv.position));
    memcpy(sv.texcoord0, vertex->texcoord0, sizeof(sv.texcoord0));
    memcpy(sv.texcoord1, vertex->texcoord1, sizeof(sv.texcoord1));
    sv.packedNormal = vkrtPackOctNormal32(vertex->normal);
    sv.packedTangent = vkrtPackTangent32(vertex->tangent);
    sv.packedColor = vkrtPackColorRGBA8(vertex->color);
    return sv;
}

/* Vose / Walker alias table in fp32 with the same bin discipline as
 * src/core/scene/lighting.c:108-164 (LIFO small/large stacks, leftovers q=1). */
int vkrtBuildAliasTable(const float* pmf, uint32_t count, float* outQ, uint32_t* outIdx) {
    if (count == 0) return 1;
    float* scaled = (float*)malloc(sizeof(float) * count);
    uint32_t* small = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint32_t* large = (uint32_t*)malloc(sizeof(uint32_t) * count);
    if (!scaled || !small || !large) {
        free(scaled);
        free(small);
        free(large);
        return 0;
    }
    uint32_t ns = 0, nl = 0;
    for (uint32_t i = 0; i < count; i++) {
        scaled[i] = pmf[i] * (float)count;
        if (scaled[i] < 1.0f)
            small[ns++] = i;
        else
            large[nl++] = i;
    }
    while (ns > 0 && nl > 0) {
        uint32_t s = small[--ns];
        uint32_t l = large[--nl];
        outQ[s] = scaled[s];
        outIdx[s] = l;
        scaled[l] = scaled[l] + scaled[s] - 1.0f;
        if (scaled[l] < 1.0f)
            small[ns++] = l;
        else
            large[nl++] = l;
    }
    while (nl > 0) {
        uint32_t l = large[--nl];
        outQ[l] = 1.0f;
        outIdx[l] = l;
    }
    while (ns > 0) {
        uint32_t s = small[--ns];
        outQ[s] = 1.0f;
        outIdx[s] = s;
    }
    free(scaled);
    free(small);
    free(large);
    return 1;
}

void vkrtFreePackedScene(PackedScene* p) {
    if (!p) return;
    free(p->vertices);
    free(p->indices);
    free(p->meshInfos);
    free(p->materials);
    free(p->emissiveMeshes);
    free(p->emissiveTriangles);
    free(p->meshAliasQ);
    free(p->meshAliasIdx);
    free(p->triAliasQ);
    free(p->triAliasIdx);
    free(p->emissiveMeshMaterial);
    free(p->worldTriangles);
    free(p->triangleMesh);
    free(p->meshTriangleBase);
    free(p->envAliasQ);
    free(p->envAliasIdx);
    free(p->envPdf);
    free(p->triangleAlpha);
    free(p->textureTexels);
    free(p->textureInfo);
    memset(p, 0, sizeof(*p));
}

static float luminance3(const float c[3]) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

/* src/core/scene/lighting.c:55-67 */
static float emissionWeight(const Material* m) {
    if (!isfinite(m->emissionLuminance) || m->emissionLuminance <= 0.0f) return 0.0f;
    float lum = luminance3(m->emissionColor);
    if (lum <= 0.0f) return 0.0f;
    return lum * m->emissionLuminance;
}

static int eligibleForLightSampling(const MeshInfo* info, const Material* m) {
    if (info->opacity < 0.999f || m->opacity < 0.999f) return 0;
    if (m->emissiveTextureIndex != VKRT_INVALID_INDEX) return 0;
    return m->alphaMode == VKRT_MATERIAL_ALPHA_MODE_OPAQUE;
}

/* Row-major 3x4 application of a column-major 4x4, as the reference's
 * transformPosition over VkTransformMatrixKHR (lighting.c:69-76). */
static void worldPoint(const float m[4][4], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) {
        out[r] = (m[0][r] * p[0]) + (m[1][r] * p[1]) + (m[2][r] * p[2]) + m[3][r];
    }
}

static VKRT_Result buildLights(VKRT* vkrt, PackedScene* p) {
    uint32_t meshCount = vkrt->meshCount;
    uint64_t emTris = 0;
    uint32_t emMeshes = 0;
    for (uint32_t i = 0; i < meshCount; i++) {
        const HostMesh* mesh = &vkrt->meshes[i];
        const Material* m = &p->materials[mesh->info.materialIndex];
        if (!eligibleForLightSampling(&mesh->info, m) || emissionWeight(m) <= 0.0f) continue;
        uint32_t tris = mesh->info.indexCount / 3u;
        if (tris == 0) continue;
        emMeshes++;
        emTris += tris;
    }
    if (emTris > UINT32_MAX) return VKRT_ERROR_OPERATION_FAILED;
    uint32_t allocMeshes = emMeshes ? emMeshes : 1u;
    uint32_t allocTris = emTris ? (uint32_t)emTris : 1u;
    p->emissiveMeshes = (EmissiveMesh*)calloc(allocMeshes, sizeof(EmissiveMesh));
    p->emissiveTriangles = (EmissiveTriangle*)calloc(allocTris, sizeof(EmissiveTriangle));
    p->meshAliasQ = (float*)calloc(allocMeshes, sizeof(float));
    p->meshAliasIdx = (uint32_t*)calloc(allocMeshes, sizeof(uint32_t));
    p->triAliasQ = (float*)calloc(allocTris, sizeof(float));
    p->triAliasIdx = (uint32_t*)calloc(allocTris, sizeof(uint32_t));
    float* meshWeights = (float*)calloc(allocMeshes, sizeof(float));
    float* pmfScratch = (float*)calloc(allocTris > allocMeshes ? allocTris : allocMeshes, sizeof(float));
    uint32_t* sourceMesh = (uint32_t*)calloc(allocMeshes, sizeof(uint32_t));
    p->emissiveMeshMaterial = (uint32_t*)calloc(allocMeshes, sizeof(uint32_t));
    if (!p->emissiveMeshes || !p->emissiveTriangles || !p->meshAliasQ || !p->meshAliasIdx || !p->triAliasQ ||
        !p->triAliasIdx || !meshWeights || !pmfScratch || !sourceMesh || !p->emissiveMeshMaterial) {
        free(meshWeights);
        free(pmfScratch);
        free(sourceMesh);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }

    uint32_t nMesh = 0, nTri = 0;
    float totalSelection = 0.0f;
    for (uint32_t mi = 0; mi < meshCount; mi++) {
        const HostMesh* mesh = &vkrt->meshes[mi];
        const Material* m = &p->materials[mesh->info.materialIndex];
        if (!eligibleForLightSampling(&mesh->info, m) || emissionWeight(m) <= 0.0f) continue;
        if (!mesh->vertices || !mesh->indices || mesh->info.indexCount / 3u == 0u) continue;

        uint32_t triOffset = nTri;
        float totalArea = 0.0f;
        uint32_t tris = mesh->info.indexCount / 3u;
        for (uint32_t t = 0; t < tris; t++) {
            uint32_t i0 = mesh->indices[t * 3u + 0u];
            uint32_t i1 = mesh->indices[t * 3u + 1u];
            uint32_t i2 = mesh->indices[t * 3u + 2u];
            if (i0 >= mesh->info.vertexCount || i1 >= mesh->info.vertexCount || i2 >= mesh->info.vertexCount) continue;
            float p0[3], p1[3], p2[3];
            worldPoint(mesh->worldTransform, mesh->vertices[i0].position, p0);
            worldPoint(mesh->worldTransform, mesh->vertices[i1].position, p1);
            worldPoint(mesh->worldTransform, mesh->vertices[i2].position, p2);
            float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            float area = 0.5f * sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (area <= 0.0f) continue;
            EmissiveTriangle* et = &p->emissiveTriangles[nTri++];
            memset(et, 0, sizeof(*et));
            et->v0Area[0] = p0[0];
            et->v0Area[1] = p0[1];
            et->v0Area[2] = p0[2];
            et->v0Area[3] = area;
            for (int k = 0; k < 3; k++) {
                et->e1Pad[k] = e1[k];
                et->e2Pad[k] = e2[k];
            }
            totalArea += area;
        }
        uint32_t valid = nTri - triOffset;
        float selection = totalArea * emissionWeight(m);
        if (selection <= 0.0f || valid == 0u) {
            nTri = triOffset;
            continue;
        }
        float invArea = 1.0f / totalArea;
        for (uint32_t t = 0; t < valid; t++) pmfScratch[t] = p->emissiveTriangles[triOffset + t].v0Area[3] * invArea;
        if (!vkrtBuildAliasTable(pmfScratch, valid, &p->triAliasQ[triOffset], &p->triAliasIdx[triOffset])) {
            free(meshWeights);
            free(pmfScratch);
            free(sourceMesh);
            return VKRT_ERROR_OUT_OF_MEMORY;
        }
        EmissiveMesh* em = &p->emissiveMeshes[nMesh];
        memset(em, 0, sizeof(*em));
        em->triOffset = triOffset;
        em->triCount = valid;
        em->invTotalArea = invArea;
        for (int k = 0; k < 3; k++) em->emission[k] = m->emissionColor[k] * m->emissionLuminance;
        meshWeights[nMesh] = selection;
        sourceMesh[nMesh] = mi;
        p->emissiveMeshMaterial[nMesh] = mesh->info.materialIndex;
        totalSelection += selection;
        nMesh++;
    }

    for (uint32_t i = 0; i < meshCount; i++) p->meshInfos[i].lightPdfArea = 0.0f;
    if (nMesh > 0 && totalSelection > 0.0f) {
        float invTotal = 1.0f / totalSelection;
        for (uint32_t i = 0; i < nMesh; i++) {
            float pmf = meshWeights[i] * invTotal;
            p->emissiveMeshes[i].pmfMesh = pmf;
            pmfScratch[i] = pmf;
            p->meshInfos[sourceMesh[i]].lightPdfArea = pmf * p->emissiveMeshes[i].invTotalArea;
        }
        if (!vkrtBuildAliasTable(pmfScratch, nMesh, p->meshAliasQ, p->meshAliasIdx)) {
            free(meshWeights);
            free(pmfScratch);
            free(sourceMesh);
            return VKRT_ERROR_OUT_OF_MEMORY;
        }
    }
    p->emissiveMeshCount = nMesh;
    p->emissiveTriangleCount = nTri;
    free(meshWeights);
    free(pmfScratch);
    free(sourceMesh);
    return VKRT_SUCCESS;
}

static VKRT_Result buildGeometry(VKRT* vkrt, PackedScene* p) {
    uint64_t vcount = 0, icount = 0;
    for (uint32_t i = 0; i < vkrt->meshCount; i++) {
        vcount += vkrt->meshes[i].info.vertexCount;
        icount += vkrt->meshes[i].info.indexCount;
    }
    if (vcount > UINT32_MAX || icount > UINT32_MAX) return VKRT_ERROR_INVALID_ARGUMENT;
    p->vertexCount = (uint32_t)vcount;
    p->indexCount = (uint32_t)icount;
    p->triangleCount = (uint32_t)(icount / 3u);
    p->vertices = (ShaderVertex*)malloc(sizeof(ShaderVertex) * (vcount ? vcount : 1));
    p->indices = (uint32_t*)malloc(sizeof(uint32_t) * (icount ? icount : 1));
    p->worldTriangles = (float*)malloc(sizeof(float) * 9u * (p->triangleCount ? p->triangleCount : 1u));
    p->triangleMesh = (uint32_t*)malloc(sizeof(uint32_t) * (p->triangleCount ? p->triangleCount : 1u));
    p->meshTriangleBase = (uint32_t*)malloc(sizeof(uint32_t) * (vkrt->meshCount ? vkrt->meshCount : 1u));
    if (!p->vertices || !p->indices || !p->worldTriangles || !p->triangleMesh || !p->meshTriangleBase)
        return VKRT_ERROR_OUT_OF_MEMORY;

    uint32_t vbase = 0, ibase = 0, tbase = 0;
    for (uint32_t mi = 0; mi < vkrt->meshCount; mi++) {
        HostMesh* mesh = &vkrt->meshes[mi];
        MeshInfo* info = &p->meshInfos[mi];
        info->vertexBase = vbase;
        info->indexBase = ibase;
        p->meshTriangleBase[mi] = tbase;
        for (uint32_t v = 0; v < info->vertexCount; v++) p->vertices[vbase + v] = vkrtPackShaderVertex(&mesh->vertices[v]);
        memcpy(&p->indices[ibase], mesh->indices, sizeof(uint32_t) * info->indexCount);
        uint32_t tris = info->indexCount / 3u;
        for (uint32_t t = 0; t < tris; t++) {
            float* dst = &p->worldTriangles[(size_t)(tbase + t) * 9u];
            for (int c = 0; c < 3; c++) {
                uint32_t vi = mesh->indices[t * 3u + (uint32_t)c];
                if (vi >= info->vertexCount) vi = 0; /* out-of-range indices collapse the triangle */
                worldPoint(mesh->worldTransform, mesh->vertices[vi].position, &dst[c * 3]);
            }
            p->triangleMesh[tbase + t] = mi;
        }
        vbase += info->vertexCount;
        ibase += info->indexCount;
        tbase += tris;
    }
    return VKRT_SUCCESS;
}

static VKRT_Result buildEnvironmentTable(VKRT* vkrt, PackedScene* p) {
    uint32_t ti = vkrt->settings.environmentTextureIndex;
    if (ti == VKRT_INVALID_INDEX || ti >= vkrt->textureCount) return VKRT_SUCCESS;
    const HostTexture* tex = &vkrt->textures[ti];
    uint32_t w = tex->width, h = tex->height;
    size_t n = (size_t)w * h;
    p->envAliasQ = (float*)malloc(sizeof(float) * n);
    p->envAliasIdx = (uint32_t*)malloc(sizeof(uint32_t) * n);
    p->envPdf = (float*)malloc(sizeof(float) * n);
    double* weight = (double*)malloc(sizeof(double) * n);
    if (!p->envAliasQ || !p->envAliasIdx || !p->envPdf || !weight) {
        free(weight);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    double total = 0.0;
    for (uint32_t y = 0; y < h; y++) {
        double sinTheta = sin(M_PI * ((double)y + 0.5) / (double)h);
        for (uint32_t x = 0; x < w; x++) {
            const float* px = &tex->rgba[((size_t)y * w + x) * 4u];
            double lum = 0.2126 * px[0] + 0.7152 * px[1] + 0.0722 * px[2];
            if (!(lum > 0.0)) lum = 0.0;
            weight[(size_t)y * w + x] = lum * sinTheta;
            total += lum * sinTheta;
        }
    }
    /* a floor keeps every direction reachable, so bilinear bleed into dark
     * texels is still covered by the light strategy */
    double floorW = total > 0.0 ? 1e-3 * total / (double)n : 1.0;
    double total2 = 0.0;
    for (size_t i = 0; i < n; i++) {
        double sinTheta = sin(M_PI * ((double)(i / w) + 0.5) / (double)h);
        weight[i] += floorW * sinTheta;
        total2 += weight[i];
    }
    for (size_t i = 0; i < n; i++) p->envPdf[i] = (float)(weight[i] / total2);
    free(weight);
    if (!vkrtBuildAliasTable(p->envPdf, (uint32_t)n, p->envAliasQ, p->envAliasIdx)) return VKRT_ERROR_OUT_OF_MEMORY;
    return VKRT_SUCCESS;
}

/* src/shaders/rt/alpha_test.slang:8-18 (alphaMayRejectHit): a mesh runs the
 * alpha test when its material is MASK or BLEND or either opacity is below
 * 0.999; src/core/render/accel/tlas.c:298-314 marks every other instance opaque. */
static int meshMayRejectHit(const MeshInfo* info, const Material* m) {
    return m->alphaMode == VKRT_MATERIAL_ALPHA_MODE_MASK || m->alphaMode == VKRT_MATERIAL_ALPHA_MODE_BLEND ||
           m->opacity < 0.999f || info->opacity < 0.999f;
}

/* The alpha test reads the base-colour texture's alpha for MASK and BLEND
 * materials (src/shaders/rt/alpha_test.slang:20-26); other materials use the
 * vertex alpha alone. */
static int meshSamplesTextureAlpha(const Material* m) {
    return m->alphaMode != VKRT_MATERIAL_ALPHA_MODE_OPAQUE && m->baseColorTextureIndex != VKRT_INVALID_INDEX;
}

static VKRT_Result buildTriangleAlpha(VKRT* vkrt, PackedScene* p) {
    int any = 0;
    for (uint32_t mi = 0; mi < vkrt->meshCount && !any; mi++)
        any = meshMayRejectHit(&p->meshInfos[mi], &p->materials[p->meshInfos[mi].materialIndex]);
    if (!any || p->triangleCount == 0) return VKRT_SUCCESS;
    p->triangleAlpha = (uint32_t*)calloc(p->triangleCount, sizeof(uint32_t));
    if (!p->triangleAlpha) return VKRT_ERROR_OUT_OF_MEMORY;
    for (uint32_t mi = 0; mi < vkrt->meshCount; mi++) {
        const MeshInfo* info = &p->meshInfos[mi];
        if (!meshMayRejectHit(info, &p->materials[info->materialIndex])) continue;
        uint32_t tris = info->indexCount / 3u, tbase = p->meshTriangleBase[mi];
        for (uint32_t t = 0; t < tris; t++) {
            uint32_t word = 0x80000000u | (meshSamplesTextureAlpha(&p->materials[info->materialIndex]) ? 0x40000000u : 0u);
            for (uint32_t c = 0; c < 3u; c++) {
                uint32_t vi = p->indices[info->indexBase + t * 3u + c];
                if (vi >= info->vertexCount) vi = 0;
                word |= (p->vertices[info->vertexBase + vi].packedColor >> 24) << (8u * c);
            }
            p->triangleAlpha[tbase + t] = word;
        }
    }
    return VKRT_SUCCESS;
}

/* Texture indices a material holds that name no texture are treated as
 * empty slots; the pool then holds each referenced texture once, row-major
 * from row 0 (texture coordinate v = 0), as the device samples it. */
static VKRT_Result buildTexturePool(VKRT* vkrt, PackedScene* p) {
    uint32_t tc = vkrt->textureCount;
    for (uint32_t i = 0; i < p->materialCount; i++) {
        Material* m = &p->materials[i];
        uint32_t* slots[4] = {&m->baseColorTextureIndex, &m->metallicRoughnessTextureIndex, &m->normalTextureIndex,
                              &m->emissiveTextureIndex};
        for (int k = 0; k < 4; k++)
            if (*slots[k] >= tc) *slots[k] = VKRT_INVALID_INDEX;
    }
    if (tc == 0) return VKRT_SUCCESS;
    uint8_t* used = (uint8_t*)calloc(tc, 1);
    p->textureInfo = (uint32_t*)calloc((size_t)tc * 4u, sizeof(uint32_t));
    if (!used || !p->textureInfo) {
        free(used);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    p->textureCount = tc;
    for (uint32_t i = 0; i < p->materialCount; i++) {
        const Material* m = &p->materials[i];
        uint32_t slots[4] = {m->baseColorTextureIndex, m->metallicRoughnessTextureIndex, m->normalTextureIndex,
                             m->emissiveTextureIndex};
        for (int k = 0; k < 4; k++)
            if (slots[k] != VKRT_INVALID_INDEX) used[slots[k]] = 1;
    }
    uint64_t total = 0;
    for (uint32_t t = 0; t < tc; t++) {
        if (!used[t]) continue;
        const HostTexture* tex = &vkrt->textures[t];
        if (total + (uint64_t)tex->width * tex->height > 0xFFFFFFFFull) {
            free(used);
            VKRT_LOG_ERROR("material textures exceed 2^32 texels");
            return VKRT_ERROR_OUT_OF_MEMORY;
        }
        p->textureInfo[4u * t + 0u] = (uint32_t)total;
        p->textureInfo[4u * t + 1u] = tex->width;
        p->textureInfo[4u * t + 2u] = tex->height;
        total += (uint64_t)tex->width * tex->height;
    }
    p->textureTexelCount = total;
    if (total) {
        p->textureTexels = (float*)malloc((size_t)total * 4u * sizeof(float));
        if (!p->textureTexels) {
            free(used);
            return VKRT_ERROR_OUT_OF_MEMORY;
        }
        for (uint32_t t = 0; t < tc; t++)
            if (used[t])
                memcpy(p->textureTexels + (size_t)p->textureInfo[4u * t] * 4u, vkrt->textures[t].rgba,
                       (size_t)vkrt->textures[t].width * vkrt->textures[t].height * 4u * sizeof(float));
    }
    free(used);
    return VKRT_SUCCESS;
}

VKRT_Result vkrtPrepareScene(VKRT* vkrt) {
    PackedScene next;
    memset(&next, 0, sizeof(next));
    uint32_t mc = vkrt->meshCount ? vkrt->meshCount : 1u;
    next.meshInfos = (MeshInfo*)calloc(mc, sizeof(MeshInfo));
    next.materialCount = vkrt->materialCount;
    next.materials = (Material*)calloc(vkrt->materialCount ? vkrt->materialCount : 1u, sizeof(Material));
    if (!next.meshInfos || !next.materials) {
        vkrtFreePackedScene(&next);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    if (vkrt->materialCount == 0) {
        next.materials[0] = VKRT_materialDefault();
        next.materialCount = 1;
    }
    for (uint32_t i = 0; i < vkrt->materialCount; i++) next.materials[i] = vkrt->materials[i].material;
    for (uint32_t i = 0; i < next.materialCount; i++) {
        if (next.materials[i].transmission > 0.0f) next.hasTransmissive = 1u;
    }
    for (uint32_t i = 0; i < vkrt->meshCount; i++) {
        next.meshInfos[i] = vkrt->meshes[i].info;
        if (next.meshInfos[i].materialIndex >= next.materialCount) next.meshInfos[i].materialIndex = 0u;
    }
    VKRT_Result r = buildGeometry(vkrt, &next);
    if (r == VKRT_SUCCESS) r = buildLights(vkrt, &next);
    if (r == VKRT_SUCCESS) r = buildEnvironmentTable(vkrt, &next);
    if (r == VKRT_SUCCESS) r = buildTexturePool(vkrt, &next);
    if (r == VKRT_SUCCESS) r = buildTriangleAlpha(vkrt, &next);
    if (r != VKRT_SUCCESS) {
        vkrtFreePackedScene(&next);
        return r;
    }
    vkrtFreePackedScene(&vkrt->packed);
    vkrt->packed = next;
    for (uint32_t i = 0; i < vkrt->meshCount; i++) vkrt->meshes[i].info.lightPdfArea = next.meshInfos[i].lightPdfArea;
    vkrt->sceneData.emissiveMeshCount = next.emissiveMeshCount;
    vkrt->sceneData.emissiveTriangleCount = next.emissiveTriangleCount;
    return VKRT_SUCCESS;
}
