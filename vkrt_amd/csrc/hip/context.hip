// GPU context management and scene uploads for the C host layer.
#include <cstdio>
#include <cstring>
#include <string>

#include "mrt_internal.hpp"

namespace mrt {

static thread_local std::string g_lastError;

void setError(const std::string& msg) { g_lastError = msg; }

bool check(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return false;
}

}  // namespace mrt

using namespace mrt;

extern "C" const char* mrt_last_error(void) { return g_lastError.c_str(); }

extern "C" int mrt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int mrt_create(int deviceIndex, mrt_context** out) {
    if (!out) return -1;
    *out = nullptr;
    int n = mrt_device_count();
    if (n <= 0) {
        setError("no HIP device available");
        return -5;
    }
    if (deviceIndex < 0 || deviceIndex >= n) {
        setError("device index out of range");
        return -1;
    }
    if (!check(hipSetDevice(deviceIndex), "hipSetDevice")) return -5;
    hipDeviceProp_t prop;
    if (!check(hipGetDeviceProperties(&prop, deviceIndex), "hipGetDeviceProperties")) return -5;
    mrt_context* ctx = new mrt_context();
    ctx->device = deviceIndex;
    ctx->cuCount = prop.multiProcessorCount;
    ctx->name = prop.name;
    if (prop.gcnArchName[0]) ctx->name += std::string(" (") + prop.gcnArchName + ")";
    ctx->vendor = 0x1002u;
    int driver = 0;
    hipDriverGetVersion(&driver);
    ctx->driver = (uint32_t)driver;
    if (!check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        !check(hipEventCreate(&ctx->evBegin), "hipEventCreate") || !check(hipEventCreate(&ctx->evEnd), "hipEventCreate")) {
        delete ctx;
        return -5;
    }
    for (auto& e : ctx->evStage) hipEventCreate(&e);
    mrt::noAllocScratchInit(ctx);  // without it only the allocation-free query refuses to run
    *out = ctx;
    return 0;
}

extern "C" void mrt_destroy(mrt_context* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    mrt::nrcRelease(ctx);
    ctx->vertices.release();
    ctx->indices.release();
    ctx->worldTris.release();
    ctx->triMesh.release();
    ctx->triAlpha.release();
    ctx->texels.release();
    ctx->texInfo.release();
    ctx->meshTriBase.release();
    ctx->tris.release();
    ctx->nodes.release();
    ctx->shadeTris.release();
    ctx->triClass.release();
    ctx->meshes.release();
    ctx->materials.release();
    ctx->emissiveMeshes.release();
    ctx->emissiveTriangles.release();
    ctx->meshAliasQ.release();
    ctx->meshAliasIdx.release();
    ctx->triAliasQ.release();
    ctx->triAliasIdx.release();
    ctx->emissiveMeshMaterial.release();
    ctx->env.release();
    ctx->envAliasQ.release();
    ctx->envAliasIdx.release();
    ctx->envPdf.release();
    ctx->rgb2spec.release();
    ctx->accum.release();
    ctx->output.release();
    ctx->frameSum.release();
    ctx->adjImage.release();
    ctx->adjGrad.release();
    PathBuffers& b = ctx->paths;
    b.rng.release();
    b.flags.release();
    b.thr.release();
    b.rad.release();
    b.radXYZ.release();
    b.tpp.release();
    b.ptpp.release();
    b.pbtp.release();
    b.nrcQuery.release();
    b.nrcRecord.release();
    b.nrcTarget.release();
    b.sortKey.release();
    b.sortPerm.release();
    b.sortCount.release();
    b.sortOffset.release();
    b.sortScratch.release();
    b.lam.release();
    b.misc.release();
    b.sigma.release();
    for (auto& q : b.rays)
        for (auto& a : q) a.release();
    b.hits.release();
    for (auto& a : b.shadow) a.release();
    for (auto& a : b.shadowEnv) a.release();
    b.sampleRad.release();
    ctx->counters.release();
    ctx->traversalStack.release();
    hipEventDestroy(ctx->evBegin);
    hipEventDestroy(ctx->evEnd);
    for (auto& e : ctx->evStage) hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int mrt_device_synchronize(mrt_context* ctx) {
    if (!ctx) return -1;
    hipSetDevice(ctx->device);
    return check(hipDeviceSynchronize(), "device synchronize") ? 0 : -4;
}

extern "C" int mrt_device_name(const mrt_context* ctx, char* out, size_t len, uint32_t* vendorId,
                               uint32_t* driverVersion) {
    if (!ctx || !out || len == 0) return -1;
    snprintf(out, len, "%s", ctx->name.c_str());
    if (vendorId) *vendorId = ctx->vendor;
    if (driverVersion) *driverVersion = ctx->driver;
    return 0;
}

static int sync(mrt_context* ctx, const char* what) {
    return check(hipStreamSynchronize(ctx->stream), what) ? 0 : -4;
}

extern "C" int mrt_upload_geometry(mrt_context* ctx, const mrt_geometry_desc* g) {
    if (!ctx || !g) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    hipStream_t s = ctx->stream;
    if (g->triangleCount >= (1u << 27)) {
        setError("more than 2^27 triangles");
        return -1;
    }
    bool ok = ctx->vertices.upload(g->vertices, g->vertexCount, s) && ctx->indices.upload(g->indices, g->indexCount, s) &&
              ctx->worldTris.upload(g->worldTriangles, (size_t)g->triangleCount * 9u, s) &&
              ctx->triMesh.upload(g->triangleMesh, g->triangleCount, s) &&
              ctx->meshTriBase.upload(g->meshTriangleBase, g->meshCount, s);
    if (!ok) return -3;
    ctx->triCount = g->triangleCount;
    ctx->shadeTrisDirty = true;
    ctx->triClassDirty = true;
    ctx->hintPaths = 0;  // the queue lengths of the old geometry no longer say where the tail may start
    return sync(ctx, "upload geometry");
}

extern "C" int mrt_upload_meshes(mrt_context* ctx, const MeshInfo* meshInfos, uint32_t meshCount) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (!ctx->meshes.upload(meshInfos, meshCount, ctx->stream)) return -3;
    ctx->shadeTrisDirty = true;
    ctx->triClassDirty = true;
    ctx->hintPaths = 0;
    return sync(ctx, "upload meshes");
}

extern "C" int mrt_upload_materials(mrt_context* ctx, const Material* materials, uint32_t materialCount) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (!ctx->materials.upload(materials, materialCount, ctx->stream)) return -3;
    ctx->sheenEnabled = false;
    for (uint32_t i = 0; i < materialCount; i++) {
        const float* tw = materials[i].sheenTintWeight;
        if (tw[3] > 0.0f && (tw[0] > 0.0f || tw[1] > 0.0f || tw[2] > 0.0f)) ctx->sheenEnabled = true;
    }
    ctx->triClassDirty = true;
    return sync(ctx, "upload materials");
}

extern "C" int mrt_upload_alpha(mrt_context* ctx, const uint32_t* triangleAlpha, uint32_t triangleCount) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    ctx->alphaEnabled = false;
    if (!triangleAlpha || triangleCount == 0) return 0;
    if (!ctx->triAlpha.upload(triangleAlpha, triangleCount, ctx->stream)) return -3;
    ctx->alphaEnabled = true;
    return sync(ctx, "upload alpha");
}

extern "C" int mrt_upload_textures(mrt_context* ctx, const float* texels, uint64_t texelCount, const uint32_t* info,
                                   uint32_t textureCount) {
    if (!ctx) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    ctx->texturesEnabled = false;
    if (!texels || texelCount == 0 || !info || textureCount == 0) return 0;
    if (!ctx->texels.upload(reinterpret_cast<const float4*>(texels), texelCount, ctx->stream) ||
        !ctx->texInfo.upload(reinterpret_cast<const uint4*>(info), textureCount, ctx->stream))
        return -3;
    ctx->texturesEnabled = true;
    return sync(ctx, "upload textures");
}

extern "C" int mrt_upload_lights(mrt_context* ctx, const mrt_light_desc* l) {
    if (!ctx || !l) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    hipStream_t s = ctx->stream;
    uint32_t m = l->emissiveMeshCount ? l->emissiveMeshCount : 1u;
    uint32_t t = l->emissiveTriangleCount ? l->emissiveTriangleCount : 1u;
    bool ok = ctx->emissiveMeshes.upload(l->emissiveMeshes, m, s) &&
              ctx->emissiveTriangles.upload(l->emissiveTriangles, t, s) && ctx->meshAliasQ.upload(l->meshAliasQ, m, s) &&
              ctx->meshAliasIdx.upload(l->meshAliasIdx, m, s) && ctx->triAliasQ.upload(l->triAliasQ, t, s) &&
              ctx->triAliasIdx.upload(l->triAliasIdx, t, s);
    if (ok && l->emissiveMeshMaterial) ok = ctx->emissiveMeshMaterial.upload(l->emissiveMeshMaterial, m, s);
    if (!ok) return -3;
    return sync(ctx, "upload lights");
}

extern "C" int mrt_upload_environment(mrt_context* ctx, const mrt_environment_desc* e) {
    if (!ctx || !e) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    hipStream_t s = ctx->stream;
    if (!e->pixels || e->width == 0 || e->height == 0) {
        ctx->envWidth = ctx->envHeight = 0;
        return 0;
    }
    size_t n = (size_t)e->width * e->height;
    bool ok = ctx->env.upload(reinterpret_cast<const float4*>(e->pixels), n, s);
    if (ok && e->aliasQ) ok = ctx->envAliasQ.upload(e->aliasQ, n, s) && ctx->envAliasIdx.upload(e->aliasIdx, n, s) &&
                               ctx->envPdf.upload(e->pdf, n, s);
    if (!ok) return -3;
    ctx->envWidth = e->width;
    ctx->envHeight = e->height;
    return sync(ctx, "upload environment");
}

extern "C" int mrt_upload_rgb2spec(mrt_context* ctx, const float* table, uint32_t floatCount) {
    if (!ctx || !table || floatCount == 0) return -1;
    if (!check(hipSetDevice(ctx->device), "hipSetDevice")) return -2;
    if (!ctx->rgb2spec.upload(table, floatCount, ctx->stream)) return -3;
    return sync(ctx, "upload rgb2spec");
}
