// Internal state of one GPU context (one HIP device, one stream).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../mrt.h"
#include "dev_common.h"

namespace mrt {

void setError(const std::string& msg);
bool check(hipError_t e, const char* what);

template <typename T>
struct DeviceArray {
    T* ptr = nullptr;
    size_t count = 0;     // elements in use
    size_t capacity = 0;  // elements allocated
    bool reserve(size_t n) {
        if (n <= capacity) return true;
        if (ptr) hipFree(ptr);
        ptr = nullptr;
        capacity = 0;
        size_t bytes = (n ? n : 1) * sizeof(T);
        if (!check(hipMalloc(&ptr, bytes), "hipMalloc")) return false;
        capacity = n;
        return true;
    }
    bool upload(const T* host, size_t n, hipStream_t s) {
        if (!reserve(n)) return false;
        count = n;
        if (n == 0) return true;
        return check(hipMemcpyAsync(ptr, host, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    }
    void release() {
        if (ptr) hipFree(ptr);
        ptr = nullptr;
        count = capacity = 0;
    }
};

// Wavefront buffers, sized for the largest chunk of paths seen so far.
struct PathBuffers {
    DeviceArray<uint4> rng;
    DeviceArray<uint32_t> flags;
    DeviceArray<float4> thr, rad, radXYZ, tpp, ptpp, pbtp, lam, misc, sigma;
    DeviceArray<float4> rays[2][2];  // [queue][origin+path | direction+tmax]
    DeviceArray<float4> hits;
    DeviceArray<float4> shadow[4];   // origin+path, direction+tmax, contribution if visible, otherwise
    DeviceArray<float4> shadowEnv[3];  // environment NEE: origin+path, direction+tmax, contribution if visible
    DeviceArray<float4> sampleRad;
    DeviceArray<float4> nrcQuery, nrcRecord, nrcTarget;  // neural cache queries, training records, targets
    DeviceArray<uint8_t> sortKey;                         // material class of every queued ray
    DeviceArray<uint32_t> sortPerm, sortCount, sortOffset;  // shading order, per-block counts, their scan
    DeviceArray<uint8_t> sortScratch;                     // rocPRIM scan temporary storage
    // denoiser features per path (only while feature images or debug views are on):
    // state bit 0 = resolved, bit 1 = followed a specular hit at depth 0;
    // albedo.w = resolve depth + 1
    DeviceArray<uint32_t> featState;
    DeviceArray<float4> featAlbedo, featNormal;
    size_t capacity = 0;
};

// neural radiance cache (nrc.hip)
bool nrcEnsure(mrt_context* ctx, uint32_t seed);
void nrcRelease(mrt_context* ctx);
struct Spec;
// spec: null for RGB queries; for hero spectral queries the rgb2spec table the
// cached XYZ is uplifted with
bool nrcInferQueries(mrt_context* ctx, const float4* queries, uint32_t stride, const uint32_t* count,
                     uint32_t capacity, float4* rad, const Spec* spec);
bool nrcTrainSteps(mrt_context* ctx, const float4* in, uint32_t stride, const float4* target, const uint32_t* count,
                   uint32_t capacity, uint32_t steps, uint32_t batch, float lr);
float nrcReadLoss(mrt_context* ctx);

// allocation-free ray query (render.hip): resolves its scratch array's device address
bool noAllocScratchInit(mrt_context* ctx);

}  // namespace mrt

struct mrt_context {
    void* nrc = nullptr;  // neural radiance cache state (nrc.hip), created on first use
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t evBegin = nullptr, evEnd = nullptr;
    hipEvent_t evStage[8] = {};
    int cuCount = 0;
    std::string name;
    uint32_t vendor = 0, driver = 0;

    // scene
    mrt::DeviceArray<ShaderVertex> vertices;
    mrt::DeviceArray<uint32_t> indices;
    mrt::DeviceArray<float> worldTris;
    mrt::DeviceArray<uint32_t> triMesh;
    mrt::DeviceArray<uint32_t> triAlpha;  // per-triangle alpha words, valid when alphaEnabled
    bool alphaEnabled = false;
    mrt::DeviceArray<float4> texels;  // material texture pool (linear RGBA)
    mrt::DeviceArray<uint4> texInfo;  // per texture: first texel, width, height
    bool texturesEnabled = false;     // some material references a texture
    bool sheenEnabled = false;        // some material has a sheen layer (selects the sheen shading variants)
    mrt::DeviceArray<uint32_t> meshTriBase;
    mrt::DeviceArray<mrt::TriRecord> tris;
    mrt::DeviceArray<mrt::BvhNode> nodes;
    mrt::DeviceArray<mrt::ShadeTri> shadeTris;
    bool shadeTrisDirty = true;  // geometry or mesh transforms changed since the records were built
    mrt::DeviceArray<uint8_t> triClass;  // material class per triangle (shading sort)
    bool triClassDirty = true;
    mrt::DeviceArray<MeshInfo> meshes;
    mrt::DeviceArray<Material> materials;
    mrt::DeviceArray<EmissiveMesh> emissiveMeshes;
    mrt::DeviceArray<EmissiveTriangle> emissiveTriangles;
    mrt::DeviceArray<float> meshAliasQ, triAliasQ;
    mrt::DeviceArray<uint32_t> meshAliasIdx, triAliasIdx;
    mrt::DeviceArray<uint32_t> emissiveMeshMaterial;
    mrt::DeviceArray<float4> env;
    uint32_t envWidth = 0, envHeight = 0;
    mrt::DeviceArray<float> envAliasQ, envPdf;
    mrt::DeviceArray<uint32_t> envAliasIdx;
    mrt::DeviceArray<float> rgb2spec;
    uint32_t triCount = 0;
    uint32_t bvhMaxDepth = 0;

    // accumulation
    mrt::DeviceArray<float4> accum;
    mrt::DeviceArray<ushort4> output;
    mrt::DeviceArray<float4> frameSum;
    // denoiser feature images (accumulation extent, .w = feature weight) and the
    // per-pixel frame records behind them and the debug views
    mrt::DeviceArray<float4> albedoImg, normalImg;
    mrt::DeviceArray<float4> frameFeat;  // 3 per pixel: albedo sum + weight, normal sum + depth sum, follow count
    mrt::DeviceArray<float4> debug0;     // per pixel: primary-surface debug value of sample 0 (w = 1 once set)
    mrt::DeviceArray<uint32_t> bounce0;  // per pixel: bounce count of sample 0
    bool featureImagesValid = false;
    uint32_t* noAllocScratch = nullptr;
    // bounding box of the triangles of the last BVH build, padded outward
    // (kExtend rejects rays that miss it before touching the tree)
    float sceneBoxLo[3] = {-INFINITY, -INFINITY, -INFINITY};
    float sceneBoxHi[3] = {INFINITY, INFINITY, INFINITY};  // device address of the allocation-free query's fixed scratch array
    uint32_t width = 0, height = 0;

    // adjoint (render_backward): loss gradient image and parameter gradients
    mrt::DeviceArray<float> adjImage;
    mrt::DeviceArray<float> adjGrad;

    // wavefront + traversal scratch
    mrt::PathBuffers paths;
    mrt::DeviceArray<uint32_t> counters;
    mrt::DeviceArray<uint32_t> traversalStack;
    uint32_t traceGrid = 0;
    uint32_t tailGrid = 0;
    // queue length of every depth in the last single-chunk frame, with the
    // chunk size and depth it was measured for: where the fused tail may start
    uint32_t depthHint[32] = {};
    uint64_t hintPaths = 0;
    uint32_t hintMaxDepth = 0;

    mrt::SceneView view() const {
        mrt::SceneView v{};
        v.nodes = nodes.ptr;
        v.tris = tris.ptr;
        v.triCount = triCount;
        v.nodeCount = triCount ? (uint32_t)nodes.count : 0u;
        v.vertices = vertices.ptr;
        v.indices = indices.ptr;
        v.triMesh = triMesh.ptr;
        v.meshTriBase = meshTriBase.ptr;
        v.meshes = meshes.ptr;
        v.materials = materials.ptr;
        v.meshCount = (uint32_t)meshes.count;
        v.materialCount = (uint32_t)materials.count;
        v.emissiveMeshes = emissiveMeshes.ptr;
        v.emissiveTriangles = emissiveTriangles.ptr;
        v.meshAliasQ = meshAliasQ.ptr;
        v.meshAliasIdx = meshAliasIdx.ptr;
        v.triAliasQ = triAliasQ.ptr;
        v.triAliasIdx = triAliasIdx.ptr;
        v.env = envWidth ? env.ptr : nullptr;
        v.envWidth = envWidth;
        v.envHeight = envHeight;
        v.envAliasQ = envAliasQ.ptr;
        v.envAliasIdx = envAliasIdx.ptr;
        v.envPdf = envPdf.ptr;
        v.rgb2spec = rgb2spec.ptr;
        v.shadeTris = shadeTris.ptr;
        v.triAlpha = alphaEnabled && triAlpha.count >= triCount ? triAlpha.ptr : nullptr;
        v.texels = texturesEnabled ? texels.ptr : nullptr;
        v.texInfo = texturesEnabled ? texInfo.ptr : nullptr;
        v.texCount = texturesEnabled ? (uint32_t)texInfo.count : 0u;
        for (int a = 0; a < 3; a++) {
            v.boxLo[a] = sceneBoxLo[a];
            v.boxHi[a] = sceneBoxHi[a];
        }
        return v;
    }
};

// traversal launch helpers shared by the integrator and the query entry points
namespace mrt {
bool ensureTraversalScratch(mrt_context* ctx);
}
