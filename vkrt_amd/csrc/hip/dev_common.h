// Device-side common definitions for the gfx950 path tracer: small vector
// math, the view of the scene the kernels read, and the wavefront queues.
//
// Floating-point policy: shading code is compiled without FMA contraction and
// uses correctly rounded division / sqrt, so that with the same random stream
// it follows the reference's arithmetic (src/shaders/**) closely enough for
// per-sample comparisons; rsqrt(x) is spelled 1/sqrt(x) for the same reason.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vkrt_shared.h"

#define DEV __device__ __forceinline__

namespace mrt {

struct F2 {
    float x, y;
};
struct F3 {
    float x, y, z;
};
struct F4 {
    float x, y, z, w;
};

DEV F3 f3(float x, float y, float z) { return F3{x, y, z}; }
DEV F3 f3(float s) { return F3{s, s, s}; }
DEV F4 f4(float x, float y, float z, float w) { return F4{x, y, z, w}; }
DEV F4 f4(float s) { return F4{s, s, s, s}; }
DEV F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEV F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DEV F3 operator/(F3 a, F3 b) { return F3{a.x / b.x, a.y / b.y, a.z / b.z}; }
DEV F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
DEV F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
DEV F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
DEV F4 operator+(F4 a, F4 b) { return F4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
DEV F4 operator*(F4 a, F4 b) { return F4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
DEV F4 operator-(F4 a, F4 b) { return F4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
DEV F4 operator*(F4 a, float s) { return F4{a.x * s, a.y * s, a.z * s, a.w * s}; }
DEV F4 operator/(F4 a, float s) { return F4{a.x / s, a.y / s, a.z / s, a.w / s}; }

DEV float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV F3 cross(F3 a, F3 b) { return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
DEV float rsq(float x) { return 1.0f / sqrtf(x); }
DEV F3 normalize(F3 a) { return a * rsq(dot(a, a)); }
DEV float sat(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
DEV F3 sat(F3 a) { return F3{sat(a.x), sat(a.y), sat(a.z)}; }
DEV F4 sat(F4 a) { return F4{sat(a.x), sat(a.y), sat(a.z), sat(a.w)}; }
DEV float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
DEV float fracf(float x) { return x - floorf(x); }
DEV float lerpf(float a, float b, float t) { return a + (b - a) * t; }
DEV F3 lerp(F3 a, F3 b, float t) { return F3{lerpf(a.x, b.x, t), lerpf(a.y, b.y, t), lerpf(a.z, b.z, t)}; }
DEV float maxc(F3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
DEV float maxc(F4 a) { return fmaxf(fmaxf(a.x, a.y), fmaxf(a.z, a.w)); }
DEV bool anyPos(F3 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }
DEV bool anyPos(F4 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f || a.w > 0.0f; }
DEV float sum(F4 a) { return a.x + a.y + a.z + a.w; }
DEV F3 maxs(F3 a, float s) { return F3{fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
DEV F3 xyz(float4 v) { return F3{v.x, v.y, v.z}; }
DEV F4 ld4(float4 v) { return F4{v.x, v.y, v.z, v.w}; }
DEV float4 st4(F4 v) { return make_float4(v.x, v.y, v.z, v.w); }
DEV float4 st4(F3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
// src/shaders/scene/constants.slang:6-15
constexpr float kRayTMin = 0.001f;
constexpr float kRayTMax = 10000.0f;
constexpr float kShadowOriginOffset = 0.001f;
constexpr float kShadowDistanceOffset = 0.002f;
constexpr float kRRMin = 0.05f;
constexpr float kRRMax = 0.95f;
constexpr float kSafeNormalizeEps = 1e-12f;

// ---------------------------------------------------------------- BVH --
// 8-wide node, 96 bytes = six 16-byte loads. Child boxes are quantised to 8
// bits per plane relative to the node's own box: a child's bound on axis a is
// origin[a] + q * 2^(e[a] - 127), with lower planes rounded down and upper
// planes rounded up when the node is built, so a decoded box always contains
// the exact one.
//   word 0..2  origin x, y, z
//   word 3     e[0] | e[1] << 8 | e[2] << 16 | childCount << 24
//   word 4..9  lower planes: 8 bytes for x, 8 for y, 8 for z
//   word 10..15 upper planes, same layout
//   word 16..23 child references (see below); slots >= childCount are unused
struct alignas(16) BvhNode {
    float ox, oy, oz;
    uint32_t packed;
    uint32_t qlo[6];
    uint32_t qhi[6];
    uint32_t child[8];
};
static_assert(sizeof(BvhNode) == 96, "BvhNode layout");
constexpr int kBvhWidth = 8;

// node reference: bit31 = leaf; leaf: bits 0..26 first triangle (BVH order),
// bits 27..30 triangle count - 1; internal: node index. kEmptyRef = empty slot.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0xFFFFFFFFu;
// Leaves hold one triangle: on the 1080p spectral workload leaves of at most
// 1, 2, 3 and 4 triangles traced 3.56, 3.48, 3.38 and 3.24 Gbounces/s (the
// tighter boxes cut triangle tests from 3.9 to 1.3 per ray for 8.4 instead of
// 7.7 node visits); the format keeps room for up to 16.
constexpr uint32_t kLeafMaxTris = 1u;
// Above kLargeScene triangles leaves take up to two: on the 20.8M-triangle
// stress workload leaves of 1, 2, 3 and 4 triangles traced 1.25, 1.30, 1.30
// and 1.27 Gbounces/s (its incoherent rays cross many small leaves).
constexpr uint32_t kLeafMaxTrisLarge = 2u;
constexpr uint32_t kLargeScene = 1u << 23;
DEV uint32_t leafFirst(uint32_t ref) { return ref & 0x07FFFFFFu; }
DEV uint32_t leafCount(uint32_t ref) { return ((ref >> 27) & 0xFu) + 1u; }

// Triangles in BVH order: three float4 (v0.xyz + global triangle id bits, v1, v2).
struct alignas(16) TriRecord {
    float4 v0;
    float4 v1;
    float4 v2;
};

// Per-triangle shading record in global triangle order, 64 bytes (one cache
// line): everything surface reconstruction needs besides the mesh transform,
// so a hit costs one line instead of gathers of the index triple and three
// 48-byte vertices. The geometric normal is computed once per scene change
// with exactly the arithmetic reconstruction would use.
struct alignas(16) ShadeTri {
    uint32_t mesh;
    float gn[3];        // world geometric normal (normalised, flipped for mirroring transforms)
    uint32_t normal[3];  // packed octahedral shading normals of the three vertices
    uint32_t tangent[3];
    uint32_t color[3];
    uint32_t pad[3];
};
static_assert(sizeof(ShadeTri) == 64, "ShadeTri layout");

// What the kernels read of the scene. Passed by value as a kernel argument.
struct SceneView {
    const BvhNode* nodes;
    const TriRecord* tris;
    uint32_t triCount;
    uint32_t nodeCount;
    const ShaderVertex* vertices;
    const uint32_t* indices;
    const uint32_t* triMesh;      // global triangle -> mesh
    const uint32_t* meshTriBase;  // mesh -> first global triangle
    const MeshInfo* meshes;
    const Material* materials;
    uint32_t meshCount;
    uint32_t materialCount;
    const EmissiveMesh* emissiveMeshes;
    const EmissiveTriangle* emissiveTriangles;
    const float* meshAliasQ;
    const uint32_t* meshAliasIdx;
    const float* triAliasQ;
    const uint32_t* triAliasIdx;
    const float4* env;
    uint32_t envWidth;
    uint32_t envHeight;
    const float* envAliasQ;
    const uint32_t* envAliasIdx;
    const float* envPdf;
    const float* rgb2spec;
    const ShadeTri* shadeTris;
    const uint32_t* triAlpha;  // per-triangle alpha words; nullptr when every mesh is opaque
    const float4* texels;      // material texture pool; nullptr when no material has a texture
    const uint4* texInfo;      // per texture: first texel, width, height
    uint32_t texCount;
    float boxLo[3], boxHi[3];  // the triangles' bounding box, padded outward (infinite before a build)
};

struct Hit {
    float t;
    uint32_t tri;  // global triangle id, kNoHit for a miss
    float u, v;
};
constexpr uint32_t kNoHit = 0xFFFFFFFFu;

}  // namespace mrt
