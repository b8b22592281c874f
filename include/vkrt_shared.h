/*
 * Byte layouts shared by the host library, the HIP kernels and the CPU oracle.
 *
 * Every struct here keeps the exact size and field offsets of the reference's
 * shared GPU layouts (src/shared/types.h:25-159) so that scene data prepared by
 * a caller of the reference API can be handed over unchanged:
 *   Vertex 80 B, ShaderVertex 48 B, MeshInfo 80 B, Material 272 B,
 *   EmissiveMesh 32 B, EmissiveTriangle 48 B, RGB2SpecTableInfo 12 B,
 *   SceneData 240 B, Selection 8 B.
 * The reference builds these from cglm types (vec4 16-byte aligned, vec3/vec2
 * unaligned); the typedefs below reproduce that alignment without cglm.
 * Constants follow src/shared/constants.h:4-76.
 */
#ifndef VKRT_SHARED_H
#define VKRT_SHARED_H

#include <stdint.h>

#if defined(__cplusplus)
#define VKRT_ALIGN16 alignas(16)
#else
#define VKRT_ALIGN16 _Alignas(16)
#endif

typedef float vkrt_f2[2];
typedef float vkrt_f3[3];
typedef float vkrt_f4[4] __attribute__((aligned(16)));
typedef uint32_t vkrt_u4[4];
typedef vkrt_f4 vkrt_m4[4]; /* column-major: m[column][row], as cglm mat4 */

/* ---- constants (src/shared/constants.h) ---- */
#define VKRT_MAX_ABSORPTION_COEFFICIENT 1000000.0f

#define VKRT_TONE_MAPPING_MODE_NONE 0u
#define VKRT_TONE_MAPPING_MODE_ACES 1u
#define VKRT_TONE_MAPPING_MODE_COUNT 2u

#define VKRT_RENDER_MODE_RGB 0u
#define VKRT_RENDER_MODE_SPECTRAL 1u
#define VKRT_RENDER_MODE_COUNT 2u

#define VKRT_SPECTRAL_SAMPLING_MODE_SINGLE 0u
#define VKRT_SPECTRAL_SAMPLING_MODE_HERO 1u
#define VKRT_SPECTRAL_SAMPLING_MODE_COUNT 2u

#define VKRT_SCENE_TONE_MAPPING_MODE_MASK 0x0000FFFFu
#define VKRT_SCENE_RENDER_MODE_BIT_SHIFT 16u
#define VKRT_SCENE_RENDER_MODE_BITS_MASK (0x000000FFu << VKRT_SCENE_RENDER_MODE_BIT_SHIFT)
#define VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BIT_SHIFT 24u
#define VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BITS_MASK (0x000000FFu << VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BIT_SHIFT)
#define VKRT_SCENE_PACK_RENDER_SETTINGS(tone, render, spectral)                     \
    (((tone) & VKRT_SCENE_TONE_MAPPING_MODE_MASK) |                                  \
     (((render) << VKRT_SCENE_RENDER_MODE_BIT_SHIFT) & VKRT_SCENE_RENDER_MODE_BITS_MASK) | \
     (((spectral) << VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BIT_SHIFT) & VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BITS_MASK))
#define VKRT_SCENE_TONE_MAPPING_MODE_FROM_RENDER_SETTINGS(v) ((v) & VKRT_SCENE_TONE_MAPPING_MODE_MASK)
#define VKRT_SCENE_RENDER_MODE_FROM_RENDER_SETTINGS(v) \
    (((v) & VKRT_SCENE_RENDER_MODE_BITS_MASK) >> VKRT_SCENE_RENDER_MODE_BIT_SHIFT)
#define VKRT_SCENE_SPECTRAL_SAMPLING_MODE_FROM_RENDER_SETTINGS(v) \
    (((v) & VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BITS_MASK) >> VKRT_SCENE_SPECTRAL_SAMPLING_MODE_BIT_SHIFT)

#define VKRT_DEBUG_MODE_NONE 0u
#define VKRT_DEBUG_MODE_NORMALS 1u
#define VKRT_DEBUG_MODE_DEPTH 2u
#define VKRT_DEBUG_MODE_BOUNCE_COUNT 3u
#define VKRT_DEBUG_MODE_NEE_ONLY 4u
#define VKRT_DEBUG_MODE_BSDF_ONLY 5u
#define VKRT_DEBUG_MODE_SELECTION_MASK 6u
#define VKRT_DEBUG_MODE_BASE_COLOR_MAP 7u
#define VKRT_DEBUG_MODE_METALLIC_MAP 8u
#define VKRT_DEBUG_MODE_ROUGHNESS_MAP 9u
#define VKRT_DEBUG_MODE_NORMAL_MAP 10u
#define VKRT_DEBUG_MODE_EMISSIVE_MAP 11u
#define VKRT_DEBUG_MODE_DENOISER_ALBEDO 12u
#define VKRT_DEBUG_MODE_DENOISER_NORMAL 13u
#define VKRT_DEBUG_MODE_DENOISER_FEATURE_VALIDITY 14u
#define VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH 15u
#define VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR 16u
#define VKRT_DEBUG_MODE_COUNT 17u

#define VKRT_INVALID_INDEX 0xFFFFFFFFu

#define VKRT_MATERIAL_ALPHA_MODE_OPAQUE 0u
#define VKRT_MATERIAL_ALPHA_MODE_MASK 1u
#define VKRT_MATERIAL_ALPHA_MODE_BLEND 2u

#define VKRT_MATERIAL_TEXTURE_SLOT_BASE_COLOR 0u
#define VKRT_MATERIAL_TEXTURE_SLOT_METALLIC_ROUGHNESS 1u
#define VKRT_MATERIAL_TEXTURE_SLOT_NORMAL 2u
#define VKRT_MATERIAL_TEXTURE_SLOT_EMISSIVE 3u
#define VKRT_MATERIAL_TEXTURE_SLOT_COUNT 4u

/* glTF sampler wrap codes; a material's packed wrap holds U in bits 0-15 and
 * V in bits 16-31 (src/shared/constants.h:72-76) */
#define VKRT_TEXTURE_WRAP_REPEAT 10497u
#define VKRT_TEXTURE_WRAP_CLAMP_TO_EDGE 33071u
#define VKRT_TEXTURE_WRAP_MIRRORED_REPEAT 33648u
#define VKRT_TEXTURE_WRAP_DEFAULT (VKRT_TEXTURE_WRAP_REPEAT | (VKRT_TEXTURE_WRAP_REPEAT << 16u))

#define VKRT_TEXTURE_COLOR_SPACE_SRGB 0u
#define VKRT_TEXTURE_COLOR_SPACE_LINEAR 1u

#define VKRT_TEXTURE_FORMAT_RGBA8_UNORM 0u
#define VKRT_TEXTURE_FORMAT_RGBA16_UNORM 1u
#define VKRT_TEXTURE_FORMAT_RGBA16_SFLOAT 2u
#define VKRT_TEXTURE_FORMAT_RGBA32_SFLOAT 3u
#define VKRT_TEXTURE_FORMAT_COUNT 4u

/* ---- layouts (src/shared/types.h:25-159) ---- */
typedef struct Vertex {
    vkrt_f4 position;
    vkrt_f4 normal;
    vkrt_f4 tangent;
    vkrt_f4 color;
    vkrt_f2 texcoord0;
    vkrt_f2 texcoord1;
} Vertex;

typedef struct ShaderVertex {
    vkrt_f4 position;
    vkrt_f2 texcoord0;
    vkrt_f2 texcoord1;
    uint32_t packedNormal;
    uint32_t packedTangent;
    uint32_t packedColor;
} ShaderVertex;

typedef struct MeshInfo {
    vkrt_f3 position;
    uint32_t vertexBase;
    vkrt_f3 rotation; /* Euler XYZ, degrees */
    uint32_t vertexCount;
    vkrt_f3 scale;
    uint32_t indexBase;
    uint32_t indexCount;
    uint32_t materialIndex;
    uint32_t renderBackfaces;
    float lightPdfArea;
    float opacity;
    uint32_t reserved0;
    uint32_t reserved1;
    uint32_t reserved2;
} MeshInfo;

typedef struct Material {
    vkrt_f3 baseColor;
    float roughness;
    vkrt_f3 emissionColor;
    float emissionLuminance;
    vkrt_f3 eta;
    float metallic;
    vkrt_f3 k;
    float anisotropic;
    float specular;
    float specularTint;
    float abbeNumber;
    float reserved0;
    vkrt_f4 sheenTintWeight;
    float clearcoat;
    float clearcoatGloss;
    float ior;
    float diffuseRoughness;
    float transmission;
    float subsurface;
    float sheenRoughness;
    float absorptionCoefficient;
    vkrt_f3 attenuationColor;
    float normalTextureScale;
    uint32_t baseColorTextureIndex;
    uint32_t metallicRoughnessTextureIndex;
    uint32_t normalTextureIndex;
    uint32_t emissiveTextureIndex;
    uint32_t baseColorTextureWrap;
    uint32_t metallicRoughnessTextureWrap;
    uint32_t normalTextureWrap;
    uint32_t emissiveTextureWrap;
    float opacity;
    float alphaCutoff;
    uint32_t alphaMode;
    uint32_t textureTexcoordSets;
    vkrt_f4 baseColorTextureTransform;
    vkrt_f4 metallicRoughnessTextureTransform;
    vkrt_f4 normalTextureTransform;
    vkrt_f4 emissiveTextureTransform;
    vkrt_f4 textureRotations;
} Material;

typedef struct EmissiveMesh {
    uint32_t triOffset;
    uint32_t triCount;
    float pmfMesh;
    float invTotalArea;
    vkrt_f3 emission;
    float reserved0;
} EmissiveMesh;

typedef struct EmissiveTriangle {
    vkrt_f4 v0Area;
    vkrt_f4 e1Pad;
    vkrt_f4 e2Pad;
} EmissiveTriangle;

typedef struct RGB2SpecTableInfo {
    uint32_t res;
    uint32_t scaleOffset;
    uint32_t dataOffset;
} RGB2SpecTableInfo;

typedef struct SceneData {
    vkrt_m4 viewInverse;
    vkrt_m4 projInverse;
    uint32_t frameNumber;
    uint32_t samplesPerPixel;
    uint32_t rrMaxDepth;
    uint32_t rrMinDepth;
    vkrt_u4 viewportRect;
    uint32_t packedRenderSettings;
    float exposure;
    float timeBase;
    float timeStep;
    vkrt_f4 environmentLight;
    uint32_t environmentTextureIndex;
    float environmentRotation;
    uint32_t debugMode;
    uint32_t misNeeEnabled;
    uint32_t emissiveMeshCount;
    uint32_t emissiveTriangleCount;
    uint32_t selectionEnabled;
    uint32_t selectedMeshIndex;
    RGB2SpecTableInfo rgb2specSRGB;
} SceneData;

typedef struct Selection {
    uint32_t pixel;
    uint32_t hitMeshIndex;
} Selection;

#if defined(__cplusplus)
static_assert(sizeof(Vertex) == 80, "Vertex layout");
static_assert(sizeof(ShaderVertex) == 48, "ShaderVertex layout");
static_assert(sizeof(MeshInfo) == 80, "MeshInfo layout");
static_assert(sizeof(Material) == 272, "Material layout");
static_assert(sizeof(EmissiveMesh) == 32, "EmissiveMesh layout");
static_assert(sizeof(EmissiveTriangle) == 48, "EmissiveTriangle layout");
static_assert(sizeof(SceneData) == 240, "SceneData layout");
#else
_Static_assert(sizeof(Vertex) == 80, "Vertex layout");
_Static_assert(sizeof(ShaderVertex) == 48, "ShaderVertex layout");
_Static_assert(sizeof(MeshInfo) == 80, "MeshInfo layout");
_Static_assert(sizeof(Material) == 272, "Material layout");
_Static_assert(sizeof(EmissiveMesh) == 32, "EmissiveMesh layout");
_Static_assert(sizeof(EmissiveTriangle) == 48, "EmissiveTriangle layout");
_Static_assert(sizeof(SceneData) == 240, "SceneData layout");
#endif

#endif
