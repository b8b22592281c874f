"""Thin Python binding of the MI355X renderer's C ABI (include/vkrt.h).

The library is libvkrt_amd.so in this directory, built by `make -C vkrt_amd`
(or `__graft_entry__.build()`). Loading fails loudly when it is missing: there
is no fallback path. Structures mirror include/vkrt_shared.h and
include/vkrt_types.h byte for byte (sizes are asserted at import).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvkrt_amd.so")

VKRT_SUCCESS = 0
VKRT_INVALID_INDEX = 0xFFFFFFFF
VKRT_ERROR_INVALID_ARGUMENT = -1
VKRT_ERROR_OPERATION_FAILED = -2
VKRT_ERROR_OUT_OF_MEMORY = -3
VKRT_ERROR_DEVICE_LOST = -4
VKRT_ERROR_INITIALIZATION_FAILED = -5
RENDER_MODE_RGB = 0
RENDER_MODE_SPECTRAL = 1
SPECTRAL_SINGLE = 0
SPECTRAL_HERO = 1
TONE_NONE = 0
TONE_ACES = 1
RNG_HASH = 0
RNG_PCG32 = 1
TEXTURE_FORMAT_RGBA8_UNORM = 0
TEXTURE_FORMAT_RGBA16_UNORM = 1
TEXTURE_FORMAT_RGBA16_SFLOAT = 2
TEXTURE_FORMAT_RGBA32_SFLOAT = 3
TEXTURE_COLOR_SPACE_SRGB = 0
TEXTURE_WRAP_REPEAT = 10497
TEXTURE_WRAP_CLAMP_TO_EDGE = 33071
TEXTURE_WRAP_MIRRORED_REPEAT = 33648
TEXTURE_COLOR_SPACE_LINEAR = 1
TEXTURE_WRAP_DEFAULT = TEXTURE_WRAP_REPEAT | (TEXTURE_WRAP_REPEAT << 16)
MATERIAL_TEXTURE_SLOT_BASE_COLOR = 0
MATERIAL_TEXTURE_SLOT_METALLIC_ROUGHNESS = 1
MATERIAL_TEXTURE_SLOT_NORMAL = 2
MATERIAL_TEXTURE_SLOT_EMISSIVE = 3
VKRT_DEBUG_MODE_NONE = 0
VKRT_DEBUG_MODE_NORMALS = 1
VKRT_DEBUG_MODE_DEPTH = 2
VKRT_DEBUG_MODE_BOUNCE_COUNT = 3
VKRT_DEBUG_MODE_NEE_ONLY = 4
VKRT_DEBUG_MODE_BSDF_ONLY = 5
VKRT_DEBUG_MODE_SELECTION_MASK = 6
VKRT_DEBUG_MODE_BASE_COLOR_MAP = 7
VKRT_DEBUG_MODE_METALLIC_MAP = 8
VKRT_DEBUG_MODE_ROUGHNESS_MAP = 9
VKRT_DEBUG_MODE_NORMAL_MAP = 10
VKRT_DEBUG_MODE_EMISSIVE_MAP = 11
VKRT_DEBUG_MODE_DENOISER_ALBEDO = 12
VKRT_DEBUG_MODE_DENOISER_NORMAL = 13
VKRT_DEBUG_MODE_DENOISER_FEATURE_VALIDITY = 14
VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH = 15
VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR = 16
RENDER_PHASE_INACTIVE = 0
RENDER_PHASE_SAMPLING = 1
RENDER_PHASE_DENOISING = 2
RENDER_PHASE_COMPLETE_RAW = 3
RENDER_PHASE_COMPLETE_DENOISED = 4


class Vertex(C.Structure):
    _fields_ = [
        ("position", C.c_float * 4),
        ("normal", C.c_float * 4),
        ("tangent", C.c_float * 4),
        ("color", C.c_float * 4),
        ("texcoord0", C.c_float * 2),
        ("texcoord1", C.c_float * 2),
    ]


class ShaderVertex(C.Structure):
    _fields_ = [
        ("position", C.c_float * 4),
        ("texcoord0", C.c_float * 2),
        ("texcoord1", C.c_float * 2),
        ("packedNormal", C.c_uint32),
        ("packedTangent", C.c_uint32),
        ("packedColor", C.c_uint32),
        ("_pad", C.c_uint32),
    ]


class MeshInfo(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("vertexBase", C.c_uint32),
        ("rotation", C.c_float * 3),
        ("vertexCount", C.c_uint32),
        ("scale", C.c_float * 3),
        ("indexBase", C.c_uint32),
        ("indexCount", C.c_uint32),
        ("materialIndex", C.c_uint32),
        ("renderBackfaces", C.c_uint32),
        ("lightPdfArea", C.c_float),
        ("opacity", C.c_float),
        ("reserved0", C.c_uint32),
        ("reserved1", C.c_uint32),
        ("reserved2", C.c_uint32),
    ]


class Material(C.Structure):
    _fields_ = [
        ("baseColor", C.c_float * 3),
        ("roughness", C.c_float),
        ("emissionColor", C.c_float * 3),
        ("emissionLuminance", C.c_float),
        ("eta", C.c_float * 3),
        ("metallic", C.c_float),
        ("k", C.c_float * 3),
        ("anisotropic", C.c_float),
        ("specular", C.c_float),
        ("specularTint", C.c_float),
        ("abbeNumber", C.c_float),
        ("reserved0", C.c_float),
        ("sheenTintWeight", C.c_float * 4),
        ("clearcoat", C.c_float),
        ("clearcoatGloss", C.c_float),
        ("ior", C.c_float),
        ("diffuseRoughness", C.c_float),
        ("transmission", C.c_float),
        ("subsurface", C.c_float),
        ("sheenRoughness", C.c_float),
        ("absorptionCoefficient", C.c_float),
        ("attenuationColor", C.c_float * 3),
        ("normalTextureScale", C.c_float),
        ("baseColorTextureIndex", C.c_uint32),
        ("metallicRoughnessTextureIndex", C.c_uint32),
        ("normalTextureIndex", C.c_uint32),
        ("emissiveTextureIndex", C.c_uint32),
        ("baseColorTextureWrap", C.c_uint32),
        ("metallicRoughnessTextureWrap", C.c_uint32),
        ("normalTextureWrap", C.c_uint32),
        ("emissiveTextureWrap", C.c_uint32),
        ("opacity", C.c_float),
        ("alphaCutoff", C.c_float),
        ("alphaMode", C.c_uint32),
        ("textureTexcoordSets", C.c_uint32),
        ("baseColorTextureTransform", C.c_float * 4),
        ("metallicRoughnessTextureTransform", C.c_float * 4),
        ("normalTextureTransform", C.c_float * 4),
        ("emissiveTextureTransform", C.c_float * 4),
        ("textureRotations", C.c_float * 4),
    ]


class EmissiveMesh(C.Structure):
    _fields_ = [
        ("triOffset", C.c_uint32),
        ("triCount", C.c_uint32),
        ("pmfMesh", C.c_float),
        ("invTotalArea", C.c_float),
        ("emission", C.c_float * 3),
        ("reserved0", C.c_float),
    ]


class EmissiveTriangle(C.Structure):
    _fields_ = [("v0Area", C.c_float * 4), ("e1Pad", C.c_float * 4), ("e2Pad", C.c_float * 4)]


class RGB2SpecTableInfo(C.Structure):
    _fields_ = [("res", C.c_uint32), ("scaleOffset", C.c_uint32), ("dataOffset", C.c_uint32)]


class SceneData(C.Structure):
    _fields_ = [
        ("viewInverse", C.c_float * 16),
        ("projInverse", C.c_float * 16),
        ("frameNumber", C.c_uint32),
        ("samplesPerPixel", C.c_uint32),
        ("rrMaxDepth", C.c_uint32),
        ("rrMinDepth", C.c_uint32),
        ("viewportRect", C.c_uint32 * 4),
        ("packedRenderSettings", C.c_uint32),
        ("exposure", C.c_float),
        ("timeBase", C.c_float),
        ("timeStep", C.c_float),
        ("environmentLight", C.c_float * 4),
        ("environmentTextureIndex", C.c_uint32),
        ("environmentRotation", C.c_float),
        ("debugMode", C.c_uint32),
        ("misNeeEnabled", C.c_uint32),
        ("emissiveMeshCount", C.c_uint32),
        ("emissiveTriangleCount", C.c_uint32),
        ("selectionEnabled", C.c_uint32),
        ("selectedMeshIndex", C.c_uint32),
        ("rgb2specSRGB", RGB2SpecTableInfo),
        ("_pad", C.c_uint32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("target", C.c_float * 3),
        ("up", C.c_float * 3),
        ("nearZ", C.c_float),
        ("farZ", C.c_float),
        ("vfov", C.c_float),
    ]


class SceneSettingsSnapshot(C.Structure):
    _fields_ = [
        ("camera", Camera),
        ("samplesPerPixel", C.c_uint32),
        ("rrMaxDepth", C.c_uint32),
        ("rrMinDepth", C.c_uint32),
        ("toneMappingMode", C.c_uint32),
        ("renderMode", C.c_uint32),
        ("spectralSamplingMode", C.c_uint32),
        ("exposure", C.c_float),
        ("autoExposureEnabled", C.c_uint8),
        ("autoSPPEnabled", C.c_uint8),
        ("autoSPPTargetFPS", C.c_uint32),
        ("environmentColor", C.c_float * 3),
        ("environmentStrength", C.c_float),
        ("environmentRotation", C.c_float),
        ("environmentTextureIndex", C.c_uint32),
        ("timeBase", C.c_float),
        ("timeStep", C.c_float),
        ("debugMode", C.c_uint32),
        ("misNeeEnabled", C.c_uint32),
        ("selectionEnabled", C.c_uint32),
        ("selectedMeshIndex", C.c_uint32),
    ]


class RenderStatusSnapshot(C.Structure):
    _fields_ = [
        ("framesPerSecond", C.c_uint32),
        ("averageFrametime", C.c_float),
        ("frametimes", C.c_float * 128),
        ("displayTimeMs", C.c_float),
        ("renderTimeMs", C.c_float),
        ("accumulationFrame", C.c_uint32),
        ("totalSamples", C.c_uint64),
        ("renderPhase", C.c_int),
        ("renderDenoiseEnabled", C.c_uint8),
        ("renderTargetSamples", C.c_uint32),
        ("displayRenderTimeMs", C.c_float),
        ("displayFrameTimeMs", C.c_float),
    ]


class CreateInfo(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("title", C.c_char_p),
        ("startMaximized", C.c_uint8),
        ("startFullscreen", C.c_uint8),
        ("headless", C.c_uint8),
        ("disableSER", C.c_uint8),
        ("preferredDeviceIndex", C.c_int32),
        ("preferredDeviceName", C.c_char_p),
    ]


class MeshUpload(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(Vertex)),
        ("vertexCount", C.c_size_t),
        ("indices", C.POINTER(C.c_uint32)),
        ("indexCount", C.c_size_t),
    ]


class TextureUpload(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("pixels", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("format", C.c_uint32),
        ("colorSpace", C.c_uint32),
    ]


class SystemInfo(C.Structure):
    _fields_ = [("deviceName", C.c_char * 256), ("vendorID", C.c_uint32), ("driverVersion", C.c_uint32)]


class CameraInput(C.Structure):
    _fields_ = [("orbitDx", C.c_float), ("orbitDy", C.c_float), ("panDx", C.c_float), ("panDy", C.c_float),
                ("scroll", C.c_float), ("orbiting", C.c_uint8), ("panning", C.c_uint8), ("captureMouse", C.c_uint8)]


class RuntimeSnapshot(C.Structure):
    _fields_ = [("displayWidth", C.c_uint32), ("displayHeight", C.c_uint32), ("swapchainWidth", C.c_uint32),
                ("swapchainHeight", C.c_uint32), ("renderWidth", C.c_uint32), ("renderHeight", C.c_uint32),
                ("displayViewportRect", C.c_uint32 * 4), ("presentMode", C.c_uint32), ("displayRefreshHz", C.c_float)]


class TextureSnapshot(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32), ("colorSpace", C.c_uint32),
                ("useCount", C.c_uint32), ("name", C.c_char * 256)]


class TraceStats(C.Structure):
    _fields_ = [
        ("extensionRays", C.c_uint64),
        ("shadowRays", C.c_uint64),
        ("paths", C.c_uint64),
        ("traceMs", C.c_float),
        ("extendMs", C.c_float),
        ("shadowMs", C.c_float),
        ("shadeMs", C.c_float),
        ("extendLaunches", C.c_uint32),
        ("bvhBuildMs", C.c_float),
        ("bvhNodeCount", C.c_uint32),
        ("bvhLeafTriangles", C.c_uint32),
        ("bvhMaxDepth", C.c_uint32),
        ("triangleCount", C.c_uint32),
        ("shadeLaunches", C.c_uint32),
        ("shadowLaunches", C.c_uint32),
        ("nodeVisits", C.c_uint64),
        ("triangleTests", C.c_uint64),
        ("backwardMs", C.c_float),
        ("extendMsDepth0", C.c_float),
        ("shadeMsDepth0", C.c_float),
        ("tailStartDepth", C.c_uint32),
        ("tailLaunches", C.c_uint32),
        ("tailMs", C.c_float),
        ("tailRays", C.c_uint64),
        ("raysBeforeTail", C.c_uint64),
        ("nodeVisitsBeforeTail", C.c_uint64),
        ("triangleTestsBeforeTail", C.c_uint64),
    ]


class MaterialSnapshot(C.Structure):
    # include/vkrt_types.h: 16-byte aligned like Material's vector members, 544 bytes
    _fields_ = [("material", Material), ("useCount", C.c_uint32), ("name", C.c_char * 256), ("_pad", C.c_uint8 * 12)]


class NeuralCacheSettings(C.Structure):
    _fields_ = [
        ("enabled", C.c_uint32),
        ("prefixBounces", C.c_uint32),
        ("trainingPaths", C.c_uint32),
        ("trainingSteps", C.c_uint32),
        ("batchSize", C.c_uint32),
        ("learningRate", C.c_float),
        ("train", C.c_uint32),
    ]


class NeuralCacheStats(C.Structure):
    _fields_ = [
        ("lastLoss", C.c_float),
        ("queries", C.c_uint32),
        ("trainingRecords", C.c_uint32),
        ("inferMs", C.c_float),
        ("trainMs", C.c_float),
    ]


class Gradients(C.Structure):
    _fields_ = [
        ("baseColor", C.POINTER(C.c_float)),
        ("roughness", C.POINTER(C.c_float)),
        ("emission", C.POINTER(C.c_float)),
        ("materialCount", C.c_uint32),
        ("environment", C.POINTER(C.c_float)),
        ("environmentWidth", C.c_uint32),
        ("environmentHeight", C.c_uint32),
    ]


class SceneBuffers(C.Structure):
    _fields_ = [
        ("sceneData", C.POINTER(SceneData)),
        ("vertices", C.POINTER(ShaderVertex)),
        ("vertexCount", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)),
        ("indexCount", C.c_uint32),
        ("meshInfos", C.POINTER(MeshInfo)),
        ("meshCount", C.c_uint32),
        ("materials", C.POINTER(Material)),
        ("materialCount", C.c_uint32),
        ("emissiveMeshes", C.POINTER(EmissiveMesh)),
        ("emissiveTriangles", C.POINTER(EmissiveTriangle)),
        ("meshAliasQ", C.POINTER(C.c_float)),
        ("meshAliasIdx", C.POINTER(C.c_uint32)),
        ("triAliasQ", C.POINTER(C.c_float)),
        ("triAliasIdx", C.POINTER(C.c_uint32)),
        ("rgb2specTable", C.POINTER(C.c_float)),
        ("rgb2specFloatCount", C.c_uint32),
        ("environmentPixels", C.POINTER(C.c_float)),
        ("environmentWidth", C.c_uint32),
        ("environmentHeight", C.c_uint32),
        ("worldTriangles", C.POINTER(C.c_float)),
        ("triangleCount", C.c_uint32),
        ("triangleMesh", C.POINTER(C.c_uint32)),
        ("rngPolicy", C.c_uint32),
        ("rngSeed", C.c_uint64),
        ("environmentImportanceSampling", C.c_uint32),
        ("environmentAliasQ", C.POINTER(C.c_float)),
        ("environmentAliasIdx", C.POINTER(C.c_uint32)),
        ("environmentPdf", C.POINTER(C.c_float)),
        ("triangleAlpha", C.POINTER(C.c_uint32)),
        ("textureTexels", C.POINTER(C.c_float)),
        ("textureTexelCount", C.c_uint64),
        ("textureInfo", C.POINTER(C.c_uint32)),
        ("textureCount", C.c_uint32),
    ]


class MeshSnapshot(C.Structure):
    _fields_ = [
        ("info", MeshInfo),
        ("material", Material),
        ("materialIndex", C.c_uint32),
        ("geometrySource", C.c_uint32),
        ("hasMaterialAssignment", C.c_uint8),
        ("ownsGeometry", C.c_uint8),
        ("name", C.c_char * 256),
        ("_tail", C.c_uint8 * 6),  # the C struct is padded to Material's 16-byte alignment
    ]


class RenderExportSettings(C.Structure):
    _fields_ = [("denoiseEnabled", C.c_uint8)]


# include/vkrt_import.h (src/app/mesh/loader.h:5-55)
class MaterialImportEntry(C.Structure):
    # Material holds 16-byte aligned vec4 members, so it starts at offset 16
    _fields_ = [("name", C.c_char_p), ("_pad", C.c_uint8 * 8), ("material", Material)]


class TextureImportEntry(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("pixels", C.c_void_p),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("format", C.c_uint32),
        ("colorSpace", C.c_uint32),
    ]


class MeshImportEntry(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("vertices", C.POINTER(Vertex)),
        ("indices", C.POINTER(C.c_uint32)),
        ("vertexCount", C.c_size_t),
        ("indexCount", C.c_size_t),
        ("nodeIndex", C.c_uint32),
        ("materialIndex", C.c_uint32),
        ("renderBackfaces", C.c_uint8),
        ("position", C.c_float * 3),
        ("rotation", C.c_float * 3),
        ("scale", C.c_float * 3),
    ]


class NodeImportEntry(C.Structure):
    _fields_ = [
        ("localTransform", C.c_float * 16),
        ("name", C.c_char_p),
        ("parentIndex", C.c_uint32),
        ("meshEntryCount", C.c_uint32),
        ("position", C.c_float * 3),
        ("rotation", C.c_float * 3),
        ("scale", C.c_float * 3),
    ]


class MeshImportData(C.Structure):
    _fields_ = [
        ("entries", C.POINTER(MeshImportEntry)),
        ("count", C.c_uint32),
        ("entryCapacity", C.c_uint32),
        ("nodes", C.POINTER(NodeImportEntry)),
        ("nodeCount", C.c_uint32),
        ("nodeCapacity", C.c_uint32),
        ("materials", C.POINTER(MaterialImportEntry)),
        ("materialCount", C.c_uint32),
        ("textures", C.POINTER(TextureImportEntry)),
        ("textureCount", C.c_uint32),
        ("textureCapacity", C.c_uint32),
    ]


_EXPECTED_SIZES = {
    Vertex: 80,
    ShaderVertex: 48,
    MeshInfo: 80,
    Material: 272,
    EmissiveMesh: 32,
    EmissiveTriangle: 48,
    SceneData: 240,
    MaterialSnapshot: 544,
    NeuralCacheSettings: 28,
    NeuralCacheStats: 20,
    MeshSnapshot: 624,
    MaterialImportEntry: 288,
    TextureImportEntry: 32,
    MeshImportEntry: 88,
    NodeImportEntry: 120,
    MeshImportData: 64,
    CameraInput: 24,
    RuntimeSnapshot: 48,
    TextureSnapshot: 276,
}
for _cls, _size in _EXPECTED_SIZES.items():
    assert C.sizeof(_cls) == _size, f"{_cls.__name__} is {C.sizeof(_cls)} bytes, expected {_size}"

_lib = None


def library() -> C.CDLL:
    """Loads libvkrt_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}`")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32, u64, f32, sz = C.c_uint32, C.c_uint64, C.c_float, C.c_size_t
    fp = C.POINTER(C.c_float)
    up = C.POINTER(C.c_uint32)
    sig = {
        "VKRT_defaultCreateInfo": (None, [C.POINTER(CreateInfo)]),
        "VKRT_create": (C.c_int, [C.POINTER(P)]),
        "VKRT_destroy": (None, [P]),
        "VKRT_initWithCreateInfo": (C.c_int, [P, C.POINTER(CreateInfo)]),
        "VKRT_init": (C.c_int, [P]),
        "VKRT_deinit": (None, [P]),
        "VKRT_draw": (C.c_int, [P]),
        "VKRT_beginFrame": (C.c_int, [P]),
        "VKRT_updateScene": (C.c_int, [P]),
        "VKRT_trace": (C.c_int, [P]),
        "VKRT_present": (C.c_int, [P]),
        "VKRT_endFrame": (C.c_int, [P]),
        "VKRT_uploadMeshData": (C.c_int, [P, C.POINTER(Vertex), sz, up, sz]),
        "VKRT_uploadMeshDataBatch": (C.c_int, [P, C.POINTER(MeshUpload), sz]),
        "VKRT_removeMesh": (C.c_int, [P, u32]),
        "VKRT_invalidateAccumulation": (C.c_int, [P]),
        "VKRT_setSamplesPerPixel": (C.c_int, [P, u32]),
        "VKRT_setPathDepth": (C.c_int, [P, u32, u32]),
        "VKRT_setAutoSPPEnabled": (C.c_int, [P, C.c_uint8]),
        "VKRT_setToneMappingMode": (C.c_int, [P, u32]),
        "VKRT_setRenderMode": (C.c_int, [P, u32]),
        "VKRT_setSpectralSamplingMode": (C.c_int, [P, u32]),
        "VKRT_setExposure": (C.c_int, [P, f32]),
        "VKRT_setEnvironmentLight": (C.c_int, [P, fp, f32]),
        "VKRT_setEnvironmentRotation": (C.c_int, [P, f32]),
        "VKRT_clearEnvironmentTexture": (C.c_int, [P]),
        "VKRT_setDebugMode": (C.c_int, [P, u32]),
        "VKRT_setMisNeeEnabled": (C.c_int, [P, C.c_uint8]),
        "VKRT_startRender": (C.c_int, [P, u32, u32, u32]),
        "VKRT_continueRender": (C.c_int, [P, u32]),
        "VKRT_stopRenderSampling": (C.c_int, [P]),
        "VKRT_stopRender": (C.c_int, [P]),
        "VKRT_getSceneSettings": (C.c_int, [P, C.POINTER(SceneSettingsSnapshot)]),
        "VKRT_getRenderStatus": (C.c_int, [P, C.POINTER(RenderStatusSnapshot)]),
        "VKRT_getSystemInfo": (C.c_int, [P, C.POINTER(SystemInfo)]),
        "VKRT_getMeshCount": (C.c_int, [P, up]),
        "VKRT_getMaterialCount": (C.c_int, [P, up]),
        "VKRT_getTextureCount": (C.c_int, [P, up]),
        "VKRT_addTextureFromPixels": (C.c_int, [P, C.POINTER(TextureUpload), up]),
        "VKRT_removeTexture": (C.c_int, [P, u32]),
        "VKRT_setMaterialSorting": (C.c_int, [P, C.c_uint8]),
        "VKRT_setFusedTail": (C.c_int, [P, C.c_uint32]),
        "VKRT_defaultNeuralCacheSettings": (None, [C.POINTER(NeuralCacheSettings)]),
        "VKRT_setNeuralCache": (C.c_int, [P, C.POINTER(NeuralCacheSettings)]),
        "VKRT_getNeuralCacheStats": (C.c_int, [P, C.POINTER(NeuralCacheStats)]),
        "VKRT_neuralCacheParameterCount": (u32, []),
        "VKRT_neuralCacheReset": (C.c_int, [P, u32]),
        "VKRT_neuralCacheGetWeights": (C.c_int, [P, fp, sz]),
        "VKRT_neuralCacheSetWeights": (C.c_int, [P, fp, sz]),
        "VKRT_neuralCacheInfer": (C.c_int, [P, fp, u32, fp]),
        "VKRT_neuralCacheTrain": (C.c_int, [P, fp, fp, u32, u32, u32, f32, fp]),
        "VKRT_neuralCacheGradient": (C.c_int, [P, fp, fp, u32, fp, fp]),
        "VKRT_addMaterial": (C.c_int, [P, C.POINTER(Material), C.c_char_p, up]),
        "VKRT_setMaterial": (C.c_int, [P, u32, C.POINTER(Material)]),
        "VKRT_setMeshMaterialIndex": (C.c_int, [P, u32, u32]),
        "VKRT_setMeshTransform": (C.c_int, [P, u32, fp, fp, fp]),
        "VKRT_setMeshTransformMatrix": (C.c_int, [P, u32, fp]),
        "VKRT_setRenderViewport": (C.c_int, [P, u32, u32, u32, u32]),
        "VKRT_cameraSetPose": (C.c_int, [P, fp, fp, fp, f32]),
        "VKRT_cameraGetPose": (C.c_int, [P, fp, fp, fp, fp]),
        "VKRT_buildMeshTransformMatrix": (None, [fp, fp, fp, fp]),
        "VKRT_decomposeMeshTransform": (None, [fp, fp, fp, fp]),
        "VKRT_setRandomPolicy": (C.c_int, [P, u32, u64]),
        "VKRT_setEnvironmentImportanceSampling": (C.c_int, [P, C.c_uint8]),
        "VKRT_setEnvironmentTexture": (C.c_int, [P, u32]),
        "VKRT_readAccumulation": (C.c_int, [P, fp, sz]),
        "VKRT_readOutputImage": (C.c_int, [P, C.POINTER(C.c_uint16), sz]),
        "VKRT_getTraceStats": (C.c_int, [P, C.POINTER(TraceStats)]),
        "VKRT_exportSceneBuffers": (C.c_int, [P, C.POINTER(SceneBuffers)]),
        "VKRT_traceRays": (C.c_int, [P, fp, sz, up, up, fp, fp]),
        "VKRT_traceShadowRays": (C.c_int, [P, fp, sz, up]),
        "VKRT_traceRaysDeviceNoAlloc": (C.c_int, [P, P, sz, P, P, P, P]),
        "VKRT_tracePrimaryHits": (C.c_int, [P, u32, up, up]),
        "VKRT_setTraversalStatistics": (C.c_int, [P, C.c_uint8]),
        "VKRT_deviceSynchronize": (C.c_int, [P]),
        "VKRT_renderBackward": (C.c_int, [P, fp, C.POINTER(Gradients)]),
        "VKRT_getMaterialSnapshot": (C.c_int, [P, u32, C.POINTER(MaterialSnapshot)]),
        "VKRT_getMeshSnapshot": (C.c_int, [P, u32, C.POINTER(MeshSnapshot)]),
        "VKRT_setMeshName": (C.c_int, [P, u32, C.c_char_p]),
        "VKRT_removeMaterial": (C.c_int, [P, u32]),
        "VKRT_setMaterialName": (C.c_int, [P, u32, C.c_char_p]),
        "VKRT_clearMeshMaterialAssignment": (C.c_int, [P, u32]),
        "VKRT_setMeshOpacity": (C.c_int, [P, u32, f32]),
        "VKRT_setMeshRenderBackfaces": (C.c_int, [P, u32, u32]),
        "VKRT_addTexturesBatch": (C.c_int, [P, C.POINTER(TextureUpload), sz, up]),
        "VKRT_buildImportedNodeTransform": (None, [fp, fp]),
        "VKRT_decomposeMeshNodeTransform": (None, [fp, fp, fp, fp]),
        "meshLoadFromFile": (C.c_int, [C.c_char_p, C.POINTER(MeshImportData)]),
        "meshReleaseImportData": (None, [C.POINTER(MeshImportData)]),
        "VKRT_importMesh": (C.c_int, [P, C.c_char_p, C.c_char_p, up, up]),
        "VKRT_loadSceneFile": (C.c_int, [P, C.c_char_p]),
        "VKRT_defaultRenderExportSettings": (None, [C.POINTER(RenderExportSettings)]),
        "VKRT_saveRenderImageEx": (C.c_int, [P, C.c_char_p, C.POINTER(RenderExportSettings)]),
        "VKRT_saveRenderImage": (C.c_int, [P, C.c_char_p]),
        "VKRT_applyCameraInput": (C.c_int, [P, C.POINTER(CameraInput)]),
        "VKRT_setEnvironmentTextureFromFile": (C.c_int, [P, C.c_char_p]),
        "VKRT_setRenderDenoiseEnabled": (C.c_int, [P, C.c_uint8]),
        "VKRT_denoiseRenderToViewport": (C.c_int, [P]),
        "VKRT_getRuntimeSnapshot": (C.c_int, [P, C.POINTER(RuntimeSnapshot)]),
        "VKRT_getTextureSnapshot": (C.c_int, [P, u32, C.POINTER(TextureSnapshot)]),
        "VKRT_addTextureFromFile": (C.c_int, [P, C.c_char_p, C.c_char_p, u32, up]),
        "VKRT_setMaterialTexture": (C.c_int, [P, u32, u32, u32]),
        "VKRT_setDenoiserFeatures": (C.c_int, [P, C.c_uint8]),
        "VKRT_readDenoiserFeatures": (C.c_int, [P, fp, fp, sz]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def load_gltf(path: str) -> dict:
    """meshLoadFromFile through the C ABI, copied into numpy arrays.

    Returns {"entries": [{"name", "vertices" (N x 20 float32, Vertex layout),
    "indices", "node", "material", "renderBackfaces"}], "nodes": [{"name",
    "parent", "local" (4 x 4, column-major rows = columns)}], "materials":
    [{"name", "material"}]}. Raises VkrtError if the file cannot be imported.
    """
    lib = library()
    data = MeshImportData()
    if lib.meshLoadFromFile(os.fsencode(path), C.byref(data)) != 0:
        raise VkrtError(f"meshLoadFromFile failed for {path}")
    try:
        entries = []
        for i in range(data.count):
            e = data.entries[i]
            verts = np.ctypeslib.as_array(C.cast(e.vertices, C.POINTER(C.c_float)), (e.vertexCount * 20,))
            idx = np.ctypeslib.as_array(e.indices, (e.indexCount,))
            entries.append({"name": e.name.decode() if e.name else "", "vertices": verts.reshape(-1, 20).copy(),
                            "indices": idx.copy(), "node": e.nodeIndex, "material": e.materialIndex,
                            "renderBackfaces": bool(e.renderBackfaces)})
        nodes = []
        for i in range(data.nodeCount):
            n = data.nodes[i]
            nodes.append({"name": n.name.decode() if n.name else "", "parent": n.parentIndex,
                          "local": np.array(n.localTransform[:], np.float32).reshape(4, 4),
                          "meshEntryCount": n.meshEntryCount})
        materials = [{"name": data.materials[i].name.decode(), "material": Material.from_buffer_copy(
            data.materials[i].material)} for i in range(data.materialCount)]
        return {"entries": entries, "nodes": nodes, "materials": materials, "textureCount": data.textureCount}
    finally:
        lib.meshReleaseImportData(C.byref(data))


def fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class VkrtError(RuntimeError):
    pass


def _ok(code: int, what: str):
    if code != VKRT_SUCCESS:
        raise VkrtError(f"{what} failed with VKRT_Result {code}")


def default_material(**kw) -> Material:
    """The reference's VKRT_materialDefault (src/core/api/vkrt_types.h:200-242) with overrides."""
    m = Material()
    m.baseColor[:] = (0.8, 0.8, 0.8)
    m.roughness = 0.5
    m.emissionColor[:] = (1.0, 1.0, 1.0)
    m.specular = 0.5
    m.sheenTintWeight[:] = (1.0, 1.0, 1.0, 0.0)
    m.clearcoatGloss = 1.0
    m.ior = 1.5
    m.sheenRoughness = 0.5
    m.attenuationColor[:] = (1.0, 1.0, 1.0)
    m.normalTextureScale = 1.0
    m.baseColorTextureIndex = VKRT_INVALID_INDEX
    m.metallicRoughnessTextureIndex = VKRT_INVALID_INDEX
    m.normalTextureIndex = VKRT_INVALID_INDEX
    m.emissiveTextureIndex = VKRT_INVALID_INDEX
    m.opacity = 1.0
    m.alphaCutoff = 0.5
    for t in ("baseColorTextureTransform", "metallicRoughnessTextureTransform", "normalTextureTransform",
              "emissiveTextureTransform"):
        getattr(m, t)[:] = (1.0, 1.0, 0.0, 0.0)
    for key, value in kw.items():
        field = getattr(m, key)
        if isinstance(field, (int, float)):
            setattr(m, key, value)
        else:
            field[:] = value
    return m


def make_vertices(positions: np.ndarray, normals: np.ndarray, colors: np.ndarray | None = None,
                  texcoord0: np.ndarray | None = None, texcoord1: np.ndarray | None = None) -> np.ndarray:
    """Packs positions/normals (N x 3) into an array of the reference's 80-byte Vertex."""
    n = len(positions)
    v = np.zeros((n, 20), dtype=np.float32)
    v[:, 0:3] = positions
    v[:, 3] = 1.0
    v[:, 4:7] = normals
    # tangent: any unit vector orthogonal to the normal, handedness +1
    nrm = normals.astype(np.float64)
    helper = np.where(np.abs(nrm[:, 2:3]) < 0.999, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(helper, nrm)
    t /= np.maximum(np.linalg.norm(t, axis=1, keepdims=True), 1e-20)
    v[:, 8:11] = t
    v[:, 11] = 1.0
    v[:, 12:16] = 1.0 if colors is None else colors
    if texcoord0 is not None:
        v[:, 16:18] = texcoord0
    if texcoord1 is not None:
        v[:, 18:20] = texcoord1
    return np.ascontiguousarray(v)


class Renderer:
    """Owns one VKRT instance. Methods map 1:1 onto the C API."""

    def __init__(self, width: int = 256, height: int = 256, init_device: bool = True):
        self.lib = library()
        h = C.c_void_p()
        _ok(self.lib.VKRT_create(C.byref(h)), "VKRT_create")
        self.h = h
        self.device_ready = False
        if init_device:
            info = CreateInfo()
            self.lib.VKRT_defaultCreateInfo(C.byref(info))
            info.width, info.height, info.headless = width, height, 1
            _ok(self.lib.VKRT_initWithCreateInfo(self.h, C.byref(info)), "VKRT_initWithCreateInfo")
            self.device_ready = True
        # a fixed sample count per frame, as the reference's offline benchmark
        # locks it (src/app/render/benchmark.c:55-67); auto_spp() turns the
        # frame-time controller back on
        _ok(self.lib.VKRT_setAutoSPPEnabled(self.h, 0), "VKRT_setAutoSPPEnabled")
        self._keep = []

    def auto_spp(self, enabled: bool, target_fps: int | None = None):
        """VKRT_setAutoSPPEnabled / VKRT_setAutoSPPTargetFPS."""
        if target_fps is not None:
            _ok(self.lib.VKRT_setAutoSPPTargetFPS(self.h, target_fps), "VKRT_setAutoSPPTargetFPS")
        _ok(self.lib.VKRT_setAutoSPPEnabled(self.h, 1 if enabled else 0), "VKRT_setAutoSPPEnabled")

    def close(self):
        if self.h:
            self.lib.VKRT_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # scene -------------------------------------------------------------
    def add_mesh(self, vertices: np.ndarray, indices: np.ndarray) -> int:
        vertices = np.ascontiguousarray(vertices, dtype=np.float32)
        indices = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
        count = C.c_uint32()
        self.lib.VKRT_getMeshCount(self.h, C.byref(count))
        _ok(self.lib.VKRT_uploadMeshData(self.h, vertices.ctypes.data_as(C.POINTER(Vertex)), len(vertices),
                                         uptr(indices), len(indices)), "VKRT_uploadMeshData")
        return count.value

    def add_material(self, material: Material, name: str = "") -> int:
        idx = C.c_uint32()
        _ok(self.lib.VKRT_addMaterial(self.h, C.byref(material), name.encode(), C.byref(idx)), "VKRT_addMaterial")
        return idx.value

    def set_mesh_material(self, mesh: int, material: int):
        _ok(self.lib.VKRT_setMeshMaterialIndex(self.h, mesh, material), "VKRT_setMeshMaterialIndex")

    def set_mesh_transform_matrix(self, mesh: int, matrix_col_major: np.ndarray):
        m = np.ascontiguousarray(matrix_col_major, dtype=np.float32).reshape(16)
        _ok(self.lib.VKRT_setMeshTransformMatrix(self.h, mesh, fptr(m)), "VKRT_setMeshTransformMatrix")

    def add_texture_rgba32f(self, pixels: np.ndarray, name: str = "env") -> int:
        pixels = np.ascontiguousarray(pixels, dtype=np.float32)
        h, w, _ = pixels.shape
        up = TextureUpload(name.encode(), pixels.ctypes.data_as(C.c_void_p), w, h, TEXTURE_FORMAT_RGBA32_SFLOAT,
                           TEXTURE_COLOR_SPACE_LINEAR)
        idx = C.c_uint32()
        _ok(self.lib.VKRT_addTextureFromPixels(self.h, C.byref(up), C.byref(idx)), "VKRT_addTextureFromPixels")
        return idx.value

    def add_texture(self, pixels: np.ndarray, fmt: int, color_space: int, name: str = "Texture") -> int:
        """VKRT_addTextureFromPixels with an H x W x 4 array of the format's
        element type (uint8, uint16, float16 or float32)."""
        dtype = {TEXTURE_FORMAT_RGBA8_UNORM: np.uint8, TEXTURE_FORMAT_RGBA16_UNORM: np.uint16,
                 TEXTURE_FORMAT_RGBA16_SFLOAT: np.float16, TEXTURE_FORMAT_RGBA32_SFLOAT: np.float32}[fmt]
        pixels = np.ascontiguousarray(pixels, dtype=dtype)
        h, w, _ = pixels.shape
        up = TextureUpload(name.encode(), pixels.ctypes.data_as(C.c_void_p), w, h, fmt, color_space)
        idx = C.c_uint32()
        _ok(self.lib.VKRT_addTextureFromPixels(self.h, C.byref(up), C.byref(idx)), "VKRT_addTextureFromPixels")
        return idx.value

    def set_environment_texture(self, index: int):
        _ok(self.lib.VKRT_setEnvironmentTexture(self.h, index), "VKRT_setEnvironmentTexture")

    def remove_texture(self, index: int):
        _ok(self.lib.VKRT_removeTexture(self.h, index), "VKRT_removeTexture")

    def texture_count(self) -> int:
        c = C.c_uint32()
        _ok(self.lib.VKRT_getTextureCount(self.h, C.byref(c)), "VKRT_getTextureCount")
        return c.value

    def save_image(self, path: str, denoise: bool | None = None):
        """VKRT_saveRenderImage(Ex): PNG / JPEG / EXR by extension."""
        if denoise is None:
            _ok(self.lib.VKRT_saveRenderImage(self.h, os.fsencode(path)), "VKRT_saveRenderImage")
        else:
            st = RenderExportSettings(1 if denoise else 0)
            _ok(self.lib.VKRT_saveRenderImageEx(self.h, os.fsencode(path), C.byref(st)), "VKRT_saveRenderImageEx")

    # scene input (include/vkrt_import.h)
    def import_mesh(self, path: str, name: str | None = None) -> tuple[int, int]:
        """VKRT_importMesh: a .gltf/.glb file into the scene; returns (first mesh, mesh count)."""
        first, count = C.c_uint32(), C.c_uint32()
        _ok(self.lib.VKRT_importMesh(self.h, os.fsencode(path), name.encode() if name else None, C.byref(first),
                                     C.byref(count)), "VKRT_importMesh")
        return first.value, count.value

    def load_scene_file(self, path: str):
        """VKRT_loadSceneFile: replaces the scene with a vkrt.scene document."""
        _ok(self.lib.VKRT_loadSceneFile(self.h, os.fsencode(path)), "VKRT_loadSceneFile")

    def mesh_count(self) -> int:
        c = C.c_uint32()
        _ok(self.lib.VKRT_getMeshCount(self.h, C.byref(c)), "VKRT_getMeshCount")
        return c.value

    def mesh(self, index: int) -> MeshSnapshot:
        snap = MeshSnapshot()
        _ok(self.lib.VKRT_getMeshSnapshot(self.h, index, C.byref(snap)), "VKRT_getMeshSnapshot")
        return snap

    def material_name(self, index: int) -> str:
        snap = MaterialSnapshot()
        _ok(self.lib.VKRT_getMaterialSnapshot(self.h, index, C.byref(snap)), "VKRT_getMaterialSnapshot")
        return snap.name.decode()

    def remove_material(self, index: int):
        _ok(self.lib.VKRT_removeMaterial(self.h, index), "VKRT_removeMaterial")

    def set_material_name(self, index: int, name: str):
        _ok(self.lib.VKRT_setMaterialName(self.h, index, name.encode()), "VKRT_setMaterialName")

    def clear_mesh_material(self, mesh: int):
        _ok(self.lib.VKRT_clearMeshMaterialAssignment(self.h, mesh), "VKRT_clearMeshMaterialAssignment")

    def set_mesh_opacity(self, mesh: int, opacity: float):
        _ok(self.lib.VKRT_setMeshOpacity(self.h, mesh, opacity), "VKRT_setMeshOpacity")

    def set_mesh_render_backfaces(self, mesh: int, enabled: bool):
        _ok(self.lib.VKRT_setMeshRenderBackfaces(self.h, mesh, 1 if enabled else 0), "VKRT_setMeshRenderBackfaces")

    def scene_settings(self) -> SceneSettingsSnapshot:
        out = SceneSettingsSnapshot()
        _ok(self.lib.VKRT_getSceneSettings(self.h, C.byref(out)), "VKRT_getSceneSettings")
        return out

    def replace_environment(self, pixels: np.ndarray) -> int:
        """Uploads `pixels` (H x W x 3 or 4, linear) as the environment texture and
        removes the texture it replaces."""
        px = np.asarray(pixels, dtype=np.float32)
        if px.shape[2] == 3:
            px = np.concatenate([px, np.ones(px.shape[:2] + (1,), np.float32)], axis=2)
        old = self.scene_settings().environmentTextureIndex
        idx = self.add_texture_rgba32f(px)
        self.set_environment_texture(idx)
        if old != VKRT_INVALID_INDEX:
            self.remove_texture(old)
            idx = self.scene_settings().environmentTextureIndex
        return idx

    def set_environment_light(self, color, strength: float):
        c = np.asarray(color, dtype=np.float32)
        _ok(self.lib.VKRT_setEnvironmentLight(self.h, fptr(c), strength), "VKRT_setEnvironmentLight")

    def set_camera(self, pos, target, up, vfov: float):
        p, t, u = (np.asarray(x, dtype=np.float32) for x in (pos, target, up))
        _ok(self.lib.VKRT_cameraSetPose(self.h, fptr(p), fptr(t), fptr(u), vfov), "VKRT_cameraSetPose")

    # settings ----------------------------------------------------------
    def configure(self, spp=None, min_depth=None, max_depth=None, mode=None, spectral=None, tone=None,
                  rng=None, seed=0, env_importance=None, nee=None):
        L = self.lib
        if spp is not None:
            _ok(L.VKRT_setAutoSPPEnabled(self.h, 0), "VKRT_setAutoSPPEnabled")  # an explicit count is kept
            _ok(L.VKRT_setSamplesPerPixel(self.h, spp), "setSamplesPerPixel")
        if max_depth is not None:
            _ok(L.VKRT_setPathDepth(self.h, min_depth if min_depth is not None else 4, max_depth), "setPathDepth")
        if mode is not None:
            _ok(L.VKRT_setRenderMode(self.h, mode), "setRenderMode")
        if spectral is not None:
            _ok(L.VKRT_setSpectralSamplingMode(self.h, spectral), "setSpectralSamplingMode")
        if tone is not None:
            _ok(L.VKRT_setToneMappingMode(self.h, tone), "setToneMappingMode")
        if rng is not None:
            _ok(L.VKRT_setRandomPolicy(self.h, rng, seed), "setRandomPolicy")
        if env_importance is not None:
            _ok(L.VKRT_setEnvironmentImportanceSampling(self.h, 1 if env_importance else 0), "envImportance")
        if nee is not None:
            _ok(L.VKRT_setMisNeeEnabled(self.h, 1 if nee else 0), "setMisNeeEnabled")

    # rendering ---------------------------------------------------------
    def start(self, width: int, height: int, target_samples: int = 0):
        _ok(self.lib.VKRT_startRender(self.h, width, height, target_samples), "VKRT_startRender")
        self.width, self.height = width, height

    def draw(self):
        _ok(self.lib.VKRT_draw(self.h), "VKRT_draw")

    def accumulation(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        _ok(self.lib.VKRT_readAccumulation(self.h, fptr(out), out.size), "VKRT_readAccumulation")
        return out

    def output_image(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), dtype=np.uint16)
        _ok(self.lib.VKRT_readOutputImage(self.h, out.ctypes.data_as(C.POINTER(C.c_uint16)), out.size),
            "VKRT_readOutputImage")
        return out

    def material(self, index: int) -> Material:
        snap = MaterialSnapshot()
        _ok(self.lib.VKRT_getMaterialSnapshot(self.h, index, C.byref(snap)), "VKRT_getMaterialSnapshot")
        return snap.material

    def set_material(self, index: int, material: Material):
        _ok(self.lib.VKRT_setMaterial(self.h, index, C.byref(material)), "VKRT_setMaterial")

    def material_count(self) -> int:
        c = C.c_uint32()
        _ok(self.lib.VKRT_getMaterialCount(self.h, C.byref(c)), "VKRT_getMaterialCount")
        return c.value

    def render_backward(self, d_image: np.ndarray, environment_shape=None) -> dict:
        """Gradients of sum(d_image * I) for the frame the next draw() renders
        (VKRT_renderBackward). d_image: H x W x 3. Returns numpy arrays."""
        d = np.ascontiguousarray(d_image, dtype=np.float32)
        m = self.material_count()
        out = {"baseColor": np.zeros((m, 3), np.float32), "roughness": np.zeros(m, np.float32),
               "emission": np.zeros((m, 3), np.float32)}
        g = Gradients()
        g.baseColor = fptr(out["baseColor"])
        g.roughness = fptr(out["roughness"])
        g.emission = fptr(out["emission"])
        g.materialCount = m
        if environment_shape is not None:
            h, w = environment_shape
            out["environment"] = np.zeros((h, w, 3), np.float32)
            g.environment = fptr(out["environment"])
            g.environmentWidth, g.environmentHeight = w, h
        _ok(self.lib.VKRT_renderBackward(self.h, fptr(d), C.byref(g)), "VKRT_renderBackward")
        return out

    # neural radiance cache --------------------------------------------
    def set_neural_cache(self, enabled=True, **kw) -> NeuralCacheSettings:
        st = NeuralCacheSettings()
        self.lib.VKRT_defaultNeuralCacheSettings(C.byref(st))
        st.enabled = 1 if enabled else 0
        for k, v in kw.items():
            setattr(st, k, v)
        _ok(self.lib.VKRT_setNeuralCache(self.h, C.byref(st)), "VKRT_setNeuralCache")
        return st

    def neural_cache_stats(self) -> NeuralCacheStats:
        out = NeuralCacheStats()
        _ok(self.lib.VKRT_getNeuralCacheStats(self.h, C.byref(out)), "VKRT_getNeuralCacheStats")
        return out

    def nrc_reset(self, seed: int = 0):
        _ok(self.lib.VKRT_neuralCacheReset(self.h, seed), "VKRT_neuralCacheReset")

    def nrc_weights(self) -> np.ndarray:
        n = self.lib.VKRT_neuralCacheParameterCount()
        out = np.zeros(n, np.float32)
        _ok(self.lib.VKRT_neuralCacheGetWeights(self.h, fptr(out), n), "VKRT_neuralCacheGetWeights")
        return out

    def nrc_set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _ok(self.lib.VKRT_neuralCacheSetWeights(self.h, fptr(w), w.size), "VKRT_neuralCacheSetWeights")

    def nrc_infer(self, inputs: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 11)
        out = np.zeros((len(x), 3), np.float32)
        _ok(self.lib.VKRT_neuralCacheInfer(self.h, fptr(x), len(x), fptr(out)), "VKRT_neuralCacheInfer")
        return out

    def nrc_train(self, inputs, targets, steps: int, batch: int, lr: float) -> float:
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 11)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        loss = np.zeros(1, np.float32)
        _ok(self.lib.VKRT_neuralCacheTrain(self.h, fptr(x), fptr(t), len(x), steps, batch, lr, fptr(loss)),
            "VKRT_neuralCacheTrain")
        return float(loss[0])

    def nrc_gradient(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 11)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(self.lib.VKRT_neuralCacheParameterCount(), np.float32)
        loss = np.zeros(1, np.float32)
        _ok(self.lib.VKRT_neuralCacheGradient(self.h, fptr(x), fptr(t), len(x), fptr(g), fptr(loss)),
            "VKRT_neuralCacheGradient")
        return g, float(loss[0])

    def synchronize(self):
        _ok(self.lib.VKRT_deviceSynchronize(self.h), "VKRT_deviceSynchronize")

    def set_denoiser_features(self, enabled: bool):
        _ok(self.lib.VKRT_setDenoiserFeatures(self.h, 1 if enabled else 0), "VKRT_setDenoiserFeatures")

    def denoiser_features(self) -> tuple[np.ndarray, np.ndarray]:
        """(albedo, normal) feature images, H x W x 4 float32 (.w = feature weight)."""
        albedo = np.zeros((self.height, self.width, 4), dtype=np.float32)
        normal = np.zeros_like(albedo)
        _ok(self.lib.VKRT_readDenoiserFeatures(self.h, fptr(albedo), fptr(normal), albedo.size),
            "VKRT_readDenoiserFeatures")
        return albedo, normal

    def set_debug_mode(self, mode: int):
        _ok(self.lib.VKRT_setDebugMode(self.h, mode), "VKRT_setDebugMode")

    def set_material_sorting(self, enabled: bool):
        _ok(self.lib.VKRT_setMaterialSorting(self.h, 1 if enabled else 0), "VKRT_setMaterialSorting")

    def set_fused_tail(self, max_rays: int):
        _ok(self.lib.VKRT_setFusedTail(self.h, max_rays), "VKRT_setFusedTail")

    def set_traversal_statistics(self, enabled: bool):
        _ok(self.lib.VKRT_setTraversalStatistics(self.h, 1 if enabled else 0), "VKRT_setTraversalStatistics")

    def stats(self) -> TraceStats:
        s = TraceStats()
        _ok(self.lib.VKRT_getTraceStats(self.h, C.byref(s)), "VKRT_getTraceStats")
        return s

    def status(self) -> RenderStatusSnapshot:
        s = RenderStatusSnapshot()
        _ok(self.lib.VKRT_getRenderStatus(self.h, C.byref(s)), "VKRT_getRenderStatus")
        return s

    def export(self) -> SceneBuffers:
        b = SceneBuffers()
        _ok(self.lib.VKRT_exportSceneBuffers(self.h, C.byref(b)), "VKRT_exportSceneBuffers")
        return b

    def primary_hits(self, sample_index: int = 0):
        n = self.width * self.height
        mesh = np.zeros(n, dtype=np.uint32)
        prim = np.zeros(n, dtype=np.uint32)
        _ok(self.lib.VKRT_tracePrimaryHits(self.h, sample_index, uptr(mesh), uptr(prim)), "VKRT_tracePrimaryHits")
        return mesh, prim

    def trace_rays(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = len(rays)
        mesh = np.zeros(n, dtype=np.uint32)
        prim = np.zeros(n, dtype=np.uint32)
        t = np.zeros(n, dtype=np.float32)
        bary = np.zeros((n, 2), dtype=np.float32)
        _ok(self.lib.VKRT_traceRays(self.h, fptr(rays), n, uptr(mesh), uptr(prim), fptr(t), fptr(bary)),
            "VKRT_traceRays")
        return mesh, prim, t, bary

    def trace_rays_device_noalloc(self, rays: int, count: int, mesh: int, prim: int, t: int = 0, bary: int = 0):
        """VKRT_traceRaysDeviceNoAlloc on device addresses (integers, e.g. a
        torch tensor's data_ptr() on the renderer's GPU): rays count x 8
        float32, mesh / prim count uint32, optional t count float32 and bary
        count x 2 float32 (0 leaves them out). Work queued on other streams
        that writes the rays must be finished first; the results are written
        when this returns."""
        _ok(self.lib.VKRT_traceRaysDeviceNoAlloc(self.h, rays or None, count, mesh or None, prim or None, t or None,
                                                 bary or None), "VKRT_traceRaysDeviceNoAlloc")

    def trace_shadow_rays(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        vis = np.zeros(len(rays), dtype=np.uint32)
        _ok(self.lib.VKRT_traceShadowRays(self.h, fptr(rays), len(rays), uptr(vis)), "VKRT_traceShadowRays")
        return vis
