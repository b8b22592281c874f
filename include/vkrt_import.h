/*
 * Scene input: glTF 2.0 / GLB import and vkrt.scene files.
 *
 * In the reference these live in the application layer: the glTF reader
 * (src/app/mesh/loader.h:5-58, built on cgltf), the import that pushes its
 * result through the core API (src/app/mesh/controller.c:494-632,
 * meshControllerImportMesh) and the scene-file loader
 * (src/app/scene/controller.c:1528-1658, sceneControllerLoadSceneFromPath).
 * This library carries them beside the core API so that a headless caller can
 * load the reference's own scenes; the editor session those functions also
 * update (scene-object tree, selection, file records) is not part of it, so
 * the two VKRT_* entry points below take no Session and apply the scene
 * objects' world transforms directly to the meshes.
 */
#ifndef VKRT_IMPORT_H
#define VKRT_IMPORT_H

#ifdef __cplusplus
extern "C" {
#endif

#include "vkrt.h"

#include <stddef.h>
#include <stdint.h>

/* loader.h:5-8 */
typedef struct MaterialImportEntry {
    char* name;
    Material material;
} MaterialImportEntry;

/* loader.h:10-17 */
typedef struct TextureImportEntry {
    char* name;
    void* pixels;
    uint32_t width;
    uint32_t height;
    uint32_t format;
    uint32_t colorSpace;
} TextureImportEntry;

/* loader.h:19-31: one triangle primitive, vertices already in engine space
 * (glTF +Y up becomes +Z up: (x, y, z) -> (x, -z, y)). */
typedef struct MeshImportEntry {
    char* name;
    Vertex* vertices;
    uint32_t* indices;
    size_t vertexCount;
    size_t indexCount;
    uint32_t nodeIndex;
    uint32_t materialIndex; /* index into MeshImportData.materials, or VKRT_INVALID_INDEX */
    uint8_t renderBackfaces;
    float position[3];
    float rotation[3];
    float scale[3];
} MeshImportEntry;

/* loader.h:33-41: one glTF node, its local transform conjugated into engine
 * space (VKRT_buildImportedNodeTransform); column-major like cglm's mat4. */
typedef struct NodeImportEntry {
    float localTransform[4][4];
    char* name;
    uint32_t parentIndex;
    uint32_t meshEntryCount;
    float position[3];
    float rotation[3];
    float scale[3];
} NodeImportEntry;

/* loader.h:43-55 */
typedef struct MeshImportData {
    MeshImportEntry* entries;
    uint32_t count;
    uint32_t entryCapacity;
    NodeImportEntry* nodes;
    uint32_t nodeCount;
    uint32_t nodeCapacity;
    MaterialImportEntry* materials;
    uint32_t materialCount;
    TextureImportEntry* textures;
    uint32_t textureCount;
    uint32_t textureCapacity;
} MeshImportData;

/* loader.h:57 (loader.c:2133-2179): reads a .gltf (JSON, external or data-URI
 * buffers) or .glb file. Returns 0 on success, -1 on failure (missing file,
 * malformed JSON, out-of-range accessor, or no triangle primitive). Image
 * textures are not decoded: material texture slots stay VKRT_INVALID_INDEX and
 * textureCount is 0. */
int meshLoadFromFile(const char* filePath, MeshImportData* outImportData);
/* loader.h:58 */
void meshReleaseImportData(MeshImportData* importData);

/* src/app/mesh/controller.c:494-632 (meshControllerImportMesh) without the
 * editor session: uploads every primitive as a mesh, adds the file's
 * materials, assigns them, names the meshes and sets each mesh's world
 * transform to the product of its node chain. On failure everything the call
 * added is removed again. outFirstMeshIndex / outMeshCount may be NULL. */
VKRT_Result VKRT_importMesh(VKRT* vkrt, const char* path, const char* importName, uint32_t* outFirstMeshIndex,
                            uint32_t* outMeshCount);

/* src/app/scene/controller.c:1599-1658 (sceneControllerLoadSceneFromPath):
 * loads a "vkrt.scene" version-1 file, replacing the current scene (meshes,
 * materials but the default one, textures, environment texture, scene
 * settings and camera). Asset paths resolve against the scene file's folder,
 * then as given. On failure the previous scene is put back and the call
 * returns VKRT_ERROR_OPERATION_FAILED (VKRT_ERROR_INVALID_ARGUMENT for a NULL
 * argument). */
VKRT_Result VKRT_loadSceneFile(VKRT* vkrt, const char* path);

#ifdef __cplusplus
}
#endif

#endif
