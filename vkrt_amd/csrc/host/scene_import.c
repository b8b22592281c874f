/*
 * glTF import into a VKRT scene and vkrt.scene file loading.
 *
 * VKRT_importMesh follows meshControllerImportMesh
 * (src/app/mesh/controller.c:494-632): upload every primitive, add the file's
 * materials, assign them, name the meshes, then give each mesh the world
 * transform of its glTF node chain (the reference builds editor scene objects
 * for the nodes and syncs their transforms, src/app/mesh/controller.c:170-340
 * and src/app/session/session.c:257-300; the resulting mesh transforms are the
 * same). VKRT_loadSceneFile follows loadSceneDocument
 * (src/app/scene/controller.c:1241-1597) step by step. The editor's undo of a
 * failed load re-serialises the previous scene; here the previous scene's
 * arrays are set aside and put back instead.
 */
#include "internal.h"
#include "json.h"
#include "vkrt_import.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void identity4(float m[4][4]) {
    memset(m, 0, sizeof(float) * 16);
    m[0][0] = m[1][1] = m[2][2] = m[3][3] = 1.0f;
}

static void mul4(const float a[4][4], const float b[4][4], float out[4][4]) {
    float r[4][4];
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++)
            r[c][row] = a[0][row] * b[c][0] + a[1][row] * b[c][1] + a[2][row] * b[c][2] + a[3][row] * b[c][3];
    memcpy(out, r, sizeof(r));
}

static void rollback(VKRT* vkrt, uint32_t meshes, uint32_t materials, uint32_t textures) {
    while (vkrt->meshCount > meshes) VKRT_removeMesh(vkrt, vkrt->meshCount - 1u);
    uint32_t keep = materials > 0u ? materials : 1u;
    while (vkrt->materialCount > keep) VKRT_removeMaterial(vkrt, vkrt->materialCount - 1u);
    while (vkrt->textureCount > textures) VKRT_removeTexture(vkrt, vkrt->textureCount - 1u);
}

VKRT_Result VKRT_importMesh(VKRT* vkrt, const char* path, const char* importName, uint32_t* outFirstMeshIndex,
                            uint32_t* outMeshCount) {
    if (outFirstMeshIndex) *outFirstMeshIndex = VKRT_INVALID_INDEX;
    if (outMeshCount) *outMeshCount = 0u;
    if (!vkrt || !path || !path[0]) return VKRT_ERROR_INVALID_ARGUMENT;
    uint32_t meshBase = vkrt->meshCount, materialBase = vkrt->materialCount, textureBase = vkrt->textureCount;

    MeshImportData data;
    if (meshLoadFromFile(path, &data) != 0) return VKRT_ERROR_OPERATION_FAILED;

    VKRT_Result result = VKRT_SUCCESS;
    uint32_t* materialMap = NULL;
    float(*world)[4][4] = NULL;
    VKRT_MeshUpload* uploads = (VKRT_MeshUpload*)calloc(data.count, sizeof(VKRT_MeshUpload));
    if (!uploads) result = VKRT_ERROR_OUT_OF_MEMORY;
    for (uint32_t i = 0; result == VKRT_SUCCESS && i < data.count; i++) {
        uploads[i].vertices = data.entries[i].vertices;
        uploads[i].vertexCount = data.entries[i].vertexCount;
        uploads[i].indices = data.entries[i].indices;
        uploads[i].indexCount = data.entries[i].indexCount;
    }
    if (result == VKRT_SUCCESS) result = VKRT_uploadMeshDataBatch(vkrt, uploads, data.count);
    free(uploads);

    /* decoded textures first; the materials' slot indices are local to the file */
    if (result == VKRT_SUCCESS && data.textureCount) {
        VKRT_TextureUpload* tu = (VKRT_TextureUpload*)calloc(data.textureCount, sizeof(VKRT_TextureUpload));
        if (!tu) result = VKRT_ERROR_OUT_OF_MEMORY;
        for (uint32_t t = 0; tu && t < data.textureCount; t++) {
            tu[t].name = data.textures[t].name;
            tu[t].pixels = data.textures[t].pixels;
            tu[t].width = data.textures[t].width;
            tu[t].height = data.textures[t].height;
            tu[t].format = data.textures[t].format;
            tu[t].colorSpace = data.textures[t].colorSpace;
        }
        if (tu) result = VKRT_addTexturesBatch(vkrt, tu, data.textureCount, NULL);
        free(tu);
        for (uint32_t m = 0; result == VKRT_SUCCESS && m < data.materialCount; m++) {
            Material* mat = &data.materials[m].material;
            uint32_t* slots[4] = {&mat->baseColorTextureIndex, &mat->metallicRoughnessTextureIndex,
                                  &mat->normalTextureIndex, &mat->emissiveTextureIndex};
            for (int k = 0; k < 4; k++)
                if (*slots[k] != VKRT_INVALID_INDEX) *slots[k] += textureBase;
        }
    }
    if (result == VKRT_SUCCESS && data.materialCount) {
        materialMap = (uint32_t*)malloc(sizeof(uint32_t) * data.materialCount);
        if (!materialMap) result = VKRT_ERROR_OUT_OF_MEMORY;
        for (uint32_t m = 0; result == VKRT_SUCCESS && m < data.materialCount; m++)
            result = VKRT_addMaterial(vkrt, &data.materials[m].material, data.materials[m].name, &materialMap[m]);
    }

    /* node world transforms: parents precede their children in the node list */
    if (result == VKRT_SUCCESS && data.nodeCount) {
        world = (float(*)[4][4])malloc(sizeof(float[4][4]) * data.nodeCount);
        if (!world) result = VKRT_ERROR_OUT_OF_MEMORY;
        for (uint32_t n = 0; result == VKRT_SUCCESS && n < data.nodeCount; n++) {
            uint32_t parent = data.nodes[n].parentIndex;
            if (parent != VKRT_INVALID_INDEX && parent < n)
                mul4(world[parent], data.nodes[n].localTransform, world[n]);
            else
                memcpy(world[n], data.nodes[n].localTransform, sizeof(world[n]));
        }
    }

    /* src/app/mesh/controller.c:421-457 */
    for (uint32_t i = 0; result == VKRT_SUCCESS && i < data.count; i++) {
        const MeshImportEntry* e = &data.entries[i];
        uint32_t meshIndex = meshBase + i;
        if (e->materialIndex != VKRT_INVALID_INDEX) {
            if (!materialMap || e->materialIndex >= data.materialCount)
                result = VKRT_ERROR_OPERATION_FAILED;
            else
                result = VKRT_setMeshMaterialIndex(vkrt, meshIndex, materialMap[e->materialIndex]);
        }
        if (result == VKRT_SUCCESS) result = VKRT_setMeshRenderBackfaces(vkrt, meshIndex, e->renderBackfaces);
        const char* name = e->name;
        if (importName && importName[0] && data.count == 1u) name = importName;
        if (!name || !name[0]) name = path;
        if (result == VKRT_SUCCESS) VKRT_setMeshName(vkrt, meshIndex, name);
        if (result == VKRT_SUCCESS) {
            float m[4][4];
            if (world && e->nodeIndex < data.nodeCount)
                memcpy(m, world[e->nodeIndex], sizeof(m));
            else
                identity4(m);
            result = VKRT_setMeshTransformMatrix(vkrt, meshIndex, m);
        }
    }
    uint32_t imported = data.count;
    free(materialMap);
    free(world);
    meshReleaseImportData(&data);
    if (result != VKRT_SUCCESS) {
        VKRT_LOG_ERROR("Mesh import failed. File: %s", path);
        rollback(vkrt, meshBase, materialBase, textureBase);
        return result;
    }
    if (outFirstMeshIndex) *outFirstMeshIndex = meshBase;
    if (outMeshCount) *outMeshCount = imported;
    return VKRT_SUCCESS;
}

/* ------------------------------------------------------------ scene files */

static char* readText(const char* path, size_t* outLength) {
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    char* text = NULL;
    if (fseek(f, 0, SEEK_END) == 0) {
        long n = ftell(f);
        if (n >= 0 && fseek(f, 0, SEEK_SET) == 0 && (text = (char*)malloc((size_t)n + 1u))) {
            if (fread(text, 1, (size_t)n, f) != (size_t)n) {
                free(text);
                text = NULL;
            } else {
                text[n] = '\0';
                *outLength = (size_t)n;
            }
        }
    }
    fclose(f);
    return text;
}

static int fileExists(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return 0;
    fclose(f);
    return 1;
}

/* src/app/scene/controller.c:125-143: absolute paths as given; relative ones
 * against the scene file's folder first, then as given. */
static int resolveAssetPath(const char* scenePath, const char* stored, char* out, size_t outSize) {
    if (!stored || !stored[0]) return 0;
    if (stored[0] == '/') {
        snprintf(out, outSize, "%s", stored);
        return fileExists(out);
    }
    const char* slash = strrchr(scenePath, '/');
    if (slash) {
        int n = snprintf(out, outSize, "%.*s/%s", (int)(slash - scenePath), scenePath, stored);
        if (n > 0 && (size_t)n < outSize && fileExists(out)) return 1;
    }
    snprintf(out, outSize, "%s", stored);
    return fileExists(out);
}

static int readBool(const JsonValue* v, uint8_t* out) {
    if (!v || v->type != JSON_BOOL) return 0;
    *out = v->boolean;
    return 1;
}

/* null -> VKRT_INVALID_INDEX (src/app/scene/controller.c:450-455) */
static int readIndexOrNull(const JsonValue* v, uint32_t* out) {
    *out = VKRT_INVALID_INDEX;
    if (!v) return 0;
    if (v->type == JSON_NULL) return 1;
    return vkrtJsonToU32(v, out);
}

static int optFloat(const JsonValue* o, const char* key, float* out) {
    const JsonValue* v = vkrtJsonGet(o, key);
    return v ? vkrtJsonToFloat(v, out) : 1;
}

static int optFloats(const JsonValue* o, const char* key, float* out, uint32_t n) {
    const JsonValue* v = vkrtJsonGet(o, key);
    return v ? vkrtJsonToFloats(v, out, n) : 1;
}

static int optU32(const JsonValue* o, const char* key, uint32_t* out) {
    const JsonValue* v = vkrtJsonGet(o, key);
    return v ? vkrtJsonToU32(v, out) : 1;
}

static int optIndex(const JsonValue* o, const char* key, uint32_t* out) {
    const JsonValue* v = vkrtJsonGet(o, key);
    if (!v) return 1;
    return readIndexOrNull(v, out);
}

static int optBool(const JsonValue* o, const char* key, uint8_t* out) {
    const JsonValue* v = vkrtJsonGet(o, key);
    return v ? readBool(v, out) : 1;
}

/* src/app/scene/controller.c:585-638: every Material field may appear;
 * absent fields keep VKRT_materialDefault(). */
static int parseMaterial(const JsonValue* o, Material* out) {
    if (!o || o->type != JSON_OBJECT) return 0;
    Material m = VKRT_materialDefault();
    int ok = optFloats(o, "baseColor", m.baseColor, 3) && optFloat(o, "roughness", &m.roughness) &&
             optFloats(o, "emissionColor", m.emissionColor, 3) &&
             optFloat(o, "emissionLuminance", &m.emissionLuminance) && optFloats(o, "eta", m.eta, 3) &&
             optFloat(o, "metallic", &m.metallic) && optFloats(o, "k", m.k, 3) &&
             optFloat(o, "anisotropic", &m.anisotropic) && optFloat(o, "specular", &m.specular) &&
             optFloat(o, "specularTint", &m.specularTint) && optFloats(o, "sheenTintWeight", m.sheenTintWeight, 4) &&
             optFloat(o, "clearcoat", &m.clearcoat) && optFloat(o, "clearcoatGloss", &m.clearcoatGloss) &&
             optFloat(o, "ior", &m.ior) && optFloat(o, "abbeNumber", &m.abbeNumber) &&
             optFloat(o, "diffuseRoughness", &m.diffuseRoughness) && optFloat(o, "transmission", &m.transmission) &&
             optFloat(o, "subsurface", &m.subsurface) && optFloat(o, "sheenRoughness", &m.sheenRoughness) &&
             optFloat(o, "absorptionCoefficient", &m.absorptionCoefficient) &&
             optFloats(o, "attenuationColor", m.attenuationColor, 3) &&
             optFloat(o, "normalTextureScale", &m.normalTextureScale) &&
             optIndex(o, "baseColorTextureIndex", &m.baseColorTextureIndex) &&
             optIndex(o, "metallicRoughnessTextureIndex", &m.metallicRoughnessTextureIndex) &&
             optIndex(o, "normalTextureIndex", &m.normalTextureIndex) &&
             optIndex(o, "emissiveTextureIndex", &m.emissiveTextureIndex) &&
             optU32(o, "baseColorTextureWrap", &m.baseColorTextureWrap) &&
             optU32(o, "metallicRoughnessTextureWrap", &m.metallicRoughnessTextureWrap) &&
             optU32(o, "normalTextureWrap", &m.normalTextureWrap) &&
             optU32(o, "emissiveTextureWrap", &m.emissiveTextureWrap) && optFloat(o, "opacity", &m.opacity) &&
             optFloat(o, "alphaCutoff", &m.alphaCutoff) && optU32(o, "alphaMode", &m.alphaMode) &&
             optU32(o, "textureTexcoordSets", &m.textureTexcoordSets) &&
             optFloats(o, "baseColorTextureTransform", m.baseColorTextureTransform, 4) &&
             optFloats(o, "metallicRoughnessTextureTransform", m.metallicRoughnessTextureTransform, 4) &&
             optFloats(o, "normalTextureTransform", m.normalTextureTransform, 4) &&
             optFloats(o, "emissiveTextureTransform", m.emissiveTextureTransform, 4) &&
             optFloats(o, "textureRotations", m.textureRotations, 4);
    if (!ok) return 0;
    *out = m;
    return 1;
}

/* Everything a failed load must put back. */
typedef struct SceneBackup {
    HostMesh* meshes;
    uint32_t meshCount, meshCapacity;
    HostMaterial* materials;
    uint32_t materialCount, materialCapacity;
    HostTexture* textures;
    uint32_t textureCount, textureCapacity;
    VKRT_SceneSettingsSnapshot settings;
} SceneBackup;

/* Sets the current scene aside and leaves an empty one holding a copy of the
 * default material, as clearCurrentScene (src/app/scene/controller.c:663-688)
 * leaves it. */
static VKRT_Result setSceneAside(VKRT* vkrt, SceneBackup* b) {
    HostMaterial* fresh = (HostMaterial*)calloc(4, sizeof(HostMaterial));
    if (!fresh) return VKRT_ERROR_OUT_OF_MEMORY;
    if (vkrt->materialCount > 0) {
        fresh[0] = vkrt->materials[0];
    } else {
        fresh[0].material = VKRT_materialDefault();
        snprintf(fresh[0].name, VKRT_NAME_LEN, "%s", "Default");
    }
    b->meshes = vkrt->meshes;
    b->meshCount = vkrt->meshCount;
    b->meshCapacity = vkrt->meshCapacity;
    b->materials = vkrt->materials;
    b->materialCount = vkrt->materialCount;
    b->materialCapacity = vkrt->materialCapacity;
    b->textures = vkrt->textures;
    b->textureCount = vkrt->textureCount;
    b->textureCapacity = vkrt->textureCapacity;
    b->settings = vkrt->settings;
    vkrt->meshes = NULL;
    vkrt->meshCount = vkrt->meshCapacity = 0;
    vkrt->materials = fresh;
    vkrt->materialCount = 1;
    vkrt->materialCapacity = 4;
    vkrt->textures = NULL;
    vkrt->textureCount = vkrt->textureCapacity = 0;
    vkrt->settings.environmentTextureIndex = VKRT_INVALID_INDEX;
    vkrt->geometryRevision++;
    vkrt->materialRevision++;
    vkrt->environmentRevision++;
    vkrtResetSceneData(vkrt);
    return VKRT_SUCCESS;
}

static void freeSceneArrays(HostMesh* meshes, uint32_t meshCount, HostMaterial* materials, HostTexture* textures,
                            uint32_t textureCount) {
    for (uint32_t i = 0; i < meshCount; i++) {
        free(meshes[i].vertices);
        free(meshes[i].indices);
    }
    for (uint32_t i = 0; i < textureCount; i++) free(textures[i].rgba);
    free(meshes);
    free(materials);
    free(textures);
}

static void restoreScene(VKRT* vkrt, SceneBackup* b) {
    freeSceneArrays(vkrt->meshes, vkrt->meshCount, vkrt->materials, vkrt->textures, vkrt->textureCount);
    vkrt->meshes = b->meshes;
    vkrt->meshCount = b->meshCount;
    vkrt->meshCapacity = b->meshCapacity;
    vkrt->materials = b->materials;
    vkrt->materialCount = b->materialCount;
    vkrt->materialCapacity = b->materialCapacity;
    vkrt->textures = b->textures;
    vkrt->textureCount = b->textureCount;
    vkrt->textureCapacity = b->textureCapacity;
    vkrt->settings = b->settings;
    vkrt->geometryRevision++;
    vkrt->materialRevision++;
    vkrt->environmentRevision++;
    vkrtWriteSceneStateUniform(vkrt);
    vkrtSyncCameraMatrices(vkrt);
    vkrtResetSceneData(vkrt);
    memset(b, 0, sizeof(*b));
}

static void dropBackup(SceneBackup* b) {
    freeSceneArrays(b->meshes, b->meshCount, b->materials, b->textures, b->textureCount);
    memset(b, 0, sizeof(*b));
}

#define SCENE_FAIL(...)                   \
    do {                                  \
        VKRT_LOG_ERROR(__VA_ARGS__);      \
        goto fail;                        \
    } while (0)

typedef struct SceneObjectRecord {
    uint32_t parent;
    uint32_t mesh; /* loaded mesh index or VKRT_INVALID_INDEX */
    float local[4][4];
} SceneObjectRecord;

/* src/app/session/session.c:257-300: depth-first from the roots in index
 * order, children in index order, world = parent world * local. */
static VKRT_Result syncObjectTransforms(VKRT* vkrt, const SceneObjectRecord* objects, uint32_t count) {
    if (count == 0) return VKRT_SUCCESS;
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * count);
    float(*world)[4][4] = (float(*)[4][4])malloc(sizeof(float[4][4]) * count);
    if (!stack || !world) {
        free(stack);
        free(world);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    VKRT_Result r = VKRT_SUCCESS;
    size_t top = 0;
    for (uint32_t i = count; i > 0; i--)
        if (objects[i - 1u].parent == VKRT_INVALID_INDEX) stack[top++] = i - 1u;
    while (top > 0 && r == VKRT_SUCCESS) {
        uint32_t o = stack[--top];
        const SceneObjectRecord* obj = &objects[o];
        if (obj->parent == VKRT_INVALID_INDEX)
            memcpy(world[o], obj->local, sizeof(world[o]));
        else
            mul4(world[obj->parent], obj->local, world[o]);
        if (obj->mesh != VKRT_INVALID_INDEX) r = VKRT_setMeshTransformMatrix(vkrt, obj->mesh, world[o]);
        /* parents precede children (checked on load), so each object is pushed once */
        for (uint32_t c = count; c > 0; c--)
            if (objects[c - 1u].parent == o) stack[top++] = c - 1u;
    }
    free(stack);
    free(world);
    return r;
}

VKRT_Result VKRT_loadSceneFile(VKRT* vkrt, const char* path) {
    if (!vkrt || !path || !path[0]) return VKRT_ERROR_INVALID_ARGUMENT;
    size_t length = 0;
    char* text = readText(path, &length);
    if (!text) {
        VKRT_LOG_ERROR("Scene file not found: %s", path);
        return VKRT_ERROR_OPERATION_FAILED;
    }
    JsonValue root;
    char err[256];
    int parsed = vkrtJsonParse(text, length, &root, err, sizeof(err));
    free(text);
    if (parsed != 0) {
        VKRT_LOG_ERROR("Scene file '%s': %s", path, err);
        return VKRT_ERROR_OPERATION_FAILED;
    }

    SceneBackup backup;
    memset(&backup, 0, sizeof(backup));
    int asideDone = 0;
    uint32_t* batchFirst = NULL;
    uint32_t* batchCount = NULL;
    uint32_t* savedToLoaded = NULL;
    uint8_t* keep = NULL;
    SceneObjectRecord* objects = NULL;
    uint32_t* textureMap = NULL;

    /* src/app/scene/controller.c:1541-1556 */
    const JsonValue* format = vkrtJsonGet(&root, "format");
    const JsonValue* meshImports = vkrtJsonGet(&root, "meshImports");
    const JsonValue* textureImports = vkrtJsonGet(&root, "textureImports");
    const JsonValue* materials = vkrtJsonGet(&root, "materials");
    const JsonValue* meshes = vkrtJsonGet(&root, "meshes");
    const JsonValue* sceneObjects = vkrtJsonGet(&root, "sceneObjects");
    const JsonValue* settings = vkrtJsonGet(&root, "sceneSettings");
    const JsonValue* envPath = vkrtJsonGet(&root, "environmentTexturePath");
    uint32_t version = 0;
    if (!format || format->type != JSON_STRING || strcmp(format->string, "vkrt.scene") != 0 ||
        !vkrtJsonToU32(vkrtJsonGet(&root, "version"), &version) || version != 1u || !meshImports ||
        meshImports->type != JSON_ARRAY || (textureImports && textureImports->type != JSON_ARRAY) || !meshes ||
        meshes->type != JSON_ARRAY || !sceneObjects || sceneObjects->type != JSON_ARRAY || !settings ||
        settings->type != JSON_OBJECT || (materials && materials->type != JSON_ARRAY) ||
        (envPath && envPath->type != JSON_STRING && envPath->type != JSON_NULL)) {
        VKRT_LOG_ERROR("'%s' is not a version-1 vkrt.scene document", path);
        vkrtJsonFree(&root);
        return VKRT_ERROR_OPERATION_FAILED;
    }

    if (setSceneAside(vkrt, &backup) != VKRT_SUCCESS) {
        vkrtJsonFree(&root);
        return VKRT_ERROR_OUT_OF_MEMORY;
    }
    asideDone = 1;

    /* loadImportedMeshBatches (controller.c:1241-1272) */
    uint32_t batches = meshImports->count;
    batchFirst = (uint32_t*)calloc(batches + 1u, sizeof(uint32_t));
    batchCount = (uint32_t*)calloc(batches + 1u, sizeof(uint32_t));
    if (!batchFirst || !batchCount) SCENE_FAIL("out of memory");
    for (uint32_t b = 0; b < batches; b++) {
        const JsonValue* item = &meshImports->items[b];
        char resolved[8192];
        if (item->type != JSON_STRING || !resolveAssetPath(path, item->string, resolved, sizeof(resolved)))
            SCENE_FAIL("Scene '%s': mesh import %u not found", path, b);
        if (VKRT_importMesh(vkrt, resolved, NULL, &batchFirst[b], &batchCount[b]) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': importing '%s' failed", path, resolved);
    }

    /* querySavedTextureIndexMapCount / loadStandaloneTextures (controller.c:690-716,
     * 1274-1298): each saved standalone texture is loaded from its file and the
     * saved index is mapped to the new one */
    uint32_t textureMapCount = 0;
    for (uint32_t t = 0; textureImports && t < textureImports->count; t++) {
        const JsonValue* to = &textureImports->items[t];
        uint32_t saved = 0, colorSpace = 0;
        const JsonValue* tp = vkrtJsonGet(to, "path");
        if (to->type != JSON_OBJECT || !vkrtJsonToU32(vkrtJsonGet(to, "index"), &saved) || !tp ||
            tp->type != JSON_STRING || !vkrtJsonToU32(vkrtJsonGet(to, "colorSpace"), &colorSpace))
            SCENE_FAIL("Scene '%s': texture import %u is malformed", path, t);
        if (saved >= (1u << 20)) SCENE_FAIL("Scene '%s': texture index %u too large", path, saved);
        if (saved + 1u > textureMapCount) textureMapCount = saved + 1u;
    }
    textureMap = (uint32_t*)malloc(sizeof(uint32_t) * (textureMapCount + 1u));
    if (!textureMap) SCENE_FAIL("out of memory");
    for (uint32_t t = 0; t < textureMapCount; t++) textureMap[t] = VKRT_INVALID_INDEX;
    for (uint32_t t = 0; textureImports && t < textureImports->count; t++) {
        const JsonValue* to = &textureImports->items[t];
        uint32_t saved = 0, colorSpace = 0, loaded = VKRT_INVALID_INDEX;
        char resolved[8192];
        vkrtJsonToU32(vkrtJsonGet(to, "index"), &saved);
        vkrtJsonToU32(vkrtJsonGet(to, "colorSpace"), &colorSpace);
        if (!resolveAssetPath(path, vkrtJsonGet(to, "path")->string, resolved, sizeof(resolved)) ||
            VKRT_addTextureFromFile(vkrt, resolved, NULL, colorSpace, &loaded) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': texture import %u could not be loaded", path, t);
        textureMap[saved] = loaded;
    }
    /* loadEnvironmentTexture (controller.c:1300-1312) */
    if (envPath && envPath->type == JSON_STRING) {
        char resolved[8192];
        if (!resolveAssetPath(path, envPath->string, resolved, sizeof(resolved)) ||
            VKRT_setEnvironmentTextureFromFile(vkrt, resolved) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': environment texture '%s' could not be loaded", path, envPath->string);
    }

    /* buildSavedMeshIndexMap (controller.c:1314-1337) */
    uint32_t savedMeshCount = meshes->count;
    savedToLoaded = (uint32_t*)malloc(sizeof(uint32_t) * (savedMeshCount + 1u));
    if (!savedToLoaded) SCENE_FAIL("out of memory");
    for (uint32_t s = 0; s < savedMeshCount; s++) {
        const JsonValue* m = &meshes->items[s];
        uint32_t importIndex = 0, localIndex = 0;
        if (!vkrtJsonToU32(vkrtJsonGet(m, "importIndex"), &importIndex) ||
            !vkrtJsonToU32(vkrtJsonGet(m, "importLocalIndex"), &localIndex) || importIndex >= batches ||
            localIndex >= batchCount[importIndex])
            SCENE_FAIL("Scene '%s': mesh %u has an invalid import reference", path, s);
        savedToLoaded[s] = batchFirst[importIndex] + localIndex;
    }

    /* pruneNonSceneMeshes (controller.c:1339-1373) */
    uint32_t loadedCount = vkrt->meshCount;
    keep = (uint8_t*)calloc(loadedCount + 1u, 1);
    if (!keep) SCENE_FAIL("out of memory");
    for (uint32_t s = 0; s < savedMeshCount; s++) keep[savedToLoaded[s]] = 1;
    for (uint32_t m = loadedCount; m > 0; m--) {
        uint32_t remove = m - 1u;
        if (keep[remove]) continue;
        if (VKRT_removeMesh(vkrt, remove) != VKRT_SUCCESS) SCENE_FAIL("removing mesh %u failed", remove);
        for (uint32_t s = 0; s < savedMeshCount; s++)
            if (savedToLoaded[s] > remove) savedToLoaded[s]--;
    }

    /* applyLoadedMaterials (controller.c:894-945,1375-1408) */
    uint32_t highest = 0, contiguous = 0;
    for (uint32_t i = 0; materials && i < materials->count; i++, contiguous++) {
        uint32_t index = contiguous;
        const JsonValue* explicitIndex = vkrtJsonGet(&materials->items[i], "index");
        if (explicitIndex && !vkrtJsonToU32(explicitIndex, &index)) SCENE_FAIL("Scene '%s': bad material index", path);
        if (index > highest) highest = index;
    }
    for (uint32_t s = 0; s < savedMeshCount; s++) {
        uint32_t index = 0;
        if (!vkrtJsonToU32(vkrtJsonGet(&meshes->items[s], "materialIndex"), &index))
            SCENE_FAIL("Scene '%s': mesh %u has no materialIndex", path, s);
        if (index > highest) highest = index;
    }
    if (highest >= (1u << 20)) SCENE_FAIL("Scene '%s': material index %u too large", path, highest);
    while (vkrt->materialCount < highest + 1u)
        if (VKRT_addMaterial(vkrt, NULL, NULL, NULL) != VKRT_SUCCESS) SCENE_FAIL("adding a material failed");
    contiguous = 0;
    for (uint32_t i = 0; materials && i < materials->count; i++, contiguous++) {
        const JsonValue* mo = &materials->items[i];
        const JsonValue* name = vkrtJsonGet(mo, "name");
        const JsonValue* explicitIndex = vkrtJsonGet(mo, "index");
        uint32_t index = contiguous;
        Material mat;
        if ((explicitIndex && !vkrtJsonToU32(explicitIndex, &index)) || !name || name->type != JSON_STRING ||
            !parseMaterial(vkrtJsonGet(mo, "material"), &mat))
            SCENE_FAIL("Scene '%s': material entry %u is malformed", path, i);
        /* remapStandaloneTextureIndices (controller.c:640-661): slots that name a
         * saved standalone texture move to its new index; the others keep the
         * index of the texture the mesh imports created */
        uint32_t* slots[4] = {&mat.baseColorTextureIndex, &mat.metallicRoughnessTextureIndex,
                              &mat.normalTextureIndex, &mat.emissiveTextureIndex};
        for (int k = 0; k < 4; k++)
            if (*slots[k] != VKRT_INVALID_INDEX && *slots[k] < textureMapCount &&
                textureMap[*slots[k]] != VKRT_INVALID_INDEX)
                *slots[k] = textureMap[*slots[k]];
        if (VKRT_setMaterial(vkrt, index, &mat) != VKRT_SUCCESS ||
            VKRT_setMaterialName(vkrt, index, name->string) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': setting material %u failed", path, index);
    }

    /* applyLoadedMeshes (controller.c:1410-1441) */
    for (uint32_t s = 0; s < savedMeshCount; s++) {
        const JsonValue* mo = &meshes->items[s];
        const JsonValue* name = vkrtJsonGet(mo, "name");
        uint32_t materialRef = 0;
        uint8_t assigned = 0, backfaces = 0;
        float opacity = 1.0f;
        uint32_t mesh = savedToLoaded[s];
        if (!name || name->type != JSON_STRING || !vkrtJsonToU32(vkrtJsonGet(mo, "materialIndex"), &materialRef) ||
            !readBool(vkrtJsonGet(mo, "hasMaterialAssignment"), &assigned) ||
            !readBool(vkrtJsonGet(mo, "renderBackfaces"), &backfaces) ||
            !vkrtJsonToFloat(vkrtJsonGet(mo, "opacity"), &opacity) ||
            VKRT_setMeshName(vkrt, mesh, name->string) != VKRT_SUCCESS ||
            (assigned ? VKRT_setMeshMaterialIndex(vkrt, mesh, materialRef)
                      : VKRT_clearMeshMaterialAssignment(vkrt, mesh)) != VKRT_SUCCESS ||
            VKRT_setMeshOpacity(vkrt, mesh, opacity) != VKRT_SUCCESS ||
            VKRT_setMeshRenderBackfaces(vkrt, mesh, backfaces) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': mesh entry %u is malformed", path, s);
    }

    /* applyLoadedSceneObjects (controller.c:1457-1526, 737-769) */
    uint32_t objectCount = sceneObjects->count;
    objects = (SceneObjectRecord*)calloc(objectCount + 1u, sizeof(SceneObjectRecord));
    if (!objects) SCENE_FAIL("out of memory");
    for (uint32_t o = 0; o < objectCount; o++) {
        const JsonValue* so = &sceneObjects->items[o];
        const JsonValue* name = vkrtJsonGet(so, "name");
        SceneObjectRecord* rec = &objects[o];
        uint32_t savedMesh = VKRT_INVALID_INDEX;
        if (!name || name->type != JSON_STRING || !readIndexOrNull(vkrtJsonGet(so, "parentIndex"), &rec->parent) ||
            !readIndexOrNull(vkrtJsonGet(so, "meshIndex"), &savedMesh))
            SCENE_FAIL("Scene '%s': scene object %u is malformed", path, o);
        if (rec->parent != VKRT_INVALID_INDEX && rec->parent >= o)
            SCENE_FAIL("Scene '%s': scene object %u names a later parent", path, o);
        if (savedMesh != VKRT_INVALID_INDEX && savedMesh >= savedMeshCount)
            SCENE_FAIL("Scene '%s': scene object %u names a missing mesh", path, o);
        rec->mesh = savedMesh == VKRT_INVALID_INDEX ? VKRT_INVALID_INDEX : savedToLoaded[savedMesh];
        const JsonValue* pos = vkrtJsonGet(so, "localPosition");
        const JsonValue* rot = vkrtJsonGet(so, "localRotation");
        const JsonValue* scl = vkrtJsonGet(so, "localScale");
        if (pos || rot || scl) {
            float p[3], r[3], s[3];
            if (!vkrtJsonToFloats(pos, p, 3) || !vkrtJsonToFloats(rot, r, 3) || !vkrtJsonToFloats(scl, s, 3))
                SCENE_FAIL("Scene '%s': scene object %u has a partial transform", path, o);
            VKRT_buildMeshTransformMatrix(p, r, s, rec->local);
        } else {
            float v[16];
            if (!vkrtJsonToFloats(vkrtJsonGet(so, "localTransform"), v, 16))
                SCENE_FAIL("Scene '%s': scene object %u has no transform", path, o);
            memcpy(rec->local, v, sizeof(v)); /* column by column, as saved */
        }
    }
    if (syncObjectTransforms(vkrt, objects, objectCount) != VKRT_SUCCESS)
        SCENE_FAIL("Scene '%s': applying scene-object transforms failed", path);

    /* applySceneSettings (controller.c:823-892): absent fields keep their values */
    {
        const JsonValue* cam = vkrtJsonGet(settings, "camera");
        if (!cam || cam->type != JSON_OBJECT) SCENE_FAIL("Scene '%s': sceneSettings.camera missing", path);
        VKRT_SceneSettingsSnapshot cur = vkrt->settings;
        float camPos[3], camTarget[3], camUp[3], vfov = cur.camera.vfov;
        memcpy(camPos, cur.camera.pos, sizeof(camPos));
        memcpy(camTarget, cur.camera.target, sizeof(camTarget));
        memcpy(camUp, cur.camera.up, sizeof(camUp));
        uint32_t rrMin = cur.rrMinDepth, rrMax = cur.rrMaxDepth, tone = cur.toneMappingMode, mode = cur.renderMode,
                 spectral = cur.spectralSamplingMode;
        float exposure = cur.exposure, envStrength = cur.environmentStrength, envRotation = cur.environmentRotation;
        float envColor[3] = {cur.environmentColor[0], cur.environmentColor[1], cur.environmentColor[2]};
        uint8_t autoExposure = cur.autoExposureEnabled ? 1u : 0u, misNee = cur.misNeeEnabled ? 1u : 0u;
        if (!optFloats(cam, "position", camPos, 3) || !optFloats(cam, "target", camTarget, 3) ||
            !optFloats(cam, "up", camUp, 3) || !optFloat(cam, "vfov", &vfov) ||
            !optU32(settings, "rrMinDepth", &rrMin) || !optU32(settings, "rrMaxDepth", &rrMax) ||
            !optU32(settings, "toneMappingMode", &tone) || !optU32(settings, "renderMode", &mode) ||
            !optU32(settings, "spectralSamplingMode", &spectral) || !optFloat(settings, "exposure", &exposure) ||
            !optBool(settings, "autoExposureEnabled", &autoExposure) ||
            !optFloats(settings, "environmentColor", envColor, 3) ||
            !optFloat(settings, "environmentStrength", &envStrength) ||
            !optFloat(settings, "environmentRotation", &envRotation) ||
            !optBool(settings, "misNeeEnabled", &misNee))
            SCENE_FAIL("Scene '%s': sceneSettings is malformed", path);
        if (VKRT_setPathDepth(vkrt, rrMin, rrMax) != VKRT_SUCCESS ||
            VKRT_setToneMappingMode(vkrt, (VKRT_ToneMappingMode)tone) != VKRT_SUCCESS ||
            VKRT_setRenderMode(vkrt, (VKRT_RenderMode)mode) != VKRT_SUCCESS ||
            VKRT_setSpectralSamplingMode(vkrt, (VKRT_SpectralSamplingMode)spectral) != VKRT_SUCCESS ||
            VKRT_setExposure(vkrt, exposure) != VKRT_SUCCESS ||
            VKRT_setAutoExposureEnabled(vkrt, autoExposure) != VKRT_SUCCESS ||
            VKRT_setEnvironmentLight(vkrt, envColor, envStrength) != VKRT_SUCCESS ||
            VKRT_setEnvironmentRotation(vkrt, envRotation) != VKRT_SUCCESS ||
            VKRT_setMisNeeEnabled(vkrt, misNee) != VKRT_SUCCESS ||
            VKRT_cameraSetPose(vkrt, camPos, camTarget, camUp, vfov) != VKRT_SUCCESS)
            SCENE_FAIL("Scene '%s': a scene setting was rejected", path);
    }

    free(batchFirst);
    free(batchCount);
    free(savedToLoaded);
    free(keep);
    free(objects);
    free(textureMap);
    vkrtJsonFree(&root);
    dropBackup(&backup);
    return VKRT_SUCCESS;

fail:
    free(batchFirst);
    free(batchCount);
    free(savedToLoaded);
    free(keep);
    free(objects);
    free(textureMap);
    vkrtJsonFree(&root);
    if (asideDone) restoreScene(vkrt, &backup);
    return VKRT_ERROR_OPERATION_FAILED;
}
