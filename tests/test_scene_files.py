"""vkrt.scene loading (include/vkrt_import.h: VKRT_loadSceneFile) and glTF
import into a scene (VKRT_importMesh), on host state without a GPU.

The reference's own scene files (assets/scenes/prism.json, caustics.json,
cornell.json with their GLB models, kept under tests/golden/assets) are loaded
and every applied value is checked against the file read here with json and
numpy: material fields, mesh assignments, scene-object world transforms
(T * Rz * Ry * Rx * S per object, parent first; src/core/scene/transform.c:26-35,
src/app/session/session.c:257-300) applied to the independently read
vertices, and the scene settings (src/app/scene/controller.c:823-892).
cornell.json names assets/models/dragon.glb, which the reference does not
ship: loading it must fail and leave the previous scene in place
(src/app/scene/controller.c:1636-1654).
"""

import json
import math
import os
import shutil

import numpy as np
import pytest

import vkrt_amd as V
import gltf_files as G

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets")
SCENES = os.path.join(ASSETS, "scenes")


def trs(position, rotation_deg, scale):
    """VKRT_buildMeshTransformMatrix in float64: T * Rz * Ry * Rx * S."""
    a, b, c = (math.radians(v) for v in rotation_deg)
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    rz = np.array([[math.cos(c), -math.sin(c), 0], [math.sin(c), math.cos(c), 0], [0, 0, 1]])
    m = np.eye(4)
    m[:3, :3] = rz @ ry @ rx @ np.diag(scale)
    m[:3, 3] = position
    return m


def expected_scene(path):
    """Per saved mesh: (world matrix, engine-space vertices, indices, material
    index) from the scene file and its GLBs, read independently."""
    doc = json.load(open(path))
    base = os.path.dirname(path)
    objects = doc["sceneObjects"]
    world = []
    for o in objects:
        local = trs(o["localPosition"], o["localRotation"], o["localScale"]) if "localPosition" in o else \
            np.array(o["localTransform"], np.float64).reshape(4, 4).T
        world.append(local if o["parentIndex"] is None else world[o["parentIndex"]] @ local)
    mesh_world = {}
    for o, w in zip(objects, world):
        if o["meshIndex"] is not None:
            mesh_world[o["meshIndex"]] = w
    out = []
    for i, m in enumerate(doc["meshes"]):
        gdoc, bufs = G.read(os.path.join(base, doc["meshImports"][m["importIndex"]]))
        prim = [p for mm in gdoc["meshes"] for p in mm["primitives"]][m["importLocalIndex"]]
        pos, _, idx = G.expected_primitive(gdoc, bufs, prim)
        node_world = G.BASIS @ G.node_matrix(gdoc["nodes"][0]) @ G.BASIS.T
        out.append((mesh_world.get(i, node_world), pos, idx, m))
    return doc, out


def world_triangles(r):
    b = r.export()
    tris = np.ctypeslib.as_array(b.worldTriangles, (b.triangleCount * 9,)).reshape(-1, 3, 3).copy()
    tri_mesh = np.ctypeslib.as_array(b.triangleMesh, (b.triangleCount,)).copy()
    return tris, tri_mesh


@pytest.mark.parametrize("scene", ["prism", "caustics"])
def test_reference_scene_file_is_applied(scene):
    path = os.path.join(SCENES, f"{scene}.json")
    doc, meshes = expected_scene(path)
    r = V.Renderer(init_device=False)
    r.load_scene_file(path)
    assert r.mesh_count() == len(doc["meshes"])
    tris, tri_mesh = world_triangles(r)
    for i, (world, pos, idx, saved) in enumerate(meshes):
        snap = r.mesh(i)
        assert snap.name.decode() == saved["name"]
        assert snap.materialIndex == saved["materialIndex"]
        assert bool(snap.hasMaterialAssignment) == saved["hasMaterialAssignment"]
        assert snap.info.renderBackfaces == int(saved["renderBackfaces"])
        assert snap.info.opacity == pytest.approx(saved["opacity"])
        expect = (np.c_[pos, np.ones(len(pos))] @ world.T)[:, :3][idx.reshape(-1, 3)]
        got = tris[tri_mesh == i]
        assert got.shape == expect.shape
        np.testing.assert_allclose(got, expect, atol=2e-5 * (1 + np.abs(expect).max()))
    for saved in doc["materials"]:
        m = r.material(saved["index"])
        assert r.material_name(saved["index"]) == saved["name"]
        d = V.default_material()
        for key, value in saved["material"].items():
            got = getattr(m, key)
            got = list(got) if hasattr(got, "__len__") else got
            assert got == pytest.approx(value, rel=1e-6), key
        for key in ("ior", "specular", "metallic"):
            if key not in saved["material"]:
                assert getattr(m, key) == pytest.approx(getattr(d, key))
    s = r.scene_settings()
    st = doc["sceneSettings"]
    assert (s.rrMinDepth, s.rrMaxDepth) == (st["rrMinDepth"], st["rrMaxDepth"])
    assert (s.renderMode, s.spectralSamplingMode, s.toneMappingMode) == \
        (st["renderMode"], st["spectralSamplingMode"], st["toneMappingMode"])
    assert list(s.environmentColor) == pytest.approx(st["environmentColor"])
    assert s.environmentStrength == pytest.approx(st["environmentStrength"])
    assert bool(s.misNeeEnabled) == st["misNeeEnabled"]
    cam = st["camera"]
    assert list(s.camera.pos) == pytest.approx(cam["position"])
    assert list(s.camera.target) == pytest.approx(cam["target"])
    assert list(s.camera.up) == pytest.approx(cam["up"])
    assert s.camera.vfov == pytest.approx(cam["vfov"])
    r.close()


def _state(r):
    s = r.scene_settings()
    return (r.mesh_count(), r.material_count(), [r.mesh(i).name for i in range(r.mesh_count())],
            [r.material_name(i) for i in range(r.material_count())], list(s.camera.pos), s.renderMode,
            world_triangles(r)[0].tobytes())


def test_failed_load_restores_the_previous_scene(tmp_path):
    """cornell.json imports dragon.glb, which is absent: the load fails after
    five meshes were already imported, and the prism scene comes back whole."""
    r = V.Renderer(init_device=False)
    r.load_scene_file(os.path.join(SCENES, "prism.json"))
    before = _state(r)
    with pytest.raises(V.VkrtError):
        r.load_scene_file(os.path.join(SCENES, "cornell.json"))
    assert _state(r) == before
    # a document of another format or version is rejected before anything changes
    bad = tmp_path / "bad.json"
    doc = json.load(open(os.path.join(SCENES, "prism.json")))
    doc["version"] = 2
    bad.write_text(json.dumps(doc))
    with pytest.raises(V.VkrtError):
        r.load_scene_file(str(bad))
    assert _state(r) == before
    r.close()


def test_scene_object_hierarchy_and_pruning(tmp_path):
    """A synthetic document: a parent object without a mesh, a child given as a
    localTransform matrix, a grandchild as TRS; one imported mesh that no saved
    mesh references is pruned; materials beyond the file's are created."""
    models = tmp_path / "models"
    models.mkdir()
    for f in ("cube.glb", "sphere.glb"):
        shutil.copy(os.path.join(ASSETS, "models", f), models / f)
    child = trs([0.5, 0, 0], [0, 0, 30], [1, 2, 1])
    doc = {
        "format": "vkrt.scene", "version": 1,
        "meshImports": ["models/cube.glb", "models/sphere.glb"],
        "materials": [{"index": 4, "name": "red", "material": {"baseColor": [0.9, 0.1, 0.1], "roughness": 0.3}}],
        "meshes": [
            {"name": "box", "materialIndex": 4, "hasMaterialAssignment": True, "renderBackfaces": False,
             "opacity": 1, "importIndex": 0, "importLocalIndex": 0},
        ],
        "sceneObjects": [
            {"name": "group", "parentIndex": None, "meshIndex": None, "localPosition": [1, 2, 3],
             "localRotation": [10, 20, 30], "localScale": [2, 2, 2]},
            {"name": "mid", "parentIndex": 0, "meshIndex": None,
             "localTransform": child.T.ravel().tolist()},
            {"name": "leaf", "parentIndex": 1, "meshIndex": 0, "localPosition": [0, 0, 1],
             "localRotation": [0, 45, 0], "localScale": [1, 1, 0.5]},
        ],
        "sceneSettings": {"camera": {"position": [5, 5, 5], "target": [0, 0, 0], "up": [0, 0, 1], "vfov": 30},
                          "renderMode": 1},
    }
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(doc))
    r = V.Renderer(init_device=False)
    r.load_scene_file(str(path))
    assert r.mesh_count() == 1  # the sphere is pruned
    assert r.material_count() == 5
    assert r.mesh(0).materialIndex == 4 and r.material_name(4) == "red"
    assert list(r.material(4).baseColor) == pytest.approx([0.9, 0.1, 0.1])
    world = trs([1, 2, 3], [10, 20, 30], [2, 2, 2]) @ child @ trs([0, 0, 1], [0, 45, 0], [1, 1, 0.5])
    gdoc, bufs = G.read(str(models / "cube.glb"))
    pos, _, idx = G.expected_primitive(gdoc, bufs, gdoc["meshes"][0]["primitives"][0])
    expect = (np.c_[pos, np.ones(len(pos))] @ world.T)[:, :3][idx.reshape(-1, 3)]
    tris, _ = world_triangles(r)
    np.testing.assert_allclose(tris, expect, atol=1e-4)
    assert r.scene_settings().renderMode == 1
    r.close()


def test_import_mesh_adds_materials_and_node_transform():
    """VKRT_importMesh: prism.glb's node scale (glTF Y) becomes an engine Z
    scale; its material is appended after the default one and assigned."""
    r = V.Renderer(init_device=False)
    first, count = r.import_mesh(os.path.join(ASSETS, "models", "prism.glb"))
    assert (first, count) == (0, 1)
    assert r.material_count() == 2 and r.material_name(1) == "Material"
    snap = r.mesh(0)
    assert snap.materialIndex == 1 and snap.hasMaterialAssignment and snap.info.renderBackfaces == 1
    assert snap.name.decode() == "Cube"
    assert list(snap.info.scale) == pytest.approx([1.0, 1.0, 0.9986131191253662], abs=1e-6)
    first, count = r.import_mesh(os.path.join(ASSETS, "models", "sphere.glb"), name="ball")
    assert (first, count) == (1, 1) and r.mesh(1).name.decode() == "ball"
    assert r.mesh(1).materialIndex == 0 and not r.mesh(1).hasMaterialAssignment
    with pytest.raises(V.VkrtError):
        r.import_mesh(os.path.join(ASSETS, "models", "dragon.glb"))
    assert r.mesh_count() == 2 and r.material_count() == 2
    r.close()


def test_material_and_mesh_setters_follow_the_reference():
    """VKRT_removeMaterial (mesh.c:296-340): the default cannot go, users fall
    back to it, later indices move down; VKRT_setMeshOpacity rejects values
    outside [0, 1]; VKRT_setMaterialName with "" restores the generated name."""
    r = V.Renderer(init_device=False)
    r.import_mesh(os.path.join(ASSETS, "models", "cube.glb"))
    r.import_mesh(os.path.join(ASSETS, "models", "prism.glb"))
    assert [r.mesh(i).materialIndex for i in range(2)] == [1, 2]
    with pytest.raises(V.VkrtError):
        r.remove_material(0)
    r.remove_material(1)
    assert r.material_count() == 2
    assert r.mesh(0).materialIndex == 0 and not r.mesh(0).hasMaterialAssignment
    assert r.mesh(1).materialIndex == 1
    r.set_material_name(1, "")
    assert r.material_name(1) == "Material 1"
    r.set_mesh_opacity(0, 0.5)
    assert r.mesh(0).info.opacity == 0.5
    for bad in (-0.1, 1.5, float("nan")):
        with pytest.raises(V.VkrtError):
            r.set_mesh_opacity(0, bad)
    r.clear_mesh_material(1)
    assert r.mesh(1).materialIndex == 0
    r.close()


def test_scene_file_keeps_imported_textures_and_loads_texture_files(tmp_path):
    """A scene file whose mesh import brings a textured GLB, plus a standalone
    texture file and an environment map file (controller.c:1274-1312): slot
    indices of the GLB's textures are kept, the standalone texture's saved
    index is remapped to the new one (remapStandaloneTextureIndices,
    controller.c:640-661), and the environment map is set from its file."""
    from test_gltf_textures import png
    from test_image_io import encode_exr

    rng = np.random.default_rng(8)
    b = G.Builder()
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    prim = {"attributes": {"POSITION": b.accessor(pos, "VEC3"), "TEXCOORD_0": b.accessor(uv, "VEC2")},
            "indices": b.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16), "SCALAR"), "material": 0}
    b.doc["images"] = [{"bufferView": b.view(png(rng.integers(0, 256, (4, 4, 4)), 6)), "mimeType": "image/png"}]
    b.doc["textures"] = [{"source": 0}]
    b.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}]
    b.doc["meshes"] = [{"primitives": [prim]}]
    b.doc["nodes"] = [{"mesh": 0}]
    b.doc["scenes"] = [{"nodes": [0]}]
    b.finish(str(tmp_path / "textured.glb"))
    (tmp_path / "glow.png").write_bytes(png(rng.integers(0, 256, (3, 5, 3)), 2))
    (tmp_path / "sky.exr").write_bytes(encode_exr([(c, 2, np.full((4, 8), 0.5, np.float32)) for c in "RGB"], 3, 8, 4))
    doc = {
        "format": "vkrt.scene", "version": 1,
        "meshImports": ["textured.glb"],
        "textureImports": [{"path": "glow.png", "index": 7, "colorSpace": 0}],
        "environmentTexturePath": "sky.exr",
        "materials": [{"index": 1, "name": "textured",
                       "material": {"baseColor": [1, 1, 1], "baseColorTextureIndex": 0,
                                    "emissiveTextureIndex": 7}}],
        "meshes": [{"name": "quad", "materialIndex": 1, "hasMaterialAssignment": True, "renderBackfaces": False,
                    "opacity": 1, "importIndex": 0, "importLocalIndex": 0}],
        "sceneObjects": [{"name": "quad", "parentIndex": None, "meshIndex": 0, "localPosition": [0, 0, 0],
                          "localRotation": [0, 0, 0], "localScale": [1, 1, 1]}],
        "sceneSettings": {"camera": {"position": [0, -3, 0.5], "target": [0, 0, 0.5], "up": [0, 0, 1], "vfov": 40}},
    }
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(doc))
    r = V.Renderer(init_device=False)
    r.load_scene_file(str(path))
    assert r.texture_count() == 3  # GLB texture 0, glow.png 1, sky.exr 2
    m = r.material(r.mesh(0).materialIndex)
    assert m.baseColorTextureIndex == 0  # the GLB's texture, not cleared
    assert m.emissiveTextureIndex == 1  # saved index 7 -> the new index of glow.png
    assert r.scene_settings().environmentTextureIndex == 2
    # a texture file that does not load fails the whole document and restores the scene
    doc["textureImports"][0]["path"] = "missing.png"
    path.write_text(json.dumps(doc))
    with pytest.raises(V.VkrtError):
        r.load_scene_file(str(path))
    assert r.texture_count() == 3 and r.scene_settings().environmentTextureIndex == 2
    r.close()
