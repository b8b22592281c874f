"""glTF import (include/vkrt_import.h: meshLoadFromFile) without a GPU.

Checks the C importer against an independent numpy reading of the same files
(tests/gltf_files.py) for the reference's own models (assets/models/*.glb) and
for synthetic files that exercise what those models do not: node hierarchies
with TRS and matrix transforms, interleaved and sparse accessors, 8-bit
indices, normalized colours, generated normals, winding alignment, skipped
non-triangle primitives, the KHR material extensions, data-URI and external
buffers, and the error paths. The conversions checked are those of
src/app/mesh/loader.c (swizzle 1816-1820, winding 1894-1907, normals
1599-1644, materials 942-1078 and 1472-1555, nodes 1557-1597 and 2027-2107).
"""

import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import vkrt_amd as V
import gltf_files as G

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "assets", "models")
F32 = np.float32


@pytest.mark.parametrize("name", ["plane", "cube", "prism", "sphere", "suzanne"])
def test_reference_models_match_independent_reading(name):
    path = os.path.join(MODELS, f"{name}.glb")
    doc, buffers = G.read(path)
    got = V.load_gltf(path)
    prims = [p for m in doc["meshes"] for p in m["primitives"]]
    assert len(got["entries"]) == len(prims) == 1
    e = got["entries"][0]
    pos, nrm, idx = G.expected_primitive(doc, buffers, prims[0])
    v = e["vertices"]
    np.testing.assert_array_equal(v[:, 0:3], pos.astype(F32))
    assert np.all(v[:, 3] == 1.0)
    np.testing.assert_array_equal(v[:, 4:7], nrm.astype(F32))
    np.testing.assert_array_equal(e["indices"], idx.astype(np.uint32))
    uv = G.accessor(doc, buffers, prims[0]["attributes"]["TEXCOORD_0"]).astype(F32)
    np.testing.assert_array_equal(v[:, 16:18], uv)
    assert np.all(v[:, 12:16] == 1.0)  # no COLOR_0: white
    # tangents: unit length, orthogonal to the normal, handedness +-1
    t = v[:, 8:11].astype(np.float64)
    assert np.allclose(np.linalg.norm(t, axis=1), 1.0, atol=1e-5)
    assert np.abs((t * v[:, 4:7]).sum(1)).max() < 1e-4
    assert set(np.unique(v[:, 11])) <= {-1.0, 1.0}
    # materials and names
    node = doc["nodes"][0]
    assert e["name"] == node["name"]
    mats = doc.get("materials", [])
    assert len(got["materials"]) == len(mats)
    if mats:
        m = got["materials"][0]["material"]
        pbr = mats[0]["pbrMetallicRoughness"]
        assert list(m.baseColor) == pytest.approx(pbr["baseColorFactor"][:3])
        assert m.metallic == pytest.approx(pbr["metallicFactor"])
        assert m.roughness == pytest.approx(pbr["roughnessFactor"])
        assert e["renderBackfaces"] == mats[0].get("doubleSided", False)
        assert e["material"] == 0
    else:
        assert e["material"] == V.VKRT_INVALID_INDEX
    # node transform: B * M * B^-1
    local = got["nodes"][0]["local"].T.astype(np.float64)
    np.testing.assert_allclose(local, G.BASIS @ G.node_matrix(node) @ G.BASIS.T, atol=1e-6)


def test_winding_alignment_flips_triangles_against_their_normals(tmp_path):
    b = G.Builder()
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0]], F32)
    # glTF +Z normals; first triangle counter-clockwise seen from +Z, second one listed clockwise
    nrm = np.tile(np.array([[0, 0, 1]], F32), (6, 1))
    idx = np.array([0, 1, 2, 3, 5, 4], np.uint16)
    b.doc["meshes"] = [{"primitives": [{"attributes": {"POSITION": b.accessor(pos, "VEC3"),
                                                        "NORMAL": b.accessor(nrm, "VEC3")},
                                         "indices": b.accessor(idx, "SCALAR")}]}]
    b.doc["nodes"] = [{"mesh": 0}]
    path = str(tmp_path / "wind.glb")
    b.finish(path)
    e = V.load_gltf(path)["entries"][0]
    assert e["indices"].tolist() == [0, 1, 2, 3, 4, 5]
    assert e["name"] == "mesh"


def _hierarchy_file(tmp_path, container):
    """Node tree root(TRS) -> child(matrix, mesh with three primitives) and a
    node outside the default scene."""
    b = G.Builder()
    # primitive 0: interleaved POSITION + TEXCOORD_0 (stride 20), 8-bit indices,
    # normalized 16-bit RGBA colours, no normals
    inter = np.zeros(4, dtype=[("p", F32, 3), ("uv", F32, 2)])
    inter["p"] = [[0, 0, 0], [2, 0, 0], [0, 2, 0], [2, 2, 1]]
    inter["uv"] = [[0, 0], [1, 0], [0, 1], [1, 1]]
    iv = b.view(inter.tobytes(), stride=20)
    p0 = b.accessor(np.zeros((0, 3), F32), "VEC3", view=iv, count=4)
    uv0 = b.accessor(np.zeros((0, 2), F32), "VEC2", view=iv, offset=12, count=4)
    col = np.array([[65535, 0, 0, 65535], [0, 65535, 0, 32768], [0, 0, 65535, 0], [13107, 26214, 39321, 52428]],
                   np.uint16)
    c0 = b.accessor(col, "VEC4", normalized=True)
    i0 = b.accessor(np.array([0, 1, 2, 2, 1, 3], np.uint8), "SCALAR")
    # primitive 1: points (skipped)
    p1 = b.accessor(np.zeros((2, 3), F32), "VEC3")
    # primitive 2: sparse positions over a base view, given normals, no indices
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32)
    sparse_idx = b.view(np.array([1], np.uint16).tobytes())
    sparse_val = b.view(np.array([[5, 0, 0]], F32).tobytes())
    p2 = b.accessor(base, "VEC3", sparse={"count": 1, "indices": {"bufferView": sparse_idx, "componentType": 5123},
                                          "values": {"bufferView": sparse_val}})
    n2 = b.accessor(np.tile(np.array([[0, 0, 1]], F32), (3, 1)), "VEC3")
    b.doc["meshes"] = [{"name": "Shape", "primitives": [
        {"attributes": {"POSITION": p0, "TEXCOORD_0": uv0, "COLOR_0": c0}, "indices": i0, "material": 0},
        {"attributes": {"POSITION": p1}, "mode": 0},
        {"attributes": {"POSITION": p2, "NORMAL": n2}, "material": 1},
    ]}]
    half = math.sqrt(0.5)
    b.doc["nodes"] = [
        {"name": "root", "translation": [1, 2, 3], "rotation": [0, half, 0, half], "scale": [2, 2, 2],
         "children": [1]},
        {"name": "child", "mesh": 0,
         "matrix": [0.5, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1.5, 0, 0, 1, 0, 1]},
        {"name": "outside", "mesh": 0},
    ]
    b.doc["scenes"] = [{"nodes": [0]}]
    b.doc["scene"] = 0
    b.doc["materials"] = [
        {"name": "coated", "doubleSided": True, "alphaMode": "MASK", "alphaCutoff": 0.3,
         "pbrMetallicRoughness": {"baseColorFactor": [0.1, 0.2, 0.3, 0.9], "metallicFactor": 0.25,
                                  "roughnessFactor": 0.6},
         "emissiveFactor": [2.0, 1.0, 0.5],
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 3.0},
                        "KHR_materials_ior": {"ior": 1.7},
                        "KHR_materials_transmission": {"transmissionFactor": 0.6},
                        "KHR_materials_volume": {"attenuationColor": [0.5, 0.6, 0.7], "attenuationDistance": 2.0},
                        "KHR_materials_clearcoat": {"clearcoatFactor": 0.3, "clearcoatRoughnessFactor": 0.2},
                        "KHR_materials_specular": {"specularFactor": 0.4}}},
        {"extensions": {"KHR_materials_pbrSpecularGlossiness": {"diffuseFactor": [0.2, 0.3, 0.4, 0.5],
                                                                "glossinessFactor": 0.8}},
         "alphaMode": "BLEND"},
    ]
    ext = {"glb": ".glb", "datauri": ".gltf", "external": ".gltf"}[container]
    path = str(tmp_path / f"tree{ext}")
    b.finish(path, container, bin_name="tree data.bin")
    return path


@pytest.mark.parametrize("container", ["glb", "datauri", "external"])
def test_hierarchy_materials_and_accessor_forms(tmp_path, container):
    path = _hierarchy_file(tmp_path, container)
    doc, buffers = G.read(path)
    got = V.load_gltf(path)
    # only the scene's root subtree is imported; the points primitive is skipped
    assert [n["name"] for n in got["nodes"]] == ["root", "child"]
    assert [n["parent"] for n in got["nodes"]] == [V.VKRT_INVALID_INDEX, 0]
    assert [n["meshEntryCount"] for n in got["nodes"]] == [0, 2]
    assert [e["name"] for e in got["entries"]] == ["child_0", "child_2"]
    assert [e["node"] for e in got["entries"]] == [1, 1]
    assert [e["material"] for e in got["entries"]] == [0, 1]

    e0, e2 = got["entries"]
    prims = doc["meshes"][0]["primitives"]
    pos0, _, idx0 = G.expected_primitive(doc, buffers, prims[0])
    np.testing.assert_array_equal(e0["vertices"][:, 0:3], pos0.astype(F32))
    np.testing.assert_array_equal(e0["indices"], idx0.astype(np.uint32))
    np.testing.assert_array_equal(e0["vertices"][:, 16:18],
                                  G.accessor(doc, buffers, prims[0]["attributes"]["TEXCOORD_0"]).astype(F32))
    np.testing.assert_allclose(e0["vertices"][:, 12:16],
                               G.accessor(doc, buffers, prims[0]["attributes"]["COLOR_0"]), rtol=1e-6)
    # generated normals: normalised sums of the adjacent face normals (engine space)
    tri = idx0.reshape(-1, 3)
    face = np.cross(pos0[tri[:, 1]] - pos0[tri[:, 0]], pos0[tri[:, 2]] - pos0[tri[:, 0]])
    acc = np.zeros_like(pos0)
    for t, f in zip(tri, face):
        acc[t] += f
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    np.testing.assert_allclose(e0["vertices"][:, 4:7], acc, atol=1e-6)
    assert np.all(e0["vertices"][:, 7] == 0.0)

    pos2, nrm2, idx2 = G.expected_primitive(doc, buffers, prims[2])
    assert pos2[1].tolist() == [5.0, 0.0, 0.0]  # the sparse substitution
    np.testing.assert_array_equal(e2["vertices"][:, 0:3], pos2.astype(F32))
    np.testing.assert_array_equal(e2["indices"], idx2.astype(np.uint32))

    # world transform of the mesh node: B * (root * child) * B^-1
    world = np.eye(4)
    for n in got["nodes"]:
        world = world @ n["local"].T.astype(np.float64)
    expect = G.BASIS @ G.node_matrix(doc["nodes"][0]) @ G.node_matrix(doc["nodes"][1]) @ G.BASIS.T
    np.testing.assert_allclose(world, expect, atol=1e-5)

    m0, m1 = got["materials"][0]["material"], got["materials"][1]["material"]
    assert got["materials"][0]["name"] == "coated" and got["materials"][1]["name"] == "Material 1"
    assert list(m0.baseColor) == pytest.approx([0.1, 0.2, 0.3]) and m0.opacity == pytest.approx(0.9)
    assert m0.metallic == pytest.approx(0.25) and m0.roughness == pytest.approx(0.6)
    assert list(m0.emissionColor) == pytest.approx([1.0, 0.5, 0.25]) and m0.emissionLuminance == pytest.approx(6.0)
    assert m0.ior == pytest.approx(1.7) and m0.transmission == pytest.approx(0.6)
    assert list(m0.attenuationColor) == pytest.approx([0.5, 0.6, 0.7])
    assert m0.absorptionCoefficient == pytest.approx(0.5)
    assert m0.clearcoat == pytest.approx(0.3) and m0.clearcoatGloss == pytest.approx(0.8)
    assert m0.specular == pytest.approx(0.4)
    assert m0.alphaMode == 1 and m0.alphaCutoff == pytest.approx(0.3)
    assert list(m1.baseColor) == pytest.approx([0.2, 0.3, 0.4]) and m1.opacity == pytest.approx(0.5)
    assert m1.roughness == pytest.approx(0.2)
    assert list(m1.emissionColor) == [1.0, 1.0, 1.0] and m1.emissionLuminance == 0.0
    assert m1.alphaMode == 2 and m1.alphaCutoff == pytest.approx(1.0 / 255.0)
    assert e0["renderBackfaces"] and not e2["renderBackfaces"]


def test_material_defaults_follow_the_reference_mapping(tmp_path):
    """Absent factors: metallic-roughness defaults (1, 1, 1), no emission;
    a volume without attenuationDistance gives 1 / FLT_MAX (the reference's
    library stores FLT_MAX for infinity)."""
    b = G.Builder()
    p = b.accessor(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32), "VEC3")
    b.doc["meshes"] = [{"primitives": [{"attributes": {"POSITION": p}, "material": 0}]}]
    b.doc["nodes"] = [{"mesh": 0}]
    b.doc["materials"] = [{"pbrMetallicRoughness": {},
                           "extensions": {"KHR_materials_volume": {"attenuationColor": [0.9, 0.9, 0.9]}}}]
    path = str(tmp_path / "defaults.glb")
    b.finish(path)
    m = V.load_gltf(path)["materials"][0]["material"]
    d = V.default_material()
    assert list(m.baseColor) == [1.0, 1.0, 1.0] and m.metallic == 1.0 and m.roughness == 1.0
    assert m.emissionLuminance == 0.0 and list(m.emissionColor) == [1.0, 1.0, 1.0]
    assert m.absorptionCoefficient == pytest.approx(1.0 / np.finfo(np.float32).max, rel=1e-6)
    assert m.ior == d.ior and m.transmission == d.transmission


def _broken(tmp_path, kind):
    b = G.Builder()
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], F32)
    p = b.accessor(pos, "VEC3")
    prim = {"attributes": {"POSITION": p}}
    nodes = [{"mesh": 0}]
    if kind == "index_range":
        prim["indices"] = b.accessor(np.array([0, 1, 9], np.uint8), "SCALAR")
    elif kind == "overrun":
        b.doc["accessors"][p]["count"] = 4
    elif kind == "points_only":
        prim["mode"] = 0
    elif kind == "cycle":
        nodes = [{"mesh": 0, "children": [1]}, {"children": [0]}]
        b.doc["scenes"] = [{"nodes": [0]}]
        b.doc["scene"] = 0
    b.doc["meshes"] = [{"primitives": [prim]}]
    b.doc["nodes"] = nodes
    path = str(tmp_path / f"{kind}.glb")
    b.finish(path)
    if kind == "short_buffer":
        raw = bytearray(open(path, "rb").read())
        doc, _ = G.read(path)
        doc["buffers"][0]["byteLength"] += 64
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        binlen = int.from_bytes(raw[12:16], "little")
        binchunk = raw[20 + binlen:]
        body = len(js).to_bytes(4, "little") + b"JSON" + js + bytes(binchunk)
        open(path, "wb").write(b"glTF" + (2).to_bytes(4, "little") + (12 + len(body)).to_bytes(4, "little") + body)
    if kind == "bad_json":
        open(path, "w").write('{"asset": {"version": "2.0"}, "nodes": [}')
    return path


@pytest.mark.parametrize("kind", ["index_range", "overrun", "points_only", "cycle", "short_buffer", "bad_json"])
def test_malformed_files_are_rejected(tmp_path, kind):
    with pytest.raises(V.VkrtError):
        V.load_gltf(_broken(tmp_path, kind))


def test_missing_file_is_rejected(tmp_path):
    with pytest.raises(V.VkrtError):
        V.load_gltf(str(tmp_path / "absent.glb"))


def test_imported_node_transform_is_the_basis_conjugate():
    """VKRT_buildImportedNodeTransform (src/core/scene/transform.c:101-111)."""
    lib = V.library()
    rng = np.random.default_rng(5)
    m = np.eye(4)
    m[:3, :3] = rng.normal(size=(3, 3))
    m[:3, 3] = rng.normal(size=3)
    src = np.ascontiguousarray(m.T, np.float32)  # column-major
    out = np.zeros((4, 4), np.float32)
    lib.VKRT_buildImportedNodeTransform(V.fptr(src), V.fptr(out))
    np.testing.assert_allclose(out.T, G.BASIS @ m @ G.BASIS.T, atol=1e-5)
    pos, rot, scl = (np.zeros(3, np.float32) for _ in range(3))
    lib.VKRT_decomposeMeshNodeTransform(V.fptr(src), V.fptr(pos), V.fptr(rot), V.fptr(scl))
    np.testing.assert_allclose(pos, (G.BASIS @ m)[:3, 3], atol=1e-5)
