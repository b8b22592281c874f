"""C-ABI boundary checks that need no GPU: the library loads, exports every
function the public headers declare, the headers compile as C and C++, and the
API's argument handling follows the reference (src/core/api/*.c)."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import vkrt_amd as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        pattern = r"^([A-Za-z_][\w\s\*]*?)\b(VKRT_\w+|meshLoadFromFile|meshReleaseImportData)\s*\("
        for m in re.finditer(pattern, text, flags=re.M):
            if "static" not in m.group(1):  # header-only inline helpers are not exported
                names.add(m.group(2))
    return sorted(names)


def test_headers_declare_the_reference_trace_surface():
    names = declared_functions()
    for required in ("meshLoadFromFile", "meshReleaseImportData", "VKRT_importMesh", "VKRT_loadSceneFile",
                     "VKRT_create", "VKRT_initWithCreateInfo", "VKRT_destroy", "VKRT_uploadMeshData",
                     "VKRT_uploadMeshDataBatch", "VKRT_addMaterial", "VKRT_setMaterial", "VKRT_setMeshMaterialIndex",
                     "VKRT_setMeshTransformMatrix", "VKRT_cameraSetPose", "VKRT_setRenderMode",
                     "VKRT_setSpectralSamplingMode", "VKRT_setPathDepth", "VKRT_setSamplesPerPixel",
                     "VKRT_setMisNeeEnabled", "VKRT_setEnvironmentLight", "VKRT_addTextureFromPixels",
                     "VKRT_startRender", "VKRT_draw", "VKRT_trace", "VKRT_getRenderStatus", "VKRT_getSceneSettings",
                     "VKRT_getSystemInfo"):
        assert required in names, required


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) > 40
    lib = C.CDLL(V.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", V.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


@pytest.mark.parametrize("compiler,flag", [("gcc", "-std=c11"), ("g++", "-std=c++17")])
def test_public_headers_compile(tmp_path, compiler, flag):
    src = tmp_path / ("t.c" if compiler == "gcc" else "t.cpp")
    src.write_text('#include "vkrt.h"\n#include "vkrt_mi355x.h"\n#include "vkrt_import.h"\n'
                   'int main(void){ VKRT* v = 0; (void)v; MeshImportData d; (void)d;'
                   ' return (int)sizeof(SceneData) - 240; }\n')
    subprocess.run([compiler, flag, "-Wall", "-Werror", "-fsyntax-only", "-I", INCLUDE, str(src)], check=True)


def test_struct_sizes_match_reference_layout():
    # src/shared/types.h: Vertex 80, ShaderVertex 48, MeshInfo 80, Material 272, SceneData 240,
    # EmissiveMesh 32, EmissiveTriangle 48
    assert C.sizeof(V.Vertex) == 80
    assert C.sizeof(V.ShaderVertex) == 48
    assert C.sizeof(V.MeshInfo) == 80
    assert C.sizeof(V.Material) == 272
    assert C.sizeof(V.SceneData) == 240
    assert C.sizeof(V.EmissiveMesh) == 32
    assert C.sizeof(V.EmissiveTriangle) == 48


@pytest.fixture
def api():
    r = V.Renderer(init_device=False)
    yield r
    r.close()


def settings(r):
    s = V.SceneSettingsSnapshot()
    assert r.lib.VKRT_getSceneSettings(r.h, C.byref(s)) == 0
    return s


def test_null_arguments_are_rejected(api):
    lib = api.lib
    assert lib.VKRT_create(None) == -1
    assert lib.VKRT_setSamplesPerPixel(None, 4) == -1
    assert lib.VKRT_getSceneSettings(api.h, None) == -1
    assert lib.VKRT_uploadMeshData(api.h, None, 3, None, 3) == -1


def test_setters_clamp_like_the_reference(api):
    # src/core/api/settings.c: spp 0 -> 1, depth clamped to the maximum, min <= max
    assert api.lib.VKRT_setSamplesPerPixel(api.h, 0) == 0
    assert settings(api).samplesPerPixel == 1
    assert api.lib.VKRT_setPathDepth(api.h, 200, 500) == 0
    s = settings(api)
    assert s.rrMaxDepth == 64 and s.rrMinDepth <= s.rrMaxDepth
    assert api.lib.VKRT_setEnvironmentRotation(api.h, C.c_float(540.0)) == 0
    assert -180.0 <= settings(api).environmentRotation < 180.0


def test_camera_rejects_degenerate_pose(api):
    F3 = C.c_float * 3
    bad = api.lib.VKRT_cameraSetPose(api.h, F3(0, 0, 0), F3(0, 0, 0), F3(0, 0, 1), C.c_float(40))
    assert bad != 0
    nan = api.lib.VKRT_cameraSetPose(api.h, F3(float("nan"), 0, 0), F3(0, 1, 0), F3(0, 0, 1), C.c_float(40))
    assert nan != 0
    ok = api.lib.VKRT_cameraSetPose(api.h, F3(0, -2, 0), F3(0, 0, 0), F3(0, 0, 1), C.c_float(40))
    assert ok == 0


def test_mesh_upload_copies_and_validates(api):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (3, 1))
    verts = V.make_vertices(pos, nrm)
    bad = np.array([0, 1, 3], np.uint32)  # index out of range
    assert api.lib.VKRT_uploadMeshData(api.h, verts.ctypes.data_as(C.POINTER(V.Vertex)), 3, V.uptr(bad), 3) != 0
    idx = np.array([0, 1, 2], np.uint32)
    m = api.add_mesh(verts, idx)
    assert m == 0
    verts[:] = 0  # the library owns a copy
    b = api.export()
    assert b.triangleCount == 1
    tri = np.ctypeslib.as_array(b.worldTriangles, shape=(9,))
    assert np.allclose(tri, [0, 0, 0, 1, 0, 0, 0, 1, 0])


def test_default_material_and_indices(api):
    # the first material operation creates the default material at index 0
    # (vkrtEnsureDefaultMaterial, src/core/api/mesh.c:426)
    idx = api.add_material(V.default_material(baseColor=(0.2, 0.3, 0.4)))
    assert idx == 1
    c = C.c_uint32()
    assert api.lib.VKRT_getMaterialCount(api.h, C.byref(c)) == 0
    assert c.value == 2
    assert api.lib.VKRT_setMeshMaterialIndex(api.h, 7, 1) != 0  # no such mesh


def test_trace_without_device_fails_loudly(api):
    # outside a render session the headless frame is not traced (src/core/api/frame.c:107-113)
    assert api.lib.VKRT_trace(api.h) == 0
    assert api.lib.VKRT_startRender(api.h, 16, 16, 0) == 0
    assert api.lib.VKRT_trace(api.h) != 0
    assert api.lib.VKRT_traceRays(api.h, None, 0, None, None, None, None) != 0


def test_sheen_is_kept_and_clamped_like_the_reference(api):
    # src/core/api/mesh.c:112-131: tint and weight clamped to [0, 1], non-finite -> 0
    m = V.default_material()
    m.sheenTintWeight[:] = (0.2, 1.5, float("nan"), 0.7)
    m.sheenRoughness = -3.0
    idx = api.add_material(m)
    b = api.export()
    mats = np.ctypeslib.as_array(C.cast(b.materials, C.POINTER(C.c_float)), shape=(b.materialCount * 68,))
    o = V.Material.sheenTintWeight.offset // 4
    np.testing.assert_array_equal(mats.reshape(-1, 68)[idx, o:o + 4], np.float32([0.2, 1.0, 0.0, 0.7]))
    assert mats.reshape(-1, 68)[idx, V.Material.sheenRoughness.offset // 4] == 0.0


REFERENCE_APP = "/root/reference/src/app"
# The editor's viewport overlay, mesh picking and zoom/pan view state: the
# reference's interactive UI, which this headless trace step does not carry.
EDITOR_ONLY = {"VKRT_consumeSelectedMesh", "VKRT_getOverlayInfo", "VKRT_getRenderViewCrop", "VKRT_getRenderViewState",
               "VKRT_getSelectedMesh", "VKRT_registerAppHooks", "VKRT_requestSelectionAtPixel",
               "VKRT_setRenderViewState", "VKRT_setSelectedMesh"}


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", V.LIB_PATH], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.split()[-1].startswith("VKRT_")}


def _inline_helpers():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            names |= set(re.findall(r"static inline [\w ]+?\b(VKRT_\w+)\s*\(", open(os.path.join(INCLUDE, fn)).read()))
    return names


@pytest.mark.skipif(not os.path.isdir(REFERENCE_APP), reason="reference sources are not mounted")
def test_reference_offline_benchmark_compiles_and_links_against_the_headers(tmp_path):
    """src/app/render/benchmark.c (the --render-headless driver) compiles
    unchanged with include/ in place of src/core/api, and every VKRT_ symbol
    its object needs is exported by the library."""
    obj = tmp_path / "benchmark.o"
    dirs = [INCLUDE, REFERENCE_APP, os.path.join(REFERENCE_APP, "render"), "/root/reference/src/core/utility"]
    subprocess.run(["gcc", "-std=gnu11", "-c", os.path.join(REFERENCE_APP, "render", "benchmark.c"), "-o", str(obj)] +
                   [f"-I{d}" for d in dirs], check=True)
    undefined = subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True, check=True).stdout
    needed = {line.split()[-1] for line in undefined.splitlines() if line.split()[-1].startswith("VKRT_")}
    assert needed and needed <= _exported(), sorted(needed - _exported())


@pytest.mark.skipif(not os.path.isdir(REFERENCE_APP), reason="reference sources are not mounted")
def test_every_api_call_of_the_reference_app_resolves():
    """The render, scene, mesh, session and CLI controllers of the reference
    application (which also include GLFW, Vulkan or imgui headers absent here,
    so they are scanned rather than compiled) call only functions the library
    exports or the headers define inline; the editor's calls outside that set
    are exactly its overlay, picking and zoom/pan functions."""
    def calls(paths):
        found = set()
        for p in paths:
            found |= set(re.findall(r"\b(VKRT_[A-Za-z0-9]+)\s*\(", open(p).read()))
        return found

    available = _exported() | _inline_helpers()
    app = [os.path.join(d, f) for d, _, fs in os.walk(REFERENCE_APP) for f in fs
           if f.endswith(".c") and "/editor" not in d]
    assert len(app) >= 8
    assert calls(app) - available == set()
    editor = [os.path.join(d, f) for d, _, fs in os.walk(os.path.join(REFERENCE_APP, "editor")) for f in fs
              if f.endswith(".c")]
    assert calls(editor) - available == EDITOR_ONLY
