"""glTF material textures on import (src/app/mesh/loader.c:362-404, 432-499,
907-1010): PNG images from a GLB buffer view and a data URI, every scanline
filter, 8- and 16-bit, RGB / RGBA / grey + alpha / palette, sampler wrap
modes, KHR_texture_transform and texture-coordinate sets, checked against
the pixel arrays the file was written from."""

import base64
import json
import struct
import zlib

import numpy as np

import vkrt_amd as V
from gltf_files import Builder


def png(pixels, ctype, bits=8, palette=None):
    """PNG writer cycling the five filter types row by row."""
    h, w = pixels.shape[:2]
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    dt = ">u2" if bits == 16 else np.uint8
    rows = np.ascontiguousarray(pixels.reshape(h, w * ch).astype(dt)).view(np.uint8).reshape(h, -1).astype(np.int32)
    bpp = ch * bits // 8
    raw = b""
    prev = np.zeros_like(rows[0])
    for y in range(h):
        cur, f = rows[y], y % 5
        a = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if f == 0:
            out = cur
        elif f == 1:
            out = cur - a
        elif f == 2:
            out = cur - prev
        elif f == 3:
            out = cur - (a + prev) // 2
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            out = cur - np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        raw += bytes([f]) + (out & 255).astype(np.uint8).tobytes()
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body))

    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bits, ctype, 0, 0, 0))
    if palette is not None:
        data += chunk(b"PLTE", palette[:, :3].astype(np.uint8).tobytes())
        data += chunk(b"tRNS", palette[:, 3].astype(np.uint8).tobytes())
    return data + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")


def srgb(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_gltf_png_textures_import(tmp_path):
    rng = np.random.default_rng(3)
    rgba8 = rng.integers(0, 256, (5, 7, 4))
    grey16 = rng.integers(0, 65536, (4, 6, 2))
    rgb8 = rng.integers(0, 256, (6, 3, 3))
    pal = rng.integers(0, 256, (5, 4))
    idx = rng.integers(0, 5, (3, 4, 1))
    b = Builder()
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    prim = {"attributes": {"POSITION": b.accessor(pos, "VEC3"), "TEXCOORD_0": b.accessor(uv, "VEC2"),
                           "TEXCOORD_1": b.accessor(uv * 2, "VEC2")},
            "indices": b.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16), "SCALAR"), "material": 0}
    uri = "data:image/png;base64," + base64.b64encode(png(rgb8, 2)).decode()
    b.doc["images"] = [{"bufferView": b.view(png(rgba8, 6)), "mimeType": "image/png", "name": "albedo"},
                       {"bufferView": b.view(png(grey16, 4, bits=16)), "mimeType": "image/png"},
                       {"uri": uri},
                       {"bufferView": b.view(png(idx, 3, palette=pal)), "mimeType": "image/png"}]
    b.doc["samplers"] = [{"wrapS": 33648, "wrapT": 33071}, {}]
    b.doc["textures"] = [{"source": 0, "sampler": 0}, {"source": 1, "sampler": 1}, {"source": 2}, {"source": 3}]
    tt = {"KHR_texture_transform": {"scale": [2.0, 3.0], "offset": [0.25, -0.5], "rotation": 0.3, "texCoord": 1}}
    b.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0},
                                                    "metallicRoughnessTexture": {"index": 1, "extensions": tt}},
                           "normalTexture": {"index": 2, "scale": 0.5},
                           "emissiveTexture": {"index": 3, "texCoord": 1}, "emissiveFactor": [1, 1, 1]}]
    b.doc["meshes"] = [{"primitives": [prim]}]
    b.doc["nodes"] = [{"mesh": 0}]
    b.doc["scenes"] = [{"nodes": [0]}]
    path = str(tmp_path / "textured.glb")
    b.finish(path)
    r = V.Renderer(init_device=False)
    r.import_mesh(path)
    assert r.texture_count() == 4
    buf = r.export()
    m = buf.materials[buf.meshInfos[0].materialIndex]
    assert (m.baseColorTextureIndex, m.metallicRoughnessTextureIndex, m.normalTextureIndex,
            m.emissiveTextureIndex) == (0, 1, 2, 3)
    assert m.baseColorTextureWrap == 33648 | (33071 << 16) and m.metallicRoughnessTextureWrap == 10497 | (10497 << 16)
    assert m.textureTexcoordSets == (1 << 8) | (1 << 24)
    np.testing.assert_allclose(list(m.metallicRoughnessTextureTransform), [2.0, 3.0, 0.25, -0.5])
    assert abs(m.textureRotations[1] - 0.3) < 1e-7 and abs(m.normalTextureScale - 0.5) < 1e-7
    info = np.ctypeslib.as_array(buf.textureInfo, (16,)).reshape(4, 4)
    tex = np.ctypeslib.as_array(buf.textureTexels, (buf.textureTexelCount * 4,)).reshape(-1, 4)

    def texels(t):
        first, w, h, _ = info[t]
        return tex[first:first + w * h].reshape(h, w, 4).astype(np.float64)

    want0 = rgba8 / 255.0
    want0[..., :3] = srgb(want0[..., :3])
    np.testing.assert_allclose(texels(0), want0, atol=2e-7)
    g = grey16 / 65535.0  # metallic-roughness: linear 16-bit
    np.testing.assert_allclose(texels(1), np.stack([g[..., 0]] * 3 + [g[..., 1]], -1), atol=2e-7)
    np.testing.assert_allclose(texels(2), np.concatenate([rgb8 / 255.0, np.ones((6, 3, 1))], -1), atol=2e-7)
    p = pal[idx[..., 0]] / 255.0
    p[..., :3] = srgb(p[..., :3])
    np.testing.assert_allclose(texels(3), p, atol=2e-7)
    r.close()


def _quad_with_texture(image, tex_coord=0):
    b = Builder()
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    prim = {"attributes": {"POSITION": b.accessor(pos, "VEC3"), "TEXCOORD_0": b.accessor(uv, "VEC2")},
            "indices": b.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16), "SCALAR"), "material": 0}
    b.doc["images"] = [image(b)]
    b.doc["textures"] = [{"source": 0}]
    b.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0, "texCoord": tex_coord}}}]
    b.doc["meshes"] = [{"primitives": [prim]}]
    b.doc["nodes"] = [{"mesh": 0}]
    b.doc["scenes"] = [{"nodes": [0]}]
    return b


def test_gltf_jpeg_texture_and_high_texcoord_set(tmp_path):
    """loader.c:432-499: a JPEG image in a buffer view decodes like the
    reference's turbojpeg call (fixture pixels from libjpeg-turbo); a texture
    coordinate set above 1 is stored in its 8 bits (loader.c:372-378)."""
    import os
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "images")
    jpg = open(os.path.join(golden, "rgb_420_progressive.jpg"), "rb").read()
    want = np.load(os.path.join(golden, "rgb_420_progressive_rgb.npy")) / 255.0
    b = _quad_with_texture(lambda b: {"bufferView": b.view(jpg), "mimeType": "image/jpeg"}, tex_coord=2)
    path = str(tmp_path / "jpeg.glb")
    b.finish(path)
    r = V.Renderer(init_device=False)
    r.import_mesh(path)
    buf = r.export()
    m = buf.materials[buf.meshInfos[0].materialIndex]
    assert m.baseColorTextureIndex == 0 and m.textureTexcoordSets == 2
    first, w, h, _ = np.ctypeslib.as_array(buf.textureInfo, (4,))
    tex = np.ctypeslib.as_array(buf.textureTexels, (buf.textureTexelCount * 4,)).reshape(-1, 4)
    got = tex[first:first + w * h].reshape(h, w, 4)
    np.testing.assert_allclose(got[..., :3], srgb(want), atol=2e-7)
    r.close()


def test_gltf_image_that_does_not_decode_fails_the_import(tmp_path):
    """loader.c:508-520: a texture image that cannot be decoded fails the
    import instead of leaving the material untextured."""
    b = _quad_with_texture(lambda b: {"bufferView": b.view(b"\x89PNG\r\n\x1a\n broken"), "mimeType": "image/png"})
    path = str(tmp_path / "broken.glb")
    b.finish(path)
    r = V.Renderer(init_device=False)
    try:
        V.load_gltf(path)
        raise AssertionError("import should fail")
    except V.VkrtError:
        pass
    r.close()
