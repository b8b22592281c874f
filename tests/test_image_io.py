"""Image files (PNG, JPEG, EXR) and the boundary functions that load them.

The reference decodes PNG with libspng, JPEG with turbojpeg and EXR with
tinyexr (src/core/utility/image.c:86-399, src/core/utility/exr.cpp). None of
them is in this image, so each decoder is pinned by something outside itself:
  - JPEG: pixels that libjpeg-turbo (through Pillow) decodes from the same
    files with the reference's settings, committed as fixtures and, where
    Pillow is importable, compared live on freshly encoded images;
  - PNG and EXR: lossless formats, written here by small encoders that follow
    the file specifications, so the decode must return the source samples.
CPU only: the C ABI is called without a GPU device.
"""

import ctypes as C
import io
import os
import struct
import zlib

import numpy as np
import pytest

import vkrt_amd as V

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "images")
JPEG_FIXTURES = ["rgb_420_progressive", "rgb_422_restart", "rgb_444", "grey_progressive"]


class LoadedImage(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("colorSpace", C.c_uint32)]


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def decode(data: bytes, color_space=V.TEXTURE_COLOR_SPACE_SRGB, mime=None):
    """vkrtLoadImageFromMemory (the library's internal decoder) -> (array, format, colour space)."""
    fn = V.library().vkrtLoadImageFromMemory
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p, C.c_uint32, C.POINTER(LoadedImage), C.c_char_p]
    img = LoadedImage()
    buf = C.create_string_buffer(data, len(data))
    if fn(buf, len(data), mime.encode() if mime else None, color_space, C.byref(img), b"test") != 0:
        return None
    dtype, ch = {V.TEXTURE_FORMAT_RGBA8_UNORM: (np.uint8, 4), V.TEXTURE_FORMAT_RGBA16_UNORM: (np.uint16, 4),
                 V.TEXTURE_FORMAT_RGBA16_SFLOAT: (np.float16, 4), V.TEXTURE_FORMAT_RGBA32_SFLOAT: (np.float32, 4)}[img.format]
    n = img.width * img.height * ch * np.dtype(dtype).itemsize
    arr = np.ctypeslib.as_array(C.cast(img.pixels, C.POINTER(C.c_uint8)), (n,)).copy().view(dtype)
    _libc.free(img.pixels)
    return arr.reshape(img.height, img.width, ch), img.format, img.colorSpace


# ----------------------------------------------------------------- JPEG --
@pytest.mark.parametrize("name", JPEG_FIXTURES)
def test_jpeg_matches_libjpeg_turbo_fixture(name):
    with open(os.path.join(GOLDEN, name + ".jpg"), "rb") as f:
        data = f.read()
    want = np.load(os.path.join(GOLDEN, name + "_rgb.npy"))
    got, fmt, space = decode(data, V.TEXTURE_COLOR_SPACE_SRGB)
    assert fmt == V.TEXTURE_FORMAT_RGBA8_UNORM and space == V.TEXTURE_COLOR_SPACE_SRGB
    np.testing.assert_array_equal(got[..., :3], want)
    assert np.all(got[..., 3] == 255)


def test_jpeg_matches_libjpeg_turbo_live():
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(5)
    for w, h in ((1, 1), (2, 9), (33, 17), (64, 48)):
        src = np.clip(rng.normal(128, 70, (h, w, 3)), 0, 255).astype(np.uint8)
        for mode in ("RGB", "L"):
            for sub in (0, 1, 2):
                for prog in (False, True):
                    bio = io.BytesIO()
                    Image.fromarray(src if mode == "RGB" else src[..., 0], mode).save(
                        bio, "JPEG", quality=75, subsampling=sub, progressive=prog)
                    data = bio.getvalue()
                    want = np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))
                    got = decode(data)[0]
                    np.testing.assert_array_equal(got[..., :3], want, err_msg=f"{w}x{h} {mode} {sub} {prog}")


def test_jpeg_rejects_truncated_and_unsupported_streams():
    with open(os.path.join(GOLDEN, "rgb_444.jpg"), "rb") as f:
        data = f.read()
    assert decode(data[:40]) is None
    # SOF3 (lossless) in place of the frame header
    i = data.index(b"\xff\xc0")
    assert decode(data[:i] + b"\xff\xc3" + data[i + 2:]) is None


# ------------------------------------------------------------------ PNG --
def png_chunk(kind: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def pack_row(samples, depth):
    if depth >= 8:
        return np.asarray(samples).astype(">u2" if depth == 16 else np.uint8).tobytes()
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def encode_png(samples, ctype, depth, palette=None, trns=None, interlace=False):
    """PNG writer (ISO/IEC 15948): samples is H x W x channels, filter type 0."""
    h, w = samples.shape[:2]
    flat = samples.reshape(h, w, -1)
    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = flat[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        for row in sub:
            raw += b"\x00" + pack_row(row.reshape(-1), depth)
    out = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    if palette is not None:
        out += png_chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        out += png_chunk(b"tRNS", np.asarray(trns, np.uint8).tobytes())
    return out + png_chunk(b"IDAT", zlib.compress(raw)) + png_chunk(b"IEND", b"")


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_grey_low_bit_depths(depth, interlace):
    rng = np.random.default_rng(depth)
    g = rng.integers(0, 1 << depth, (13, 11), dtype=np.uint16)
    got, fmt, space = decode(encode_png(g[..., None], 0, depth, interlace=interlace), V.TEXTURE_COLOR_SPACE_LINEAR)
    assert fmt == V.TEXTURE_FORMAT_RGBA8_UNORM and space == V.TEXTURE_COLOR_SPACE_LINEAR
    want = (g * (255 // ((1 << depth) - 1))).astype(np.uint8)
    for c in range(3):
        np.testing.assert_array_equal(got[..., c], want)
    assert np.all(got[..., 3] == 255)


@pytest.mark.parametrize("depth", [1, 2, 4, 8])
def test_png_palette_with_alpha(depth):
    n = 1 << depth
    rng = np.random.default_rng(10 + depth)
    pal = rng.integers(0, 256, (n, 3))
    alpha = rng.integers(0, 256, n // 2 + 1)  # tRNS may cover only the first entries
    idx = rng.integers(0, n, (9, 14))
    got = decode(encode_png(idx[..., None], 3, depth, palette=pal, trns=alpha))[0]
    want_a = np.full(n, 255)
    want_a[:len(alpha)] = alpha
    np.testing.assert_array_equal(got[..., :3], pal[idx])
    np.testing.assert_array_equal(got[..., 3], want_a[idx])


@pytest.mark.parametrize("interlace", [False, True])
def test_png_rgba16_interlaced_and_srgb_conversion(interlace):
    rng = np.random.default_rng(3)
    px = rng.integers(0, 65536, (10, 7, 4), dtype=np.uint16)
    data = encode_png(px, 6, 16, interlace=interlace)
    got, fmt, space = decode(data, V.TEXTURE_COLOR_SPACE_LINEAR)
    assert fmt == V.TEXTURE_FORMAT_RGBA16_UNORM and space == V.TEXTURE_COLOR_SPACE_LINEAR
    np.testing.assert_array_equal(got, px)
    # image.c:65-84,183-187: 16-bit sRGB files are stored linear, rounded to 16 bits
    got, fmt, space = decode(data, V.TEXTURE_COLOR_SPACE_SRGB)
    assert fmt == V.TEXTURE_FORMAT_RGBA16_UNORM and space == V.TEXTURE_COLOR_SPACE_LINEAR
    e = px[..., :3].astype(np.float32) / np.float32(65535.0)
    lin = np.where(e <= 0.04045, e / np.float32(12.92), ((e + np.float32(0.055)) / np.float32(1.055)) ** np.float32(2.4))
    want = np.clip(np.rint(lin.astype(np.float32) * np.float32(65535.0)), 0, 65535)
    assert np.abs(got[..., :3].astype(np.int64) - want).max() <= 1
    np.testing.assert_array_equal(got[..., 3], px[..., 3])


def test_png_grey_alpha_and_rgb8_interlaced():
    rng = np.random.default_rng(4)
    ga = rng.integers(0, 256, (17, 9, 2), dtype=np.uint8)
    got = decode(encode_png(ga, 4, 8, interlace=True))[0]
    np.testing.assert_array_equal(got[..., 0], ga[..., 0])
    np.testing.assert_array_equal(got[..., 3], ga[..., 1])
    rgb = rng.integers(0, 256, (5, 19, 3), dtype=np.uint8)
    got = decode(encode_png(rgb, 2, 8, interlace=True))[0]
    np.testing.assert_array_equal(got[..., :3], rgb)


# ------------------------------------------------------------------ EXR --
def exr_rle(data: bytes) -> bytes:
    out = bytearray()
    i = 0
    while i < len(data):
        j = i
        while j + 1 < len(data) and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j - i >= 2:  # a run of at least 3 equal bytes
            out += bytes([j - i, data[i]])
            i = j + 1
            continue
        k = i
        while k < len(data) and k - i < 127 and not (k + 2 < len(data) and data[k] == data[k + 1] == data[k + 2]):
            k += 1
        out += bytes([(256 - (k - i)) & 0xFF]) + data[i:k]
        i = k
    return bytes(out)


def exr_pre(data: bytes) -> bytes:
    """OpenEXR ZIP/RLE pre-pass: even bytes then odd bytes, then a byte delta."""
    t = data[0::2] + data[1::2]
    d = bytearray(len(t))
    prev = 0
    for i, v in enumerate(t):
        d[i] = v if i == 0 else (v - prev + 128) & 0xFF
        prev = v
    return bytes(d)


def encode_exr(channels, compression, width, height, y_min=0):
    """Scanline OpenEXR 2 writer. channels: list of (name, pixel type, H x W array)."""
    def attr(name, kind, value):
        return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value

    chl = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", t, 0, 1, 1) for n, t, _ in sorted(channels)) + b"\0"
    box = struct.pack("<iiii", 0, y_min, width - 1, y_min + height - 1)
    header = (b"\x76\x2f\x31\x01" + struct.pack("<i", 2) + attr("channels", "chlist", chl) +
              attr("compression", "compression", bytes([compression])) + attr("dataWindow", "box2i", box) +
              attr("displayWindow", "box2i", box) + attr("lineOrder", "lineOrder", b"\0") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    lines = {0: 1, 1: 1, 2: 1, 3: 16}[compression]
    chunks = []
    for y0 in range(0, height, lines):
        raw = b""
        for y in range(y0, min(height, y0 + lines)):
            for _, t, a in sorted(channels):
                raw += a[y].astype({0: "<u4", 1: "<f2", 2: "<f4"}[t]).tobytes()
        if compression in (2, 3):
            body = zlib.compress(exr_pre(raw))
        elif compression == 1:
            body = exr_rle(exr_pre(raw))
        else:
            body = raw
        if len(body) >= len(raw):
            body = raw  # stored raw when compression does not help
        chunks.append(struct.pack("<ii", y_min + y0, len(body)) + body)
    table_at = len(header)
    offsets, at = [], table_at + 8 * len(chunks)
    for c in chunks:
        offsets.append(at)
        at += len(c)
    return header + b"".join(struct.pack("<Q", o) for o in offsets) + b"".join(chunks)


@pytest.mark.parametrize("compression", [0, 1, 2, 3])
def test_exr_half_rgba_is_rgba16f(compression):
    rng = np.random.default_rng(compression)
    h, w = 37, 23
    planes = {c: rng.normal(0, 10, (h, w)).astype(np.float16) for c in "RGBA"}
    planes["G"][:5] = 1.5  # runs for RLE
    data = encode_exr([(c, 1, planes[c]) for c in "RGBA"], compression, w, h, y_min=-3)
    got, fmt, space = decode(data)
    assert fmt == V.TEXTURE_FORMAT_RGBA16_SFLOAT and space == V.TEXTURE_COLOR_SPACE_LINEAR
    for i, c in enumerate("RGBA"):
        np.testing.assert_array_equal(got[..., i].view(np.uint16), planes[c].view(np.uint16))


@pytest.mark.parametrize("compression", [0, 3])
def test_exr_float_and_mixed_channels_are_rgba32f_with_unit_alpha(compression):
    rng = np.random.default_rng(7)
    h, w = 20, 33
    r = rng.normal(5, 3, (h, w)).astype(np.float32)
    g = rng.normal(5, 3, (h, w)).astype(np.float16)
    b = rng.integers(0, 1000, (h, w)).astype(np.uint32)
    data = encode_exr([("diffuse.R", 2, r), ("diffuse.G", 1, g), ("diffuse.B", 0, b), ("Z", 2, r * 0)],
                      compression, w, h)
    got, fmt, space = decode(data)
    assert fmt == V.TEXTURE_FORMAT_RGBA32_SFLOAT and space == V.TEXTURE_COLOR_SPACE_LINEAR
    np.testing.assert_array_equal(got[..., 0], r)
    np.testing.assert_array_equal(got[..., 1], g.astype(np.float32))
    np.testing.assert_array_equal(got[..., 2], b.astype(np.float32))
    assert np.all(got[..., 3] == 1.0)


def test_exr_luminance_only_fills_rgb():
    y = np.linspace(0, 4, 6 * 5, dtype=np.float32).reshape(6, 5)
    got = decode(encode_exr([("Y", 2, y)], 2, 5, 6))[0]
    for c in range(3):
        np.testing.assert_array_equal(got[..., c], y)
    assert np.all(got[..., 3] == 1.0)


def test_exr_rejects_unsupported_files():
    y = np.ones((4, 4), np.float32)
    piz = bytearray(encode_exr([("Y", 2, y)], 0, 4, 4))
    i = piz.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    piz[i] = 4  # PIZ
    assert decode(bytes(piz)) is None
    tiled = bytearray(encode_exr([("Y", 2, y)], 0, 4, 4))
    tiled[5] |= 0x02  # version flag bit 9: tiled
    assert decode(bytes(tiled)) is None
    assert decode(encode_exr([("Z", 2, y)], 0, 4, 4)) is None  # no RGB or Y channel
    assert decode(b"not an image at all") is None


def test_codec_follows_mime_type_then_signature():
    px = np.full((2, 2, 3), 200, np.uint8)
    png = encode_png(px, 2, 8)
    assert decode(png, mime="image/png") is not None
    assert decode(png, mime="image/jpeg") is None  # the MIME type wins over the bytes
    assert decode(png, mime="application/octet-stream") is not None  # unknown type: sniff the signature


# ------------------------------------------------- boundary: texture files --
@pytest.fixture
def renderer():
    r = V.Renderer(init_device=False)
    yield r
    r.close()


def add_file(r, path, space, name=None):
    out = C.c_uint32()
    code = r.lib.VKRT_addTextureFromFile(r.h, os.fsencode(path), name.encode() if name else None, space, C.byref(out))
    return code, out.value


def snapshot(r, index):
    s = V.TextureSnapshot()
    assert r.lib.VKRT_getTextureSnapshot(r.h, index, C.byref(s)) == V.VKRT_SUCCESS
    return s


def test_add_texture_from_file_and_snapshot(renderer, tmp_path):
    r = renderer
    jpg = os.path.join(GOLDEN, "rgb_444.jpg")
    code, t = add_file(r, jpg, V.TEXTURE_COLOR_SPACE_SRGB)
    assert code == V.VKRT_SUCCESS and t == 0
    s = snapshot(r, t)
    assert (s.width, s.height, s.format, s.colorSpace, s.useCount) == (17, 13, V.TEXTURE_FORMAT_RGBA8_UNORM,
                                                                        V.TEXTURE_COLOR_SPACE_SRGB, 0)
    assert s.name == b"rgb_444.jpg"
    exr = tmp_path / "sky.exr"
    exr.write_bytes(encode_exr([(c, 1, np.ones((4, 8), np.float16)) for c in "RGB"], 3, 8, 4))
    code, t2 = add_file(r, exr, V.TEXTURE_COLOR_SPACE_SRGB, name="named")
    s = snapshot(r, t2)
    assert code == V.VKRT_SUCCESS and s.name == b"named" and s.colorSpace == V.TEXTURE_COLOR_SPACE_LINEAR
    assert s.format == V.TEXTURE_FORMAT_RGBA16_SFLOAT
    # texture.c:16-28 / textures.c:484-497: bad arguments, missing and undecodable files
    assert add_file(r, tmp_path / "missing.png", 0) == (V.VKRT_ERROR_INVALID_ARGUMENT, V.VKRT_INVALID_INDEX)
    assert add_file(r, jpg, 2)[0] == V.VKRT_ERROR_INVALID_ARGUMENT
    junk = tmp_path / "junk.png"
    junk.write_bytes(b"\x89PNG but not really")
    assert add_file(r, junk, 0) == (V.VKRT_ERROR_OPERATION_FAILED, V.VKRT_INVALID_INDEX)
    assert r.texture_count() == 2
    assert r.lib.VKRT_getTextureSnapshot(r.h, 2, C.byref(V.TextureSnapshot())) == V.VKRT_ERROR_INVALID_ARGUMENT


def test_environment_texture_from_file_replaces_and_releases(renderer, tmp_path):
    r = renderer
    a, b = tmp_path / "a.exr", tmp_path / "b.exr"
    a.write_bytes(encode_exr([(c, 2, np.full((2, 4), 1.0, np.float32)) for c in "RGB"], 0, 4, 2))
    b.write_bytes(encode_exr([(c, 2, np.full((3, 6), 2.0, np.float32)) for c in "RGB"], 0, 6, 3))
    keep = r.add_texture(np.ones((2, 2, 4), np.uint8), V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_SRGB)
    assert r.lib.VKRT_setEnvironmentTextureFromFile(r.h, os.fsencode(a)) == V.VKRT_SUCCESS
    assert r.scene_settings().environmentTextureIndex == 1 and snapshot(r, 1).useCount == 1
    assert r.lib.VKRT_setEnvironmentTextureFromFile(r.h, os.fsencode(b)) == V.VKRT_SUCCESS
    # environment.c:11-38: the unused previous map is released, indices shift down
    assert r.texture_count() == 2 and r.scene_settings().environmentTextureIndex == 1
    assert snapshot(r, 1).width == 6
    # environment.c:46: files load as linear, so an 8-bit PNG is a linear RGBA8 map
    png = tmp_path / "c.png"
    png.write_bytes(encode_png(np.zeros((5, 3, 3), np.uint8), 2, 8))
    assert r.lib.VKRT_setEnvironmentTextureFromFile(r.h, os.fsencode(png)) == V.VKRT_SUCCESS
    s = snapshot(r, 1)
    assert r.texture_count() == 2 and r.scene_settings().environmentTextureIndex == 1
    assert (s.width, s.format, s.colorSpace) == (3, V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_LINEAR)
    assert r.lib.VKRT_setEnvironmentTextureFromFile(r.h, os.fsencode(tmp_path / "none.exr")) != V.VKRT_SUCCESS
    assert r.texture_count() == 2 and r.scene_settings().environmentTextureIndex == 1
    assert r.lib.VKRT_clearEnvironmentTexture(r.h) == V.VKRT_SUCCESS
    assert r.texture_count() == 1 and r.scene_settings().environmentTextureIndex == V.VKRT_INVALID_INDEX
    assert keep == 0 and snapshot(r, 0).width == 2


def test_set_material_texture(renderer):
    r = renderer
    srgb = r.add_texture(np.ones((2, 2, 4), np.uint8), V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_SRGB)
    lin = r.add_texture(np.ones((2, 2, 4), np.uint8), V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_LINEAR)
    lin2 = r.add_texture(np.ones((4, 2, 4), np.uint8), V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_LINEAR)
    m = V.default_material()
    m.normalTextureWrap = 7
    m.textureTexcoordSets = 0x00010101  # set 1 for slots 0, 1 and 2
    m.normalTextureTransform[:] = (2, 3, 4, 5)
    m.textureRotations[2] = 1.0
    mi = r.add_material(m, "m")
    f = r.lib.VKRT_setMaterialTexture
    # textures.c:94-105: sRGB only in colour slots
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_NORMAL, srgb) == V.VKRT_ERROR_INVALID_ARGUMENT
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_BASE_COLOR, srgb) == V.VKRT_SUCCESS
    assert f(r.h, mi, 4, lin) == V.VKRT_ERROR_INVALID_ARGUMENT
    assert f(r.h, 99, 0, lin) == V.VKRT_ERROR_INVALID_ARGUMENT
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_NORMAL, 17) == V.VKRT_ERROR_INVALID_ARGUMENT
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_NORMAL, lin) == V.VKRT_SUCCESS
    got = r.material(mi)
    # textures.c:614-618: binding resets the slot's wrap, set, transform and rotation
    assert got.normalTextureIndex == lin and got.normalTextureWrap == V.TEXTURE_WRAP_DEFAULT
    assert got.textureTexcoordSets == 0x00000100 and list(got.normalTextureTransform) == [1, 1, 0, 0]
    assert got.textureRotations[2] == 0.0
    assert snapshot(r, lin).useCount == 1
    # replacing leaves the previous texture without users: it is released (textures.c:621)
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_NORMAL, lin2) == V.VKRT_SUCCESS
    assert r.texture_count() == 2 and r.material(mi).normalTextureIndex == lin2 - 1
    assert r.material(mi).baseColorTextureIndex == srgb
    assert f(r.h, mi, V.MATERIAL_TEXTURE_SLOT_NORMAL, V.VKRT_INVALID_INDEX) == V.VKRT_SUCCESS
    assert r.texture_count() == 1


# ------------------------------------- boundary: camera, runtime, denoise --
def pose(r):
    p, t, u, f = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 3)(), C.c_float()
    r.lib.VKRT_cameraGetPose(r.h, p, t, u, C.byref(f))
    return np.array(p[:]), np.array(t[:])


def camera_input(**kw):
    ci = V.CameraInput()
    for k, v in kw.items():
        setattr(ci, k, v)
    return ci


def test_apply_camera_input_closed_forms(renderer):
    r = renderer
    r.set_camera((0.0, -4.0, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 1.0), 40.0)
    apply = lambda **kw: r.lib.VKRT_applyCameraInput(r.h, C.byref(camera_input(**kw)))  # noqa: E731
    assert r.lib.VKRT_applyCameraInput(r.h, None) == V.VKRT_ERROR_INVALID_ARGUMENT
    # camera.c:36-52: view along +y, right = view x z = +x (normalised), up = right x view = +z
    assert apply(panDx=10.0, panDy=5.0, panning=1) == V.VKRT_SUCCESS
    p, t = pose(r)
    np.testing.assert_allclose(p, (-0.04, -4.0, 1.02), atol=1e-6)
    np.testing.assert_allclose(t, (-0.04, 0.0, 1.02), atol=1e-6)
    # no motion while the mouse is captured, and none without the orbit flag
    assert apply(panDx=10.0, panning=1, captureMouse=1) == V.VKRT_SUCCESS
    assert apply(orbitDx=50.0) == V.VKRT_SUCCESS
    np.testing.assert_allclose(pose(r)[0], p, atol=0)
    # camera.c:55-80: orbit by -0.004 rad per unit of yaw about the target, distance kept
    assert apply(orbitDx=100.0, orbiting=1) == V.VKRT_SUCCESS
    p2, t2 = pose(r)
    yaw = np.pi / 2 - 0.4
    np.testing.assert_allclose(p2, t - 4.0 * np.array([np.cos(yaw), np.sin(yaw), 0.0]), atol=2e-6)
    np.testing.assert_allclose(t2, t, atol=0)
    # pitch is held 0.001 rad short of the pole
    assert apply(orbitDy=-1e5, orbiting=1) == V.VKRT_SUCCESS
    p3, _ = pose(r)
    d = t2 - p3
    assert abs(np.arcsin(d[2] / np.linalg.norm(d)) - (np.pi / 2 - 0.001)) < 1e-4
    # camera.c:83-88: scroll moves the position along the view vector by 0.075 per unit
    r.set_camera((0.0, -4.0, 1.0), (0.0, 0.0, 1.0), (0.0, 0.0, 1.0), 40.0)
    assert apply(scroll=2.0) == V.VKRT_SUCCESS
    np.testing.assert_allclose(pose(r)[0], (0.0, -4.0 + 0.15 * 4.0, 1.0), atol=1e-6)


def test_runtime_snapshot_is_the_render_target(renderer):
    r = renderer
    r.start(64, 48, 0)
    s = V.RuntimeSnapshot()
    assert r.lib.VKRT_getRuntimeSnapshot(r.h, C.byref(s)) == V.VKRT_SUCCESS
    assert (s.renderWidth, s.renderHeight, s.displayWidth, s.swapchainHeight) == (64, 48, 64, 48)
    assert list(s.displayViewportRect) == [0, 0, 64, 48]
    assert r.lib.VKRT_getRuntimeSnapshot(r.h, None) == V.VKRT_ERROR_INVALID_ARGUMENT


def test_render_denoise_flag_and_phases(renderer):
    r = renderer
    lib = r.lib
    # render.c:73-81: nothing to denoise before a render has finished sampling
    assert lib.VKRT_denoiseRenderToViewport(r.h) == V.VKRT_ERROR_OPERATION_FAILED
    r.start(8, 8, 0)
    assert r.status().renderDenoiseEnabled == 1  # lifecycle.c:338 default, copied at start
    assert lib.VKRT_setRenderDenoiseEnabled(r.h, 0) == V.VKRT_SUCCESS
    assert r.status().renderDenoiseEnabled == 0  # render.c:31-34: applies while sampling
    assert lib.VKRT_denoiseRenderToViewport(r.h) == V.VKRT_ERROR_OPERATION_FAILED
    assert lib.VKRT_stopRenderSampling(r.h) == V.VKRT_SUCCESS
    assert r.status().renderPhase == V.RENDER_PHASE_COMPLETE_RAW
    assert lib.VKRT_setRenderDenoiseEnabled(r.h, 1) == V.VKRT_SUCCESS
    assert r.status().renderDenoiseEnabled == 0  # not sampling: the finished render keeps its flag
    # no denoiser in this build: the request ends as the raw result, as when the reference's fails
    assert lib.VKRT_denoiseRenderToViewport(r.h) == V.VKRT_SUCCESS
    assert r.status().renderPhase == V.RENDER_PHASE_COMPLETE_RAW
    assert lib.VKRT_continueRender(r.h, 0) == V.VKRT_SUCCESS
    assert r.status().renderDenoiseEnabled == 1
