"""Render image export (VKRT_saveRenderImage / VKRT_saveRenderImageEx;
src/core/utility/export/{api,image}.c) read back independently: PNG chunks
and zlib data decoded here, EXR headers and scanlines parsed here from the
OpenEXR layout, JPEG decoded by Pillow."""

import io
import os
import struct
import zlib

import numpy as np
import pytest

import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu


def rendered(mode=V.RENDER_MODE_RGB, w=72, h=40):
    r = V.Renderer(w, h)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=16, max_depth=8, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=2)
    r.start(w, h, 0)
    r.draw()
    return r


def read_png16(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    idat = b""
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(kind + body)
        if kind == b"IDAT":
            idat += body
        chunks[kind] = body
        pos += 12 + n
    w, h, depth, ctype = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    assert (depth, ctype) == (16, 6) and b"IEND" in chunks
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * 8)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].copy().view(">u2").astype(np.uint16).reshape(h, w, 4)


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack("<ii", data[:8])
    assert magic == 20000630 and version & 0xFF == 2
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        tend = data.index(b"\0", end + 1)
        size = struct.unpack("<i", data[tend + 1:tend + 5])[0]
        attrs[name] = data[tend + 5:tend + 5 + size]
        pos = tend + 5 + size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs["compression"] == b"\0"
    chl, channels, p = attrs["channels"], [], 0
    while chl[p] != 0:
        end = chl.index(b"\0", p)
        channels.append((chl[p:end].decode(), struct.unpack("<i", chl[end + 1:end + 5])[0]))
        p = end + 17
    assert all(t == 2 for _, t in channels)  # FLOAT
    offsets = struct.unpack(f"<{h}Q", data[pos:pos + 8 * h])
    img = {name: np.zeros((h, w), np.float32) for name, _ in channels}
    for off in offsets:
        y, n = struct.unpack("<ii", data[off:off + 8])
        line = np.frombuffer(data[off + 8:off + 8 + n], "<f4").reshape(len(channels), w)
        for k, (name, _) in enumerate(channels):
            img[name][y] = line[k]
    return np.stack([img["R"], img["G"], img["B"], img["A"]], axis=-1)


def test_png_is_the_display_image(tmp_path):
    r = rendered()
    path = str(tmp_path / "frame.png")
    r.save_image(path, denoise=False)
    np.testing.assert_array_equal(read_png16(path), r.output_image())
    # the linear path (taken whenever denoising is asked for) tone-maps on the host
    # exactly as the device does, up to one step of the 16-bit encoding
    r.save_image(str(tmp_path / "linear.png"))
    diff = np.abs(read_png16(str(tmp_path / "linear.png")).astype(np.int32) - r.output_image().astype(np.int32))
    print(f"host vs device tone mapping: max {diff.max()} LSB")
    assert diff.max() <= 1
    r.close()


def test_path_without_extension_gets_png_and_unknown_extension_fails(tmp_path):
    r = rendered()
    r.save_image(str(tmp_path / "plain"), denoise=False)
    assert os.path.exists(tmp_path / "plain.png")
    with pytest.raises(V.VkrtError):
        r.save_image(str(tmp_path / "frame.bmp"))
    assert not os.path.exists(tmp_path / "frame.bmp")
    r.close()


@pytest.mark.parametrize("mode", [V.RENDER_MODE_RGB, V.RENDER_MODE_SPECTRAL])
def test_exr_is_the_linear_accumulation(tmp_path, mode):
    r = rendered(mode)
    path = str(tmp_path / "frame.EXR")
    r.save_image(path)
    got = read_exr(path)
    acc = r.accumulation().astype(np.float64)
    if mode == V.RENDER_MODE_SPECTRAL:
        # XYZ (equal-energy white) -> Bradford adaptation to D65 -> linear sRGB
        brad = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
        scale = np.diag([0.9413344, 1.0404175, 1.0895327])
        to_srgb = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                            [0.0556434, -0.2040259, 1.0572252]])
        m = to_srgb @ np.linalg.inv(brad) @ scale @ brad
        expect = acc[..., :3] @ m.T
        np.testing.assert_allclose(got[..., :3], expect, rtol=2e-4, atol=2e-6)
    else:
        np.testing.assert_array_equal(got[..., :3], acc[..., :3].astype(np.float32))
    assert np.all(got[..., 3] == 1.0)
    r.close()


def test_jpeg_decodes_to_the_display_image(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    r = rendered(w=77, h=45)  # not multiples of 8: edge blocks
    path = str(tmp_path / "frame.jpeg")
    r.save_image(path, denoise=False)
    with Image.open(path) as im:
        assert im.format == "JPEG" and im.size == (77, 45)
        got = np.asarray(im.convert("RGB"), dtype=np.float64)
    disp = r.output_image()[..., :3].astype(np.uint32)
    ref = ((disp * 255 + 32767) // 65535).astype(np.uint8)

    def psnr(x):
        mse = np.mean((x - ref.astype(np.float64)) ** 2)
        return 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))

    # libjpeg (through Pillow) at the same quality and 4:4:4 sampling as the yardstick
    buf = io.BytesIO()
    Image.fromarray(ref).save(buf, format="JPEG", quality=95, subsampling=0)
    with Image.open(buf) as im:
        lib = np.asarray(im.convert("RGB"), dtype=np.float64)
    print(f"JPEG quality 95: PSNR {psnr(got):.1f} dB, libjpeg {psnr(lib):.1f} dB")
    assert psnr(got) > psnr(lib) - 1.0
    r.close()
