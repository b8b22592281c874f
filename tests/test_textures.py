"""Material textures (src/shaders/material/textures.slang:8-155, the samplers of
src/core/scene/textures.c:141-178, texture alpha in
src/shaders/rt/alpha_test.slang:20-26).

CPU: the host texture pool decodes every upload format to the linear values
a sampled image view returns; the oracle's sampler is pinned by what bilinear
filtering of mip level 0 fixes (texel centres return the texel, affine texel
data is reproduced exactly, each wrap mode's symmetry, the UV transform's
quarter turn, coordinate-set selection, empty-slot fallbacks); the oracle's
texture stage is pinned by reductions to untextured materials (a constant
texture of 0.5 scales base colour, roughness, metallic and emission exactly as
halving the material's own factor does; a flat normal map changes nothing),
and texture alpha by the alpha test's acceptance probabilities. GPU: primary
hits and same-seed images of a Cornell box that samples every kind of
texture against the oracle.
"""

import math

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes


def srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def pool(b):
    info = np.ctypeslib.as_array(b.textureInfo, (b.textureCount * 4,)).reshape(-1, 4)
    texels = np.ctypeslib.as_array(b.textureTexels, (b.textureTexelCount * 4,)).reshape(-1, 4)
    return info, texels


def test_texture_pool_decodes_every_format():
    s = scenes.textured_cornell()
    r = V.Renderer(init_device=False)
    scenes.load(r, s)
    unused = r.add_texture(np.zeros((3, 5, 4), np.uint8), V.TEXTURE_FORMAT_RGBA8_UNORM, V.TEXTURE_COLOR_SPACE_LINEAR)
    m = V.default_material()
    m.baseColorTextureIndex = 99  # names no texture: an empty slot
    r.set_mesh_material(0, r.add_material(m, "dangling"))
    b = r.export()
    info, texels = pool(b)
    assert b.textureCount == len(s.textures) + 1 and info[unused][1] == 0
    for t, (pixels, fmt, space) in enumerate(s.textures):
        first, w, h, _ = info[t]
        assert (w, h) == (pixels.shape[1], pixels.shape[0])
        got = texels[first:first + w * h].reshape(h, w, 4)
        if fmt == V.TEXTURE_FORMAT_RGBA8_UNORM:
            want = pixels / 255.0
            if space == V.TEXTURE_COLOR_SPACE_SRGB:
                want[..., :3] = srgb_to_linear(want[..., :3])  # alpha stays linear
        elif fmt == V.TEXTURE_FORMAT_RGBA16_UNORM:
            want = pixels / 65535.0
        else:
            want = pixels.astype(np.float64)
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-7)
    assert b.materials[len(s.materials)].baseColorTextureIndex == V.VKRT_INVALID_INDEX
    r.close()


def sampler_scene(pixels, fmt=V.TEXTURE_FORMAT_RGBA32_SFLOAT, space=V.TEXTURE_COLOR_SPACE_LINEAR):
    r = V.Renderer(init_device=False)
    scenes.load(r, scenes.cornell_box())
    t = r.add_texture(pixels, fmt, space)
    r.set_material(0, r.material(0))  # the pool is rebuilt on material changes
    return r, t


def slot_material(tex, slot=0, wrap=V.TEXTURE_WRAP_REPEAT | (V.TEXTURE_WRAP_REPEAT << 16)):
    m = V.default_material()
    names = ["baseColor", "metallicRoughness", "normal", "emissive"]
    setattr(m, names[slot] + "TextureIndex", tex)
    setattr(m, names[slot] + "TextureWrap", wrap)
    return m


def sample(r, m, uv, slot=0):
    r.set_material(0, m)
    b = r.export()
    uv = np.asarray(uv, np.float32)
    if uv.shape[1] == 2:
        uv = np.concatenate([uv, np.zeros_like(uv)], axis=1)
    return O.sample_material_slot(b, b.materials[0], slot, uv)


def test_oracle_sampler_texel_centres_and_affine_data():
    rng = np.random.default_rng(4)
    w, h = 7, 5
    px = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    r, t = sampler_scene(px)
    m = slot_material(t)
    jj, ii = np.mgrid[0:h, 0:w]
    uv = np.stack([(ii.ravel() + 0.5) / w, (jj.ravel() + 0.5) / h], 1)
    np.testing.assert_allclose(sample(r, m, uv), px.reshape(-1, 4), rtol=0, atol=1e-6)
    # bilinear filtering reproduces affine texel data between texel centres
    aff = np.zeros((h, w, 4), np.float32)
    aff[..., 0] = 0.25 * ii + 0.125 * jj
    aff[..., 1] = 2.0 - 0.25 * jj
    aff[..., 2] = 1.0
    aff[..., 3] = 0.0625 * ii * 1.0
    r2, t2 = sampler_scene(aff)
    m2 = slot_material(t2)
    x = rng.uniform(0.0, w - 1.0, 500)
    y = rng.uniform(0.0, h - 1.0, 500)
    got = sample(r2, m2, np.stack([(x + 0.5) / w, (y + 0.5) / h], 1))
    xs, ys = (np.float32((x + 0.5) / w) * w - 0.5), (np.float32((y + 0.5) / h) * h - 0.5)
    np.testing.assert_allclose(got[:, 0], 0.25 * xs + 0.125 * ys, atol=2e-6)
    np.testing.assert_allclose(got[:, 1], 2.0 - 0.25 * ys, atol=2e-6)
    np.testing.assert_allclose(got[:, 3], 0.0625 * xs, atol=2e-6)
    r.close()
    r2.close()


def test_oracle_sampler_wrap_modes():
    rng = np.random.default_rng(5)
    w, h = 6, 4
    px = rng.uniform(0, 1, (h, w, 4)).astype(np.float32)
    r, t = sampler_scene(px)
    u = rng.uniform(-3, 3, 400) / 1.0
    v = rng.uniform(-3, 3, 400)
    rep = V.TEXTURE_WRAP_REPEAT | (V.TEXTURE_WRAP_REPEAT << 16)
    mir = V.TEXTURE_WRAP_MIRRORED_REPEAT | (V.TEXTURE_WRAP_MIRRORED_REPEAT << 16)
    clamp = V.TEXTURE_WRAP_CLAMP_TO_EDGE | (V.TEXTURE_WRAP_CLAMP_TO_EDGE << 16)
    quarter = lambda a: np.round(a * 4) / 4  # coordinates exact in float after shifts by whole periods
    u, v = quarter(u) + 0.0625, quarter(v) + 0.125
    base = sample(r, slot_material(t, wrap=rep), np.stack([u, v], 1))
    np.testing.assert_allclose(sample(r, slot_material(t, wrap=rep), np.stack([u + 2, v - 1], 1)), base, atol=1e-6)
    mb = sample(r, slot_material(t, wrap=mir), np.stack([u, v], 1))
    np.testing.assert_allclose(sample(r, slot_material(t, wrap=mir), np.stack([-u, v], 1)), mb, atol=1e-6)
    np.testing.assert_allclose(sample(r, slot_material(t, wrap=mir), np.stack([2 - u, 4 + v], 1)), mb, atol=1e-6)
    # clamp to edge: everything left of the first texel centre is column 0
    inside = np.stack([np.full(50, 0.5 / w), rng.uniform(0.5 / h, 1 - 0.5 / h, 50)], 1)
    left = inside.copy()
    left[:, 0] = rng.uniform(-5, 0.5 / w, 50)
    np.testing.assert_allclose(sample(r, slot_material(t, wrap=clamp), left),
                               sample(r, slot_material(t, wrap=clamp), inside), atol=1e-6)
    # per-axis wrap: repeat in u, clamp in v
    mixed = V.TEXTURE_WRAP_REPEAT | (V.TEXTURE_WRAP_CLAMP_TO_EDGE << 16)
    a = sample(r, slot_material(t, wrap=mixed), np.stack([u, np.full_like(v, 9.0)], 1))
    b = sample(r, slot_material(t, wrap=mixed), np.stack([u + 1, np.full_like(v, 1 - 0.5 / h)], 1))
    np.testing.assert_allclose(a, b, atol=1e-6)
    r.close()


def test_oracle_sampler_transform_sets_and_fallbacks():
    rng = np.random.default_rng(6)
    px = rng.uniform(0, 1, (8, 8, 4)).astype(np.float32)
    r, t = sampler_scene(px)
    uv = rng.uniform(0, 1, (64, 2))
    plain = slot_material(t)
    m = slot_material(t)
    m.baseColorTextureTransform[:] = (2.0, 0.5, 0.25, -0.125)
    m.textureRotations[0] = math.pi / 2  # a quarter turn: (x, y) -> (-y, x)
    got = sample(r, m, uv)
    su, sv = uv[:, 0] * 2.0, uv[:, 1] * 0.5
    want = sample(r, plain, np.stack([-sv + 0.25, su - 0.125], 1))
    np.testing.assert_allclose(got, want, atol=2e-5)
    # coordinate set 1 is read when the slot's byte is 1
    m1 = slot_material(t, slot=3)
    m1.textureTexcoordSets = 1 << 24
    both = np.concatenate([np.zeros_like(uv), uv], 1)
    np.testing.assert_allclose(sample(r, m1, both, slot=3), sample(r, slot_material(t, slot=3), uv, slot=3), atol=0)
    # empty slots: 1 everywhere, the flat normal (0.5, 0.5, 1) for the normal slot
    empty = V.default_material()
    np.testing.assert_array_equal(sample(r, empty, uv, slot=0), np.ones((64, 4), np.float32))
    np.testing.assert_array_equal(sample(r, empty, uv, slot=2), np.tile([0.5, 0.5, 1.0, 1.0], (64, 1)))
    r.close()


def oracle_image(scene, w=24, h=24, spp=4):
    r = V.Renderer(init_device=False)
    scenes.load(r, scene)
    r.lib.VKRT_setRenderViewport(r.h, 0, 0, w, h)
    r.configure(spp=spp, max_depth=5, min_depth=3, mode=V.RENDER_MODE_RGB, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32,
                seed=3)
    img, _ = O.OracleScene(r.export(), use_bvh=False).render(w, h, precision=32)
    r.close()
    return img[..., :3]


def half_texture_scene(slot, alter):
    """The Cornell box with the floor, the back wall and the light reading a
    constant 0.5 texture in one slot (the walls carry no texture coordinates,
    so every lookup is (0, 0)); alter(material) applies the untextured
    equivalent."""
    tex = np.full((2, 2, 4), 0.5, np.float32)
    a, b = scenes.cornell_box(), scenes.cornell_box()
    a.textures = [(tex, V.TEXTURE_FORMAT_RGBA32_SFLOAT, V.TEXTURE_COLOR_SPACE_LINEAR)]
    for s in (a, b):
        s.materials[0] = V.default_material(baseColor=(0.7, 0.6, 0.5), roughness=0.6, metallic=0.5, specular=0.5)
    names = ["baseColor", "metallicRoughness", "normal", "emissive"]
    for i in (0, 3):
        setattr(a.materials[i], names[slot] + "TextureIndex", 0)
        alter(b.materials[i])
    return a, b


@pytest.mark.parametrize("slot", [0, 1, 3])
def test_constant_half_texture_equals_halved_material(slot):
    def alter(m):
        if slot == 0:
            m.baseColor[:] = [c * 0.5 for c in m.baseColor]
        elif slot == 1:
            m.roughness *= 0.5
            m.metallic *= 0.5
        else:
            # an emissive texture takes a mesh out of light sampling
            # (src/core/scene/lighting.c), so the equivalent reads a texture of ones
            m.emissionLuminance *= 0.5
            m.emissiveTextureIndex = 0

    a, b = half_texture_scene(slot, alter)
    if slot == 3:
        b.textures = [(np.ones((2, 2, 4), np.float32), V.TEXTURE_FORMAT_RGBA32_SFLOAT, V.TEXTURE_COLOR_SPACE_LINEAR)]
    ia, ib = oracle_image(a), oracle_image(b)
    assert ib.mean() > 0.01
    np.testing.assert_allclose(ia, ib, rtol=1e-6, atol=1e-7)


def test_flat_normal_map_changes_nothing():
    tex = np.tile(np.array([0.5, 0.5, 1.0, 1.0], np.float32), (2, 2, 1))
    a, b = scenes.cornell_box(), scenes.cornell_box()
    a.textures = [(tex, V.TEXTURE_FORMAT_RGBA32_SFLOAT, V.TEXTURE_COLOR_SPACE_LINEAR)]
    a.materials[0].normalTextureIndex = 0
    a.materials[0].normalTextureScale = 3.0  # scales x and y, which are 0
    ia, ib = oracle_image(a), oracle_image(b)
    close = np.mean(np.abs(ia - ib) <= 1e-5 * (1 + np.abs(ib)))
    assert O.relmse(ia, ib) < 1e-6 and close > 0.99


def _alpha_wall(alpha_mode, tex_alpha, cutoff=0.5):
    s = scenes.Scene()
    tex = np.ones((2, 2, 4), np.float32)
    tex[..., 3] = tex_alpha
    s.textures = [(tex, V.TEXTURE_FORMAT_RGBA32_SFLOAT, V.TEXTURE_COLOR_SPACE_LINEAR)]
    m = V.default_material()
    m.alphaMode, m.alphaCutoff, m.baseColorTextureIndex = alpha_mode, cutoff, 0
    s.materials = [m]
    pos, nrm, idx = scenes._quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, -1, 0))
    s.meshes.append(scenes.MeshData(V.make_vertices(pos, nrm), idx, 0))
    return s


def _hit_fraction(scene, n=40000):
    rng = np.random.default_rng(8)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2], rays[:, 3] = rng.uniform(0.05, 0.95, n), -1.0, rng.uniform(0.05, 0.95, n), 1e-3
    rays[:, 5], rays[:, 7] = 1.0, 1e4
    r = V.Renderer(init_device=False)
    scenes.load(r, scene)
    b = r.export()
    words = np.ctypeslib.as_array(b.triangleAlpha, (2,)).copy()
    mesh, _, _, _ = O.OracleScene(b, use_bvh=False).trace_rays(rays)
    r.close()
    assert np.all(words >> 30 == 3)  # the mesh runs the test and reads the texture's alpha
    return float(np.mean(mesh != V.VKRT_INVALID_INDEX))


def test_texture_alpha_in_the_alpha_test():
    n = 40000
    # MASK: the cutoff compares texture alpha x vertex alpha; surviving hits use the vertex alpha (1)
    assert _hit_fraction(_alpha_wall(1, 0.3)) == 0.0
    assert _hit_fraction(_alpha_wall(1, 0.7)) == 1.0
    # BLEND: texture alpha is the opacity
    f = _hit_fraction(_alpha_wall(2, 0.25))
    assert abs(f - 0.25) <= 4.0 * math.sqrt(0.25 * 0.75 / n), f


def make_gpu(w, h, spp, mode=V.RENDER_MODE_RGB):
    r = V.Renderer(w, h)
    scenes.load(r, scenes.textured_cornell())
    r.configure(spp=spp, max_depth=8, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=5)
    r.start(w, h, 0)
    return r


@pytest.mark.gpu
def test_textured_primary_hits_match_the_oracle():
    w, h = 256, 256
    r = make_gpu(w, h, 1)
    gm, gp = r.primary_hits(0)
    om, op = O.OracleScene(r.export(), use_bvh=True).primary_hits(w, h, 0)
    agree = np.mean((gm == om) & (gp == op))
    print(f"primary-hit agreement {agree:.6f}; {np.isin(om, [7, 8]).mean():.3f} of pixels on texture-alpha quads")
    assert np.isin(om, [7, 8]).mean() > 0.05
    assert agree >= 0.9999
    r.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [V.RENDER_MODE_RGB, V.RENDER_MODE_SPECTRAL])
def test_textured_same_seed_parity(mode):
    w, h = 128, 128
    r = make_gpu(w, h, 16, mode)
    orc = O.OracleScene(r.export(), use_bvh=False)
    r.draw()
    gpu = r.accumulation()
    st = r.stats()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-4 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-4: {close:.5f}, bounces gpu {st.extensionRays} oracle {bounces}")
    assert st.extensionRays == bounces
    assert err < 1e-9  # measured 2e-11 .. 5e-11
    assert close > 0.999
    r.close()


@pytest.mark.gpu
def test_textured_scene_with_the_neural_cache_and_backward_refusal():
    w, h = 64, 64
    r = make_gpu(w, h, 4)
    r.set_neural_cache(True, trainingPaths=4096, train=1)
    r.draw()
    img = r.accumulation()
    assert np.isfinite(img).all() and img[..., :3].mean() > 0.0
    with pytest.raises(V.VkrtError):
        r.render_backward(np.ones((h, w, 3), np.float32))
    r.close()
