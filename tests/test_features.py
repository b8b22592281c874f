"""Denoiser feature images and debug views in the oracle (CPU).

The oracle restates src/shaders/integrator/path/loop.slang:83-127,
debug.slang, utility/debug.slang, bsdf/principled/denoiser.slang and
writeback.slang:9-123; each view is pinned here by a closed form on a scene
whose answer is known: Lambert walls (albedo = base colour, normal = wall
normal, feature depth 1, nothing to follow), a smooth mirror (features
resolve one bounce later, on what the mirror shows), a single quad under an
empty sky (exactly one bounce per path), and the empty selection mask.
"""

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

SPP = 4


def render(scene, w, h, debug=V.VKRT_DEBUG_MODE_NONE, depth=8, spp=SPP,
           mode=V.RENDER_MODE_RGB, frames=1):
    r = V.Renderer(init_device=False)
    scenes.load(r, scene)
    r.configure(spp=spp, max_depth=depth, min_depth=4, mode=mode, rng=V.RNG_PCG32, seed=5)
    r.set_debug_mode(debug)
    r.start(w, h, 0)
    osc = O.OracleScene(r.export(), use_bvh=False)
    acc = np.zeros((h, w, 4), np.float32)
    alb = np.zeros_like(acc)
    nrm = np.zeros_like(acc)
    for _ in range(frames):
        acc, out, _ = osc.render(w, h, accum=acc, features=(alb, nrm), output=True)
    osc.close()
    r.close()
    return acc, alb, nrm, out


def back_wall_pixels(w, h):
    """Pixels of the Cornell view that see only the back wall (centre band above the boxes)."""
    ys, xs = np.mgrid[0:h, 0:w]
    return (np.abs(xs + 0.5 - w / 2) < w * 0.15) & (ys >= h * 0.26) & (ys <= h * 0.40)


def test_lambert_walls_give_base_colour_normal_and_full_weight():
    w = h = 48
    acc, alb, nrm, _ = render(scenes.cornell_box(), w, h, frames=2)
    assert np.all(alb[..., 3] == 2 * SPP) and np.all(nrm[..., 3] == 2 * SPP)  # every path resolves at depth 0
    wall = back_wall_pixels(w, h)
    assert wall.sum() > 20
    np.testing.assert_allclose(alb[wall][:, :3], 0.73, rtol=0, atol=1e-6)  # denoiser.slang:24-46 for Lambert
    np.testing.assert_allclose(nrm[wall][:, :3], np.tile([0.0, -1.0, 0.0], (wall.sum(), 1)), atol=1e-6)
    assert np.all(acc[..., 3] == 2 * SPP)


@pytest.mark.parametrize("mode,expect", [
    (V.VKRT_DEBUG_MODE_DENOISER_FEATURE_VALIDITY, 1.0),  # every sample resolved
    (V.VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH, 0.0),     # resolved at the first vertex
    (V.VKRT_DEBUG_MODE_SELECTION_MASK, 0.05),            # no mesh selected
])
def test_uniform_debug_views(mode, expect):
    w = h = 24
    acc, alb, nrm, out = render(scenes.cornell_box(), w, h, debug=mode)
    np.testing.assert_allclose(acc[..., :3], expect, atol=1e-6)
    assert np.all(acc[..., 3] == 0.0)  # writeback.slang:59-66: a debug frame leaves no sample count
    assert np.all(alb == 0.0) and np.all(nrm == 0.0)


def test_follow_specular_view_is_the_floor_colour_without_mirrors():
    acc = render(scenes.cornell_box(), 16, 16, debug=V.VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR)[0]
    np.testing.assert_allclose(acc[..., :3], 0.05, atol=1e-6)


def test_normal_and_albedo_views_match_the_feature_images():
    w = h = 32
    _, alb, nrm, _ = render(scenes.cornell_box(), w, h)
    va = render(scenes.cornell_box(), w, h, debug=V.VKRT_DEBUG_MODE_DENOISER_ALBEDO)[0]
    vn = render(scenes.cornell_box(), w, h, debug=V.VKRT_DEBUG_MODE_DENOISER_NORMAL)[0]
    np.testing.assert_allclose(va[..., :3], alb[..., :3], atol=1e-6)
    np.testing.assert_allclose(vn[..., :3], nrm[..., :3] * 0.5 + 0.5, atol=1e-6)
    # the primary-surface normal view (debug.slang:45-49) on the back wall
    v1 = render(scenes.cornell_box(), w, h, debug=V.VKRT_DEBUG_MODE_NORMALS)[0]
    wall = back_wall_pixels(w, h)
    np.testing.assert_allclose(v1[wall][:, :3], np.tile([0.5, 0.0, 0.5], (wall.sum(), 1)), atol=1e-6)


def mirror_in_front_of_wall():
    """A smooth mirror quad facing the camera, tilted 45 degrees so it shows a
    red Lambert wall above it; a green wall behind, nothing else."""
    s = scenes.Scene()
    mirror = V.default_material(baseColor=(0.9, 0.9, 0.9), roughness=0.0, metallic=1.0)
    s.materials = [mirror, scenes.lambert((0.8, 0.1, 0.1)), scenes.lambert((0.1, 0.7, 0.1))]
    c = 0.7071
    quad = scenes._quad((-1, -1, -1), (1, -1, -1), (1, 1, 1), (-1, 1, 1), (0, -c, c))
    s.meshes.append(scenes.MeshData(*scenes._merge([quad]), 0))
    top = scenes._quad((-5, -5, 3), (-5, 5, 3), (5, 5, 3), (5, -5, 3), (0, 0, -1))
    s.meshes.append(scenes.MeshData(*scenes._merge([top]), 1))
    s.camera = ((0.0, -6.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 8.0)
    return s


def test_mirror_features_resolve_on_what_it_reflects():
    w = h = 16
    _, alb, nrm, _ = render(mirror_in_front_of_wall(), w, h, depth=6)
    centre = alb[6:10, 6:10]
    assert np.all(centre[..., 3] == SPP)
    np.testing.assert_allclose(centre[..., :3], np.tile([0.8, 0.1, 0.1], (4, 4, 1)), atol=1e-6)
    np.testing.assert_allclose(nrm[6:10, 6:10, :3], np.tile([0.0, 0.0, -1.0], (4, 4, 1)), atol=1e-5)
    depth = render(mirror_in_front_of_wall(), w, h, depth=6, debug=V.VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH)[0]
    np.testing.assert_allclose(depth[6:10, 6:10, :3], (2 - 1) / (6 - 1), atol=1e-6)
    follow = render(mirror_in_front_of_wall(), w, h, depth=6, debug=V.VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR)[0]
    np.testing.assert_allclose(follow[6:10, 6:10, :3], np.tile([1.0, 0.6, 0.0], (4, 4, 1)), atol=1e-6)


def single_quad():
    s = scenes.Scene()
    s.materials = [scenes.lambert((0.5, 0.5, 0.5))]
    s.meshes.append(scenes.MeshData(*scenes._merge([scenes._quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1),
                                                                 (0, -1, 0))]), 0))
    s.camera = ((0.0, -4.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 60.0)
    return s


def test_bounce_count_view_and_depth_view():
    w = h = 24
    depth = 8
    acc = render(single_quad(), w, h, depth=depth, debug=V.VKRT_DEBUG_MODE_BOUNCE_COUNT)[0]
    # loop.slang:113-127: one surface vertex on the quad, then the sky: t = 1 / 8; a miss: t = 0
    c = acc[12, 12, :3]
    np.testing.assert_allclose(c, [1 / depth, 1 - 1 / depth, 0.0], atol=1e-6)
    np.testing.assert_allclose(acc[0, 0, :3], [0.0, 1.0, 0.0], atol=1e-6)
    # depth view (debug.slang:51-55): 1 / (1 + t) for the quad 4 units away; misses render normally
    dv = render(single_quad(), w, h, depth=depth, debug=V.VKRT_DEBUG_MODE_DEPTH)[0]
    t = 1.0 / dv[11:13, 11:13, 0] - 1.0
    assert np.all((t > 3.99) & (t < 4.02))
    assert dv[0, 0, 3] == SPP and np.all(dv[0, 0, :3] == 0.0)  # a sky pixel keeps the (black) render


def test_texture_map_views_without_textures_show_the_fallback():
    w = h = 12
    for mode, want in ((V.VKRT_DEBUG_MODE_BASE_COLOR_MAP, 1.0), (V.VKRT_DEBUG_MODE_METALLIC_MAP, 1.0),
                       (V.VKRT_DEBUG_MODE_ROUGHNESS_MAP, 1.0), (V.VKRT_DEBUG_MODE_EMISSIVE_MAP, 1.0)):
        acc = render(scenes.cornell_box(), w, h, debug=mode)[0]
        np.testing.assert_allclose(acc[..., :3], want, atol=0)
    acc = render(scenes.cornell_box(), w, h, debug=V.VKRT_DEBUG_MODE_NORMAL_MAP)[0]
    np.testing.assert_allclose(acc[..., :3], np.tile([0.5, 0.5, 1.0], (h, w, 1)), atol=0)
