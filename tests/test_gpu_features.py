"""Denoiser feature images and debug views on the GPU against the oracle
(same random streams): the feature pass, its per-pixel reduction and the
accumulation kernel that applies the views (src/shaders/integrator/path/
loop.slang:83-127, debug.slang, writeback.slang:9-123)."""

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu

MIRROR = None


def mirror_scene():
    from test_features import mirror_in_front_of_wall
    return mirror_in_front_of_wall()


def run(scene, w, h, mode=V.RENDER_MODE_RGB, debug=V.VKRT_DEBUG_MODE_NONE, features=True, spp=4, depth=8, frames=2):
    r = V.Renderer(w, h)
    scenes.load(r, scene)
    r.configure(spp=spp, max_depth=depth, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=9)
    r.set_debug_mode(debug)
    r.set_denoiser_features(features)
    r.start(w, h, 0)
    acc = np.zeros((h, w, 4), np.float32)
    alb, nrm = np.zeros_like(acc), np.zeros_like(acc)
    for _ in range(frames):
        osc = O.OracleScene(r.export(), use_bvh=False)
        r.draw()
        acc, out, _ = osc.render(w, h, accum=acc, features=(alb, nrm), output=True)
        osc.close()
    gpu = dict(acc=r.accumulation(), out=r.output_image())
    if features or debug != V.VKRT_DEBUG_MODE_NONE:
        gpu["alb"], gpu["nrm"] = r.denoiser_features()
    r.close()
    return gpu, dict(acc=acc, out=out, alb=alb, nrm=nrm)


def compare(gpu, ref, what, tol=1e-5):
    for k in ("acc", "alb", "nrm"):
        if k in gpu:
            d = np.abs(gpu[k] - ref[k]).max()
            close = np.mean(np.abs(gpu[k] - ref[k]) <= tol * (1.0 + np.abs(ref[k])))
            print(f"{what} {k}: max |diff| {d:.3e}, within tol {close:.5f}", flush=True)
            assert close > 0.999, (what, k)
    assert np.mean(np.abs(gpu["out"].astype(np.int64) - ref["out"].astype(np.int64)) <= 1) > 0.999, what


@pytest.mark.parametrize("mode", [V.RENDER_MODE_RGB, V.RENDER_MODE_SPECTRAL])
def test_feature_images_match_the_oracle(mode):
    gpu, ref = run(scenes.sheen_cornell(), 64, 64, mode=mode)
    assert np.all(gpu["alb"][..., 3] == 8) and np.all(ref["alb"][..., 3] == 8)
    compare(gpu, ref, f"features mode {mode}")


def test_mirror_features_resolve_one_bounce_later():
    gpu, ref = run(mirror_scene(), 32, 32, depth=6)
    compare(gpu, ref, "mirror features")
    np.testing.assert_allclose(gpu["alb"][13:19, 13:19, :3], np.tile([0.8, 0.1, 0.1], (6, 6, 1)), atol=1e-6)


@pytest.mark.parametrize("debug", [V.VKRT_DEBUG_MODE_NORMALS, V.VKRT_DEBUG_MODE_DEPTH, V.VKRT_DEBUG_MODE_BOUNCE_COUNT,
                                   V.VKRT_DEBUG_MODE_SELECTION_MASK, V.VKRT_DEBUG_MODE_DENOISER_ALBEDO,
                                   V.VKRT_DEBUG_MODE_DENOISER_NORMAL, V.VKRT_DEBUG_MODE_DENOISER_FEATURE_VALIDITY,
                                   V.VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH,
                                   V.VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR])
def test_debug_views_match_the_oracle(debug):
    scene = mirror_scene() if debug in (V.VKRT_DEBUG_MODE_DENOISER_FEATURE_DEPTH,
                                        V.VKRT_DEBUG_MODE_DENOISER_FOLLOW_SPECULAR) else scenes.cornell_box()
    gpu, ref = run(scene, 48, 48, debug=debug, features=False, depth=6)
    compare(gpu, ref, f"debug {debug}")


@pytest.mark.parametrize("debug", [V.VKRT_DEBUG_MODE_BASE_COLOR_MAP, V.VKRT_DEBUG_MODE_METALLIC_MAP,
                                   V.VKRT_DEBUG_MODE_ROUGHNESS_MAP, V.VKRT_DEBUG_MODE_NORMAL_MAP,
                                   V.VKRT_DEBUG_MODE_EMISSIVE_MAP])
def test_texture_map_views_match_the_oracle(debug):
    gpu, ref = run(scenes.textured_cornell(), 48, 48, debug=debug, features=False)
    compare(gpu, ref, f"texture view {debug}")
    assert np.std(gpu["acc"][..., :3]) > 0.0


def test_features_off_leave_the_plain_render_unchanged():
    a, _ = run(scenes.cornell_box(), 32, 32, features=False, frames=1)
    b, _ = run(scenes.cornell_box(), 32, 32, features=True, frames=1)
    np.testing.assert_array_equal(a["acc"], b["acc"])
