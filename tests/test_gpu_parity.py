"""GPU vs oracle parity through the C ABI (include/vkrt.h).

Same-seed tests render with identical random streams on the GPU and in the
FP32 oracle; paths then follow the same vertices except where a float
difference flips a branch, so images agree far below Monte-Carlo noise.
"""

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu


def make(scene, w, h, spp, mode=V.RENDER_MODE_RGB, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=1, depth=8):
    r = V.Renderer(w, h)
    scenes.load(r, scene)
    r.configure(spp=spp, max_depth=depth, min_depth=4, mode=mode, spectral=spectral, rng=rng, seed=seed)
    r.start(w, h, 0)
    return r


def small_sphere_scene(freq=40, env=(256, 128)):
    s = scenes.displaced_sphere(freq)
    s.environment = scenes.sky_environment(env[0], env[1], sun_radiance=50.0, sun_diameter_deg=6.0)
    return s


@pytest.mark.parametrize("rng", [V.RNG_HASH, V.RNG_PCG32])
def test_primary_hit_ids_cornell(rng):
    w, h = 256, 256
    r = make(scenes.cornell_box(), w, h, 1, rng=rng)
    gm, gp = r.primary_hits(0)
    om, op = O.OracleScene(r.export(), use_bvh=False).primary_hits(w, h, 0)
    agree = np.mean((gm == om) & (gp == op))
    print(f"primary-hit agreement {agree:.6f}")
    assert agree >= 0.9999
    r.close()


def test_primary_hit_ids_displaced_sphere():
    w, h = 320, 180
    r = make(small_sphere_scene(freq=80), w, h, 1, mode=V.RENDER_MODE_SPECTRAL)
    gm, gp = r.primary_hits(0)
    om, op = O.OracleScene(r.export(), use_bvh=True).primary_hits(w, h, 0)
    hit = om != V.VKRT_INVALID_INDEX
    agree = np.mean((gm == om) & (gp == op))
    print(f"primary-hit agreement {agree:.6f} over {hit.sum()} hits")
    assert hit.sum() > 0.2 * w * h
    assert agree >= 0.9999
    r.close()


def test_primary_hit_ids_instances_and_tiny_triangles():
    """A scaled-down stress scene: transformed instances of a displaced sphere
    and 200,000 random triangles with edges of 1e-3..1e-2 (the boxes that
    8-bit plane quantisation must still bound conservatively)."""
    w, h = 320, 180
    s = scenes.stress_scene(instances=4, freq=40, random_count=200_000)
    s.camera = ((0.5, -1.2, 0.6), (0.5, 0.5, 0.45), (0.0, 0.0, 1.0), 60.0)  # into the cube of small triangles
    r = V.Renderer(w, h)
    scenes.load(r, s)
    r.configure(spp=1, max_depth=4, min_depth=4, rng=V.RNG_PCG32, seed=3)
    r.start(w, h, 0)
    gm, gp = r.primary_hits(0)
    om, op = O.OracleScene(r.export(), use_bvh=True).primary_hits(w, h, 0)
    hit = om != V.VKRT_INVALID_INDEX
    agree = np.mean((gm == om) & (gp == op))
    print(f"primary-hit agreement {agree:.6f} over {hit.sum()} hits, {s.triangle_count} triangles")
    assert hit.sum() > 0.1 * w * h
    assert agree >= 0.9999
    r.close()


@pytest.mark.parametrize(
    "mode,spectral,rng",
    [
        (V.RENDER_MODE_RGB, V.SPECTRAL_HERO, V.RNG_PCG32),
        (V.RENDER_MODE_RGB, V.SPECTRAL_HERO, V.RNG_HASH),
        (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_HERO, V.RNG_PCG32),
        (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_SINGLE, V.RNG_PCG32),
    ],
)
def test_cornell_same_seed_parity(mode, spectral, rng):
    # the RGB / PCG32 case is configuration C1 itself: 256 x 256, 16 spp, depth 8, seed 1
    w = h = 256 if (mode, rng) == (V.RENDER_MODE_RGB, V.RNG_PCG32) else 128
    r = make(scenes.cornell_box(), w, h, 16, mode=mode, spectral=spectral, rng=rng)
    orc = O.OracleScene(r.export(), use_bvh=False)
    r.draw()
    gpu = r.accumulation()
    st = r.stats()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-4 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-4: {close:.5f}, bounces gpu {st.extensionRays} oracle {bounces}")
    assert np.all(gpu[..., 3] == 16)
    assert abs(st.extensionRays - bounces) <= 1e-5 * bounces  # a float branch flip may move one path
    assert err < 1e-12  # measured 1e-14 .. 8e-14
    assert close > 0.9999
    r.close()


@pytest.mark.parametrize(
    "mode,spectral",
    [(V.RENDER_MODE_RGB, V.SPECTRAL_HERO), (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_SINGLE),
     (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_HERO)],
)
def test_sheen_same_seed_parity(mode, spectral):
    """Microfiber sheen (sheen.slang) on both boxes: the sheen kernel variants
    against the oracle's restatement, same random streams."""
    w, h = 128, 128
    r = make(scenes.sheen_cornell(), w, h, 16, mode=mode, spectral=spectral, rng=V.RNG_PCG32)
    orc = O.OracleScene(r.export(), use_bvh=False)
    r.draw()
    gpu = r.accumulation()
    st = r.stats()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-4 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-4: {close:.5f}, bounces gpu {st.extensionRays} oracle {bounces}")
    assert np.all(gpu[..., 3] == 16)
    assert st.extensionRays == bounces
    assert err < 1e-12  # measured 1e-14 .. 9e-14
    assert close > 0.9999
    r.close()


def test_sheen_changes_the_image():
    """The sheen layer is live: removing it from the boxes changes the render
    well beyond noise at the same seed."""
    w, h = 64, 64
    imgs = []
    for sheen in (True, False):
        s = scenes.sheen_cornell()
        if not sheen:
            for m in s.materials:
                m.sheenTintWeight[3] = 0.0
        r = make(s, w, h, 16)
        r.draw()
        imgs.append(r.accumulation()[..., :3])
        r.close()
    assert O.relmse(imgs[0], imgs[1]) > 1e-3


def test_multi_frame_accumulation_parity():
    w, h = 64, 64
    r = make(scenes.cornell_box(), w, h, 4)
    acc = None
    orc = None
    for frame in range(3):
        orc = O.OracleScene(r.export(), use_bvh=False)
        r.draw()
        acc, _ = orc.render(w, h, accum=acc, precision=32)
    gpu = r.accumulation()
    assert np.all(gpu[..., 3] == 12) and np.all(acc[..., 3] == 12)
    err = O.relmse(gpu[..., :3], acc[..., :3])
    print(f"3-frame relMSE {err:.3e}")
    assert err < 1e-12  # measured 7e-15
    r.close()


@pytest.mark.parametrize("mode", [V.RENDER_MODE_RGB, V.RENDER_MODE_SPECTRAL])
def test_sphere_dielectric_ggx_environment_parity(mode):
    """64 GGX / dielectric materials under an HDR sky: exercises the rough
    dielectric, GGX, env lookup and (spectral) rgb2spec paths."""
    w, h = 96, 54
    r = make(small_sphere_scene(), w, h, 8, mode=mode, depth=12)
    orc = O.OracleScene(r.export(), use_bvh=True)
    r.draw()
    gpu = r.accumulation()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-3 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-3: {close:.5f}, bounces gpu {r.stats().extensionRays} oracle {bounces}")
    assert np.isfinite(gpu).all()
    assert close > 0.999
    assert err < 1e-10  # measured 9e-13 .. 1.1e-12
    r.close()


@pytest.mark.parametrize("mode,spectral", [(V.RENDER_MODE_RGB, V.SPECTRAL_HERO),
                                           (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_HERO),
                                           (V.RENDER_MODE_SPECTRAL, V.SPECTRAL_SINGLE)])
def test_environment_nee_same_seed_parity(mode, spectral):
    """Environment NEE (alias-table sampler, MIS with BSDF misses, its own
    shadow pass) against the oracle's restatement with the same streams."""
    w, h = 96, 54
    r = make(small_sphere_scene(), w, h, 8, mode=mode, spectral=spectral, depth=12)
    r.configure(env_importance=True)
    orc = O.OracleScene(r.export(), use_bvh=True)
    r.draw()
    gpu = r.accumulation()
    st = r.stats()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-3 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-3: {close:.5f}, shadow rays {st.shadowRays}, "
          f"bounces gpu {st.extensionRays} oracle {bounces}")
    assert st.shadowRays > 0
    assert np.isfinite(gpu).all()
    assert close > 0.999
    assert err < 2e-10  # measured 9e-13 .. 1.8e-11
    r.close()


def test_environment_nee_converges_to_the_reference_estimator():
    """With and without environment NEE the GPU converges to the same image:
    the per-frame image means of the two estimators agree within their
    standard errors."""
    w, h = 64, 36
    frames = 12
    means = {}
    for nee in (False, True):
        r = make(small_sphere_scene(), w, h, 256, depth=8, seed=11 if nee else 5)
        r.configure(env_importance=nee)
        m = []
        prev_sum = 0.0
        for k in range(1, frames + 1):
            r.draw()  # each frame draws new streams (the frame number is part of the stream id)
            acc = r.accumulation().astype(np.float64)
            assert np.all(acc[..., 3] == 256 * k)
            total = acc[..., :3].mean() * k  # running mean x frames = sum of per-frame means
            m.append(total - prev_sum)
            prev_sum = total
        r.close()
        means[nee] = np.array(m)
    se = np.hypot(means[False].std(ddof=1), means[True].std(ddof=1)) / np.sqrt(frames)
    diff = means[True].mean() - means[False].mean()
    print(f"image mean without NEE {means[False].mean():.6f}, with {means[True].mean():.6f}, diff {diff:.2e}, se {se:.2e}")
    assert abs(diff) < 4 * se + 1e-4 * means[False].mean()


def test_material_sorted_shading_gives_identical_images():
    """Shading the hits of every depth grouped by material class changes only the
    order of the work: every path keeps its stream, so the image is bit-identical."""
    s = small_sphere_scene()
    imgs = []
    for sort in (False, True):
        r = make(s, 64, 48, 16, mode=V.RENDER_MODE_SPECTRAL, seed=4)
        r.set_material_sorting(sort)
        r.start(64, 48, 0)
        r.draw()
        imgs.append(r.accumulation())
        r.close()
    assert np.array_equal(imgs[0], imgs[1])


@pytest.fixture(scope="module")
def spectral_workload_scene():
    """The 1080p spectral workload's scene: 1,048,820 displaced triangles, 64
    materials, 2048x1024 sky with the 1e4 sun."""
    s = scenes.displaced_sphere(229)
    s.environment = scenes.sky_environment(2048, 1024)
    return s


def test_primary_hit_ids_spectral_workload_scene(spectral_workload_scene):
    w, h = 480, 270  # the workload's 16:9 view at a quarter of its resolution
    r = make(spectral_workload_scene, w, h, 1, mode=V.RENDER_MODE_SPECTRAL, depth=12)
    gm, gp = r.primary_hits(0)
    om, op = O.OracleScene(r.export(), use_bvh=True).primary_hits(w, h, 0)
    hit = om != V.VKRT_INVALID_INDEX
    agree = np.mean((gm == om) & (gp == op))
    print(f"primary-hit agreement {agree:.6f} over {hit.sum()} hits of {r.stats().triangleCount} triangles")
    assert r.stats().triangleCount == 1048820
    assert hit.sum() > 0.2 * w * h
    assert agree >= 0.9999
    r.close()


def test_spectral_workload_same_seed_parity(spectral_workload_scene):
    """Hero-wavelength spectral mode, depth 12, on the full 1M-triangle scene:
    GPU and FP32 oracle (its own binned-SAH BVH) on the same streams."""
    w, h = 192, 108
    r = make(spectral_workload_scene, w, h, 4, mode=V.RENDER_MODE_SPECTRAL, depth=12)
    orc = O.OracleScene(r.export(), use_bvh=True)
    r.draw()
    gpu = r.accumulation()
    ref, bounces = orc.render(w, h, precision=32)
    err = O.relmse(gpu[..., :3], ref[..., :3])
    close = np.mean(np.abs(gpu[..., :3] - ref[..., :3]) <= 1e-3 * (1.0 + np.abs(ref[..., :3])))
    print(f"relMSE {err:.3e}, pixels within 1e-3: {close:.5f}, bounces gpu {r.stats().extensionRays} oracle {bounces}")
    assert np.isfinite(gpu).all()
    assert close > 0.999
    assert err < 1e-5  # measured 5.5e-7
    r.close()
