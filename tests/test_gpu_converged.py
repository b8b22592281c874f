"""Converged GPU images against converged FP64 oracle images.

The fixtures (tests/golden/make_converged_fixtures.py) are the Cornell box
rendered by the FP64 oracle at 32x32 with 65,536 samples per pixel. The GPU
renders the same scene with an independent seed at the same sample count; the
two estimates must agree within relMSE 1e-4 (their combined Monte Carlo error
is about 1.5e-5).
"""

import os

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name,mode", [("cornell_rgb_fp64.npy", V.RENDER_MODE_RGB),
                                       ("cornell_hero_fp64.npy", V.RENDER_MODE_SPECTRAL)])
def test_converged_cornell_matches_fp64_oracle(name, mode):
    ref = np.load(os.path.join(GOLDEN, name))
    size = ref.shape[0]
    r = V.Renderer(size, size)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=65536, max_depth=8, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32, seed=7)
    r.start(size, size, 0)
    r.draw()
    gpu = r.accumulation()
    assert np.all(gpu[..., 3] == 65536)
    err = O.relmse(gpu[..., :3], ref)
    print(f"{name}: relMSE(GPU 65536 spp, FP64 oracle 65536 spp) = {err:.3e}")
    assert err < 1e-4
    r.close()


def test_converged_spectral_workload_scene_matches_fp64_oracle():
    """The 1,048,820-triangle spectral scene (sky without the sun disc), 32x18
    pixels of its view, hero spectral, depth 12, 65,536 spp on the GPU with an
    independent seed against the FP64 oracle's image (tests/golden/
    make_converged_fixtures.py)."""
    ref = np.load(os.path.join(GOLDEN, "sphere_sky_hero_fp64.npy"))
    h, w = ref.shape[:2]
    s = scenes.displaced_sphere(229)
    s.environment = scenes.sky_environment(2048, 1024, sun_radiance=0.0)
    r = V.Renderer(w, h)
    scenes.load(r, s)
    r.configure(spp=65536, max_depth=12, min_depth=4, mode=V.RENDER_MODE_SPECTRAL, spectral=V.SPECTRAL_HERO,
                rng=V.RNG_PCG32, seed=7)
    r.start(w, h, 0)
    r.draw()
    gpu = r.accumulation()
    assert np.all(gpu[..., 3] == 65536)
    err = O.relmse(gpu[..., :3], ref)
    print(f"sphere_sky_hero_fp64.npy: relMSE(GPU 65536 spp, FP64 oracle 65536 spp) = {err:.3e}")
    assert err < 1e-4
    r.close()


def test_converged_spectral_workload_with_the_sun_is_as_close_as_the_oracle_to_itself():
    """The same scene under the full sky with its 0.5-degree, 1e4 sun and
    environment importance sampling, 64x36 pixels, 65,536 spp. Sun light
    carried through the near-specular dielectrics makes the estimator
    heavy-tailed (two FP64 oracle runs with different seeds differ by relMSE
    ~5e-3, and the gap shrinks far slower than 1/N), so the GPU image (its own
    seed) is held to the pair of FP64 images: no farther from either than
    they are from each other, with a factor 3 for the spread of such
    heavy-tailed distances, and its image mean equally close."""
    a = np.load(os.path.join(GOLDEN, "sphere_sun_hero_fp64_a.npy"))
    b = np.load(os.path.join(GOLDEN, "sphere_sun_hero_fp64_b.npy"))
    h, w = a.shape[:2]
    s = scenes.displaced_sphere(229)
    s.environment = scenes.sky_environment(2048, 1024)
    r = V.Renderer(w, h)
    scenes.load(r, s)
    r.configure(spp=65536, max_depth=12, min_depth=4, mode=V.RENDER_MODE_SPECTRAL, spectral=V.SPECTRAL_HERO,
                rng=V.RNG_PCG32, seed=7, env_importance=True)
    r.start(w, h, 0)
    r.draw()
    gpu = r.accumulation()[..., :3]
    r.close()
    d_ab, d_ba = O.relmse(a, b), O.relmse(b, a)
    d_ga, d_gb = O.relmse(gpu, a), O.relmse(gpu, b)
    ma, mb, mg = a.mean(), b.mean(), gpu.mean()
    print(f"relMSE GPU-vs-A {d_ga:.3e}, GPU-vs-B {d_gb:.3e}, A-vs-B {d_ba:.3e}, B-vs-A {d_ab:.3e}; "
          f"image means GPU {mg:.6f} A {ma:.6f} B {mb:.6f}", flush=True)
    assert d_ga <= 3.0 * d_ba and d_gb <= 3.0 * d_ab
    assert abs(mg - 0.5 * (ma + mb)) <= 3.0 * abs(ma - mb) + 1e-4 * ma
