"""The auto-SPP controller (src/core/scene/timing.c:121-178, run from
VKRT_endFrame as src/core/api/frame.c:380-381 does) and the offline
benchmark's locked sampling (src/app/render/benchmark.c:55-67).

A small Cornell box traces in far less than the 33 ms budget of 30 FPS, so
the controller raises the sample count from 1: never by more than the x1.25
step limit (or one sample), never down, and every frame's samples are
counted. Turning it off locks the count; a render start with it on restarts
at one sample per pixel."""

import math

import pytest

import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu


def test_controller_raises_the_sample_count_within_its_step_limit():
    r = V.Renderer(64, 64)
    scenes.load(r, scenes.cornell_box())
    r.configure(max_depth=4, min_depth=2, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=1)
    r.auto_spp(True, target_fps=30)
    r.start(64, 64, 0)
    assert r.scene_settings().samplesPerPixel == 1
    seen, total = [], 0
    for _ in range(12):
        spp = r.scene_settings().samplesPerPixel
        r.draw()
        total += spp
        seen.append(spp)
        assert r.status().totalSamples == total
    after = r.scene_settings().samplesPerPixel
    print("samples per frame:", seen, "->", after)
    steps = seen + [after]
    for a, b in zip(steps, steps[1:]):
        assert b >= a
        assert b <= max(a + 1, math.ceil(a * 1.25))
    assert after > 4
    # locked: the count no longer moves
    r.auto_spp(False)
    for _ in range(3):
        r.draw()
    assert r.scene_settings().samplesPerPixel == after
    # a new render with the controller on starts again from one sample
    r.auto_spp(True)
    r.start(64, 64, 0)
    assert r.scene_settings().samplesPerPixel == 1
    r.close()
