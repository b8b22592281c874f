"""Path-replay backpropagation (VKRT_renderBackward) against finite differences.

The loss is L = sum_px w[px] . I[px] with fixed random weights w, so
dL/dI = w. With Lambertian materials, one light mesh and Russian roulette out
of reach (min depth > max depth) every path's vertices and random decisions
are independent of base colour, emission and environment texels, so the
renderer's estimate is a smooth function of those parameters for fixed random
numbers: central differences of frames rendered with the same streams (common
random numbers) converge to the adjoint's gradient as the step shrinks. The
same differences of the FP64 oracle (identical streams and paths) give the
CPU reference gradient.
"""

import numpy as np
import pytest

import oracle as O
import vkrt_amd as V
from vkrt_amd import scenes

pytestmark = pytest.mark.gpu

SIZE = 32
SPP = 64


def cornell():
    r = V.Renderer(SIZE, SIZE)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=SPP, max_depth=4, min_depth=8, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=3)
    r.start(SIZE, SIZE, 0)
    return r


def frame(r):
    r.start(SIZE, SIZE, 0)  # fresh accumulation, frame 0: the same streams every time
    r.draw()
    return r.accumulation()[..., :3].astype(np.float64)


def weights(seed=5):
    return np.random.default_rng(seed).normal(size=(SIZE, SIZE, 3)).astype(np.float32)


# (material index, field, channel, step); the Cornell materials are
# 1 white, 2 red, 3 green, 4 light (0 is the default material)
PARAMS = [(1, "baseColor", 0, 0.01), (2, "baseColor", 0, 0.01), (3, "baseColor", 1, 0.01),
          (4, "emissionColor", 0, 0.01)]


def adjoint_value(grads, r, mat, field, ch):
    if field == "baseColor":
        return float(grads["baseColor"][mat, ch])
    # emission gradient is with respect to emissionColor * emissionLuminance
    return float(grads["emission"][mat, ch]) * r.material(mat).emissionLuminance


def perturbed(r, mat, field, ch, delta, render):
    m = r.material(mat)
    orig = getattr(m, field)[ch]
    getattr(m, field)[ch] = orig + delta
    r.set_material(mat, m)
    out = render()
    getattr(m, field)[ch] = orig
    r.set_material(mat, m)
    return out


def test_adjoint_matches_gpu_finite_differences():
    r = cornell()
    w = weights()
    r.start(SIZE, SIZE, 0)
    g = r.render_backward(w)
    print(f"backward {r.stats().backwardMs:.2f} ms")
    for mat, field, ch, h in PARAMS:
        lp = float(np.sum(w * perturbed(r, mat, field, ch, h, lambda: frame(r))))
        lm = float(np.sum(w * perturbed(r, mat, field, ch, -h, lambda: frame(r))))
        fd = (lp - lm) / (2 * h)
        ad = adjoint_value(g, r, mat, field, ch)
        print(f"material {mat} {field}[{ch}]: adjoint {ad:.6f} finite difference {fd:.6f}")
        assert abs(ad - fd) <= 1e-2 * abs(fd) + 1e-6
    r.close()


def test_adjoint_matches_fp64_cpu_gradient():
    """The FP64 oracle follows the same paths; its central differences are the
    CPU gradient of the same estimator."""
    r = cornell()
    w = weights(7).astype(np.float64)
    r.start(SIZE, SIZE, 0)
    g = r.render_backward(w.astype(np.float32))

    def cpu_frame():
        r.start(SIZE, SIZE, 0)
        osc = O.OracleScene(r.export(), use_bvh=False)
        acc, _ = osc.render(SIZE, SIZE, precision=64)
        osc.close()
        return acc[..., :3]

    for mat, field, ch, h in PARAMS[1:3]:
        lp = float(np.sum(w * perturbed(r, mat, field, ch, h, cpu_frame)))
        lm = float(np.sum(w * perturbed(r, mat, field, ch, -h, cpu_frame)))
        fd = (lp - lm) / (2 * h)
        ad = adjoint_value(g, r, mat, field, ch)
        print(f"material {mat} {field}[{ch}]: adjoint {ad:.6f} FP64 CPU {fd:.6f}")
        assert abs(ad - fd) <= 1e-2 * abs(fd) + 1e-6
    r.close()


def test_environment_texel_gradients():
    """A Lambertian sphere under a small RGB environment (gathered on BSDF
    misses): the image is linear in the texels."""
    env = np.random.default_rng(2).uniform(0.2, 1.0, (16, 32, 4)).astype(np.float32)
    env[..., 3] = 1.0
    s = scenes.Scene()
    v, i = scenes.geodesic_sphere(8)
    s.meshes.append(scenes.MeshData(V.make_vertices(v.astype(np.float32), v.astype(np.float32)),
                                    i.astype(np.uint32), 0))
    s.materials = [scenes.lambert((0.6, 0.5, 0.4))]
    s.camera = ((0.0, -3.0, 0.5), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 50.0)
    r = V.Renderer(SIZE, SIZE)
    scenes.load(r, s)
    tex = r.add_texture_rgba32f(env)
    r.set_environment_texture(tex)
    r.set_environment_light((1.0, 1.0, 1.0), 1.0)
    r.configure(spp=SPP, max_depth=3, min_depth=8, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=9)
    w = weights(11)
    r.start(SIZE, SIZE, 0)
    g = r.render_backward(w, environment_shape=(16, 32))["environment"]
    assert np.abs(g).sum() > 0
    h = 0.05
    for (y, x, c) in [(4, 7, 0), (8, 16, 1), (12, 30, 2)]:
        vals = []
        for delta in (h, -h):
            e2 = env.copy()
            e2[y, x, c] += delta
            r.set_environment_texture(r.add_texture_rgba32f(e2))
            vals.append(float(np.sum(w * frame(r))))
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"texel ({y},{x},{c}): adjoint {g[y, x, c]:.6f} finite difference {fd:.6f}")
        assert abs(g[y, x, c] - fd) <= 1e-2 * abs(fd) + 1e-4
    r.close()


def test_backward_leaves_the_forward_frame_unchanged():
    r = cornell()
    before = frame(r)
    r.start(SIZE, SIZE, 0)
    r.render_backward(weights())
    r.draw()
    after = r.accumulation()[..., :3].astype(np.float64)
    assert np.array_equal(before, after)
    assert r.status().totalSamples == SPP
    r.close()
