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


def test_emission_gradient_of_surfaces_that_do_not_emit():
    """With NEE off, emission reaches the camera only through BSDF hits, so the image
    is linear in every material's emission under common random numbers, also for
    materials that emit nothing (the red wall here)."""
    r = cornell()
    r.configure(nee=False)
    w = weights(13)
    r.start(SIZE, SIZE, 0)
    g = r.render_backward(w)
    m = r.material(2)
    assert m.emissionLuminance == 0.0
    h = 0.05

    def image_with(e):
        mm = r.material(2)
        mm.emissionColor[:] = (1.0, 0.0, 0.0)
        mm.emissionLuminance = e
        r.set_material(2, mm)
        return frame(r)

    fd = (np.sum(w * image_with(h)) - np.sum(w * image_with(0.0))) / h  # linear in e: one-sided is exact
    r.set_material(2, m)
    ad = float(g["emission"][2, 0])
    print(f"red wall emission[0]: adjoint {ad:.6f} finite difference {fd:.6f}")
    assert abs(ad - fd) <= 1e-2 * abs(fd) + 1e-6
    r.close()


def prb_renderer(size, spp, depth=4, nee=True, seed=3, min_depth=8):
    r = V.Renderer(size, size)
    s = scenes.differentiable_cornell()
    scenes.load(r, s)
    r.configure(spp=spp, max_depth=depth, min_depth=min_depth, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=seed,
                nee=nee)
    r.start(size, size, 0)
    return r, s


def test_roughness_gradient_matches_the_detached_estimator():
    """Roughness moves sampled directions and lobe choices, so finite differences
    of frames under common random numbers are not the derivative of the expected
    image (they miss the lobe-switch jumps), and the reference's GGX sampler draws
    bounded visible normals while its pdf is the unbounded VNDF pdf, so the
    forward estimator over-weights glossy samples by about 2 %. The adjoint is the
    detached path-replay estimator E[L (df/dr) cos / pdf]; this test checks it on a
    plane of coated diffuse material under a constant sky (depth 2: one BSDF sample
    per path that escapes), against the same estimator computed on the CPU from the
    FP64 oracle's own sampler and BSDF evaluation with independent samples."""
    import math

    size, spp = 32, 4096
    s = scenes.Scene()
    L = 50.0
    pos = np.array([[-L, -L, 0], [L, -L, 0], [L, L, 0], [-L, L, 0]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
    s.meshes.append(scenes.MeshData(V.make_vertices(pos, nrm), np.array([[0, 1, 2], [0, 2, 3]], np.uint32), 0))
    mat = V.default_material(baseColor=(0.5, 0.5, 0.5), roughness=0.45, specular=0.5, metallic=0.0)
    s.materials = [mat]
    th = math.radians(60)
    s.camera = ((3 * math.sin(th), 0.0, 3 * math.cos(th)), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0)
    s.environment_color, s.environment_strength = (1.0, 1.0, 1.0), 1.0
    r = V.Renderer(size, size)
    scenes.load(r, s)
    r.configure(spp=spp, max_depth=2, min_depth=8, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=5)
    mi = r.material_count() - 1
    r.start(size, size, 0)
    g = r.render_backward(np.ones((size, size, 3), np.float32))
    gpu = float(g["roughness"][mi]) / (size * size * 3)  # per pixel and channel (grey material)
    r.close()

    wo = (math.sin(th), 0.0, math.cos(th))  # the view direction (1 degree field of view)
    n, step = 200_000, 1e-3
    smp = O.bsdf_sample_batch(None, mat, wo, n, seed=17)
    est = np.zeros(n)
    for k, row in enumerate(smp):
        if row[6] <= 0:
            continue
        mat.roughness = 0.45 + step
        fp = O.bsdf_eval(None, mat, wo, row[:3])[0][0]
        mat.roughness = 0.45 - step
        fm = O.bsdf_eval(None, mat, wo, row[:3])[0][0]
        est[k] = (fp - fm) / (2 * step) * abs(row[2]) / row[6]
    cpu, se = float(est.mean()), float(est.std() / math.sqrt(n))
    print(f"d albedo / d roughness: GPU adjoint {gpu:.5f}, CPU detached estimator {cpu:.5f} +- {se:.5f}")
    assert abs(gpu - cpu) <= 4 * se * math.sqrt(1 + n / (size * size * spp))


def test_torch_binding_gradients_equal_render_backward():
    torch = pytest.importorskip("torch")
    from vkrt_amd import autograd as A

    size = 32
    r, s = prb_renderer(size, 16)
    p = A.SceneParameters.from_renderer(r, s.environment)
    for t in p.tensors():
        t.requires_grad_(True)
    w = torch.tensor(weights(21)[:size, :size], dtype=torch.float32)
    view = scenes.prb_views()[1]
    loss = (A.render(r, p, view) * w).sum()
    loss.backward()
    r.set_camera(*view)
    r.start(size, size, 0)
    g = r.render_backward(w.numpy(), environment_shape=s.environment.shape[:2])

    def same(a, b):
        # the two passes replay the same paths, but gradients are summed with
        # float atomics whose order varies between launches: compare to the
        # rounding of such sums, relative to the array's scale
        return np.abs(a - b).max() <= 1e-4 * np.abs(b).max() + 1e-7

    assert same(p.base_color.grad.numpy(), g["baseColor"])
    assert same(p.roughness.grad.numpy(), g["roughness"])
    assert same(p.emission.grad.numpy(), g["emission"])
    assert same(p.environment.grad.numpy(), g["environment"])
    assert np.abs(g["environment"]).sum() > 0 and np.abs(g["roughness"]).sum() > 0
    r.close()


def test_inverse_rendering_reduces_the_loss():
    """Targets rendered with material albedos perturbed by +-20 % (common random
    numbers with the optimised renders); Adam on the albedos from the unperturbed
    values over three views lowers the L2 image loss and the albedo error of the
    materials the views see."""
    torch = pytest.importorskip("torch")
    from vkrt_amd import autograd as A

    size, iters = 64, 40
    r, s = prb_renderer(size, 16)
    views = scenes.prb_views()
    start = A.SceneParameters.from_renderer(r, s.environment)
    rng = np.random.default_rng(4)
    target = A.SceneParameters(start.base_color * torch.tensor(rng.uniform(0.8, 1.2, (start.base_color.shape)),
                                                               dtype=torch.float32),
                               start.roughness, start.emission, start.environment)
    target.clamp_()
    with torch.no_grad():
        ref = [A.render(r, target, v) for v in views]
    p = A.SceneParameters(start.base_color.clone().requires_grad_(True), start.roughness, start.emission,
                          start.environment)
    opt = torch.optim.Adam([p.base_color], lr=0.01)
    seen = None
    losses = []
    for it in range(iters):
        opt.zero_grad()
        loss = sum(((A.render(r, p, v) - t) ** 2).mean() for v, t in zip(views, ref))
        loss.backward()
        if it == 0:
            gm = p.base_color.grad.abs().amax(dim=1)
            seen = gm > 1e-3 * float(gm.max())
            e0 = float((p.base_color - target.base_color).detach().abs()[seen].mean())
        opt.step()
        p.clamp_()
        losses.append(float(loss.detach()))
    e1 = float((p.base_color - target.base_color).detach().abs()[seen].mean())
    print(f"loss {losses[0]:.5f} -> {losses[-1]:.5f}, albedo error of {int(seen.sum())} seen materials "
          f"{e0:.4f} -> {e1:.4f}")
    assert losses[-1] < 0.5 * losses[0]
    assert e1 < 0.8 * e0
    r.close()


def test_adjoint_on_the_differentiable_configuration_at_h_1e_3():
    """The inverse-rendering configuration's own parameters (64 materials and a
    512x256 sky, depth 8, 32 spp) against central differences with h = 1e-3
    under common random numbers, at 64x64: albedo of three wall patches, the
    light's emission, the emission of a patch that emits nothing, and the six
    sky texels with the largest gradients. The patches are made Lambertian
    (specular 0, roughness 1) because the coated materials' lobe choice depends
    on albedo, so their frames jump where a sample switches lobe and finite
    differences of one realisation are not the derivative of the expected image
    (roughness is covered by the detached-estimator test above). Russian
    roulette is kept out of reach (min depth 9 > max depth 8)."""
    size, spp, h = 64, 32, 1e-3
    r = V.Renderer(size, size)
    s = scenes.differentiable_cornell()
    for m in s.materials[:-1]:
        m.specular, m.roughness = 0.0, 1.0
    scenes.load(r, s)
    r.configure(spp=spp, max_depth=8, min_depth=9, mode=V.RENDER_MODE_RGB, rng=V.RNG_PCG32, seed=3)
    env = np.ascontiguousarray(s.environment, dtype=np.float32)
    w = np.random.default_rng(31).normal(size=(size, size, 3)).astype(np.float32)

    def loss():
        r.start(size, size, 0)
        r.draw()
        return float(np.sum(w.astype(np.float64) * r.accumulation()[..., :3].astype(np.float64)))

    r.start(size, size, 0)
    g = r.render_backward(w, environment_shape=env.shape[:2])
    checks = []
    light = r.material_count() - 1
    gb = np.abs(g["baseColor"][1:light]).max(axis=1)
    for mi in (1 + np.argsort(gb)[-3:]).tolist():
        ch = int(np.argmax(np.abs(g["baseColor"][mi])))
        checks.append((f"albedo m{mi}[{ch}]", float(g["baseColor"][mi, ch]),
                       lambda d, mi=mi, ch=ch: perturbed(r, mi, "baseColor", ch, d, loss)))
    lum = r.material(light).emissionLuminance
    checks.append(("light emission[1]", float(g["emission"][light, 1]) * lum,
                   lambda d: perturbed(r, light, "emissionColor", 1, d, loss)))
    dark = int(1 + np.argmax(np.abs(g["emission"][1:light, 0])))

    def dark_emission(d):
        m = r.material(dark)
        orig = (tuple(m.emissionColor), m.emissionLuminance)
        m.emissionColor[:] = (1.0, 0.0, 0.0)
        m.emissionLuminance = d
        r.set_material(dark, m)
        out = loss()
        m.emissionColor[:], m.emissionLuminance = orig
        r.set_material(dark, m)
        return out

    ge = g["environment"]
    for flat in np.argsort(np.abs(ge).ravel())[-6:].tolist():
        y, x, c = np.unravel_index(flat, ge.shape)

        def texel(d, y=y, x=x, c=c):
            e2 = env.copy()
            e2[y, x, c] += d
            r.set_environment_texture(r.add_texture_rgba32f(e2))
            return loss()

        checks.append((f"sky texel ({y},{x},{c})", float(ge[y, x, c]), texel))
    worst = 0.0
    results = []
    for name, ad, f in checks:
        results.append((name, ad, (f(h) - f(-h)) / (2 * h)))
    # a patch that starts to emit becomes a light for next-event estimation, which
    # changes every path's light choice: its emission is checked with NEE off, where
    # it reaches the camera only through BSDF hits and enters linearly from 0
    r.set_environment_texture(r.add_texture_rgba32f(env))
    r.configure(nee=False)
    r.start(size, size, 0)
    g2 = r.render_backward(w, environment_shape=env.shape[:2])
    results.append((f"emission of non-emitting m{dark}[0], NEE off", float(g2["emission"][dark, 0]),
                    (dark_emission(h) - dark_emission(0.0)) / h))
    for name, ad, fd in results:
        rel = abs(ad - fd) / max(abs(fd), 1e-12)
        worst = max(worst, rel)
        print(f"{name}: adjoint {ad:.6f} finite difference {fd:.6f} relative {rel:.2e}", flush=True)
        assert abs(fd) > 0
        assert abs(ad - fd) <= 1e-2 * abs(fd)
    print(f"worst relative difference {worst:.2e} over {len(results)} parameters")
    r.close()
