"""Neural radiance cache network on the GPU against a numpy reference of the
same operations: fp16 encoded inputs, fp16 weights and activations, fp32
accumulation (the reference accumulates in float64 on fp16-rounded values)."""

import math

import numpy as np
import pytest

import vkrt_amd as V

pytestmark = pytest.mark.gpu

F16 = np.float16
OFF_HIDDEN, OFF_OUT, PARAMS = 12288, 28672, 29696


def encode(v):
    """12-frequency sin/cos of the 7 periodic inputs, then roughness and albedo, a 1, zero padding."""
    v = np.asarray(v, np.float32)
    x = np.zeros((len(v), 192), np.float32)
    for d in range(7):
        for f in range(12):
            s = np.float32(0.5 if f == 0 else 2.0 ** (f - 1))
            a = v[:, d] * s
            a = (a - np.floor(a)).astype(np.float64) * 2 * math.pi
            x[:, d * 24 + 2 * f] = np.sin(a)
            x[:, d * 24 + 2 * f + 1] = np.cos(a)
    x[:, 168:172] = v[:, 7:11]
    x[:, 172] = 1.0
    return x


def layers(w):
    w = np.asarray(w, np.float32)
    mats = [w[:OFF_HIDDEN].reshape(64, 192)]
    mats += [w[OFF_HIDDEN + k * 4096:OFF_HIDDEN + (k + 1) * 4096].reshape(64, 64) for k in range(4)]
    mats.append(w[OFF_OUT:].reshape(16, 64)[:3])
    return [m.astype(F16).astype(np.float64) for m in mats]


def forward(w, v):
    mats = layers(w)
    acts = [encode(v).astype(F16).astype(np.float64)]
    for m in mats[:-1]:
        acts.append(np.maximum(acts[-1] @ m.T, 0.0).astype(F16).astype(np.float64))
    return acts, acts[-1] @ mats[-1].T


def gradient(w, v, t):
    mats = layers(w)
    acts, y = forward(w, v)
    den = y * y + 0.01
    loss = float(np.sum((y - t) ** 2 / den))
    d = np.clip(2 * (y - t) / den, -1000, 1000)
    grads = [None] * 6
    for k in range(5, -1, -1):
        grads[k] = d.T @ acts[k]
        if k > 0:
            d = (d @ mats[k]) * (acts[k] > 0)
    flat = np.zeros(PARAMS)
    flat[:OFF_HIDDEN] = grads[0].ravel()
    for k in range(4):
        flat[OFF_HIDDEN + k * 4096:OFF_HIDDEN + (k + 1) * 4096] = grads[k + 1].ravel()
    flat[OFF_OUT:OFF_OUT + 3 * 64] = grads[5].ravel()
    return flat, loss


def random_inputs(n, seed):
    return np.random.default_rng(seed).uniform(0, 1, (n, 11)).astype(np.float32)


@pytest.fixture(scope="module")
def renderer():
    r = V.Renderer(16, 16)
    yield r
    r.close()


def test_parameter_layout(renderer):
    renderer.nrc_reset(3)
    w = renderer.nrc_weights()
    assert w.size == PARAMS
    assert np.all(w[OFF_OUT + 3 * 64:] == 0.0)  # padded output rows
    assert np.all(w[:OFF_HIDDEN].reshape(64, 192)[:, 173:] == 0.0)  # zero-padded inputs
    assert np.abs(w[:OFF_HIDDEN]).max() <= math.sqrt(6 / 173) + 1e-6


def test_inference_matches_reference(renderer):
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.15, PARAMS).astype(np.float32)
    renderer.nrc_set_weights(w)
    w = renderer.nrc_weights()
    for n in (1, 15, 16, 1000, 4099):  # ragged counts: partial blocks of 16 samples
        v = random_inputs(n, n)
        gpu = renderer.nrc_infer(v)
        _, ref = forward(w, v)
        scale = np.abs(ref).max() + 1e-3
        err = np.abs(gpu - ref).max() / scale
        print(f"n={n}: max |gpu - ref| / max |ref| = {err:.2e}")
        assert err < 1e-2


def test_gradient_matches_reference(renderer):
    rng = np.random.default_rng(2)
    renderer.nrc_set_weights(rng.normal(0, 0.15, PARAMS).astype(np.float32))
    w = renderer.nrc_weights()
    n = 777
    v = random_inputs(n, 5)
    t = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    g, loss = renderer.nrc_gradient(v, t)
    ref, ref_loss = gradient(w, v, t)
    assert loss == pytest.approx(ref_loss, rel=1e-2)
    spans = [(0, OFF_HIDDEN)] + [(OFF_HIDDEN + k * 4096, OFF_HIDDEN + (k + 1) * 4096) for k in range(4)]
    spans.append((OFF_OUT, OFF_OUT + 192))
    for a, b in spans:
        rel = np.linalg.norm(g[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        print(f"params {a}:{b}: relative gradient error {rel:.2e}")
        assert rel < 2e-2
    assert np.all(g[OFF_OUT + 192:] == 0.0)


def test_gradient_of_a_batch_larger_than_one_slice(renderer):
    """66,770 samples: the training passes run over two slices of the batch
    (65,536 and 1,234 samples) and the second adds to the first's partials."""
    rng = np.random.default_rng(4)
    renderer.nrc_set_weights(rng.normal(0, 0.15, PARAMS).astype(np.float32))
    w = renderer.nrc_weights()
    n = 65536 + 1234
    v = random_inputs(n, 9)
    t = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    g, loss = renderer.nrc_gradient(v, t)
    ref, ref_loss = gradient(w, v, t)
    assert loss == pytest.approx(ref_loss, rel=1e-2)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print(f"relative gradient error {rel:.2e}")
    assert rel < 2e-2
    # the second slice alone: its gradient is the difference of the two
    g1, _ = renderer.nrc_gradient(v[:65536], t[:65536])
    ref2, _ = gradient(w, v[65536:], t[65536:])
    rel2 = np.linalg.norm((g - g1) - ref2) / np.linalg.norm(ref2)
    print(f"second slice: relative gradient error {rel2:.2e}")
    assert rel2 < 5e-2


def test_training_fits_a_smooth_function(renderer):
    renderer.nrc_reset(7)
    n = 1 << 15
    v = random_inputs(n, 11)
    t = np.stack([0.5 + 0.3 * np.sin(2 * math.pi * v[:, 0]) * v[:, 8],
                  0.4 + 0.2 * v[:, 1] * v[:, 9],
                  0.3 + 0.25 * np.cos(2 * math.pi * v[:, 5]) * v[:, 10]], 1).astype(np.float32)
    first = renderer.nrc_train(v, t, 1, 4096, 1e-2) / (4096 * 3)
    for _ in range(30):
        last = renderer.nrc_train(v, t, 10, 4096, 1e-2) / (10 * 4096 * 3)
    pred = renderer.nrc_infer(v[:2000])
    mse = float(np.mean((pred - t[:2000]) ** 2))
    var = float(np.mean(np.var(t[:2000], axis=0)))
    print(f"relative L2 loss {first:.3e} -> {last:.3e}; mse after training {mse:.2e} (target variance {var:.2e})")
    assert last < 0.05 * first
    assert mse < 0.4 * var


def cornell(size, spp, depth, nee=True, seed=1, mode=V.RENDER_MODE_RGB):
    from vkrt_amd import scenes

    r = V.Renderer(size, size)
    scenes.load(r, scenes.cornell_box())
    r.configure(spp=spp, max_depth=depth, min_depth=4, mode=mode, spectral=V.SPECTRAL_HERO, rng=V.RNG_PCG32,
                seed=seed, nee=nee)
    r.start(size, size, 0)
    return r


def test_zero_cache_equals_the_truncated_tracer():
    """A cache that predicts 0 and NEE off: an image path gathers exactly the
    emission of its first prefix+1 vertices, which is the plain tracer with depth
    prefix+1; the training paths use streams of their own and leave it unchanged."""
    size, spp = 48, 8
    a = cornell(size, spp, 8, nee=False)
    a.nrc_set_weights(np.zeros(PARAMS, np.float32))
    a.set_neural_cache(True, prefixBounces=2, trainingPaths=4096, train=0)
    a.start(size, size, 0)
    a.draw()
    img = a.accumulation()
    st = a.neural_cache_stats()
    a.close()
    b = cornell(size, spp, 3, nee=False)
    b.draw()
    ref = b.accumulation()
    b.close()
    print(f"queries {st.queries}, training records {st.trainingRecords}")
    assert st.queries > 0.5 * size * size * spp and st.trainingRecords > 4096
    assert np.array_equal(img, ref)


def test_cache_output_is_added_with_the_path_throughput():
    """A cache that predicts the constant c adds c times the throughput of every
    query; with c = 1 and c = 2 the two images differ from the zero-cache image by
    amounts in the ratio 1 : 2."""
    size, spp = 32, 8
    w = np.zeros(PARAMS, np.float32)
    w[:OFF_HIDDEN].reshape(64, 192)[0, 172] = 1.0  # neuron 0 of the first layer = the constant input 1
    for k in range(4):
        w[OFF_HIDDEN + k * 4096:OFF_HIDDEN + (k + 1) * 4096].reshape(64, 64)[0, 0] = 1.0
    imgs = []
    for c in (0.0, 1.0, 2.0):
        w[OFF_OUT:OFF_OUT + 192].reshape(3, 64)[:, 0] = c
        r = cornell(size, spp, 8, nee=True)
        r.nrc_set_weights(w)
        r.set_neural_cache(True, trainingPaths=1024, train=0)
        r.start(size, size, 0)
        r.draw()
        imgs.append(r.accumulation()[..., :3].astype(np.float64))
        r.close()
    d1, d2 = imgs[1] - imgs[0], imgs[2] - imgs[0]
    assert d1.sum() > 0
    assert np.allclose(d2, 2 * d1, rtol=1e-4, atol=1e-6)


def test_online_training_improves_the_image():
    """The cache trained online for 60 frames, then frozen, against the same cache
    untrained: 64 accumulated frames of each against a 4096-spp plain image."""
    size = 64

    def relmse(x, ref):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    ref_r = cornell(size, 4096, 8)
    ref_r.draw()
    ref = ref_r.accumulation()[..., :3].astype(np.float64)
    ref_r.close()
    errs = {}
    for trained in (False, True):
        r = cornell(size, 1, 8)
        r.nrc_reset(1)
        if trained:
            r.set_neural_cache(True, trainingPaths=16384, trainingSteps=4, batchSize=4096, learningRate=1e-2)
            for _ in range(60):
                r.start(size, size, 0)
                r.draw()
            st = r.neural_cache_stats()
            print(f"last loss {st.lastLoss:.3f}, queries {st.queries}, records {st.trainingRecords}, "
                  f"inference {st.inferMs:.3f} ms, training {st.trainMs:.3f} ms")
        r.set_neural_cache(True, trainingPaths=16384, train=0)
        r.start(size, size, 0)
        for _ in range(64):
            r.draw()
        img = r.accumulation()[..., :3].astype(np.float64)
        assert np.isfinite(img).all()
        errs[trained] = relmse(img, ref)
        r.close()
    print(f"relMSE of 64 frames: untrained cache {errs[False]:.4e}, trained cache {errs[True]:.4e}")
    assert errs[True] < 0.5 * errs[False]


# ------------------------------------------------- hero spectral mode --
HERO = V.RENDER_MODE_SPECTRAL


def test_zero_cache_equals_the_truncated_tracer_in_hero_mode():
    """As in RGB: with a cache that predicts 0 and NEE off, hero spectral image
    paths gather exactly the emission of their first prefix+1 vertices."""
    size, spp = 48, 8
    a = cornell(size, spp, 8, nee=False, mode=HERO)
    a.nrc_set_weights(np.zeros(PARAMS, np.float32))
    a.set_neural_cache(True, prefixBounces=2, trainingPaths=4096, train=0)
    a.start(size, size, 0)
    a.draw()
    img = a.accumulation()
    st = a.neural_cache_stats()
    a.close()
    b = cornell(size, spp, 3, nee=False, mode=HERO)
    b.draw()
    ref = b.accumulation()
    b.close()
    print(f"queries {st.queries}, training records {st.trainingRecords}")
    assert st.queries > 0.5 * size * size * spp and st.trainingRecords > 4096
    assert np.array_equal(img, ref)


def test_hero_cache_output_is_uplifted_and_added_with_the_path_throughput():
    """A cache that predicts the constant XYZ (c, c, c): each query adds its four
    lanes' throughput times a quarter of the rgb2spec uplift of that colour's
    linear sRGB. For c = 1 and c = 2 the brightest channel exceeds 1, where the
    uplift scales linearly, so the two images differ from the zero-cache image
    by amounts in the ratio 1 : 2; the difference is grey-ish, as an
    equal-energy XYZ is."""
    size, spp = 32, 8
    w = np.zeros(PARAMS, np.float32)
    w[:OFF_HIDDEN].reshape(64, 192)[0, 172] = 1.0
    for k in range(4):
        w[OFF_HIDDEN + k * 4096:OFF_HIDDEN + (k + 1) * 4096].reshape(64, 64)[0, 0] = 1.0
    imgs = []
    for c in (0.0, 1.0, 2.0):
        w[OFF_OUT:OFF_OUT + 192].reshape(3, 64)[:, 0] = c
        r = cornell(size, spp, 8, nee=True, mode=HERO)
        r.nrc_set_weights(w)
        r.set_neural_cache(True, trainingPaths=1024, train=0)
        r.start(size, size, 0)
        r.draw()
        imgs.append(r.accumulation()[..., :3].astype(np.float64))
        r.close()
    d1, d2 = imgs[1] - imgs[0], imgs[2] - imgs[0]
    assert d1.sum() > 0
    assert np.allclose(d2, 2 * d1, rtol=1e-4, atol=1e-6)
    m = d1.reshape(-1, 3).sum(0)
    print(f"XYZ added by the unit cache: {m / m[1]}")
    assert np.all(m > 0) and np.all(np.abs(m / m[1] - 1.0) < 0.5)


def test_online_training_improves_the_image_in_hero_mode():
    """Hero spectral: the cache trained online for 60 frames on XYZ targets,
    then frozen, against the same cache untrained; 64 accumulated frames of each
    against a 4096-spp plain spectral image (XYZ)."""
    size = 64

    def relmse(x, ref):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    ref_r = cornell(size, 4096, 8, mode=HERO)
    ref_r.draw()
    ref = ref_r.accumulation()[..., :3].astype(np.float64)
    ref_r.close()
    errs = {}
    for trained in (False, True):
        r = cornell(size, 1, 8, mode=HERO)
        r.nrc_reset(1)
        if trained:
            r.set_neural_cache(True, trainingPaths=16384, trainingSteps=4, batchSize=4096, learningRate=1e-2)
            for _ in range(60):
                r.start(size, size, 0)
                r.draw()
            st = r.neural_cache_stats()
            print(f"last loss {st.lastLoss:.3f}, queries {st.queries}, records {st.trainingRecords}")
            assert st.trainingRecords > 16384
        r.set_neural_cache(True, trainingPaths=16384, train=0)
        r.start(size, size, 0)
        for _ in range(64):
            r.draw()
        img = r.accumulation()[..., :3].astype(np.float64)
        assert np.isfinite(img).all()
        errs[trained] = relmse(img, ref)
        r.close()
    print(f"hero relMSE of 64 frames: untrained cache {errs[False]:.4e}, trained cache {errs[True]:.4e}")
    assert errs[True] < 0.5 * errs[False]
